"""Synthetic atmospheric columns for the two-stream flux path.

Generates, with fixed seeds, the column inputs the `runes` interface takes
(socrates_runes.F90:47-168): layer/level pressure and temperature, layer mass,
density and heat capacity, H2O/O3 mass mixing ratios, well-mixed CO2, surface
temperature and albedo, solar geometry, liquid cloud fraction / mixing ratio /
effective radius, and McICA seeds. Arrays use the layout the interface reads
for full fields, (n_profile, n_layer) in Fortran order, i.e. numpy shape
(n_layer, n_profile) here: the profile index is fastest, layer 1 is at the top.

Profile recipe (sizes and distributions are the ones the workloads name):
  * p_level(i) = 1e5 * (1e-5)**((L-i)/L) Pa, i=0 at the top of atmosphere;
    p_layer is the geometric mean of its bounding levels.
  * Heights from the hypsometric equation; T = 288 K - 6.5 K/km * z down to
    216 K, isothermal above, plus N(0, 2 K) per layer; level temperatures
    interpolated linearly in ln p between layer centres.
  * H2O = 1e-2 exp(-z / 2 km); CO2 = 6e-4 (well mixed);
    O3 = 1e-5 exp(-0.5 (ln(p / 1000 Pa) / 0.75)**2).
  * mass = dp / g, density = p / (R_dry T), heat capacity = mass * cp.
Clouds: fraction ~ Beta(0.5, 0.5) for 1 km <= z <= 12 km, in-cloud liquid
water content lognormal (median 0.1 g/m3, sigma_ln = 1), effective radius
U(5, 20) um, exponential-random overlap with a 2 km decorrelation length
expressed in pressure (rho g dz at 5 km).
"""
import numpy as np

from .case_io import Case

GRAV = 9.80665
R_DRY = 287.026
CP_DRY = 1005.0
SOLAR_IRRAD = 1361.0
DECORR_PA = 0.74 * GRAV * 2000.0

# parametrisation indices of the synthetic spectral files (block 10 / 12)
DROP_TYPE = 5
ICE_TYPE = 8

# option codes (rad_pcf.F90)
IP_SOLAR = 1
IP_INFRA_RED = 2
IP_CLOUD_HOMOGEN = 1
IP_CLOUD_CONV_STRAT = 3
IP_EXPONENTIAL_RAND = 3
IP_MAX_RAND = 0
IP_MCICA = 2
IP_MCICA_OPTIMAL = 2


def _temperature(z_km):
    return np.maximum(288.0 - 6.5 * z_km, 216.0)


def profiles(n_profile, n_layer, seed, t_noise=2.0):
    """Clear-sky column state. Returns a dict of arrays shaped (L[+1], P)."""
    rng = np.random.default_rng(seed)
    L, P = n_layer, n_profile
    i = np.arange(L + 1, dtype=np.float64)
    p_lev1 = 1.0e5 * (1.0e-5) ** ((L - i) / L)
    p_lay1 = np.sqrt(p_lev1[:-1] * p_lev1[1:])
    # heights of levels by upward integration of the hypsometric equation
    z_lev = np.zeros(L + 1)
    for k in range(L - 1, -1, -1):
        t_est = _temperature(z_lev[k + 1] / 1000.0)
        dz = R_DRY * t_est / GRAV * np.log(p_lev1[k + 1] / p_lev1[k])
        t_mid = _temperature((z_lev[k + 1] + 0.5 * dz) / 1000.0)
        z_lev[k] = z_lev[k + 1] + R_DRY * t_mid / GRAV * np.log(p_lev1[k + 1] / p_lev1[k])
    z_lay1 = 0.5 * (z_lev[:-1] + z_lev[1:])
    p_level = np.repeat(p_lev1[:, None], P, axis=1)
    p_layer = np.repeat(p_lay1[:, None], P, axis=1)
    z_layer = np.repeat(z_lay1[:, None], P, axis=1)
    t_layer = _temperature(z_layer / 1000.0) + rng.normal(0.0, t_noise, (L, P))
    lnp_lay = np.log(p_layer)
    lnp_lev = np.log(p_level)
    t_level = np.empty((L + 1, P))
    w = (lnp_lev[1:-1] - lnp_lay[:-1]) / (lnp_lay[1:] - lnp_lay[:-1])
    t_level[1:-1] = t_layer[:-1] + w * (t_layer[1:] - t_layer[:-1])
    t_level[0] = t_layer[0]
    wb = (lnp_lev[-1] - lnp_lay[-1]) / (lnp_lay[-1] - lnp_lay[-2])
    t_level[-1] = t_layer[-1] + wb * (t_layer[-1] - t_layer[-2])
    mass = (p_level[1:] - p_level[:-1]) / GRAV
    density = p_layer / (R_DRY * t_layer)
    h2o = 1.0e-2 * np.exp(-z_layer / 2000.0)
    o3 = 1.0e-5 * np.exp(-0.5 * (np.log(p_layer / 1000.0) / 0.75) ** 2)
    return dict(p_layer=p_layer, t_layer=t_layer, p_level=p_level, t_level=t_level,
                mass=mass, density=density, layer_heat_capacity=mass * CP_DRY,
                h2o=h2o, o3=o3, z_layer=z_layer, t_ground=t_level[-1].copy())


def clouds(state, seed):
    """Liquid cloud fields on top of a column state (dict from profiles)."""
    rng = np.random.default_rng(seed)
    z = state["z_layer"]
    L, P = z.shape
    in_range = (z >= 1000.0) & (z <= 12000.0)
    frac = np.where(in_range, rng.beta(0.5, 0.5, (L, P)), 0.0)
    lwc = 1.0e-4 * np.exp(rng.normal(0.0, 1.0, (L, P)))          # kg/m3 in cloud
    liq_mmr = lwc / state["density"] * frac                        # grid-box mean
    liq_dim = rng.uniform(5.0e-6, 20.0e-6, (L, P))
    return frac, liq_mmr, liq_dim


def make_case(n_profile, n_layer, source, n_band, seed, cloudy=False,
              overlap="mcica", first_column_id=0):
    """Build a Case for the `runes` interface.

    source: "lw" or "sw". overlap: "mcica" (homogeneous cloud, McICA
    sub-columns, exponential-random) or "triple" (clear/stratiform/convective
    regions, maximum-random).
    """
    st = profiles(n_profile, n_layer, seed)
    rng = np.random.default_rng(seed + 7)
    P, L = n_profile, n_layer
    c = Case(n_profile=P, n_layer=L, n_band=n_band)
    for k in ("p_layer", "t_layer", "p_level", "t_level", "mass", "density",
              "layer_heat_capacity", "h2o", "o3", "t_ground"):
        setattr(c, k, st[k])
    c.co2_mix_ratio = 6.0e-4
    if source == "lw":
        c.i_source = IP_INFRA_RED
        c.l_rayleigh = 0
        alb = np.full(P, 0.02)                                    # emissivity 0.98
        c.cos_zen = np.ones(P)
        c.solar_irrad = np.zeros(P)
    else:
        c.i_source = IP_SOLAR
        c.l_rayleigh = 1
        alb = rng.uniform(0.05, 0.8, P)
        c.cos_zen = rng.uniform(0.05, 1.0, P)
        c.solar_irrad = np.full(P, SOLAR_IRRAD)
    c.albedo_diff = np.repeat(alb[None, :], n_band, axis=0)
    c.albedo_dir = c.albedo_diff.copy()
    c.rand_seed = (first_column_id + np.arange(P)).astype(np.int32)
    c.i_st_water = DROP_TYPE
    c.i_st_ice = ICE_TYPE
    if cloudy:
        frac, mmr, dim = clouds(st, seed + 13)
        c.l_cloud = 1
        c.cloud_frac, c.liq_mmr, c.liq_dim = frac, mmr, dim
        c.cloud_vertical_decorr = DECORR_PA
        c.cloud_horizontal_rsd = 0.75
        if overlap == "mcica":
            c.i_cloud_repr = IP_CLOUD_HOMOGEN
            c.i_overlap = IP_EXPONENTIAL_RAND
            c.i_inhom = IP_MCICA
            c.i_mcica_sampling = IP_MCICA_OPTIMAL
            c.n_subcol_gen = 64
            c.i_solver_kind = 0
        else:
            # split the cloud into stratiform and convective parts
            split = rng.uniform(0.0, 1.0, (L, P)) * (rng.uniform(size=(L, P)) < 0.3)
            c.conv_frac = frac * split
            c.cloud_frac = frac * (1.0 - split)
            c.liq_conv_mmr = mmr * split
            c.liq_mmr = mmr * (1.0 - split)
            c.liq_conv_dim = dim.copy()
            c.conv_vertical_decorr = DECORR_PA
            c.i_cloud_repr = IP_CLOUD_CONV_STRAT
            c.i_overlap = IP_MAX_RAND
            c.i_inhom = 0
            c.i_mcica_sampling = 0
            c.n_subcol_gen = 0
            c.i_solver_kind = 1
    return c


def add_aerosols(c, n_mode, seed, source=None):
    """Modal aerosols on a Case (runes' aer_* arguments): per mode a mass
    mixing ratio decaying from the surface with a 1.5-3 km scale height
    (surface values log-uniform in 1e-9..2e-8 kg/kg), and per band and mode
    specific absorption and scattering (m2/kg) and an asymmetry, constant in
    the horizontal with a few per cent noise per layer: column optical
    depths of order 0.1 in the solar region, scattering weaker and less
    forward-peaked in the infra-red."""
    rng = np.random.default_rng(seed)
    L, P, B = c.n_layer, c.n_profile, c.n_band
    src = source or ("sw" if c.i_source == IP_SOLAR else "lw")
    z = -7.0 * np.log(np.maximum(c.p_layer, 1e-3) / 1.0e5)       # km, approximate
    mix = np.empty((n_mode, L, P))
    for m in range(n_mode):
        h = rng.uniform(1.5, 3.0)
        q0 = np.exp(rng.uniform(np.log(1e-9), np.log(2e-8), P))
        mix[m] = q0[None, :] * np.exp(-z / h)
    shape = (B, n_mode, L, P)
    noise = lambda: 1.0 + 0.05 * rng.standard_normal(shape)       # noqa: E731
    if src == "sw":
        ka = rng.uniform(50.0, 800.0, (B, n_mode))[:, :, None, None] * noise()
        ks = rng.uniform(1.0e3, 6.0e3, (B, n_mode))[:, :, None, None] * noise()
        g = rng.uniform(0.55, 0.8, (B, n_mode))[:, :, None, None] * noise()
    else:
        ka = rng.uniform(50.0, 600.0, (B, n_mode))[:, :, None, None] * noise()
        ks = rng.uniform(10.0, 300.0, (B, n_mode))[:, :, None, None] * noise()
        g = rng.uniform(0.2, 0.6, (B, n_mode))[:, :, None, None] * noise()
    c.n_aer_mode = n_mode
    c.aer_mix_ratio = mix
    c.aer_absorption = np.abs(ka)
    c.aer_scattering = np.abs(ks)
    c.aer_asymmetry = np.clip(g, 0.0, 0.95)
    return c


SIGMA_SB = 5.670374419e-8


def make_two_stream_case(n_profile, n_layer, isolir, seed, n_term=1, l_scale_solar=True,
                         l_ir_source_quad=True, omc_range=(-8.0, -1.0),
                         tau_range=(-6.0, 3.0)):
    """Prescribed optical properties for the kernel-level entry (the
    two-stream solve without gas or cloud optics): per (term, layer,
    profile) tau ~ logU(10^tau_range), omega = 1 - 10^U(omc_range),
    g ~ U(0, 0.95); per profile sec(theta_0) from cos ~ U(0.05, 1),
    albedos ~ U(0.05, 0.8), incident beam 1361 cos(theta_0) W m-2;
    infra-red: band Planck fluxes B = sigma T^4 / 4 at the levels, layers and
    surface of a profiles() column (zero incident radiation, so the
    differential incident flux is -B(TOA)).
    Arrays are [term][layer][profile] for the optical properties and
    [layer][profile] / [profile] for the rest."""
    rng = np.random.default_rng(seed)
    P, L, G = n_profile, n_layer, n_term
    tau = 10.0 ** rng.uniform(tau_range[0], tau_range[1], (G, L, P))
    omega = 1.0 - 10.0 ** rng.uniform(omc_range[0], omc_range[1], (G, L, P))
    asym = rng.uniform(0.0, 0.95, (G, L, P))
    cosz = rng.uniform(0.05, 1.0, P)
    st = profiles(P, L, seed + 1)
    b_lev = 0.25 * SIGMA_SB * st["t_level"] ** 4
    b_lay = 0.25 * SIGMA_SB * st["t_layer"] ** 4
    b_gnd = 0.25 * SIGMA_SB * (st["t_ground"] + rng.normal(0.0, 2.0, P)) ** 4
    solar = isolir == IP_SOLAR
    return dict(
        n_profile=P, n_layer=L, n_term=G, isolir=isolir, l_scale_solar=int(l_scale_solar),
        l_ir_source_quad=int(l_ir_source_quad), tau=tau, omega=omega, asymmetry=asym,
        sec_0=1.0 / cosz,
        flux_inc_direct=(1361.0 * cosz) if solar else np.zeros(P),
        flux_inc_down=(1361.0 * cosz) if solar else -b_lev[0].copy(),
        diffuse_albedo=rng.uniform(0.05, 0.8, P), direct_albedo=rng.uniform(0.05, 0.8, P),
        diff_planck=b_lev[1:] - b_lev[:-1],
        diff_planck_2=2.0 * (b_lev[1:] + b_lev[:-1] - 2.0 * b_lay),
        d_planck_flux_surface=b_gnd - b_lev[-1],
        adjust_solar_ke=np.exp(rng.uniform(-0.02, 0.02, (L, P))) if solar else np.ones((L, P)))


def two_stream_term(tsc, g):
    """The single-term view of a kernel-level case (arrays [layer][profile])."""
    out = dict(tsc)
    for k in ("tau", "omega", "asymmetry"):
        out[k] = np.ascontiguousarray(tsc[k][g])
    out["n_term"] = 1
    return out


def write_two_stream_case(tsc, path, g=0):
    """Binary input of the reference harness oracle/harness/soc_two_stream_case.f90
    (one term); arrays written in Fortran (profile, layer) order, which is the
    C [layer][profile] memory order of these arrays."""
    t = two_stream_term(tsc, g)
    with open(path, "wb") as f:
        np.array([t["n_profile"], t["n_layer"], t["isolir"], t["l_scale_solar"],
                  t["l_ir_source_quad"]], dtype=np.int32).tofile(f)
        for k in ("tau", "omega", "asymmetry", "sec_0", "flux_inc_down", "flux_inc_direct",
                  "diffuse_albedo", "direct_albedo", "diff_planck", "diff_planck_2",
                  "d_planck_flux_surface", "adjust_solar_ke"):
            np.ascontiguousarray(t[k], dtype=np.float64).tofile(f)
