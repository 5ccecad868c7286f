"""Replay of the reference's file-based driver on the GPU path.

The reference's general driver reads one column set from a family of files
`<base>.<suffix>` (CDL text for l_run_cdl, netCDF for l_run_cdf), runs the
radiation code with options given on the command line of its scripts
(Cl_run_cdl / Cl_run_cdf) and writes the fluxes and heating rates as CDL.
This module does the same for the two-stream flux path the library covers,
without a netCDF library: CDL text is parsed here, netCDF classic files
(the reference's examples) through scipy's reader.

Inputs (src/aux/l_run_cdf.F90:350-1180, input_cloud_cdl.f:600-800):
  .t / .tl   layer and level temperatures; their `plev` coordinates are the
             layer and level pressures (Pa);
  .q .co2 .o3 .n2o .ch4 .o2 ... gas mass mixing ratios (gas_list_pcf.F90
             gas_suffix); a gas without a file has none;
  .tstar     surface temperature (infra-red);
  .szen .stoa solar zenith angle (degrees) and irradiance (solar);
  .surf      surface albedo `alb` (Lambertian basis functions: diffuse, then
             direct when a second is given);
  .clfr .lwm .re .iwm .ire  cloud fraction and in-cloud liquid and ice mass
             mixing ratios and sizes (homogeneous cloud).
Derived as l_run_cdl.F90:1017-1044 does: mass = dp / g, density with the
water-vapour correction, heat capacity = mass * cp_air_dry.

Options (sbin/Cl_run_cdl:30-90 and their defaults): -s spectral file,
-R first last band (the whole file), -I / -S, -g 4 (equivalent extinction
with scaling, the only gas overlap of the path), -c continuum, -r Rayleigh,
-C 5 clear | 2 maximum-random | 4 random | 7 exponential-random | 10 McICA,
-K 1 homogeneous cloud, -d droplet type, -i ice type, -m condensate RSD
(0.75), +m McICA sampling (2, optimal), -dp decorrelation pressure (1e4 Pa),
+R rescaling (required), -q quadratic infra-red source (linear otherwise),
-z scattering method (1 full, 2 none, 3 extinction as absorption, 4
approximate, 5 per k-term; the library takes the ones other than full in
clear sky and McICA only, and refuses the rest). The McICA data come from `mcica_data` beside
the spectral file (l_run_cdf.F90:269-270); the generator's seeds are the
reference's defaults, 10 + profile (socrates_set_cld_mcica.F90:159-166);
the generator's overlap is exponential (l_run_cdl.F90:1065, i_overlap = 2).

Outputs (l_run_cdl.F90:1100-1155, output_flux_cdl): .vflx total downward,
.uflx upward, .nflx net downward, .hrts heating rate (K/day), and for the
solar region .sflx direct and .dflx diffuse downward flux.
"""
import dataclasses
import os
import re

import numpy as np

from .case_io import WM_GASES, Case

GRAV_ACC = 9.80665          # rad_ccf.F90
R_GAS_DRY = 287.026
CP_AIR_DRY = 1.005e+03
MOL_WEIGHT_AIR = 28.966e-03
MOLAR_WEIGHT_H2O = 18.0153  # gas_list_pcf.F90 molar_weight(ip_h2o), g/mol
SECONDS_PER_DAY = 8.64e+04

# option codes (rad_pcf.F90)
IP_CLOUD_MIX_MAX, IP_CLOUD_MIX_RANDOM, IP_CLOUD_CLEAR = 2, 4, 5
IP_CLOUD_PART_CORR, IP_CLOUD_MCICA = 7, 10
IP_OVERLAP_K_EQV_SCL = 4


class ReplayError(ValueError):
    """An input or option the replay (or the path it drives) does not take."""


@dataclasses.dataclass
class Field:
    """One variable of a CDL/netCDF file: data as (n_vert, n_lat * n_lon) for
    fields with a vertical axis, (1, n_lat * n_lon) otherwise (profiles run
    longitude fastest, as the reference's assign_input routines number them)."""
    name: str
    data: np.ndarray
    plev: np.ndarray = None
    lat: np.ndarray = None
    lon: np.ndarray = None


_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][-+]?\d+)?"


def _parse_cdl(text):
    dims, order, data = {}, {}, {}
    m = re.search(r"dimensions:(.*?)variables:(.*?)data:(.*)\}", text, re.S)
    if not m:
        raise ReplayError("not a CDL file")
    for name, size in re.findall(r"(\w+)\s*=\s*(\d+)\s*;", m.group(1)):
        dims[name] = int(size)
    for name, shape in re.findall(r"^\s*(?:float|double|int|short|long)\s+(\w+)\s*\(([^)]*)\)\s*;",
                                  m.group(2), re.M):
        order[name] = [s.strip() for s in shape.split(",")]
    for name, vals in re.findall(r"(\w+)\s*=\s*([^;]*);", m.group(3)):
        nums = re.findall(_NUM, vals)
        data[name] = np.array([float(v.replace("d", "e").replace("D", "e")) for v in nums])
    return dims, order, data


def read_field(path):
    """The last-declared variable of a CDL text or netCDF classic file (the
    reference's field files hold their coordinates, then the field)."""
    with open(path, "rb") as f:
        head = f.read(4)
    if head[:3] == b"CDF":
        from scipy.io import netcdf_file
        nc = netcdf_file(path, "r", mmap=False)
        try:
            coords = {"lon", "lat", "plev", "basis", "band"}
            names = [k for k in nc.variables if k not in coords]
            if len(names) != 1:
                raise ReplayError("%s: expected one field, found %s" % (path, names))
            v = nc.variables[names[0]]
            dims = list(v.dimensions)
            arr = np.array(v.data, dtype=np.float64)
            get = lambda k: np.array(nc.variables[k].data, dtype=np.float64) if k in nc.variables else None
            plev, lat, lon = get("plev"), get("lat"), get("lon")
        finally:
            nc.close()
    else:
        with open(path) as f:
            dims_, order, data = _parse_cdl(f.read())
        coords = {"lon", "lat", "plev", "basis", "band"}
        names = [k for k in order if k not in coords]
        if len(names) != 1:
            raise ReplayError("%s: expected one field, found %s" % (path, names))
        name = names[0]
        dims = order[name]
        arr = data[name].reshape([dims_[d] for d in dims])
        plev, lat, lon = data.get("plev"), data.get("lat"), data.get("lon")
    # (vertical, lat, lon) -> (vertical, profiles)
    if dims[-2:] != ["lat", "lon"]:
        raise ReplayError("%s: field is not on a (lat, lon) grid: %s" % (path, dims))
    arr = arr.reshape((-1,) + arr.shape[-2:]) if arr.ndim >= 3 else arr.reshape((1,) + arr.shape)
    return Field(names[0], arr.reshape(arr.shape[0], -1), plev, lat, lon)


def _fmt(x):
    """A value as the reference's CDL writer prints it: 0.dddddd E+ee."""
    if x == 0.0 or not np.isfinite(x):
        return " 0.000000E+00" if x == 0.0 else " %12s" % x
    e = int(np.floor(np.log10(abs(x)))) + 1
    m = x / 10.0 ** e
    if abs(round(m, 6)) >= 1.0:
        m /= 10.0
        e += 1
    return "%s0.%06dE%+03d" % ("-" if m < 0 else " ", int(round(abs(m) * 1e6)), e)


def write_cdl(path, name, units, title, plev, lat, lon, values):
    """A (n_vert, n_profile) field in the reference's CDL layout
    (output_flux_cdl: dimensions lon, lat, plev; values 4 to a line)."""
    nv = values.shape[0]
    lines = ["netcdf %s{" % os.path.basename(path), "", "dimensions:",
             "    lon            = %3d;" % len(lon), "    lat            = %3d;" % len(lat),
             "    plev           = %3d;" % nv, "", "", "variables:",
             "    float lon(lon);", '             lon:units = "degree";', '             lon:title = "longitude";',
             "    float lat(lat);", '             lat:units = "degree";', '             lat:title = "latitude";',
             "    float plev(plev);", '             plev:units = "pa";', '             plev:title = "pressure";',
             "", "    float %s(plev,lat,lon);" % name, '             %s:units = "%s";' % (name, units),
             '             %s:title = "%s";' % (name, title), "", "data:"]

    def block(label, vals):
        vals = list(vals)
        out = []
        for i in range(0, len(vals), 4):
            chunk = ",".join(_fmt(v) for v in vals[i:i + 4])
            end = ";" if i + 4 >= len(vals) else ","
            prefix = ("%17s =" % label) if i == 0 else " " * 19
            out.append(prefix + chunk + end)
        return out

    lines += block("lon", lon) + block("lat", lat) + block("plev", plev)
    lines += block(name, values.reshape(-1))
    lines += ["}"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@dataclasses.dataclass
class Options:
    spectral_file: str = ""
    isolir: int = 0                 # 1 solar, 2 infra-red
    bands: tuple = None
    gas_overlap: int = 0
    continuum: bool = False
    rayleigh: bool = False
    cloud_scheme: int = IP_CLOUD_CLEAR
    cloud_representation: int = 0
    drop_type: int = 0
    ice_type: int = 0
    rsd: float = 0.75
    sampling: int = 2
    dp_corr: float = 1.0e4
    rescale: bool = False
    quadratic: bool = False
    scatter: int = 1
    two_stream: int = 0
    solver: int = 0


def parse_args(argv):
    """Cl_run_cdl / Cl_run_cdf options (sbin/Cl_run_cdl:30-90) for the paths
    this replay runs; anything else is refused with ReplayError."""
    o = Options()
    a = list(argv)
    i = 0

    def take(n):
        nonlocal i
        v = a[i + 1:i + 1 + n]
        if len(v) < n:
            raise ReplayError("option %s needs %d values" % (a[i], n))
        i += 1 + n
        return v
    while i < len(a):
        k = a[i]
        if k == "-s": o.spectral_file = take(1)[0]
        elif k == "-B": take(1)
        elif k == "-R": o.bands = tuple(int(v) for v in take(2))
        elif k == "-I": o.isolir = 2; i += 1
        elif k == "-S": o.isolir = 1; i += 1
        elif k == "-g": o.gas_overlap = int(take(1)[0])
        elif k == "-c": o.continuum = True; i += 1
        elif k == "-r": o.rayleigh = True; i += 1
        elif k == "-C": o.cloud_scheme = int(take(1)[0])
        elif k == "-K": o.cloud_representation = int(take(1)[0])
        elif k == "-d": o.drop_type = int(take(1)[0])
        elif k == "-i": o.ice_type = int(take(1)[0])
        elif k == "-m": o.rsd = float(take(1)[0])
        elif k == "+m": o.sampling = int(take(1)[0])
        elif k == "-dp": o.dp_corr = float(take(1)[0])
        elif k == "+R": o.rescale = True; i += 1
        elif k == "-q": o.quadratic = True; i += 1
        elif k == "-z": o.scatter = int(take(1)[0])
        elif k == "-t": o.two_stream = int(take(1)[0])
        elif k == "-v": o.solver = int(take(1)[0])
        else:
            raise ReplayError("option %s is not on the two-stream flux path" % k)
    return o


def _check(o, n_band):
    if not o.spectral_file:
        raise ReplayError("spectral file not set (-s)")
    if o.isolir not in (1, 2):
        raise ReplayError("spectral region not set (-S or -I)")
    if o.bands is not None and tuple(o.bands) != (1, n_band):
        raise ReplayError("band range %s: only the whole file (1 %d) is run" % (o.bands, n_band))
    if o.gas_overlap != IP_OVERLAP_K_EQV_SCL:
        raise ReplayError("gas overlap %d: only equivalent extinction with scaling (-g 4)" % o.gas_overlap)
    if not o.rescale:
        raise ReplayError("the path rescales the optical properties: +R is required")
    if o.two_stream not in (0, 12 if o.isolir == 2 else 16):
        raise ReplayError("two-stream scheme %d: only Elsasser (LW) and PIFM80 (SW)" % o.two_stream)
    if o.cloud_scheme not in (IP_CLOUD_CLEAR, IP_CLOUD_MIX_MAX, IP_CLOUD_MIX_RANDOM, IP_CLOUD_PART_CORR,
                              IP_CLOUD_MCICA):
        raise ReplayError("cloud scheme %d is not on the path" % o.cloud_scheme)
    if o.cloud_scheme != IP_CLOUD_CLEAR and o.cloud_representation != 1:
        raise ReplayError("cloud representation %d: only homogeneous cloud (-K 1)" % o.cloud_representation)
    default_solver = 13 if o.cloud_scheme in (IP_CLOUD_CLEAR, IP_CLOUD_MCICA) else 16
    if o.solver not in (0, default_solver):
        raise ReplayError("solver %d: only the default %d" % (o.solver, default_solver))
    if o.scatter not in (1, 2, 3, 4, 5):
        raise ReplayError("scattering method %d unknown" % o.scatter)


def _n_band(path):
    """Number of bands of a spectral file (block 0's band count line)."""
    with open(path) as f:
        for ln in f:
            if "Number of spectral bands" in ln or "number of spectral bands" in ln.lower():
                return int(re.findall(r"\d+", ln.split("=")[-1])[0])
    raise ReplayError("%s: band count not found" % path)


def _n_subcol(mcica_path):
    with open(mcica_path) as f:
        lines = f.read().split("\n")
    return int(lines[lines.index("tot_subcol_gen") + 1])


def build(base, options, surface="surf"):
    """The runes call of a replay: (Case, spectral file, McICA data file or
    None, grid) from the files <base>.* and parsed options. surface: the
    suffix of the surface file (the examples keep .surflw and .surfsw and
    copy one to .surf before each run)."""
    o = options if isinstance(options, Options) else parse_args(options)
    n_band = _n_band(o.spectral_file)
    _check(o, n_band)

    def fld(sfx, required=True):
        p = "%s.%s" % (base, sfx)
        if not os.path.exists(p):
            if required:
                raise ReplayError("missing input file %s" % p)
            return None
        return read_field(p)
    t = fld("t")
    tl = fld("tl")
    L, P = t.data.shape
    if tl.data.shape != (L + 1, P):
        raise ReplayError("level temperatures do not match the layers")
    p_layer = np.repeat(t.plev[:, None], P, axis=1)
    p_level = np.repeat(tl.plev[:, None], P, axis=1)
    c = Case(n_profile=P, n_layer=L, n_band=n_band)
    c.i_source = o.isolir
    c.p_layer, c.t_layer = p_layer, t.data
    c.p_level, c.t_level = p_level, tl.data
    c.mass = (p_level[1:] - p_level[:-1]) / GRAV_ACC
    gas = {}
    for sfx in ("q", "co2", "o3") + WM_GASES:
        g = fld(sfx, required=False)
        if g is not None:
            if g.data.shape != (L, P):
                raise ReplayError("%s.%s does not match the layers" % (base, sfx))
            gas[sfx] = g.data
    zero = np.zeros((L, P))
    c.h2o = gas.get("q", zero)
    c.o3 = gas.get("o3", zero)
    c.co2 = gas.get("co2", zero)
    c.wm_field = {j: gas[s] for j, s in enumerate(WM_GASES) if s in gas}
    c.wm_mix_ratio = np.zeros(len(WM_GASES))
    c.density = p_layer / (R_GAS_DRY * c.t_layer *
                           (1.0 + (MOL_WEIGHT_AIR / (MOLAR_WEIGHT_H2O * 1.0e-3) - 1.0) * c.h2o))
    c.layer_heat_capacity = c.mass * CP_AIR_DRY
    # Lambertian surface: basis function 1 is the diffuse albedo, 2 (when
    # given) the direct one (rad_pcf.F90 ip_surf_alb_diff, ip_surf_alb_dir)
    alb = fld(surface).data
    c.albedo_diff = np.repeat(alb[0][None, :], n_band, axis=0)
    c.albedo_dir = np.repeat(alb[1 if alb.shape[0] > 1 else 0][None, :], n_band, axis=0)
    if o.isolir == 2:
        c.t_ground = fld("tstar").data[0]
        c.cos_zen = np.ones(P)
        c.solar_irrad = np.zeros(P)
        c.l_ir_source_linear = 0 if o.quadratic else 1
    else:
        c.t_ground = np.zeros(P)
        c.cos_zen = np.cos(np.pi * fld("szen").data[0] / 180.0)
        c.solar_irrad = fld("stoa").data[0]
        c.l_rayleigh = int(o.rayleigh)
    c.l_no_continuum = 0 if o.continuum else 1
    c.i_scatter_method = o.scatter
    mcica = None
    if o.cloud_scheme != IP_CLOUD_CLEAR:
        cf = fld("clfr").data
        # runes takes grid-box mean condensate and divides by the fraction
        # (socrates_set_cld.F90); the driver's files hold in-cloud values
        lwm, re_ = fld("lwm", False), fld("re", False)
        iwm, ire = fld("iwm", False), fld("ire", False)
        if (cf > 0.0).any() and (cf[cf > 0.0] < 1.0e-3).any():
            raise ReplayError("cloud fractions below runes' 0.001 threshold")
        c.l_cloud = 1
        c.i_cloud_repr = 1
        c.cloud_frac = cf
        c.liq_mmr = (lwm.data if lwm is not None else zero) * cf
        c.liq_dim = re_.data if re_ is not None else np.full((L, P), 1.0e-5)
        c.i_st_water = o.drop_type
        if o.ice_type:
            c.i_st_ice = o.ice_type
            c.ice_mmr = (iwm.data if iwm is not None else zero) * cf
            c.ice_dim = ire.data if ire is not None else np.full((L, P), 3.0e-5)
        c.cloud_vertical_decorr = o.dp_corr
        if o.cloud_scheme == IP_CLOUD_MCICA:
            mcica = os.path.join(os.path.dirname(o.spectral_file), "mcica_data")
            c.i_inhom = 2
            c.i_overlap = 2
            c.i_mcica_sampling = o.sampling
            c.cloud_horizontal_rsd = o.rsd
            c.n_subcol_gen = _n_subcol(mcica)
            c.rand_seed = (10 + np.arange(1, P + 1)).astype(np.int32)
        else:
            c.i_inhom = 0
            c.i_overlap = {IP_CLOUD_MIX_MAX: 0, IP_CLOUD_MIX_RANDOM: 1, IP_CLOUD_PART_CORR: 3}[o.cloud_scheme]
    c.fill_defaults()
    grid = dict(lat=t.lat if t.lat is not None else np.zeros(1),
                lon=t.lon if t.lon is not None else np.zeros(1),
                p_layer=t.plev, p_level=tl.plev)
    return c, o.spectral_file, mcica, grid


def diagnostics(case, out, grid, isolir):
    """The driver's output fields from a call's fluxes (l_run_cdl.F90:
    1100-1130): net downward flux, heating rate in K/day from the net flux
    divergence over mass * cp_air_dry, and for the solar region the diffuse
    downward flux."""
    up, down = out["flux_up"], out["flux_down"]
    net = down - up
    hr = (net[:-1] - net[1:]) / (case.mass * CP_AIR_DRY)
    hr = SECONDS_PER_DAY * hr
    res = {"vflx": down, "uflx": up, "nflx": net, "hrts": hr}
    if isolir == 1:
        res["sflx"] = out["flux_direct"]
        res["dflx"] = down - out["flux_direct"]
    return res


_META = {"vflx": ("wm-2", "total downward flux"), "uflx": ("wm-2", "upward flux"),
         "nflx": ("wm-2", "net downward flux"), "hrts": ("k.d-1", "heating rates"),
         "sflx": ("wm-2", "direct flux"), "dflx": ("wm-2", "diffuse downward flux")}


def run(base, argv, out_base=None, precision="fp64", surface="surf"):
    """Replay one driver run on the GPU: read <base>.*, run the call, write
    <out_base>.{vflx,uflx,nflx,hrts[,sflx,dflx]} as CDL when out_base is
    given; returns the output fields ((n_vert, n_profile) arrays)."""
    from . import runes as R
    case, spec, mcica, grid = build(base, argv, surface)
    R.init()
    name = os.path.abspath(spec)
    R.set_spectrum(name, spec)
    if mcica:
        kw = {"sw_spectrum_name": name} if case.i_source == 1 else {"lw_spectrum_name": name}
        R.set_mcica(mcica, **kw)
    out = R.runes_case(case, name, precision=precision)
    res = diagnostics(case, out, grid, case.i_source)
    if out_base:
        for k, v in res.items():
            units, title = _META[k]
            plev = grid["p_layer"] if k == "hrts" else grid["p_level"]
            write_cdl("%s.%s" % (out_base, k), k, units, title, plev, grid["lat"], grid["lon"], v)
    return res


def read_kgo(path):
    """A CDL output file of the reference (values and, for each, the unit of
    its last printed digit): (name, values, ulp) with arrays (n_vert, n_profile)."""
    with open(path) as f:
        text = f.read()
    dims, order, _ = _parse_cdl(text)
    name = [k for k in order if k not in ("lon", "lat", "plev")][0]
    body = text.split("data:")[1]
    m = re.search(r"\b%s\s*=\s*([^;]*);" % name, body)
    toks = re.findall(r"[-+]?0?\.(\d+)E([-+]\d+)", m.group(1))
    raw = re.findall(r"[-+]?0?\.\d+E[-+]\d+", m.group(1))
    vals = np.array([float(v) for v in raw])
    ulp = np.array([10.0 ** (int(e) - len(d)) for d, e in toks])
    shape = [dims[d] for d in order[name]]
    vals = vals.reshape(shape).reshape(shape[0], -1)
    ulp = ulp.reshape(shape).reshape(shape[0], -1)
    return name, vals, ulp
