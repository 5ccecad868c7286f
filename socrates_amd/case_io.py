"""Column batch container and its binary file form.

A Case holds one `runes` call's inputs (socrates_runes.F90:47-168) as numpy
arrays in the interface's full-field layout: Fortran (n_profile, n_layer),
i.e. numpy shape (n_layer, n_profile), profile fastest, layer 1 at the top;
level fields have n_layer+1 rows (level 0 = top of atmosphere); per-band
surface albedos have shape (n_band, n_profile).

The binary form is read by the reference harness in oracle/harness and by
the tests; all reals are float64, integers int32, native byte order.
"""
import dataclasses

import numpy as np

MAGIC = 1397703491
VERSION = 8     # 2 adds the well-mixed mixing ratios of WM_GASES after the seeds;
                # 3 adds l_flux_ground and flux_ground (n_band, n_profile) after them;
                # 4 adds a flag and, when it is set, a CO2 field (n_layer, n_profile);
                # 5 adds a flag per WM_GASES gas and, for each one set, its field;
                # 6 adds a flag and, when it is set, ice_mmr, ice_dim, ice_conv_mmr,
                #   ice_conv_dim (n_layer, n_profile);
                # 7 adds i_scatter_method;
                # 8 adds n_aer_mode and, when it is positive, aer_mix_ratio
                #   (n_aer_mode, n_layer, n_profile), aer_absorption,
                #   aer_scattering, aer_asymmetry (n_band, n_aer_mode, n_layer, n_profile)

# further gases runes takes as well-mixed mixing ratios (the *_mix_ratio
# arguments, socrates_runes.F90:69-83), in this order
WM_GASES = ("n2o", "ch4", "o2", "so2", "ocs", "cfc11", "cfc12", "cfc113", "hcfc22", "hfc134a")

_HDR = ("n_profile", "n_layer", "n_band", "i_source", "l_cloud", "i_cloud_repr",
        "i_overlap", "i_inhom", "i_mcica_sampling", "l_rayleigh", "i_st_water",
        "i_st_ice", "n_subcol_gen", "i_solver_kind")
_SCALARS = ("co2_mix_ratio", "cloud_vertical_decorr", "cloud_horizontal_rsd",
            "conv_vertical_decorr")
_LAYER = ("p_layer", "t_layer", "mass", "density", "layer_heat_capacity", "h2o", "o3")
_LEVEL = ("p_level", "t_level")
_COLUMN = ("t_ground", "cos_zen", "solar_irrad")
_BAND = ("albedo_diff", "albedo_dir")
_CLOUD = ("cloud_frac", "liq_mmr", "liq_dim")
_CONV = ("conv_frac", "liq_conv_mmr", "liq_conv_dim")
_ICE = ("ice_mmr", "ice_dim", "ice_conv_mmr", "ice_conv_dim")
_AER = ("aer_mix_ratio", "aer_absorption", "aer_scattering", "aer_asymmetry")


@dataclasses.dataclass
class Case:
    n_profile: int
    n_layer: int
    n_band: int
    i_source: int = 2
    l_cloud: int = 0
    i_cloud_repr: int = 0
    i_overlap: int = 0
    i_inhom: int = 0
    i_mcica_sampling: int = 0
    l_rayleigh: int = 0
    i_st_water: int = 5
    i_st_ice: int = 8
    n_subcol_gen: int = 0
    i_solver_kind: int = 0
    co2_mix_ratio: float = 0.0
    cloud_vertical_decorr: float = 0.0
    cloud_horizontal_rsd: float = 0.0
    conv_vertical_decorr: float = 0.0
    p_layer: np.ndarray = None
    t_layer: np.ndarray = None
    mass: np.ndarray = None
    density: np.ndarray = None
    layer_heat_capacity: np.ndarray = None
    h2o: np.ndarray = None
    o3: np.ndarray = None
    p_level: np.ndarray = None
    t_level: np.ndarray = None
    t_ground: np.ndarray = None
    cos_zen: np.ndarray = None
    solar_irrad: np.ndarray = None
    albedo_diff: np.ndarray = None
    albedo_dir: np.ndarray = None
    cloud_frac: np.ndarray = None
    liq_mmr: np.ndarray = None
    liq_dim: np.ndarray = None
    conv_frac: np.ndarray = None
    liq_conv_mmr: np.ndarray = None
    liq_conv_dim: np.ndarray = None
    rand_seed: np.ndarray = None
    wm_mix_ratio: np.ndarray = None   # by WM_GASES, 0 = absent
    l_flux_ground: int = 0            # surface emission per band given (flux_ground)
    flux_ground: np.ndarray = None    # (n_band, n_profile) W m-2
    co2: np.ndarray = None            # (n_layer, n_profile) CO2 mmr; None: co2_mix_ratio
    wm_field: dict = None             # {WM_GASES index: (n_layer, n_profile) field}
    ice_mmr: np.ndarray = None        # (n_layer, n_profile); None: no ice condensate
    ice_dim: np.ndarray = None
    ice_conv_mmr: np.ndarray = None
    ice_conv_dim: np.ndarray = None
    l_ir_source_linear: int = 0       # infra-red source linear in tau (not written to files)
    l_no_continuum: int = 0           # continuum absorption left out (not written to files)
    i_scatter_method: int = 0         # treatment of scattering (0: runes' default, full)
    n_aer_mode: int = 0               # modal aerosols (runes' n_aer_mode); 0: none
    aer_mix_ratio: np.ndarray = None  # (n_aer_mode, n_layer, n_profile) kg/kg
    aer_absorption: np.ndarray = None  # (n_band, n_aer_mode, n_layer, n_profile) m2/kg
    aer_scattering: np.ndarray = None  # (n_band, n_aer_mode, n_layer, n_profile) m2/kg
    aer_asymmetry: np.ndarray = None   # (n_band, n_aer_mode, n_layer, n_profile)

    def fill_defaults(self):
        L, P = self.n_layer, self.n_profile
        for k in _CLOUD + _CONV:
            if getattr(self, k) is None:
                setattr(self, k, np.zeros((L, P)) if k != "liq_dim" and k != "liq_conv_dim"
                        else np.full((L, P), 1.0e-5))
        if self.rand_seed is None:
            self.rand_seed = np.arange(P, dtype=np.int32)
        if self.wm_mix_ratio is None:
            self.wm_mix_ratio = np.zeros(len(WM_GASES))
        if self.flux_ground is None:
            self.flux_ground = np.zeros((self.n_band, P))
        if self.t_ground is None:
            self.t_ground = np.zeros(P)
        return self

    def subset(self, cols):
        """The same case restricted to the given column indices."""
        cols = np.asarray(cols)
        out = dataclasses.replace(self, n_profile=int(cols.size))
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if isinstance(v, np.ndarray) and f.name != "wm_mix_ratio":
                setattr(out, f.name, np.ascontiguousarray(v[..., cols]))
            elif isinstance(v, dict):
                setattr(out, f.name, {j: np.ascontiguousarray(x[..., cols]) for j, x in v.items()})
        return out

    def write(self, path):
        self.fill_defaults()
        with open(path, "wb") as f:
            np.array([MAGIC, VERSION] + [getattr(self, k) for k in _HDR],
                     dtype=np.int32).tofile(f)
            np.array([getattr(self, k) for k in _SCALARS], dtype=np.float64).tofile(f)
            for k in _LAYER + _LEVEL + _COLUMN + _BAND + _CLOUD + _CONV:
                np.ascontiguousarray(getattr(self, k), dtype=np.float64).tofile(f)
            np.ascontiguousarray(self.rand_seed, dtype=np.int32).tofile(f)
            np.ascontiguousarray(self.wm_mix_ratio, dtype=np.float64).tofile(f)
            np.array([int(self.l_flux_ground)], dtype=np.int32).tofile(f)
            np.ascontiguousarray(self.flux_ground, dtype=np.float64).tofile(f)
            np.array([int(self.co2 is not None)], dtype=np.int32).tofile(f)
            if self.co2 is not None:
                np.ascontiguousarray(self.co2, dtype=np.float64).tofile(f)
            fields = self.wm_field or {}
            np.array([int(j in fields) for j in range(len(WM_GASES))], dtype=np.int32).tofile(f)
            for j in range(len(WM_GASES)):
                if j in fields:
                    np.ascontiguousarray(fields[j], dtype=np.float64).tofile(f)
            has_ice = self.ice_mmr is not None
            np.array([int(has_ice)], dtype=np.int32).tofile(f)
            if has_ice:
                L, P = self.n_layer, self.n_profile
                for k in _ICE:
                    v = getattr(self, k)
                    v = np.zeros((L, P)) if v is None else v
                    np.ascontiguousarray(v, dtype=np.float64).tofile(f)
            np.array([int(self.i_scatter_method)], dtype=np.int32).tofile(f)
            np.array([int(self.n_aer_mode)], dtype=np.int32).tofile(f)
            if self.n_aer_mode > 0:
                for k in _AER:
                    np.ascontiguousarray(getattr(self, k), dtype=np.float64).tofile(f)

    @classmethod
    def read(cls, path):
        raw = open(path, "rb").read()
        hdr = np.frombuffer(raw, dtype=np.int32, count=2 + len(_HDR))
        if hdr[0] != MAGIC:
            raise ValueError("not a case file: %s" % path)
        kw = dict(zip(_HDR, (int(x) for x in hdr[2:])))
        c = cls(**kw)
        off = 4 * (2 + len(_HDR))
        sc = np.frombuffer(raw, dtype=np.float64, count=len(_SCALARS), offset=off)
        off += 8 * len(_SCALARS)
        for k, v in zip(_SCALARS, sc):
            setattr(c, k, float(v))
        L, P, B = c.n_layer, c.n_profile, c.n_band
        shapes = {**{k: (L, P) for k in _LAYER + _CLOUD + _CONV},
                  **{k: (L + 1, P) for k in _LEVEL}, **{k: (P,) for k in _COLUMN},
                  **{k: (B, P) for k in _BAND}}
        for k in _LAYER + _LEVEL + _COLUMN + _BAND + _CLOUD + _CONV:
            n = int(np.prod(shapes[k]))
            setattr(c, k, np.frombuffer(raw, dtype=np.float64, count=n,
                                        offset=off).reshape(shapes[k]).copy())
            off += 8 * n
        c.rand_seed = np.frombuffer(raw, dtype=np.int32, count=P, offset=off).copy()
        off += 4 * P
        if hdr[1] >= 2:
            c.wm_mix_ratio = np.frombuffer(raw, dtype=np.float64, count=len(WM_GASES), offset=off).copy()
            off += 8 * len(WM_GASES)
        if hdr[1] >= 3:
            c.l_flux_ground = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=off)[0])
            off += 4
            c.flux_ground = np.frombuffer(raw, dtype=np.float64, count=B * P, offset=off).reshape(B, P).copy()
            off += 8 * B * P
        if hdr[1] >= 4:
            has_co2 = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=off)[0])
            off += 4
            if has_co2:
                c.co2 = np.frombuffer(raw, dtype=np.float64, count=L * P, offset=off).reshape(L, P).copy()
                off += 8 * L * P
        if hdr[1] >= 5:
            flags = np.frombuffer(raw, dtype=np.int32, count=len(WM_GASES), offset=off)
            off += 4 * len(WM_GASES)
            c.wm_field = {}
            for j in range(len(WM_GASES)):
                if flags[j]:
                    c.wm_field[j] = np.frombuffer(raw, dtype=np.float64, count=L * P,
                                                  offset=off).reshape(L, P).copy()
                    off += 8 * L * P
        if hdr[1] >= 6:
            has_ice = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=off)[0])
            off += 4
            if has_ice:
                for k in _ICE:
                    setattr(c, k, np.frombuffer(raw, dtype=np.float64, count=L * P,
                                                offset=off).reshape(L, P).copy())
                    off += 8 * L * P
        if hdr[1] >= 7:
            c.i_scatter_method = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=off)[0])
            off += 4
        if hdr[1] >= 8:
            M = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=off)[0])
            off += 4
            c.n_aer_mode = M
            if M > 0:
                for k in _AER:
                    shape = (M, L, P) if k == "aer_mix_ratio" else (B, M, L, P)
                    n = int(np.prod(shape))
                    setattr(c, k, np.frombuffer(raw, dtype=np.float64, count=n,
                                                offset=off).reshape(shape).copy())
                    off += 8 * n
        return c


def read_output(path, n_profile, n_layer, clear=False):
    """Outputs of the reference harness: flux_up, flux_down, flux_direct
    (n_layer+1, n_profile) and heating_rate (n_layer, n_profile) [K/s]; with
    clear, also the clear-sky flux_up_clear, flux_down_clear and
    flux_direct_clear (n_layer+1, n_profile) the harness writes after them."""
    a = np.fromfile(path, dtype=np.float64)
    P, L = n_profile, n_layer
    n1 = P * (L + 1)
    up = a[:n1].reshape(L + 1, P)
    down = a[n1:2 * n1].reshape(L + 1, P)
    direct = a[2 * n1:3 * n1].reshape(L + 1, P)
    hr = a[3 * n1:3 * n1 + P * L].reshape(L, P)
    out = dict(flux_up=up, flux_down=down, flux_direct=direct, heating_rate=hr)
    if clear:
        o = 3 * n1 + P * L
        for k in ("flux_up_clear", "flux_down_clear", "flux_direct_clear"):
            out[k] = a[o:o + n1].reshape(L + 1, P)
            o += n1
    return out


def case_to_arrays(case):
    """Flat dict of a Case (for npz storage)."""
    case.fill_defaults()
    out = {}
    for f in dataclasses.fields(case):
        v = getattr(case, f.name)
        if v is None:       # optional field not given (co2)
            continue
        if isinstance(v, dict):
            for j, x in v.items():
                out["case_%s_%d" % (f.name, j)] = np.asarray(x)
            continue
        out["case_" + f.name] = np.asarray(v)
    return out


def case_from_arrays(d):
    kw = {}
    for f in dataclasses.fields(Case):
        if f.name == "wm_field":
            fields = {j: d["case_wm_field_%d" % j] for j in range(len(WM_GASES))
                      if "case_wm_field_%d" % j in d}
            kw[f.name] = fields or None
            continue
        if "case_" + f.name not in d:   # fields added after the arrays were written
            continue
        v = d["case_" + f.name]
        if v.ndim == 0:
            v = v.item()
        kw[f.name] = v
    return Case(**kw)
