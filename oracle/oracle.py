"""ctypes front end of the C restatement (oracle/soc_oracle.c).

Test infrastructure only: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libsoc_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class _Case(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "n_profile", "n_layer", "n_band", "i_source", "l_cloud", "i_cloud_repr",
        "i_overlap", "i_inhom", "i_mcica_sampling", "l_rayleigh", "i_st_water",
        "n_subcol_gen")] + [(n, ctypes.c_double) for n in (
        "co2_mix_ratio", "cloud_vertical_decorr", "cloud_horizontal_rsd",
        "conv_vertical_decorr")] + [(n, _D) for n in (
        "p_layer", "t_layer", "mass", "density", "layer_heat_capacity", "h2o", "o3",
        "p_level", "t_level", "t_ground", "cos_zen", "solar_irrad", "albedo_diff",
        "albedo_dir", "cloud_frac", "liq_mmr", "liq_dim", "conv_frac", "liq_conv_mmr",
        "liq_conv_dim")] + [("rand_seed", _I), ("wm_mix_ratio", ctypes.c_double * 10),
                            ("l_flux_ground", ctypes.c_int), ("flux_ground", _D), ("co2", _D),
                            ("wm_field", _D * 10), ("i_st_ice", ctypes.c_int),
                            ("ice_mmr", _D), ("ice_dim", _D), ("ice_conv_mmr", _D),
                            ("ice_conv_dim", _D), ("l_ir_source_linear", ctypes.c_int),
                            ("l_no_continuum", ctypes.c_int), ("i_scatter_method", ctypes.c_int),
                            ("n_aer_mode", ctypes.c_int), ("aer_mix_ratio", _D),
                            ("aer_absorption", _D), ("aer_scattering", _D), ("aer_asymmetry", _D)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(variant=None):
    """The restatement library; variant "jitter" is the exp-jittered build
    used to measure the conditioning of the reference formulas."""
    global _lib
    if variant == "jitter":
        path = LIB.replace("libsoc_oracle.so", "libsoc_oracle_jitter.so")
        if not os.path.exists(path):
            build()
        j = ctypes.CDLL(path)
        j.soc_oracle_runes.restype = ctypes.c_int
        return j
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.soc_oracle_runes.restype = ctypes.c_int
    return _lib


def _ptr(a, keep, dtype=np.float64):
    a = np.ascontiguousarray(a, dtype=dtype)
    keep.append(a)
    return a.ctypes.data_as(_I if dtype == np.int32 else _D)


def runes(case, spectral_file, mcica_file=None, variant=None):
    """Fluxes [level][profile] and heating rate [layer][profile] (K/s)."""
    case.fill_defaults()
    keep = []
    cs = _Case()
    for f, _ in _Case._fields_:
        v = getattr(case, f)
        if f == "rand_seed":
            setattr(cs, f, _ptr(v, keep, np.int32))
        elif f == "wm_mix_ratio":
            setattr(cs, f, (ctypes.c_double * 10)(*[float(x) for x in v]))
        elif f == "wm_field":
            ptrs = (_D * 10)()
            for j, x in (v or {}).items():
                ptrs[j] = _ptr(x, keep)
            setattr(cs, f, ptrs)
        elif isinstance(v, np.ndarray):
            setattr(cs, f, _ptr(v, keep))
        elif v is None:
            setattr(cs, f, None)
        else:
            setattr(cs, f, v)
    L, P = case.n_layer, case.n_profile
    up = np.zeros((L + 1, P)); down = np.zeros((L + 1, P))
    direct = np.zeros((L + 1, P)); hr = np.zeros((L, P))
    rc = lib(variant).soc_oracle_runes(spectral_file.encode(),
                                (mcica_file or "").encode(), ctypes.byref(cs),
                                _ptr(up, keep), _ptr(down, keep), _ptr(direct, keep),
                                _ptr(hr, keep))
    if rc != 0:
        raise RuntimeError("soc_oracle_runes failed: %d" % rc)
    return dict(flux_up=up, flux_down=down, flux_direct=direct, heating_rate=hr)


def cloud_gen(rand_seed, rlc_cf, rlc_cw, sigma_qcw, avg_cf, p_layer, xcw, n_subcol=64,
              n_req=64, ipph=1, ioverlap=3):
    L, P = avg_cf.shape
    keep = []
    c_sub = np.zeros((n_subcol, L, P))
    ncld = np.zeros(P, dtype=np.int32)
    n1, n2 = xcw.shape
    lib().soc_oracle_cloud_gen(
        P, L, n_subcol, n_req, ipph, ioverlap, _ptr(rand_seed, keep, np.int32),
        _ptr(rlc_cf, keep), _ptr(rlc_cw, keep), _ptr(sigma_qcw, keep), _ptr(avg_cf, keep),
        _ptr(p_layer, keep), n1, n2, _ptr(np.asfortranarray(xcw).ravel(order="F"), keep),
        _ptr(c_sub, keep), _ptr(ncld, keep, np.int32))
    return c_sub, ncld


def two_coeff(isolir, tau, omega, asymmetry, sec_0=None):
    n = len(tau)
    keep = []
    outs = [np.zeros(n) for _ in range(5)]
    lib().soc_oracle_two_coeff(n, isolir, _ptr(tau, keep), _ptr(omega, keep),
                               _ptr(asymmetry, keep),
                               _ptr(sec_0 if sec_0 is not None else np.ones(n), keep),
                               *[_ptr(o, keep) for o in outs])
    return dict(zip(("trans", "reflect", "trans_0", "sc1", "sc2"), outs))


def solver_homogen_direct(trans, reflect, s_down, s_up, albedo, flux_inc_down, source_ground):
    L = len(trans)
    keep = []
    ft = np.zeros(2 * L + 2)
    lib().soc_oracle_solver_homogen_direct(
        L, _ptr(trans, keep), _ptr(reflect, keep), _ptr(s_down, keep), _ptr(s_up, keep),
        ctypes.c_double(albedo), ctypes.c_double(flux_inc_down), ctypes.c_double(source_ground),
        _ptr(ft, keep))
    return ft


def two_stream(tsc):
    """soc_oracle_two_stream on a kernel-level case (dict from
    socrates_amd.synth.make_two_stream_case, one term: arrays [layer][profile]).
    Returns (flux_direct [level][profile], flux_up, flux_down [level][profile])."""
    keep = []
    P, L = tsc["n_profile"], tsc["n_layer"]
    fdir = np.zeros((L + 1, P))
    ft = np.zeros((2 * L + 2, P))
    lib().soc_oracle_two_stream(
        P, L, tsc["isolir"], tsc["l_scale_solar"], tsc["l_ir_source_quad"],
        *[_ptr(tsc[k], keep) for k in ("tau", "omega", "asymmetry", "sec_0", "flux_inc_down",
                                       "flux_inc_direct", "diffuse_albedo", "direct_albedo",
                                       "diff_planck", "diff_planck_2", "d_planck_flux_surface",
                                       "adjust_solar_ke")],
        _ptr(fdir, keep), _ptr(ft, keep))
    return fdir, ft[0::2], ft[1::2]
