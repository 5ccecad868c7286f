"""Python mirror of the SOCRATES `runes` interface over the socgpu C ABI.

Names and argument meanings follow the reference interface:
  set_spectrum   -> socrates_set_spectrum::set_spectrum (socrates_set_spectrum.F90:29-149)
  set_mcica      -> socrates_set_spectrum::set_mcica (socrates_set_spectrum.F90:551-694)
  runes          -> socrates_runes::runes (socrates_runes.F90:47-1079), the
                    arguments on the two-stream flux path
Fatal errors of the reference (ereport with i_err_fatal, which STOPs the
process, ereport_mod.F90:27-40) raise SocgpuError here instead.

Arrays may be numpy arrays (host memory) or torch tensors on the GPU (device
memory, no copies); layouts are the reference's full-field layouts:
(n_layer, n_profile) for layer fields, (n_layer+1, n_profile) for levels,
(n_band, n_profile) for albedos. The HIP library must be built
(socrates_amd/build.py); nothing here falls back to a CPU path.
"""
import ctypes
import os

import numpy as np

from .case_io import WM_GASES, Case

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsocgpu.so")
DATA = os.path.join(HERE, "data")

FP64 = 0
MIXED = 1
_PREC = {"fp64": FP64, "mixed": MIXED, FP64: FP64, MIXED: MIXED}

ip_source_illuminate = 1
ip_source_thermal = 2


class SocgpuError(RuntimeError):
    pass


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class RunesArgs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "n_profile", "n_layer", "i_source", "l_rayleigh", "i_cloud_representation",
        "i_overlap", "i_inhom", "i_mcica_sampling", "i_st_water", "i_cnv_water",
        "n_subcol_gen", "l_inputs_on_device")] + [(n, ctypes.c_double) for n in (
        "co2_mix_ratio", "cloud_vertical_decorr", "conv_vertical_decorr",
        "cloud_horizontal_rsd")] + [("l_grey_albedo", ctypes.c_int),
                                    ("grey_albedo", ctypes.c_double)] + [(n, _D) for n in (
        "p_layer", "t_layer", "mass", "density", "layer_heat_capacity", "h2o", "o3",
        "p_level", "t_level", "t_ground", "cos_zenith_angle", "solar_irrad",
        "albedo_diff", "albedo_dir", "cloud_frac", "liq_mmr", "liq_dim", "conv_frac",
        "liq_conv_mmr", "liq_conv_dim")] + [("rand_seed", _I),
                                             ("gas_mix_ratio", ctypes.c_double * len(WM_GASES)),
                                             ("l_flux_ground", ctypes.c_int),
                                             ("l_co2_well_mixed", ctypes.c_int),
                                             ("flux_ground", _D), ("co2", _D),
                                             ("h2o_mix_ratio", ctypes.c_double),
                                             ("o3_mix_ratio", ctypes.c_double),
                                             ("gas_field", _D * len(WM_GASES)),
                                             ("i_st_ice", ctypes.c_int), ("i_cnv_ice", ctypes.c_int),
                                             ("ice_mmr", _D), ("ice_dim", _D), ("ice_conv_mmr", _D),
                                             ("ice_conv_dim", _D), ("l_ir_source_linear", ctypes.c_int),
                                             ("l_no_continuum", ctypes.c_int),
                                             ("i_scatter_method", ctypes.c_int),
                                             ("n_aer_mode", ctypes.c_int), ("aer_mix_ratio", _D),
                                             ("aer_absorption", _D), ("aer_scattering", _D),
                                             ("aer_asymmetry", _D)]


class Diag(ctypes.Structure):
    _fields_ = [("flux_up", _D), ("flux_down", _D), ("flux_direct", _D),
                ("heating_rate", _D), ("l_outputs_on_device", ctypes.c_int),
                ("flux_up_clear", _D), ("flux_down_clear", _D), ("flux_direct_clear", _D)]


# outputs of a call: the StrDiag fields of the path (socrates_def_diag.F90:
# 34-52); the clear-sky ones only when asked for (l_clear)
_OUTPUTS = ("flux_up", "flux_down", "flux_direct", "heating_rate")
_CLEAR_OUTPUTS = ("flux_up_clear", "flux_down_clear", "flux_direct_clear")


_lib = None
_initialised = False


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SocgpuError("socgpu library not built: %s (run socrates_amd/build.py)" % LIB_PATH)
        # torch ships its own HIP runtime: loaded first, it is the one the
        # library binds to, so that torch tensors and the library share one
        # runtime (loaded after the library, torch finds no device)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.socgpu_last_error.restype = ctypes.c_char_p
        for fn in ("socgpu_init", "socgpu_spectrum_register", "socgpu_mcica_register",
                   "socgpu_runes", "socgpu_cloud_gen", "socgpu_two_stream",
                   "socgpu_last_timing", "socgpu_finalize"):
            getattr(_lib, fn).restype = ctypes.c_int
    return _lib


def _check(rc):
    if rc != 0:
        raise SocgpuError("socgpu error %d: %s" % (rc, lib().socgpu_last_error().decode()))


def init(device=0):
    global _initialised
    if not _initialised:
        _check(lib().socgpu_init(int(device)))
        _initialised = True


def set_spectrum(spectrum_name, spectral_file):
    init()
    _check(lib().socgpu_spectrum_register(spectrum_name.encode(), spectral_file.encode()))


def set_mcica(mcica_data_file, sw_spectrum_name=None, lw_spectrum_name=None):
    init()
    _check(lib().socgpu_mcica_register(
        mcica_data_file.encode(),
        sw_spectrum_name.encode() if sw_spectrum_name else None,
        lw_spectrum_name.encode() if lw_spectrum_name else None))


def last_timing():
    t, s = ctypes.c_double(), ctypes.c_double()
    lib().socgpu_last_timing(ctypes.byref(t), ctypes.byref(s))
    return t.value, s.value


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ptr(x, keep, dtype=np.float64):
    if x is None:
        return None
    if _is_torch(x):
        keep.append(x)
        return ctypes.cast(ctypes.c_void_p(x.data_ptr()), _I if dtype == np.int32 else _D)
    a = np.ascontiguousarray(x, dtype=dtype)
    keep.append(a)
    return a.ctypes.data_as(_I if dtype == np.int32 else _D)


def _prepare(n_profile, n_layer, spectrum_name, i_source, p_layer, t_layer, mass,
          density=None, layer_heat_capacity=None, h2o=None, o3=None, co2_mix_ratio=0.0,
          p_level=None, t_level=None, t_ground=None, cos_zenith_angle=None,
          solar_irrad=None, albedo_diff=None, albedo_dir=None, l_grey_albedo=False,
          grey_albedo=0.0, l_rayleigh=False,
          cloud_frac=None, liq_mmr=None, liq_dim=None, conv_frac=None, liq_conv_mmr=None,
          liq_conv_dim=None, cloud_vertical_decorr=0.0, conv_vertical_decorr=0.0,
          cloud_horizontal_rsd=0.0, i_cloud_representation=0, i_overlap=0, i_inhom=0,
          i_mcica_sampling=0, i_st_water=5, i_cnv_water=5, n_subcol_gen=0,
          rand_seed=None, l_flux_ground=False, flux_ground=None, co2=None,
          l_co2_well_mixed=False, h2o_mix_ratio=0.0, o3_mix_ratio=0.0, i_st_ice=0, i_cnv_ice=0,
          ice_mmr=None, ice_dim=None, ice_conv_mmr=None, ice_conv_dim=None, l_ir_source_linear=False,
          l_no_continuum=False, i_scatter_method=0, n_aer_mode=0, l_aerosol_mode=True,
          aer_mix_ratio=None, aer_absorption=None, aer_scattering=None, aer_asymmetry=None,
          l_clear=False, out=None, **gas_amounts):
    """ctypes argument and output structures of one call (and the objects
    that must stay alive while it runs). mix_ratios: the well-mixed mixing
    ratios of further gases, as runes names them (n2o_mix_ratio,
    ch4_mix_ratio, o2_mix_ratio, so2_mix_ratio, ocs_mix_ratio,
    cfc11_mix_ratio, cfc12_mix_ratio, cfc113_mix_ratio, hcfc22_mix_ratio,
    hfc134a_mix_ratio) or, under the gas's own name (n2o, ch4, ...), its
    [layer][profile] field. l_flux_ground / flux_ground [band][profile]:
    surface emission per band; co2 [layer][profile]: CO2 field (co2_mix_ratio
    applies when it is None or l_co2_well_mixed); h2o_mix_ratio /
    o3_mix_ratio: well-mixed H2O / O3 where h2o / o3 is None; i_st_ice,
    i_cnv_ice, ice_mmr, ice_dim, ice_conv_mmr, ice_conv_dim: the ice
    component (none when i_st_ice is 0); l_ir_source_linear, l_no_continuum:
    the general driver's linear infra-red source and continuum switch (runes
    has neither); i_scatter_method: runes' treatment of scattering
    (SOCGPU_SCATTER_*: 1 full, 2 none, 3 extinction as absorption, 4
    approximate, 5 per k-term; 0 the default, full); n_aer_mode,
    aer_mix_ratio [mode][layer][profile], aer_absorption, aer_scattering,
    aer_asymmetry [band][mode][layer][profile]: modal aerosols (off when
    l_aerosol_mode is False, as in runes); l_clear: also the
    clear-sky fluxes flux_up_clear, flux_down_clear, flux_direct_clear (an
    out dict holding any of these keys asks for them as well)."""
    init()
    unknown = set(gas_amounts) - {g + "_mix_ratio" for g in WM_GASES} - set(WM_GASES)
    if unknown:
        raise TypeError("unexpected arguments: %s" % ", ".join(sorted(unknown)))
    inputs = dict(p_layer=p_layer, t_layer=t_layer, mass=mass, density=density,
                  layer_heat_capacity=layer_heat_capacity, h2o=h2o, o3=o3, p_level=p_level,
                  t_level=t_level, t_ground=t_ground, cos_zenith_angle=cos_zenith_angle,
                  solar_irrad=solar_irrad, albedo_diff=albedo_diff, albedo_dir=albedo_dir,
                  cloud_frac=cloud_frac, liq_mmr=liq_mmr, liq_dim=liq_dim, conv_frac=conv_frac,
                  liq_conv_mmr=liq_conv_mmr, liq_conv_dim=liq_conv_dim,
                  flux_ground=flux_ground if l_flux_ground else None, co2=co2, ice_mmr=ice_mmr,
                  ice_dim=ice_dim, ice_conv_mmr=ice_conv_mmr, ice_conv_dim=ice_conv_dim)
    n_aer = int(n_aer_mode) if l_aerosol_mode else 0
    if n_aer > 0:
        inputs.update(aer_mix_ratio=aer_mix_ratio, aer_absorption=aer_absorption,
                      aer_scattering=aer_scattering, aer_asymmetry=aer_asymmetry)
    on_dev = _is_torch(p_layer)
    keep = []
    a = RunesArgs()
    a.n_profile, a.n_layer, a.i_source = int(n_profile), int(n_layer), int(i_source)
    a.l_rayleigh = int(bool(l_rayleigh))
    a.i_cloud_representation, a.i_overlap = int(i_cloud_representation), int(i_overlap)
    a.i_inhom, a.i_mcica_sampling = int(i_inhom), int(i_mcica_sampling)
    a.i_st_water, a.i_cnv_water, a.n_subcol_gen = int(i_st_water), int(i_cnv_water), int(n_subcol_gen)
    a.l_inputs_on_device = int(on_dev)
    a.co2_mix_ratio = float(co2_mix_ratio)
    for j, g in enumerate(WM_GASES):
        a.gas_mix_ratio[j] = float(gas_amounts.get(g + "_mix_ratio", 0.0) or 0.0)
        f = gas_amounts.get(g)
        if f is not None:
            if _is_torch(f) != on_dev:
                raise SocgpuError("all inputs must be in the same memory space (%s)" % g)
            a.gas_field[j] = _ptr(f, keep)
    a.h2o_mix_ratio, a.o3_mix_ratio = float(h2o_mix_ratio), float(o3_mix_ratio)
    a.cloud_vertical_decorr, a.conv_vertical_decorr = float(cloud_vertical_decorr), float(conv_vertical_decorr)
    a.cloud_horizontal_rsd = float(cloud_horizontal_rsd)
    a.l_grey_albedo, a.grey_albedo = int(bool(l_grey_albedo)), float(grey_albedo)
    a.l_flux_ground, a.l_co2_well_mixed = int(bool(l_flux_ground)), int(bool(l_co2_well_mixed))
    a.i_st_ice, a.i_cnv_ice = int(i_st_ice), int(i_cnv_ice)
    a.l_ir_source_linear, a.l_no_continuum = int(bool(l_ir_source_linear)), int(bool(l_no_continuum))
    a.i_scatter_method = int(i_scatter_method)
    a.n_aer_mode = max(n_aer, 0)
    for k, v in inputs.items():
        if v is not None and _is_torch(v) != on_dev:
            raise SocgpuError("all inputs must be in the same memory space (%s)" % k)
        setattr(a, k, _ptr(v, keep))
    a.rand_seed = _ptr(rand_seed, keep, np.int32)
    L, P = int(n_layer), int(n_profile)
    if out is None:
        names = _OUTPUTS + (_CLEAR_OUTPUTS if l_clear else ())
        if on_dev:
            import torch
            dev = p_layer.device
            out = {k: torch.empty((L if k == "heating_rate" else L + 1, P), dtype=torch.float64,
                                  device=dev) for k in names}
        else:
            out = {k: np.zeros((L if k == "heating_rate" else L + 1, P)) for k in names}
    d = Diag()
    for k in _OUTPUTS + _CLEAR_OUTPUTS:
        setattr(d, k, _ptr(out.get(k), keep))
    d.l_outputs_on_device = int(on_dev)
    if not on_dev:
        for k in out:
            if out[k] is not None and not out[k].flags["C_CONTIGUOUS"]:
                raise SocgpuError("output %s must be contiguous" % k)
    return a, d, out, keep


def runes(n_profile, n_layer, spectrum_name, i_source, p_layer, t_layer, mass, precision="fp64",
          **kw):
    """One radiation call (socgpu_runes). Returns dict(flux_up, flux_down,
    flux_direct, heating_rate) in the input memory space (numpy or torch)."""
    inv = bool(kw.pop("l_invert", False))
    if inv:
        # bottom-up profiles (socrates_set_atm.F90:218-225): flipped into
        # top-down copies, the outputs flipped back (socrates_set_diag.F90:113-151)
        p_layer, t_layer, mass = _flip(p_layer), _flip(t_layer), _flip(mass)
        for k in _VERTICAL:
            if kw.get(k) is not None:
                kw[k] = _flip(kw[k])
        user_out = kw.pop("out", None)
    a, d, out, keep = _prepare(n_profile, n_layer, spectrum_name, i_source, p_layer, t_layer, mass,
                               **kw)
    _check(lib().socgpu_runes(spectrum_name.encode(), ctypes.byref(a), ctypes.byref(d),
                              _PREC[precision]))
    if inv:
        flipped = {k: _flip(v) for k, v in out.items()}
        if user_out:
            for k, v in user_out.items():
                if v is not None:
                    v.copy_(flipped[k]) if _is_torch(v) else np.copyto(v, flipped[k])
            return user_out
        return flipped
    return out


def runes_noalloc(n_profile, n_layer, spectrum_name, i_source, p_layer, t_layer, mass,
                  precision="fp64", **kw):
    """runes() through the allocation-free entry point
    (socgpu_runes_noalloc): inputs and outputs must be torch tensors on the
    GPU; the call's workspace is the library's fixed device array, so calls
    above its size fail with SocgpuError."""
    if not _is_torch(p_layer):
        raise SocgpuError("runes_noalloc takes device (torch) arrays")
    a, d, out, keep = _prepare(n_profile, n_layer, spectrum_name, i_source, p_layer, t_layer, mass,
                               **kw)
    _check(lib().socgpu_runes_noalloc(spectrum_name.encode(), ctypes.byref(a), ctypes.byref(d),
                                      _PREC[precision]))
    return out


# vertical fields of runes(), [layer or level][profile]
_VERTICAL = ("density", "layer_heat_capacity", "h2o", "o3", "p_level", "t_level", "cloud_frac",
             "liq_mmr", "liq_dim", "conv_frac", "liq_conv_mmr", "liq_conv_dim", "co2", "ice_mmr",
             "ice_dim", "ice_conv_mmr", "ice_conv_dim") + WM_GASES


def _flip(a):
    """A copy of a with its first (vertical) axis reversed."""
    if _is_torch(a):
        import torch
        return torch.flip(a, [0]).contiguous()
    return np.ascontiguousarray(np.asarray(a)[::-1])


def runes_case(case: Case, spectrum_name, precision="fp64", device_arrays=None, out=None):
    """runes() on a Case (socrates_amd.case_io). device_arrays: optional dict
    of torch tensors replacing the case's numpy arrays."""
    name, kw = case_call(case, spectrum_name, device_arrays, out)
    P, L, src = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
    return runes(P, L, name, src, precision=precision, **kw)


def register_default_spectra():
    """Register the synthetic LW/SW spectral files and McICA data shipped in
    socrates_amd/data as 'lw' and 'sw'."""
    set_spectrum("sw", os.path.join(DATA, "sp_sw_synth"))
    set_spectrum("lw", os.path.join(DATA, "sp_lw_synth"))
    set_mcica(os.path.join(DATA, "mcica_synth"), sw_spectrum_name="sw", lw_spectrum_name="lw")


def cloud_gen(rand_seed, rlc_cf, rlc_cw, sigma_qcw, avg_cf, p_layer, xcw, n_subcol=64,
              n_req=64, ipph=1, ioverlap=3):
    """Device McICA generator on host arrays; returns (c_sub[sub,layer,prof], n_subcol_cld)."""
    init()
    L, P = avg_cf.shape
    keep = []
    c_sub = np.zeros((n_subcol, L, P))
    ncld = np.zeros(P, dtype=np.int32)
    n1, n2 = xcw.shape
    _check(lib().socgpu_cloud_gen(
        P, L, n_subcol, n_req, ipph, ioverlap, _ptr(rand_seed, keep, np.int32),
        _ptr(rlc_cf, keep), _ptr(rlc_cw, keep), _ptr(sigma_qcw, keep), _ptr(avg_cf, keep),
        _ptr(p_layer, keep), n1, n2, _ptr(np.asfortranarray(xcw).ravel(order="F"), keep),
        _ptr(c_sub, keep), _ptr(ncld, keep, np.int32)))
    return c_sub, ncld


_TS_INPUTS = ("tau", "omega", "asymmetry", "sec_0", "flux_inc_down", "flux_inc_direct",
              "diffuse_albedo", "direct_albedo", "diff_planck", "diff_planck_2",
              "d_planck_flux_surface", "adjust_solar_ke")


def two_stream(n_profile, n_layer, n_term, isolir, tau, omega, asymmetry, sec_0=None,
               flux_inc_down=None, flux_inc_direct=None, diffuse_albedo=None,
               direct_albedo=None, diff_planck=None, diff_planck_2=None,
               d_planck_flux_surface=None, adjust_solar_ke=None, l_scale_solar=True,
               l_ir_source_quad=True, precision="fp64", out=None):
    """Kernel-level two-stream solve (socgpu_two_stream; the reference's
    two_stream, two_stream.F90:160-205) on prescribed optical properties.
    Arrays [term][layer][profile] for tau/omega/asymmetry, [layer][profile]
    or [profile] for the rest, numpy or torch device tensors (all alike).
    Returns dict(flux_direct, flux_up, flux_down), each [term][level][profile]
    (infra-red fluxes differential)."""
    init()
    G, L, P = int(n_term), int(n_layer), int(n_profile)
    vals = dict(tau=tau, omega=omega, asymmetry=asymmetry, sec_0=sec_0,
                flux_inc_down=flux_inc_down, flux_inc_direct=flux_inc_direct,
                diffuse_albedo=diffuse_albedo, direct_albedo=direct_albedo,
                diff_planck=diff_planck, diff_planck_2=diff_planck_2,
                d_planck_flux_surface=d_planck_flux_surface, adjust_solar_ke=adjust_solar_ke)
    on_dev = _is_torch(tau)
    for k, v in vals.items():
        if v is not None and _is_torch(v) != on_dev:
            raise SocgpuError("all inputs must be in the same memory space (%s)" % k)
    if out is None:
        if on_dev:
            import torch
            mk = lambda: torch.empty((G, L + 1, P), dtype=torch.float64, device=tau.device)  # noqa: E731
        else:
            mk = lambda: np.zeros((G, L + 1, P))  # noqa: E731
        out = dict(flux_direct=mk(), flux_up=mk(), flux_down=mk())
    keep = []
    _check(lib().socgpu_two_stream(
        P, L, G, int(isolir), int(bool(l_scale_solar)), int(bool(l_ir_source_quad)),
        *[_ptr(vals[k], keep) for k in _TS_INPUTS],
        _ptr(out["flux_direct"], keep), _ptr(out["flux_up"], keep), _ptr(out["flux_down"], keep),
        int(on_dev), _PREC[precision]))
    return out


def two_stream_case(tsc, precision="fp64", out=None):
    """two_stream() on a case dict of socrates_amd.synth.make_two_stream_case."""
    return two_stream(tsc["n_profile"], tsc["n_layer"], tsc["n_term"], tsc["isolir"],
                      **{k: tsc[k] for k in _TS_INPUTS},
                      l_scale_solar=tsc["l_scale_solar"],
                      l_ir_source_quad=tsc["l_ir_source_quad"], precision=precision, out=out)


class Tile(ctypes.Structure):
    """Launch shape of the column-solve kernel (SocgpuTile)."""
    _fields_ = [("recompute", ctypes.c_int), ("waves_per_block", ctypes.c_int),
                ("waves_per_simd", ctypes.c_int)]


def set_tile(precision, recompute, waves_per_block, waves_per_simd=4):
    init()
    t = Tile(int(bool(recompute)), int(waves_per_block), int(waves_per_simd))
    _check(lib().socgpu_set_tile(_PREC[precision], ctypes.byref(t)))


def get_tile(precision):
    init()
    t = Tile()
    _check(lib().socgpu_get_tile(_PREC[precision], ctypes.byref(t)))
    return dict(recompute=t.recompute, waves_per_block=t.waves_per_block,
                waves_per_simd=t.waves_per_simd)



class Request(ctypes.Structure):
    _fields_ = [("spectrum_name", ctypes.c_char_p), ("args", ctypes.POINTER(RunesArgs)),
                ("diag", ctypes.POINTER(Diag))]


class Requests:
    """A list of radiation calls marshalled once into the C request array
    (SocgpuRequest[]) with its outputs, for callers that run the same calls
    (the same input and output buffers) every time step."""

    def __init__(self, calls):
        self.arr, self.outs, self._keep = _requests(calls)
        self.n = len(calls)


def _requests(calls):
    """calls: list of (spectrum_name, kwargs of runes()) -> ctypes array, outputs, keep-alive."""
    if isinstance(calls, Requests):
        return calls.arr, calls.outs, calls._keep
    keep, outs = [], []
    arr = (Request * len(calls))()
    for k, (name, kw) in enumerate(calls):
        kw = dict(kw)
        P, L, src = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
        a, d, out, kp = _prepare(P, L, name, src, **kw)
        nm = name.encode()
        keep += [a, d, kp, nm]
        arr[k].spectrum_name = nm
        arr[k].args = ctypes.pointer(a)
        arr[k].diag = ctypes.pointer(d)
        outs.append(out)
    return arr, outs, keep


def runes_batch(calls, precision="fp64"):
    """Several radiation calls in one packed launch (socgpu_runes_batch).
    calls: list of (spectrum_name, dict of runes() keyword arguments
    including n_profile, n_layer and i_source). Returns the list of output
    dicts, in order."""
    arr, outs, keep = _requests(calls)
    _check(lib().socgpu_runes_batch(len(outs), arr, _PREC[precision]))
    return outs


def case_call(case, spectrum_name, device_arrays=None, out=None):
    """(spectrum_name, kwargs) of runes() for a Case, as runes_case builds them."""
    src = dict(device_arrays) if device_arrays else {}

    def g(name, attr=None):
        return src.get(name, getattr(case, attr or name))

    cloudy = bool(case.l_cloud)
    conv = cloudy and case.i_cloud_repr == 3   # stratiform + convective (triple column)
    kw = dict(
        n_profile=case.n_profile, n_layer=case.n_layer, i_source=case.i_source,
        p_layer=g("p_layer"), t_layer=g("t_layer"), mass=g("mass"), density=g("density"),
        layer_heat_capacity=g("layer_heat_capacity"), h2o=g("h2o"), o3=g("o3"),
        co2_mix_ratio=case.co2_mix_ratio, p_level=g("p_level"), t_level=g("t_level"),
        t_ground=g("t_ground"), cos_zenith_angle=g("cos_zenith_angle", "cos_zen"),
        solar_irrad=g("solar_irrad"), albedo_diff=g("albedo_diff"), albedo_dir=g("albedo_dir"),
        l_rayleigh=case.l_rayleigh,
        cloud_frac=g("cloud_frac") if cloudy else None,
        liq_mmr=g("liq_mmr") if cloudy else None, liq_dim=g("liq_dim") if cloudy else None,
        conv_frac=g("conv_frac") if conv else None,
        liq_conv_mmr=g("liq_conv_mmr") if conv else None,
        liq_conv_dim=g("liq_conv_dim") if conv else None,
        cloud_vertical_decorr=case.cloud_vertical_decorr,
        conv_vertical_decorr=case.conv_vertical_decorr,
        cloud_horizontal_rsd=case.cloud_horizontal_rsd,
        i_cloud_representation=case.i_cloud_repr if cloudy else 0, i_overlap=case.i_overlap,
        i_inhom=case.i_inhom, i_mcica_sampling=case.i_mcica_sampling,
        i_st_water=case.i_st_water, i_cnv_water=case.i_st_water,
        n_subcol_gen=case.n_subcol_gen, rand_seed=g("rand_seed") if cloudy else None, out=out)
    if case.l_flux_ground:
        kw.update(l_flux_ground=True, flux_ground=g("flux_ground"))
    if case.co2 is not None:
        kw["co2"] = g("co2")
    if case.l_ir_source_linear:
        kw["l_ir_source_linear"] = True
    if case.l_no_continuum:
        kw["l_no_continuum"] = True
    if case.i_scatter_method:
        kw["i_scatter_method"] = case.i_scatter_method
    if case.n_aer_mode > 0:
        kw.update(n_aer_mode=case.n_aer_mode, aer_mix_ratio=g("aer_mix_ratio"),
                  aer_absorption=g("aer_absorption"), aer_scattering=g("aer_scattering"),
                  aer_asymmetry=g("aer_asymmetry"))
    if cloudy and case.ice_mmr is not None:
        kw.update(i_st_ice=case.i_st_ice, i_cnv_ice=case.i_st_ice, ice_mmr=g("ice_mmr"),
                  ice_dim=g("ice_dim"))
        if conv:
            kw.update(ice_conv_mmr=g("ice_conv_mmr"), ice_conv_dim=g("ice_conv_dim"))
    for j, x in (case.wm_field or {}).items():
        kw[WM_GASES[j]] = src.get(WM_GASES[j], x)
    if case.wm_mix_ratio is not None:
        for j, gname in enumerate(WM_GASES):
            if case.wm_mix_ratio[j] != 0.0:
                kw[gname + "_mix_ratio"] = float(case.wm_mix_ratio[j])
    return spectrum_name, kw


class Batch:
    """Persistent graph-captured batch (socgpu_batch_*): replays one captured
    launch sequence per radiation call while the batch fits its capacity."""

    def __init__(self, precision="fp64"):
        init()
        self._h = ctypes.c_void_p()
        _check(lib().socgpu_batch_create(ctypes.byref(self._h), _PREC[precision]))

    def run(self, calls):
        """calls: a list of (spectrum_name, kwargs) or a Requests built once."""
        arr, outs, keep = _requests(calls)
        _check(lib().socgpu_batch_run(self._h, len(outs), arr))
        return outs

    def stats(self):
        ms, cap, rep = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        _check(lib().socgpu_batch_stats(self._h, ctypes.byref(ms), ctypes.byref(cap),
                                        ctypes.byref(rep)))
        return dict(last_ms=ms.value, captures=cap.value, replays=rep.value)

    def close(self):
        if self._h:
            lib().socgpu_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(calls, precision="fp64", repeats=2):
    """Pick the fastest solve tile for a representative batch of calls
    (socgpu_autotune); it stays in effect for the precision form."""
    arr, outs, keep = _requests(calls)
    t = Tile()
    _check(lib().socgpu_autotune(len(outs), arr, _PREC[precision], int(repeats), ctypes.byref(t)))
    return dict(recompute=t.recompute, waves_per_block=t.waves_per_block,
                waves_per_simd=t.waves_per_simd)
