"""Per-output error statistics between two sets of `runes` outputs.

The outputs compared are the ones the diagnostics carry (socrates_def_diag.F90:
31-76): upward, downward and net (down - up) flux at the levels in W m-2,
and the heating rate at the layers, reported in K/day (the diagnostic is in
K s-1; runes_driver.f90:188 scales it by 86400 for printing).

`stats(a, b)` returns, per output, the largest and the 99th-percentile
absolute difference and the largest relative difference (|a - b| / |b| where
|b| exceeds the absolute floor). Inputs are numpy arrays or torch tensors
(device tensors are reduced on the device).
"""
import numpy as np

SECONDS_PER_DAY = 86400.0


def _np(x):
    if type(x).__module__.startswith("torch"):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _one(d, ref, floor):
    d = np.abs(d).ravel()
    r = np.abs(ref).ravel()
    if d.size == 0:
        return dict(max=0.0, p99=0.0, max_rel=0.0)
    rel = np.where(r > floor, d / np.maximum(r, floor), 0.0)
    return dict(max=float(d.max()), p99=float(np.percentile(d, 99.0)), max_rel=float(rel.max()))


def stats(out, ref, flux_floor=1e-9, hr_floor_kday=1e-9):
    """Error of `out` against `ref` (dicts with flux_up, flux_down and
    heating_rate [K s-1]; flux_direct too when both carry it)."""
    up, dn = _np(out["flux_up"]), _np(out["flux_down"])
    rup, rdn = _np(ref["flux_up"]), _np(ref["flux_down"])
    res = {
        "flux_up": _one(up - rup, rup, flux_floor),
        "flux_down": _one(dn - rdn, rdn, flux_floor),
        "flux_net": _one((dn - up) - (rdn - rup), rdn - rup, flux_floor),
    }
    if out.get("flux_direct") is not None and ref.get("flux_direct") is not None:
        fx, rfx = _np(out["flux_direct"]), _np(ref["flux_direct"])
        res["flux_direct"] = _one(fx - rfx, rfx, flux_floor)
    hr = _np(out["heating_rate"]) * SECONDS_PER_DAY
    rhr = _np(ref["heating_rate"]) * SECONDS_PER_DAY
    res["heating_rate_kday"] = _one(hr - rhr, rhr, hr_floor_kday)
    return res


def line(label, st):
    """One printable line of the statistics."""
    parts = []
    for k, v in st.items():
        unit = "K/day" if k.startswith("heating") else "W m-2"
        parts.append("%s max %.3e p99 %.3e %s (rel %.2e)" % (k, v["max"], v["p99"], unit, v["max_rel"]))
    return "%s: %s" % (label, "; ".join(parts))
