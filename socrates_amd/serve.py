"""Serving front end (socgpu_serve_*): submit radiation calls from any
thread; the library's server thread admits them by device-memory footprint
and runs everything admitted as packed batches, several batches in flight.

    srv = Server(precision="mixed", hbm_budget_gb=64)
    t = srv.submit(*runes.case_call(case, "lw"))   # returns a ticket
    out = srv.wait(t)                              # dict of output arrays
    srv.stats()                                    # p50/p99 latency, throughput
"""
import ctypes

from socrates_amd import runes as R


class ServeConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("n_slots", ctypes.c_int),
                ("max_batch_requests", ctypes.c_int), ("pad", ctypes.c_int),
                ("hbm_budget_bytes", ctypes.c_size_t)]


class ServeStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_longlong), ("batches", ctypes.c_longlong),
                ("solves", ctypes.c_longlong), ("mean_batch", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("elapsed_s", ctypes.c_double)]


class Server:
    def __init__(self, precision="fp64", hbm_budget_gb=32.0, n_slots=2, max_batch_requests=256):
        R.init()
        cfg = ServeConfig(R._PREC[precision], int(n_slots), int(max_batch_requests), 0,
                          int(hbm_budget_gb * (1 << 30)))
        self._h = ctypes.c_void_p()
        R._check(R.lib().socgpu_serve_start(ctypes.byref(self._h), ctypes.byref(cfg)))
        self._live = {}

    def submit(self, spectrum_name, kw):
        """Queue one call (spectrum_name, runes() keyword arguments with
        n_profile, n_layer, i_source). Returns a ticket."""
        kw = dict(kw)
        P, L, src = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
        a, d, out, keep = R._prepare(P, L, spectrum_name, src, **kw)
        name = spectrum_name.encode()
        req = R.Request(name, ctypes.pointer(a), ctypes.pointer(d))
        t = ctypes.c_longlong()
        R._check(R.lib().socgpu_serve_submit(self._h, ctypes.byref(req), ctypes.byref(t)))
        self._live[t.value] = (out, keep, a, d, name, req)
        return t.value

    def wait(self, ticket, timeout_s=120.0):
        R._check(R.lib().socgpu_serve_wait(self._h, ctypes.c_longlong(ticket),
                                          ctypes.c_double(timeout_s)))
        return self._live.pop(ticket)[0]

    def replay(self, calls, arrival_s, picks, timeout_s=600.0):
        """Submit calls[picks[i]] at arrival_s[i] seconds after the start,
        from the library's own submitter thread (socgpu_serve_replay), and
        wait for all of them. calls: list of (spectrum_name, kwargs) as for
        submit; a call picked several times writes the same output arrays
        each time. Returns the wall time from the start to the last
        completion (s)."""
        import numpy as np
        n = len(calls)
        keep, reqs = [], (R.Request * n)()
        for k, (name, kw) in enumerate(calls):
            kw = dict(kw)
            P, L, src = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
            a, d, out, kp = R._prepare(P, L, name, src, **kw)
            nm = name.encode()
            keep += [a, d, kp, nm]
            reqs[k] = R.Request(nm, ctypes.pointer(a), ctypes.pointer(d))
        t = np.ascontiguousarray(arrival_s, dtype=np.float64)
        p = np.ascontiguousarray(picks, dtype=np.int32)
        if p.size and (p.min() < 0 or p.max() >= n):
            raise ValueError("picks out of range")
        wall = ctypes.c_double()
        R._check(R.lib().socgpu_serve_replay(
            self._h, n, reqs, int(t.size), t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            p.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.c_double(timeout_s),
            ctypes.byref(wall)))
        return wall.value

    def stats(self):
        st = ServeStats()
        R._check(R.lib().socgpu_serve_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServeStats._fields_}

    def reset_stats(self):
        R._check(R.lib().socgpu_serve_reset_stats(self._h))

    def close(self):
        if self._h:
            R.lib().socgpu_serve_stop(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
