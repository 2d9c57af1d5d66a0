"""Continuous-batching server for many independent DG solves on one GPU
(hgdg_server_* in include/hgdg.h)."""
import ctypes
import time

import numpy as np

from . import _lib
from ._lib import check, dptr

KINDS = {"advection": _lib.OP_ADVECTION_UPWIND, "acoustic": _lib.OP_ACOUSTIC_CENTRAL}


class Server:
    def __init__(self, precision="fp64", hbm_budget_gb=64.0, max_active=64, device=0):
        self.lib = _lib.load()
        prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED_FP32}[precision]
        h = ctypes.c_void_p()
        check(self.lib.hgdg_server_create(ctypes.byref(h), device, prec, int(hbm_budget_gb * (1 << 30)), max_active))
        self.h = h
        self._keep = {}

    def submit(self, dim, kind, order, cells, amp, k, dt, n_steps, velocity=(1.0, 0.0, 0.0), wave_speed=1.0,
               keep_state=False):
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        kk = np.zeros((len(amp), 3))
        kk[:, : np.asarray(k).shape[1]] = k
        cells3 = (list(cells) + [1, 1, 1])[:3]
        vel = (list(velocity) + [0.0, 0.0, 0.0])[:3]
        d = _lib.SolveDesc(dim, KINDS[kind], order, (ctypes.c_int32 * 3)(*cells3), (ctypes.c_double * 3)(*vel),
                           wave_speed, len(amp), dptr(amp), dptr(kk), dt, n_steps, int(keep_state))
        t = ctypes.c_int64()
        check(self.lib.hgdg_server_submit(self.h, ctypes.byref(d), ctypes.byref(t)))
        return t.value

    def wait(self, ticket=-1, timeout=600.0):
        done = ctypes.c_int()
        check(self.lib.hgdg_server_wait(self.h, ticket, timeout, ctypes.byref(done)))
        return bool(done.value)

    def result(self, ticket, with_state=False):
        n = ctypes.c_int64()
        norm = ctypes.c_double()
        lat = ctypes.c_double()
        check(self.lib.hgdg_server_result(self.h, ticket, ctypes.byref(n), None, ctypes.byref(norm),
                                          ctypes.byref(lat)))
        state = None
        if with_state:
            state = np.zeros(n.value)
            check(self.lib.hgdg_server_result(self.h, ticket, None, dptr(state), None, None))
        return {"n_dof": n.value, "norm": norm.value, "latency_s": lat.value, "state": state}

    def stats(self):
        st = _lib.ServerStats()
        check(self.lib.hgdg_server_get_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def set_profiling(self, enable=True):
        """Time each kernel family with HIP events (separate launches instead
        of the step graph; results unchanged)."""
        check(self.lib.hgdg_server_set_profiling(self.h, 1 if enable else 0))

    def family_times(self):
        """{(kind, dim, order): (kernel ms, launches)} and the separate update
        passes' ms (0: the RK updates run inside each family's last stage)."""
        n = ctypes.c_int()
        check(self.lib.hgdg_server_family_times(self.h, 0, None, None, None, None, None, ctypes.byref(n), None))
        cap = n.value
        kind, dim, order = (np.zeros(cap, dtype=np.int32) for _ in range(3))
        ms, cnt = np.zeros(cap), np.zeros(cap, dtype=np.int64)
        upd = ctypes.c_double()
        i32 = lambda a: a.ctypes.data_as(_lib._I32)
        check(self.lib.hgdg_server_family_times(self.h, cap, i32(kind), i32(dim), i32(order), dptr(ms),
                                                cnt.ctypes.data_as(_lib._I64), ctypes.byref(n), ctypes.byref(upd)))
        names = {v: k for k, v in KINDS.items()}
        fams = {(names[int(kind[i])], int(dim[i]), int(order[i])): (float(ms[i]), int(cnt[i])) for i in range(cap)}
        return fams, upd.value

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.hgdg_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serving_stream(n_requests=512, mean_interarrival_s=0.005, side_min=8, side_max=64, steps=200, seed=0,
                   max_order=4):
    """The serving workload: seeded Poisson arrivals; per request 2D or 3D,
    advection or acoustic, side log-uniform in [side_min, side_max] per
    dimension, p uniform in 1..max_order, seeded sine-mode initial data, RK4
    with dt = 0.2 h / (speed (2p + 1))."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n_requests):
        t += rng.exponential(mean_interarrival_s)
        dim = int(rng.integers(2, 4))
        kind = "advection" if rng.random() < 0.5 else "acoustic"
        cells = [int(round(np.exp(rng.uniform(np.log(side_min), np.log(side_max))))) for _ in range(dim)]
        order = int(rng.integers(1, max_order + 1))
        n_modes = 4
        amp = rng.uniform(-1, 1, n_modes)
        k = rng.integers(1, 5, size=(n_modes, dim)).astype(float)
        vel = list(rng.uniform(0.5, 1.0, dim) * rng.choice([-1, 1], dim))
        speed = float(np.linalg.norm(vel)) if kind == "advection" else 1.0
        h = 1.0 / max(cells)
        dt = 0.2 * h / (speed * (2 * order + 1))
        reqs.append(dict(arrival=t, dim=dim, kind=kind, order=order, cells=cells, amp=amp, k=k, dt=dt,
                         n_steps=steps, velocity=vel, wave_speed=1.0))
    return reqs


def run_stream(server, reqs, keep_state=False):
    """Submit requests at their arrival times (relative to now) and wait for all."""
    t0 = time.perf_counter()
    tickets = []
    for r in reqs:
        delay = r["arrival"] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tickets.append(server.submit(r["dim"], r["kind"], r["order"], r["cells"], r["amp"], r["k"], r["dt"],
                                     r["n_steps"], velocity=r["velocity"], wave_speed=r["wave_speed"],
                                     keep_state=keep_state))
    if not server.wait(-1, timeout=3600):
        raise TimeoutError("serving stream did not finish")
    return tickets
