"""Python handle over the serving loop (gp_server_*): continuous batching of
rollout requests with HBM-resident state and p50/p99 latency."""
import ctypes as C

import numpy as np

from . import abi


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Server:
    STATS = ("rounds", "env_steps", "busy_s", "admitted", "completed", "pools", "pool_bytes", "budget_bytes",
             "admit_s", "retire_s")

    def __init__(self, precision=abi.GP_PREC_MIXED, chunk_steps=100, slots_per_pool=4096, block_size=0,
                 mem_budget_bytes=0, lanes_per_mech=0, lanes_per_tree=0):
        self.lib = abi.load_library()
        o = abi.gp_server_opts()
        o.precision, o.chunk_steps, o.slots_per_pool = precision, chunk_steps, slots_per_pool
        o.block_size, o.mem_budget_bytes = block_size, mem_budget_bytes
        o.lanes_per_mech, o.lanes_per_tree = lanes_per_mech, lanes_per_tree
        h = C.c_void_p()
        abi.check(self.lib.gp_server_create(C.byref(o), C.byref(h)), self.lib)
        self.handle = h
        self._models = {}  # id -> Model, kept alive until the result is read

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.gp_server_destroy(h)
            self.handle = None

    def clock(self):
        t = C.c_double()
        abi.check(self.lib.gp_server_clock(self.handle, C.byref(t)), self.lib)
        return t.value

    def submit(self, model, q0, v0, tau=None, horizon=1000, dt=1e-3, arrival=-1.0):
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        assert q0.size == model.nq and v0.size == model.nv
        t = None if tau is None else np.ascontiguousarray(tau, dtype=np.float64)
        rid = C.c_int64()
        abi.check(self.lib.gp_server_submit(self.handle, model.handle, _dp(q0), _dp(v0),
                                            _dp(t) if t is not None else None, int(horizon), float(dt),
                                            float(arrival), C.byref(rid)), self.lib)
        self._models[rid.value] = model
        return rid.value

    def run(self, max_rounds=0):
        n = C.c_int32()
        abi.check(self.lib.gp_server_run(self.handle, int(max_rounds), C.byref(n)), self.lib)
        return n.value

    def pending(self):
        a, b = C.c_int64(), C.c_int64()
        abi.check(self.lib.gp_server_pending(self.handle, C.byref(a), C.byref(b)), self.lib)
        return a.value, b.value

    def completed(self, max_ids=1 << 20):
        ids = np.zeros(max_ids, dtype=np.int64)
        n = C.c_int64()
        abi.check(self.lib.gp_server_completed(self.handle, ids.ctypes.data_as(C.POINTER(C.c_int64)), max_ids,
                                               C.byref(n)), self.lib)
        return ids[:n.value]

    def result(self, rid):
        m = self._models[rid]
        q = np.empty(max(1, m.nq))
        v = np.empty(max(1, m.nv))
        lat = C.c_double()
        fault = C.c_int32()
        abi.check(self.lib.gp_server_result(self.handle, int(rid), _dp(q), _dp(v), C.byref(lat), C.byref(fault)),
                  self.lib)
        del self._models[rid]
        return q[:m.nq], v[:m.nv], lat.value, bool(fault.value)

    def stats(self):
        s = np.zeros(len(self.STATS))
        abi.check(self.lib.gp_server_stats(self.handle, _dp(s)), self.lib)
        return dict(zip(self.STATS, s.tolist()))
