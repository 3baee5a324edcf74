"""TEST INFRASTRUCTURE ONLY — ctypes binding of the CPU oracle (liboracle.so).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product path never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

CRBA, ABA, CRBA_LD, ABA_LD = 0, 1, 2, 3

_D = C.POINTER(C.c_double)


def _dp(a):
    return a.ctypes.data_as(_D)


class Oracle:
    def __init__(self, path=LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library {path} missing: run `make -C oracle`")
        lib = C.CDLL(path)
        P = C.c_void_p
        lib.oracle_dims.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.oracle_dynamics.argtypes = [P, _D, _D, _D, _D, C.c_int]
        lib.oracle_mass_matrix.argtypes = [P, _D, _D, _D, _D]
        lib.oracle_step.argtypes = [P, C.c_int, C.c_int, _D, _D, _D, C.c_double, _D, _D, _D]
        lib.oracle_step_spring.argtypes = [P, _D, _D, _D, C.c_double, _D, _D, _D, _D]
        lib.oracle_vs_step.argtypes = [P, C.c_int, _D, _D, _D, C.c_double, _D, _D, C.POINTER(C.c_int)]
        lib.oracle_gravitational_energy.argtypes = [P, _D, _D]
        lib.oracle_cone_program.argtypes = [C.c_int, _D, _D, _D, C.POINTER(C.c_int)]
        lib.oracle_rollout.argtypes = [C.POINTER(P), C.c_int64, _D, _D, _D, C.c_double, C.c_int, C.c_int,
                                       C.c_int64, C.c_int, _D]
        lib.oracle_simulate.argtypes = [P, C.c_double, C.c_double, C.c_int, C.c_int, _D, _D, _D, _D, _D,
                                        C.c_int64, C.POINTER(C.c_int64)]
        lib.oracle_energy.argtypes = [P, _D, _D, _D, _D]
        lib.oracle_contact_distances.argtypes = [P, _D, _D]
        lib.oracle_opcount.argtypes = [P, _D, _D, _D, C.c_int, C.POINTER(C.c_uint64)]
        self.lib = lib

    @staticmethod
    def _desc(d):
        return C.cast(d.ptr(), C.c_void_p)

    def dims(self, d):
        nq, nv = C.c_int(), C.c_int()
        self.lib.oracle_dims(self._desc(d), C.byref(nq), C.byref(nv))
        return nq.value, nv.value

    def dynamics(self, d, q, v, tau=None, method=CRBA):
        nq, nv = self.dims(d)
        q = np.ascontiguousarray(q, float); v = np.ascontiguousarray(v, float)
        t = np.zeros(nv) if tau is None else np.ascontiguousarray(tau, float)
        out = np.empty(nv)
        st = self.lib.oracle_dynamics(self._desc(d), _dp(q), _dp(v), _dp(t), _dp(out), method)
        if st:
            raise RuntimeError("oracle: singular mass matrix")
        return out

    def mass_matrix(self, d, q, v):
        nq, nv = self.dims(d)
        q = np.ascontiguousarray(q, float); v = np.ascontiguousarray(v, float)
        M = np.empty((nv, nv)); c = np.empty(nv)
        self.lib.oracle_mass_matrix(self._desc(d), _dp(q), _dp(v), _dp(M), _dp(c))
        return M, c

    def step(self, d, q, v, tau=None, dt=1e-3, integrator=0, method=CRBA):
        nq, nv = self.dims(d)
        q = np.ascontiguousarray(q, float); v = np.ascontiguousarray(v, float)
        t = np.zeros(nv) if tau is None else np.ascontiguousarray(tau, float)
        qo, vo, ao = np.empty(nq), np.empty(nv), np.empty(nv)
        st = self.lib.oracle_step(self._desc(d), integrator, method, _dp(q), _dp(v), _dp(t), dt, _dp(qo), _dp(vo),
                                  _dp(ao))
        if st:
            raise RuntimeError("oracle: singular mass matrix")
        return qo, vo, ao

    @staticmethod
    def spring_state(d):
        """Initial spring-contact state of a model (unregistered, as described):
        9 values per spring contact in body order, see oracle_step_spring."""
        rows = []
        sc = [d.desc.spring_contacts[i] for i in range(d.desc.n_spring_contacts)]
        for s in sorted(sc, key=lambda s: s.body):  # stable: description order within a body
            rows.append([0, 0, 0, 0, 0, s.direction[0], s.direction[1], s.direction[2], s.l_rest])
        return np.array(rows, dtype=float).reshape(-1)

    def step_spring(self, d, q, v, state, tau=None, dt=1e-3):
        """One semi-implicit Euler step of a model with spring contacts; state
        (from spring_state) is updated in place. Returns (q', v', vdot)."""
        nq, nv = self.dims(d)
        q = np.ascontiguousarray(q, float); v = np.ascontiguousarray(v, float)
        t = np.zeros(nv) if tau is None else np.ascontiguousarray(tau, float)
        qo, vo, ao = np.empty(nq), np.empty(nv), np.empty(nv)
        st = self.lib.oracle_step_spring(self._desc(d), _dp(q), _dp(v), _dp(t), dt, _dp(qo), _dp(vo), _dp(ao),
                                         _dp(state))
        if st:
            raise RuntimeError("oracle: singular mass matrix" if st == 1 else "oracle: spring pulls into the half-space")
        return qo, vo, ao

    def vs_step(self, d, q, v, tau=None, dt=1e-3, ccd=False):
        """One velocity-stepping step (ccd: with continuous collision
        detection). Returns (q', v', rounds of the collision loop)."""
        nq, nv = self.dims(d)
        q = np.ascontiguousarray(q, float); v = np.ascontiguousarray(v, float)
        t = np.zeros(nv) if tau is None or len(tau) == 0 else np.ascontiguousarray(tau, float)
        qo, vo, rounds = np.empty(nq), np.empty(nv), C.c_int()
        st = self.lib.oracle_vs_step(self._desc(d), int(bool(ccd)), _dp(q), _dp(v), _dp(t), dt, _dp(qo), _dp(vo),
                                     C.byref(rounds))
        if st:
            raise RuntimeError({1: "oracle: singular mass matrix", 3: "oracle: no finite contact impulse",
                                4: "oracle: collision loop did not finish"}.get(st, f"oracle: status {st}"))
        return qo, vo, rounds.value

    def cone_program(self, G, g):
        """min 1/2 x^T G x + g^T x, (0.95 x_n, x_t1, x_t2) in the second-order
        cone per contact. Returns (x, iterations)."""
        G = np.ascontiguousarray(G, float); g = np.ascontiguousarray(g, float)
        nc = g.size // 3
        x, it = np.empty(g.size), C.c_int()
        st = self.lib.oracle_cone_program(nc, _dp(G), _dp(g), _dp(x), C.byref(it))
        if st:
            raise RuntimeError("oracle: cone program failed")
        return x, it.value

    def rollout(self, descs, q, v, tau, dt, steps, integrator=0, method=CRBA, threads=1, want_vdot=False):
        """descs: one ModelDesc per instance; flattened state; returns new
        (q, v, failures) or (q, v, vdot_of_last_step, failures)."""
        n = len(descs)
        arr = (C.c_void_p * n)(*[C.cast(d.ptr(), C.c_void_p).value for d in descs])
        q = np.array(q, dtype=float, copy=True); v = np.array(v, dtype=float, copy=True)
        t = np.ascontiguousarray(tau, float) if tau is not None else None
        a = np.zeros(v.size) if want_vdot else None
        fails = self.lib.oracle_rollout(arr, n, _dp(q), _dp(v), _dp(t) if t is not None else None, dt, integrator,
                                        method, int(steps), int(threads), _dp(a) if want_vdot else None)
        return (q, v, a, fails) if want_vdot else (q, v, fails)

    def simulate(self, d, final_time, dt, q0, v0, tau=None, integrator=0, method=CRBA):
        nq, nv = self.dims(d)
        rec = int(final_time / dt) + 3
        qs, vs = np.zeros((rec, nq)), np.zeros((rec, nv))
        q0 = np.ascontiguousarray(q0, float); v0 = np.ascontiguousarray(v0, float)
        t = np.ascontiguousarray(tau, float) if tau is not None else None
        n = C.c_int64()
        st = self.lib.oracle_simulate(self._desc(d), final_time, dt, integrator, method, _dp(q0), _dp(v0),
                                      _dp(t) if t is not None else None, _dp(qs), _dp(vs), rec, C.byref(n))
        if st:
            raise RuntimeError("oracle: singular mass matrix")
        return qs[:n.value], vs[:n.value]

    def energy(self, d, q, v):
        ke, pe = C.c_double(), C.c_double()
        q = np.ascontiguousarray(q, float); v = np.ascontiguousarray(v, float)
        self.lib.oracle_energy(self._desc(d), _dp(q), _dp(v), C.byref(ke), C.byref(pe))
        return ke.value, pe.value

    def gravitational_energy(self, d, q):
        """sum m g z of every body frame's origin (mechanism.rs:352-362)."""
        pe = C.c_double()
        q = np.ascontiguousarray(q, float)
        self.lib.oracle_gravitational_energy(self._desc(d), _dp(q), C.byref(pe))
        return pe.value

    def contact_distances(self, d, q):
        q = np.ascontiguousarray(q, float)
        out = np.empty(max(1, d.desc.n_contacts * d.desc.n_halfspaces))
        self.lib.oracle_contact_distances(self._desc(d), _dp(q), _dp(out))
        return out[:d.desc.n_contacts * d.desc.n_halfspaces]

    def opcount(self, d, q, v, tau=None, integrator=0):
        nq, nv = self.dims(d)
        q = np.ascontiguousarray(q, float); v = np.ascontiguousarray(v, float)
        t = np.zeros(nv) if tau is None else np.ascontiguousarray(tau, float)
        c = (C.c_uint64 * 12)()
        self.lib.oracle_opcount(self._desc(d), _dp(q), _dp(v), _dp(t), integrator, c)
        names = ("add", "mul", "div", "sqrt", "trig", "pow")
        out = {k: c[i] for i, k in enumerate(names)}
        out.update({k + "64": c[6 + i] for i, k in enumerate(names)})
        return out
