"""Python handles over the C ABI's model and batch objects (ctypes)."""
import ctypes as C

import numpy as np

from . import abi


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Model:
    """gp_model_create (MechanismState::new, mechanism.rs:62)."""

    def __init__(self, desc):
        self.lib = abi.load_library()
        self.desc = desc  # keeps the POD arrays alive
        h = C.c_void_p()
        abi.check(self.lib.gp_model_create(desc.ptr(), C.byref(h)), self.lib)
        self.handle = h
        nq, nv, nb = C.c_int32(), C.c_int32(), C.c_int32()
        abi.check(self.lib.gp_model_dims(h, C.byref(nq), C.byref(nv), C.byref(nb)), self.lib)
        self.nq, self.nv, self.n_bodies = nq.value, nv.value, nb.value

    def signature(self):
        s = C.c_uint64()
        abi.check(self.lib.gp_model_signature(self.handle, C.byref(s)), self.lib)
        return s.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.gp_model_destroy(h)
            self.handle = None


def make_opts(precision=abi.GP_PREC_F64, kernel=abi.GP_KERNEL_AUTO, steps_per_launch=0, block_size=0,
              lanes_per_mech=0, use_graph=0, separate_launches=0, lanes_per_tree=0):
    o = abi.gp_batch_opts()
    o.precision, o.kernel, o.steps_per_launch = precision, kernel, steps_per_launch
    o.block_size, o.lanes_per_mech, o.use_graph = block_size, lanes_per_mech, use_graph
    o.separate_launches = separate_launches
    o.lanes_per_tree = lanes_per_tree
    return o


class Batch:
    """A batch of instances resident on the GPU (gp_batch_*)."""

    def __init__(self, models, counts, opts=None, **kw):
        self.lib = abi.load_library()
        self.models = list(models)
        n = len(self.models)
        hs = (C.c_void_p * n)(*[m.handle.value for m in self.models])
        cs = (C.c_int64 * n)(*[int(c) for c in counts])
        if opts is None:
            opts = make_opts(**kw)
        h = C.c_void_p()
        abi.check(self.lib.gp_batch_create(hs, cs, n, C.byref(opts), C.byref(h)), self.lib)
        self.handle = h
        ni, nq, nv, nb = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        abi.check(self.lib.gp_batch_info(h, C.byref(ni), C.byref(nq), C.byref(nv), C.byref(nb)), self.lib)
        self.n_instances, self.nq_total, self.nv_total, self.n_buckets = ni.value, nq.value, nv.value, nb.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.gp_batch_destroy(h)
            self.handle = None

    def offsets(self):
        qo = np.zeros(self.n_instances + 1, dtype=np.int64)
        vo = np.zeros(self.n_instances + 1, dtype=np.int64)
        abi.check(self.lib.gp_batch_offsets(self.handle, qo.ctypes.data_as(C.POINTER(C.c_int64)),
                                            vo.ctypes.data_as(C.POINTER(C.c_int64))), self.lib)
        return qo, vo

    def set_opts(self, **kw):
        o = make_opts(**kw)
        abi.check(self.lib.gp_batch_set_opts(self.handle, C.byref(o)), self.lib)

    def get_opts(self):
        o = abi.gp_batch_opts()
        abi.check(self.lib.gp_batch_get_opts(self.handle, C.byref(o)), self.lib)
        return o

    def set_state(self, q, v):
        q = np.ascontiguousarray(q, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert q.size == self.nq_total and v.size == self.nv_total
        abi.check(self.lib.gp_batch_set_state(self.handle, _dp(q), _dp(v)), self.lib)

    def get_state(self):
        q = np.empty(self.nq_total)
        v = np.empty(self.nv_total)
        abi.check(self.lib.gp_batch_get_state(self.handle, _dp(q), _dp(v)), self.lib)
        return q, v

    def get_vdot(self):
        out = np.empty(self.nv_total)
        abi.check(self.lib.gp_batch_get_vdot(self.handle, _dp(out)), self.lib)
        return out

    def set_tau(self, tau):
        if tau is None:
            abi.check(self.lib.gp_batch_set_tau(self.handle, None), self.lib)
            return
        t = np.ascontiguousarray(tau, dtype=np.float64)
        assert t.size == self.nv_total
        abi.check(self.lib.gp_batch_set_tau(self.handle, _dp(t)), self.lib)

    def step(self, dt, integrator=abi.GP_SEMI_IMPLICIT_EULER, n_steps=1, stream=None):
        abi.check(self.lib.gp_step(self.handle, float(dt), int(integrator), int(n_steps), stream), self.lib)

    def prepare(self, dt, integrator=abi.GP_SEMI_IMPLICIT_EULER, n_steps=1, stream=None):
        """gp_step_prepare: build what step(dt, integrator, n_steps) launches
        (group table, scratch, graphs) without stepping."""
        abi.check(self.lib.gp_step_prepare(self.handle, float(dt), int(integrator), int(n_steps), stream), self.lib)

    def sync(self):
        abi.check(self.lib.gp_batch_sync(self.handle), self.lib)

    def mass_matrices(self):
        """gp_mass_matrix: M(q) of every instance (a list of nv_i x nv_i arrays)."""
        _, vo = self.offsets()
        sizes = [int(n) for n in np.diff(vo)]
        out = np.empty(max(1, sum(n * n for n in sizes)))
        abi.check(self.lib.gp_mass_matrix(self.handle, _dp(out)), self.lib)
        res, o = [], 0
        for n in sizes:
            res.append(out[o:o + n * n].reshape(n, n).copy())
            o += n * n
        return res

    def dynamics_bias(self):
        """gp_dynamics_bias: c(q, v) - tau_c of every instance, flattened like v."""
        out = np.empty(max(1, self.nv_total))
        abi.check(self.lib.gp_dynamics_bias(self.handle, _dp(out)), self.lib)
        return out[:self.nv_total]

    def dynamics(self, tau=None):
        out = np.empty(self.nv_total)
        if tau is None:
            abi.check(self.lib.gp_dynamics(self.handle, None, _dp(out)), self.lib)
        else:
            t = np.ascontiguousarray(tau, dtype=np.float64)
            abi.check(self.lib.gp_dynamics(self.handle, _dp(t), _dp(out)), self.lib)
        return out

    def simulate(self, final_time, dt, control, integrator=abi.GP_SEMI_IMPLICIT_EULER, max_records=None):
        """control(q, v, t) -> tau (flattened, nv_total) or None for zeros."""
        nq, nv = self.nq_total, self.nv_total
        if max_records is None:
            max_records = int(final_time / dt) + 3
        qs = np.zeros((max_records, nq))
        vs = np.zeros((max_records, nv))
        err = []

        def cb(user, qp, vp, t, tp):
            try:
                q = np.ctypeslib.as_array(qp, shape=(nq,)).copy()
                v = np.ctypeslib.as_array(vp, shape=(nv,)).copy()
                tau = control(q, v, t) if control is not None else None
                out = np.ctypeslib.as_array(tp, shape=(nv,))
                out[:] = 0.0 if tau is None else np.asarray(tau, dtype=float)
                return 0
            except Exception as e:  # reported after the call returns
                err.append(e)
                return 1

        fn = abi.GP_CONTROL_FN(cb)
        n = C.c_int64()
        code = self.lib.gp_simulate(self.handle, float(final_time), float(dt), fn, None, int(integrator),
                                    _dp(qs), _dp(vs), max_records, C.byref(n))
        if err:
            raise err[0]
        abi.check(code, self.lib)
        k = min(n.value, max_records)
        return qs[:k], vs[:k]

    def energy(self):
        """(kinetic, gravitational) energy of every instance at the current
        state (gp_batch_energy, mechanism.rs:334-362)."""
        ke, pe = np.empty(self.n_instances), np.empty(self.n_instances)
        abi.check(self.lib.gp_batch_energy(self.handle, _dp(ke), _dp(pe)), self.lib)
        return ke, pe

    def faults(self):
        f = np.zeros(self.n_instances, dtype=np.uint8)
        abi.check(self.lib.gp_batch_faults(self.handle, f.ctypes.data_as(C.POINTER(C.c_uint8))), self.lib)
        return f

    def clear_faults(self):
        abi.check(self.lib.gp_batch_clear_faults(self.handle), self.lib)

    def collision_rounds(self):
        """Rounds of the collision loop in each instance's last
        velocity-stepping step (1 = not cut; 0 = no such step yet)."""
        r = np.zeros(self.n_instances, dtype=np.int32)
        abi.check(self.lib.gp_batch_collision_rounds(self.handle, r.ctypes.data_as(C.POINTER(C.c_int32))), self.lib)
        return r

    # device-array transfers: q, v, tau, vdot are device addresses (ints), e.g.
    # torch tensors' data_ptr(); stream is a hipStream_t address or None
    def set_state_dev(self, q_ptr, v_ptr, stream=None):
        abi.check(self.lib.gp_batch_set_state_dev(self.handle, q_ptr, v_ptr, stream), self.lib)

    def get_state_dev(self, q_ptr, v_ptr, stream=None):
        abi.check(self.lib.gp_batch_get_state_dev(self.handle, q_ptr, v_ptr, stream), self.lib)

    def set_tau_dev(self, tau_ptr, stream=None):
        abi.check(self.lib.gp_batch_set_tau_dev(self.handle, tau_ptr, stream), self.lib)

    def get_vdot_dev(self, vdot_ptr, stream=None):
        abi.check(self.lib.gp_batch_get_vdot_dev(self.handle, vdot_ptr, stream), self.lib)

    def set_spring_contact(self, instance, index, direction, l_rest):
        """gp_batch_set_spring_contact: direction (body frame) and rest length."""
        d = np.ascontiguousarray(direction, dtype=np.float64)
        abi.check(self.lib.gp_batch_set_spring_contact(self.handle, int(instance), int(index), _dp(d),
                                                       float(l_rest)), self.lib)

    def get_spring_contact(self, instance, index, with_halfspace=False):
        """(registered, world point, direction in the body frame, rest length),
        with the registered half-space's index (-1 if none) after `registered`
        when with_halfspace."""
        reg, h = C.c_int32(), C.c_int32()
        p, d, lr = np.zeros(3), np.zeros(3), C.c_double()
        abi.check(self.lib.gp_batch_get_spring_contact(self.handle, int(instance), int(index), C.byref(reg),
                                                       C.byref(h), _dp(p), _dp(d), C.byref(lr)), self.lib)
        if with_halfspace:
            return bool(reg.value), h.value, p, d, lr.value
        return bool(reg.value), p, d, lr.value

    def autotune(self, dt, integrator=abi.GP_SEMI_IMPLICIT_EULER, trial_steps=0):
        """gp_batch_autotune: time trial rollouts over lanes / workgroup size /
        steps per launch, apply the fastest options (state left unchanged).
        Returns (chosen gp_batch_opts, trial milliseconds)."""
        o, ms = abi.gp_batch_opts(), C.c_double()
        abi.check(self.lib.gp_batch_autotune(self.handle, float(dt), int(integrator), int(trial_steps),
                                             C.byref(o), C.byref(ms)), self.lib)
        return o, ms.value

    def last_timing(self):
        ms, n = C.c_double(), C.c_int64()
        abi.check(self.lib.gp_batch_last_timing(self.handle, C.byref(ms), C.byref(n)), self.lib)
        return ms.value, n.value


def batch_from_workload(w, **opts):
    """Build models and a batch for a workloads.* dict; returns (batch, models)."""
    models = [Model(d) for d in w["models"]]
    b = Batch(models, w["counts"], **opts)
    b.set_state(w["q"], w["v"])
    b.set_tau(w["tau"])
    return b, models
