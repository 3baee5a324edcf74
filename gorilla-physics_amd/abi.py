"""ctypes view of the C ABI declared in include/gorilla_amd.h.

The structures mirror the header field by field; `load_library()` opens the
HIP library built in-tree (gorilla-physics_amd/lib/libgorilla_amd.so) and
raises if it is missing — there is no CPU fallback for the product path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgorilla_amd.so")

GP_OK = 0
GP_ERR_INVALID = 1
GP_ERR_TOPOLOGY = 2
GP_ERR_NOMEM = 3
GP_ERR_DEVICE = 4
GP_ERR_UNSUPPORTED = 5
GP_ERR_FAULT = 6
GP_ERR_BUSY = 7

GP_JOINT_FIXED = 0
GP_JOINT_REVOLUTE = 1
GP_JOINT_PRISMATIC = 2
GP_JOINT_FLOATING = 3

GP_SEMI_IMPLICIT_EULER = 0
GP_RUNGE_KUTTA_2 = 1
GP_RUNGE_KUTTA_4 = 2
GP_VELOCITY_STEPPING = 3
GP_CCD_VELOCITY_STEPPING = 4

GP_PREC_F64 = 0
GP_PREC_MIXED = 1

GP_KERNEL_AUTO = 0
GP_KERNEL_PLAIN = 1
GP_KERNEL_FAST = 2


class gp_body_desc(C.Structure):
    _fields_ = [
        ("parent", C.c_int32),
        ("jtype", C.c_int32),
        ("axis", C.c_double * 3),
        ("x_tree_quat", C.c_double * 4),
        ("x_tree_trans", C.c_double * 3),
        ("moment", C.c_double * 6),
        ("mc", C.c_double * 3),
        ("mass", C.c_double),
        ("spring_k", C.c_double),
        ("spring_l", C.c_double),
    ]


class gp_contact_desc(C.Structure):
    _fields_ = [("body", C.c_int32), ("_pad", C.c_int32), ("p", C.c_double * 3), ("k", C.c_double)]


class gp_halfspace_desc(C.Structure):
    _fields_ = [
        ("point", C.c_double * 3),
        ("normal", C.c_double * 3),
        ("alpha", C.c_double),
        ("mu", C.c_double),
    ]


class gp_spring_contact_desc(C.Structure):
    _fields_ = [("body", C.c_int32), ("_pad", C.c_int32), ("direction", C.c_double * 3), ("l_rest", C.c_double),
                ("k", C.c_double)]


class gp_sphere_desc(C.Structure):
    _fields_ = [("body", C.c_int32), ("enabled", C.c_int32), ("center", C.c_double * 3), ("radius", C.c_double)]


class gp_model_desc(C.Structure):
    _fields_ = [
        ("n_bodies", C.c_int32),
        ("n_contacts", C.c_int32),
        ("n_halfspaces", C.c_int32),
        ("_pad", C.c_int32),
        ("bodies", C.POINTER(gp_body_desc)),
        ("contacts", C.POINTER(gp_contact_desc)),
        ("halfspaces", C.POINTER(gp_halfspace_desc)),
        ("gravity", C.c_double),
        ("n_spring_contacts", C.c_int32),
        ("_pad2", C.c_int32),
        ("spring_contacts", C.POINTER(gp_spring_contact_desc)),
        ("n_spheres", C.c_int32),
        ("_pad3", C.c_int32),
        ("spheres", C.POINTER(gp_sphere_desc)),
    ]


class gp_batch_opts(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("kernel", C.c_int32),
        ("steps_per_launch", C.c_int32),
        ("block_size", C.c_int32),
        ("lanes_per_mech", C.c_int32),
        ("use_graph", C.c_int32),
        ("separate_launches", C.c_int32),
        ("lanes_per_tree", C.c_int32),
    ]


class gp_server_opts(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("chunk_steps", C.c_int32),
        ("slots_per_pool", C.c_int32),
        ("block_size", C.c_int32),
        ("mem_budget_bytes", C.c_int64),
        ("lanes_per_mech", C.c_int32),
        ("lanes_per_tree", C.c_int32),
        ("_reserved", C.c_int32 * 2),
    ]


GP_CONTROL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                            C.c_double, C.POINTER(C.c_double))

# name -> (restype, argtypes); every symbol include/gorilla_amd.h declares
_P = C.c_void_p
_D = C.POINTER(C.c_double)
SIGNATURES = {
    "gp_last_error": (C.c_char_p, []),
    "gp_abi_version": (C.c_int, []),
    "gp_model_create": (C.c_int, [C.POINTER(gp_model_desc), C.POINTER(_P)]),
    "gp_model_destroy": (None, [_P]),
    "gp_model_dims": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "gp_model_signature": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "gp_batch_create": (C.c_int, [C.POINTER(_P), C.POINTER(C.c_int64), C.c_int32,
                                  C.POINTER(gp_batch_opts), C.POINTER(_P)]),
    "gp_batch_destroy": (None, [_P]),
    "gp_batch_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int32)]),
    "gp_batch_offsets": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "gp_batch_set_opts": (C.c_int, [_P, C.POINTER(gp_batch_opts)]),
    "gp_batch_get_opts": (C.c_int, [_P, C.POINTER(gp_batch_opts)]),
    "gp_batch_set_state": (C.c_int, [_P, _D, _D]),
    "gp_batch_get_state": (C.c_int, [_P, _D, _D]),
    "gp_batch_set_tau": (C.c_int, [_P, _D]),
    "gp_step": (C.c_int, [_P, C.c_double, C.c_int32, C.c_int32, _P]),
    "gp_step_prepare": (C.c_int, [_P, C.c_double, C.c_int32, C.c_int32, _P]),
    "gp_step_noalloc": (C.c_int, [_P, C.c_double, C.c_int32, C.c_int32, _P]),
    "gp_mass_matrix": (C.c_int, [_P, _D]),
    "gp_dynamics_bias": (C.c_int, [_P, _D]),
    "gp_batch_sync": (C.c_int, [_P]),
    "gp_batch_get_vdot": (C.c_int, [_P, _D]),
    "gp_dynamics": (C.c_int, [_P, _D, _D]),
    "gp_simulate": (C.c_int, [_P, C.c_double, C.c_double, GP_CONTROL_FN, _P, C.c_int32, _D, _D,
                              C.c_int64, C.POINTER(C.c_int64)]),
    "gp_batch_faults": (C.c_int, [_P, C.POINTER(C.c_uint8)]),
    "gp_batch_clear_faults": (C.c_int, [_P]),
    "gp_batch_collision_rounds": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "gp_batch_last_timing": (C.c_int, [_P, _D, C.POINTER(C.c_int64)]),
    "gp_batch_set_state_dev": (C.c_int, [_P, C.c_void_p, C.c_void_p, _P]),
    "gp_batch_get_state_dev": (C.c_int, [_P, C.c_void_p, C.c_void_p, _P]),
    "gp_batch_set_tau_dev": (C.c_int, [_P, C.c_void_p, _P]),
    "gp_batch_get_vdot_dev": (C.c_int, [_P, C.c_void_p, _P]),
    "gp_batch_energy": (C.c_int, [_P, _D, _D]),
    "gp_batch_set_spring_contact": (C.c_int, [_P, C.c_int64, C.c_int32, _D, C.c_double]),
    "gp_batch_get_spring_contact": (C.c_int, [_P, C.c_int64, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                              _D, _D, _D]),
    "gp_batch_autotune": (C.c_int, [_P, C.c_double, C.c_int32, C.c_int32, C.POINTER(gp_batch_opts), _D]),
    "gp_server_create": (C.c_int, [C.POINTER(gp_server_opts), C.POINTER(_P)]),
    "gp_server_destroy": (None, [_P]),
    "gp_server_clock": (C.c_int, [_P, _D]),
    "gp_server_submit": (C.c_int, [_P, _P, _D, _D, _D, C.c_int64, C.c_double, C.c_double,
                                   C.POINTER(C.c_int64)]),
    "gp_server_run": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "gp_server_pending": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "gp_server_completed": (C.c_int, [_P, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int64)]),
    "gp_server_result": (C.c_int, [_P, C.c_int64, _D, _D, _D, C.POINTER(C.c_int32)]),
    "gp_server_stats": (C.c_int, [_P, _D]),
}

_lib = None


class GPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gorilla_amd error {code}: {msg}")
        self.code = code


def load_library(path=None):
    """Open the in-tree HIP library; raise if it was not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP library {p} not found: build it with `python __graft_entry__.py build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code, lib=None):
    if code != GP_OK:
        lib = lib or load_library()
        msg = lib.gp_last_error()
        raise GPError(code, msg.decode() if msg else "")
    return code
