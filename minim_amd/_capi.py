"""ctypes declarations of the C ABI in include/minim_hip.h (libminim_amd.so)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libminim_amd.so")

MINIM_OK, MINIM_EINVAL, MINIM_ERUNTIME, MINIM_EHIP, MINIM_ENOMEM = range(5)
MINIM_FP64, MINIM_FP32 = 0, 1
MINIM_LBFGS, MINIM_FIRE, MINIM_GRADDESCENT = 0, 1, 2
MINIM_LS_BACKTRACKING, MINIM_LS_NONE = 0, 1
MINIM_PREC_STATE, MINIM_PREC_FP64, MINIM_PREC_FP32, MINIM_PREC_MIXED = 0, 1, 2, 3

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p


class PFParams(C.Structure):
    _fields_ = [("grid", C.c_int * 3), ("n_fluid", C.c_int), ("resolution", C.c_double),
                ("surface_tension", dp), ("n_surface_tension", C.c_int),
                ("interface_size", dp), ("n_interface_size", C.c_int),
                ("density_constraint", C.c_int), ("density_const", C.c_double),
                ("fix_fluid", C.POINTER(C.c_ubyte)), ("stencil_xchunk", C.c_int),
                ("stencil_target_blocks", C.c_int),
                ("solid", C.POINTER(C.c_ubyte)), ("contact_angle", dp), ("n_contact_angle", C.c_size_t),
                ("pressure", dp), ("force", dp), ("force_fluids", ip), ("n_force_fluids", C.c_int),
                ("confinement", dp), ("volume_fixed", C.c_int), ("vol_const", C.c_double), ("volume", dp),
                ("stencil_kernel", C.c_int)]


class BaHParams(C.Structure):
    _fields_ = [("n_node", C.c_int), ("triangles", ip), ("n_triangle", C.c_int),
                ("n_bar", C.c_int), ("bars", ip), ("n_hinge", C.c_int), ("hinges", ip),
                ("bar_k", dp), ("hinge_k", dp), ("k_bar", C.c_double), ("k_hinge", C.c_double),
                ("bar_l0", dp), ("hinge_theta0", dp), ("l0", C.c_double), ("theta0", C.c_double),
                ("force", dp), ("force_per_node", C.c_int)]


class MinimiserParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("max_iter", C.c_int), ("linesearch", C.c_int), ("m", C.c_int),
                ("max_step", C.c_double), ("init_step", C.c_double), ("precision_mode", C.c_int),
                ("switch_rms", C.c_double), ("dt_max", C.c_double), ("alpha", C.c_double),
                ("dt0", C.c_double), ("max_force_dim", C.c_int), ("batch_control", C.c_int),
                ("batch_budget", C.c_long), ("batch_block", C.c_int)]


class Result(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("n_energy", C.c_long),
                ("n_gradient", C.c_long), ("n_energy_reused", C.c_long), ("iter_fp32", C.c_int),
                ("iter_fp64", C.c_int), ("energy", C.c_double), ("grad_rms", C.c_double),
                ("dt_max", C.c_double), ("grad_max", C.c_double), ("seconds", C.c_double),
                ("cross_seconds", C.c_double * 8), ("cross_iter", C.c_int * 8), ("start_seconds", C.c_double)]


ADJUST_FN = C.CFUNCTYPE(None, C.c_int, vp, vp)

# name -> (restype, argtypes)
SIGNATURES = {
    "minim_last_error": (C.c_char_p, []),
    "minim_abi_version": (C.c_int, []),
    "minim_pf_create": (C.c_int, [C.POINTER(PFParams), C.POINTER(vp)]),
    "minim_pf_coefficients": (C.c_int, [vp, dp, dp]),
    "minim_pfu_create": (C.c_int, [C.POINTER(PFParams), C.POINTER(vp)]),
    "minim_pfu_element_list": (C.c_int, [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "minim_pf_energy_components": (C.c_int, [vp, dp, C.c_size_t, dp]),
    "minim_pf_node_volumes": (C.c_int, [vp, dp, dp]),
    "minim_pf_diffuse_solid": (C.c_int, [vp, C.POINTER(C.c_ubyte), C.c_int, C.c_int, dp]),
    "minim_pf_site_types": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_ubyte), C.POINTER(C.c_int)]),
    "minim_lj_create": (C.c_int, [C.c_int, C.c_double, C.c_double, C.POINTER(vp)]),
    "minim_lj_set_cutoff": (C.c_int, [vp, C.c_double, dp, C.c_double]),
    "minim_lj_list_builds": (C.c_int, [vp, C.POINTER(C.c_long)]),
    "minim_lj_max_neighbours": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "minim_bah_create": (C.c_int, [C.POINTER(BaHParams), C.POINTER(vp)]),
    "minim_potential_set_fixed": (C.c_int, [vp, ip, C.c_size_t]),
    "minim_potential_init": (C.c_int, [vp, dp, C.c_size_t]),
    "minim_potential_convergence": (C.c_int, [vp, dp]),
    "minim_potential_destroy": (C.c_int, [vp]),
    "minim_potential_eval": (C.c_int, [vp, C.c_int, vp, C.c_size_t, vp, C.c_double, dp, vp, vp]),
    "minim_potential_eval_async": (C.c_int, [vp, C.c_int, vp, C.c_size_t, vp, C.c_double, C.c_int, vp, vp, vp]),
    "minim_potential_eval_mixed": (C.c_int, [vp, vp, C.c_size_t, vp, C.c_double, dp, vp, vp]),
    "minim_potential_eval_step": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_size_t, vp, C.c_double, vp, dp, vp, vp]),
    "minim_potential_time_kernel": (C.c_int, [vp, C.c_int, vp, C.c_size_t, C.c_int, vp, C.c_int, vp, dp]),
    "minim_potential_eval_async_noalloc": (C.c_int, [vp, C.c_int, vp, C.c_size_t, vp, C.c_double, C.c_int, vp, vp,
                                                     vp]),
    "minim_pf_autotune": (C.c_int, [vp, C.c_int, dp, C.c_size_t, ip, C.c_int, ip, C.c_int, C.c_int, dp, ip, ip]),
    "minim_pf_tuned_shape": (C.c_int, [vp, C.c_int, ip, ip]),
    "minim_lbfgs_autotune": (C.c_int, [vp, vp, ip, C.c_int, C.c_int, dp, ip, ip]),
    "minim_autotune_clear": (C.c_int, []),
    "minim_potential_time_steps": (C.c_int, [vp, C.c_int, vp, C.c_size_t, vp, vp, C.c_int, C.c_double, vp, dp]),
    "minim_potential_constrain": (C.c_int, [vp, C.c_int, vp, C.c_size_t, vp]),
    "minim_dot": (C.c_int, [C.c_int, vp, vp, C.c_size_t, dp, vp]),
    "minim_state_create": (C.c_int, [vp, dp, C.c_size_t, C.c_int, C.POINTER(vp)]),
    "minim_state_copy": (C.c_int, [vp, C.POINTER(vp)]),
    "minim_state_destroy": (C.c_int, [vp]),
    "minim_state_energy": (C.c_int, [vp, dp, dp]),
    "minim_state_gradient": (C.c_int, [vp, dp, dp]),
    "minim_state_energy_gradient": (C.c_int, [vp, dp, dp, dp]),
    "minim_state_get_coords": (C.c_int, [vp, dp]),
    "minim_state_set_coords": (C.c_int, [vp, dp]),
    "minim_state_apply_constraints": (C.c_int, [vp, dp]),
    "minim_state_failed": (C.c_int, [vp]),
    "minim_state_info": (C.c_int, [vp, C.POINTER(C.c_size_t), ip, dp, ip]),
    "minim_state_set_convergence": (C.c_int, [vp, C.c_double]),
    "minim_state_set_precision": (C.c_int, [vp, C.c_int]),
    "minim_state_set_stream": (C.c_int, [vp, vp]),
    "minim_state_device_coords": (vp, [vp]),
    "minim_state_mark_coords_changed": (C.c_int, [vp]),
    "minim_state_component_energy": (C.c_int, [vp, C.c_int, dp]),
    "minim_batch_report": (C.c_int, [vp, ip, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_long),
                                     C.POINTER(C.c_long), ip]),
    "minim_lbfgs_autotune_batch": (C.c_int, [vp, C.POINTER(vp), C.c_int, ip, C.c_int, C.c_int, dp, ip]),
    "minim_minimiser_create": (C.c_int, [C.POINTER(MinimiserParams), C.POINTER(vp)]),
    "minim_minimiser_destroy": (C.c_int, [vp]),
    "minim_minimise": (C.c_int, [vp, vp, ADJUST_FN, vp, dp, C.POINTER(Result)]),
    "minim_minimise_log": (C.c_int, [vp, vp, C.c_char_p, dp, C.POINTER(Result)]),
    "minim_minimise_batch": (C.c_int, [vp, C.POINTER(vp), C.c_int, C.POINTER(Result)]),
}

_lib = None


ABI_VERSION = 10  # layout of the parameter/result structs below


def load():
    """Load libminim_amd.so.  There is no fallback: a missing library is an error."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("minim_amd: %s is not built; run `python -m minim_amd.build` "
                              "(or __graft_entry__.build())" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.minim_abi_version() != ABI_VERSION:
            raise ImportError("minim_amd: %s has ABI version %d, this package expects %d; rebuild it"
                              % (LIB_PATH, L.minim_abi_version(), ABI_VERSION))
        _lib = L
    return _lib


class MinimError(Exception):
    pass


def check(rc):
    if rc != MINIM_OK:
        msg = load().minim_last_error().decode(errors="replace")
        if rc == MINIM_EINVAL:
            raise ValueError(msg)
        if rc == MINIM_ENOMEM:
            raise MemoryError(msg)
        raise MinimError("minim_amd error %d: %s" % (rc, msg))
