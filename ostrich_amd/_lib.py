"""Loads the gfx950 solver library (ostrich_amd/_lib/libostrich_fb.so) with ctypes.

There is no fallback: if the library is missing or cannot be loaded the import fails
with the reason, so nothing silently runs on another path.
"""
import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libostrich_fb.so")

EXPORTS = (
    "ofb_create", "ofb_destroy", "ofb_dims", "ofb_step", "ofb_get_stats", "ofb_get_lambda",
    "ofb_get_packed_contacts", "ofb_set_lambda_seed", "ofb_last_solve_ms", "ofb_last_error",
    "ofb_get_lambda_prev", "ofb_set_shape_mu", "ofb_step_backward", "ofb_warm_start_reset",
    "ofb_get_lambda_seed", "ofb_mlp_create", "ofb_mlp_destroy", "ofb_mlp_num_params", "ofb_mlp_set_params",
    "ofb_mlp_get_params", "ofb_mlp_forward", "ofb_mlp_gradient", "ofb_mlp_train_step",
    "ofb_set_warm_start_net", "ofb_get_contact_features", "ofb_get_linesearch_steps",
    "ofb_last_kernel_ms", "ofb_set_fused_launch", "ofb_get_fused_launch", "ofb_trajectory_adam_step",
    "ofb_collider_create", "ofb_collider_destroy", "ofb_collide", "ofb_set_iteration_history",
    "ofb_log_create", "ofb_log_capture", "ofb_log_destroy", "ofb_set_class_timing", "ofb_get_class_times",
    "ofb_get_class_shapes", "ofb_set_keep_lambda_prev",
    "ofb_warm_start_state_size", "ofb_warm_start_state_get", "ofb_warm_start_state_set",
    "ofb_mlp_gradient_noalloc", "ofb_set_fused_layout", "ofb_get_fused_layout",
)

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m ostrich_amd.build`")
    lib = C.CDLL(LIB_PATH)
    vp, i32 = C.c_void_p, C.c_int32
    lib.ofb_create.argtypes = [vp, vp, i32, C.POINTER(vp)]
    lib.ofb_destroy.argtypes = [vp]
    lib.ofb_dims.argtypes = [vp, vp, vp, vp, vp]
    lib.ofb_step.argtypes = [vp, vp, vp, vp, vp, C.c_float, vp]
    lib.ofb_get_stats.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.ofb_get_lambda.argtypes = [vp, vp, vp]
    lib.ofb_get_packed_contacts.argtypes = [vp] * 9 + [vp]
    lib.ofb_set_lambda_seed.argtypes = [vp, vp, vp]
    lib.ofb_last_solve_ms.argtypes = [vp, C.POINTER(C.c_float)]
    lib.ofb_last_kernel_ms.argtypes = [vp, C.POINTER(C.c_float)]
    lib.ofb_set_fused_launch.argtypes = [vp, i32, i32]
    lib.ofb_set_class_timing.argtypes = [vp, i32]
    lib.ofb_set_keep_lambda_prev.argtypes = [vp, i32]
    lib.ofb_warm_start_state_size.argtypes = [vp, C.POINTER(C.c_int64)]
    lib.ofb_warm_start_state_get.argtypes = [vp, vp, vp]
    lib.ofb_warm_start_state_set.argtypes = [vp, vp, vp]
    lib.ofb_get_class_times.argtypes = [vp, i32, vp, vp, vp, vp]
    lib.ofb_get_class_shapes.argtypes = [vp, i32, vp, vp, vp, vp]
    lib.ofb_get_fused_launch.argtypes = [vp, vp, vp]
    lib.ofb_set_fused_layout.argtypes = [vp, i32]
    lib.ofb_get_fused_layout.argtypes = [vp, vp]
    f32 = C.c_float
    lib.ofb_trajectory_adam_step.argtypes = [vp, vp, vp, vp] + [i32] * 5 + [f32] * 5 + [i32, vp]
    lib.ofb_collider_create.argtypes = [vp, C.POINTER(vp)]
    lib.ofb_collider_destroy.argtypes = [vp]
    lib.ofb_collide.argtypes = [vp, vp, vp, vp]
    lib.ofb_set_iteration_history.argtypes = [vp, vp, vp, i32]
    lib.ofb_log_create.argtypes = [i32, vp, vp, vp, i32, C.POINTER(vp)]
    lib.ofb_log_capture.argtypes = [vp, vp, i32, vp]
    lib.ofb_log_destroy.argtypes = [vp]
    lib.ofb_last_error.restype = C.c_char_p
    lib.ofb_get_lambda_prev.argtypes = [vp, vp, vp]
    lib.ofb_set_shape_mu.argtypes = [vp, vp, vp]
    lib.ofb_step_backward.argtypes = [vp] * 9 + [C.c_float, vp]
    lib.ofb_warm_start_reset.argtypes = [vp, vp]
    lib.ofb_get_lambda_seed.argtypes = [vp, vp, vp]
    lib.ofb_mlp_create.argtypes = [C.POINTER(vp)]
    lib.ofb_mlp_destroy.argtypes = [vp]
    lib.ofb_mlp_num_params.argtypes = []
    lib.ofb_mlp_set_params.argtypes = [vp, vp, vp]
    lib.ofb_mlp_get_params.argtypes = [vp, vp, vp]
    lib.ofb_mlp_forward.argtypes = [vp, vp, C.c_int, vp, vp]
    lib.ofb_mlp_gradient.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp]
    lib.ofb_mlp_gradient_noalloc.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp]
    lib.ofb_mlp_train_step.argtypes = [vp, vp, vp, C.c_int, C.c_float, vp, vp]
    lib.ofb_set_warm_start_net.argtypes = [vp, vp]
    lib.ofb_get_contact_features.argtypes = [vp, vp, vp, vp]
    lib.ofb_get_linesearch_steps.argtypes = [vp, vp, vp]
    for name in EXPORTS:
        if name != "ofb_last_error":
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        msg = load().ofb_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")
