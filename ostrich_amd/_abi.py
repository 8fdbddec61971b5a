"""ctypes mirror of the C ABI declared in include/ostrich_fb.h.

The structs here are the boundary layout shared by the HIP library
(ostrich_amd/csrc) and every caller.  Field order must match the header exactly.
"""
import ctypes as C

OFB_PRECOND_JACOBI = 0
OFB_PRECOND_PER_BODY_PAIR = 1
OFB_FRICTION_CONE = 0
OFB_FRICTION_BILATERAL = 1
OFB_FRICTION_BILATERAL_PATCH = 2
OFB_CONTACT_COMPLEMENTARITY = 0
OFB_CONTACT_EQUALITY = 1
OFB_TERM_GLOBAL = 0
OFB_TERM_PER_WORLD = 1
OFB_WS_POSITION_MATCH = 0
OFB_WS_PAIR_AGGREGATE = 1
OFB_PATH_FUSED = 0
OFB_PATH_PLAIN = 1

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class OfbModel(C.Structure):
    _fields_ = [
        ("num_worlds", C.c_int32),
        ("bodies_per_world", C.c_int32),
        ("joints_per_world", C.c_int32),
        ("dofs_per_world", C.c_int32),
        ("shapes_per_world", C.c_int32),
        ("num_global_shapes", C.c_int32),
        ("gravity", C.c_float * 3),
        ("body_mass", C.c_void_p),
        ("body_inv_mass", C.c_void_p),
        ("body_inertia", C.c_void_p),
        ("body_inv_inertia", C.c_void_p),
        ("body_com", C.c_void_p),
        ("joint_type", C.c_void_p),
        ("joint_parent", C.c_void_p),
        ("joint_child", C.c_void_p),
        ("joint_qd_start", C.c_void_p),
        ("joint_enabled", C.c_void_p),
        ("joint_X_p", C.c_void_p),
        ("joint_X_c", C.c_void_p),
        ("joint_compliance", C.c_void_p),
        ("joint_axis", C.c_void_p),
        ("joint_dof_mode", C.c_void_p),
        ("joint_target_ke", C.c_void_p),
        ("joint_target_kd", C.c_void_p),
        ("shape_body", C.c_void_p),
        ("shape_mu", C.c_void_p),
        ("shape_friction_axis", C.c_void_p),
        ("shape_mu_perp", C.c_void_p),
        ("shape_mu_stiction_scale", C.c_void_p),
        ("shape_v_stribeck", C.c_void_p),
        ("shape_stribeck_lateral_only", C.c_void_p),
        ("shape_world", C.c_void_p),
    ]


class OfbConfig(C.Structure):
    _fields_ = [
        ("nr_max_iters", C.c_int32),
        ("nr_backtrack_min_iter", C.c_int32),
        ("nr_atol", C.c_float),
        ("lin_max_iters", C.c_int32),
        ("lin_tol", C.c_float),
        ("lin_atol", C.c_float),
        ("preconditioner", C.c_int32),
        ("regularization", C.c_float),
        ("compliance_joint", C.c_float),
        ("compliance_contact", C.c_float),
        ("compliance_friction", C.c_float),
        ("contact_fb_eps_sq", C.c_float),
        ("contact_fb_alpha", C.c_float),
        ("gyro_scale", C.c_float),
        ("friction_load_gate", C.c_float),
        ("friction_mode", C.c_int32),
        ("contact_mode", C.c_int32),
        ("max_contacts_per_world", C.c_int32),
        ("termination", C.c_int32),
        ("seed_iterate", C.c_int32),
        ("rigid_contact_max", C.c_int32),
        ("adjoint_regularization", C.c_float),
        ("adjoint_pose_vjp", C.c_int32),
        ("adjoint_pose_eps", C.c_float),
        ("warm_start", C.c_int32),
        ("warm_start_cold_gravity", C.c_int32),
        ("warm_start_cold_impact", C.c_int32),
        ("warm_start_cold_friction_v", C.c_float),
        ("linesearch", C.c_int32),
        ("linesearch_conservative_count", C.c_int32),
        ("linesearch_optimistic_count", C.c_int32),
        ("linesearch_min_step", C.c_double),
        ("linesearch_conservative_upper", C.c_double),
        ("linesearch_optimistic_window", C.c_double),
        ("pair_small_worlds", C.c_int32),
        ("contact_reduction", C.c_int32),
        ("contact_max_per_pair", C.c_int32),
        ("cluster_normal_dot", C.c_float),
        ("cluster_pos", C.c_float),
        ("warm_start_method", C.c_int32),
        ("adjoint_friction_frozen", C.c_int32),
        ("adjoint_soft_blending", C.c_int32),
        ("adjoint_soft_temperature", C.c_float),
        ("large_world_slots", C.c_int32),
    ]


class OfbStateIn(C.Structure):
    _fields_ = [("body_q", C.c_void_p), ("body_qd", C.c_void_p), ("body_f", C.c_void_p)]


class OfbStateOut(C.Structure):
    _fields_ = [("body_q", C.c_void_p), ("body_qd", C.c_void_p)]


class OfbControl(C.Structure):
    _fields_ = [("joint_target_pos", C.c_void_p), ("joint_target_vel", C.c_void_p)]


class OfbContacts(C.Structure):
    _fields_ = [
        ("count", C.c_void_p),
        ("max_contacts", C.c_int32),
        ("point0", C.c_void_p),
        ("point1", C.c_void_p),
        ("normal", C.c_void_p),
        ("shape0", C.c_void_p),
        ("shape1", C.c_void_p),
        ("margin0", C.c_void_p),
        ("margin1", C.c_void_p),
    ]


class OfbColliderDesc(C.Structure):
    _fields_ = [
        ("num_worlds", C.c_int32), ("bodies_per_world", C.c_int32), ("shape_cols", C.c_int32),
        ("num_pairs", C.c_int32),
        ("pairs", C.c_void_p),
        ("shape_geo", C.c_void_p),
        ("shape_body", C.c_void_p),
        ("shape_xform", C.c_void_p),
        ("shape_size", C.c_void_p),
        ("flat_shape_id", C.c_void_p),
        ("cylinder_samples", C.c_int32),
        ("gap", C.c_float),
        ("lateral_slack", C.c_float),
        ("max_contacts_per_world", C.c_int32),
        ("num_hf_spheres", C.c_int32),
        ("hf_spheres", C.c_void_p),
        ("hf_col", C.c_int32), ("hf_nx", C.c_int32), ("hf_ny", C.c_int32),
        ("hf_origin", C.c_float * 2), ("hf_spacing", C.c_float * 2),
        ("hf_heights", C.c_void_p),
        ("hf_points", C.c_int32),
    ]


class OfbContactBuffers(C.Structure):
    _fields_ = OfbContacts._fields_


class OfbAdjointIn(C.Structure):
    _fields_ = [("body_q_grad", C.c_void_p), ("body_qd_grad", C.c_void_p)]


class OfbAdjointOut(C.Structure):
    _fields_ = [
        ("body_q_grad", C.c_void_p),
        ("body_qd_grad", C.c_void_p),
        ("body_f_grad", C.c_void_p),
        ("joint_target_pos_grad", C.c_void_p),
        ("joint_target_vel_grad", C.c_void_p),
        ("joint_target_ke_grad", C.c_void_p),
        ("joint_target_kd_grad", C.c_void_p),
        ("shape_mu_grad", C.c_void_p),
        ("adjoint_lambda", C.c_void_p),
        ("adjoint_body", C.c_void_p),
    ]


MODEL_ARRAYS = [
    # (field, dtype, per-row width, owner) ; owner in {"body","joint","dof","shape","shape_flat"}
    ("body_mass", "f4", 1, "body"),
    ("body_inv_mass", "f4", 1, "body"),
    ("body_inertia", "f4", 9, "body"),
    ("body_inv_inertia", "f4", 9, "body"),
    ("body_com", "f4", 3, "body"),
    ("joint_type", "i4", 1, "joint"),
    ("joint_parent", "i4", 1, "joint"),
    ("joint_child", "i4", 1, "joint"),
    ("joint_qd_start", "i4", 1, "joint"),
    ("joint_enabled", "i4", 1, "joint"),
    ("joint_X_p", "f4", 7, "joint"),
    ("joint_X_c", "f4", 7, "joint"),
    ("joint_compliance", "f4", 1, "joint"),
    ("joint_axis", "f4", 3, "dof"),
    ("joint_dof_mode", "i4", 1, "dof"),
    ("joint_target_ke", "f4", 1, "dof"),
    ("joint_target_kd", "f4", 1, "dof"),
    ("shape_body", "i4", 1, "shape"),
    ("shape_mu", "f4", 1, "shape"),
    ("shape_friction_axis", "f4", 3, "shape"),
    ("shape_mu_perp", "f4", 1, "shape"),
    ("shape_mu_stiction_scale", "f4", 1, "shape"),
    ("shape_v_stribeck", "f4", 1, "shape"),
    ("shape_stribeck_lateral_only", "f4", 1, "shape"),
    ("shape_world", "i4", 1, "shape_flat"),
]
