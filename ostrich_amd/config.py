"""Solver configuration, mirroring OstrichEngineConfig (src/ostrich/core/engine_config.py).

Only the sub-configs that act on the solver path are kept; field names, defaults and
validation messages follow engine_config.py:103-577 so reference YAML overrides map 1:1.
`execution` is new: it selects the MI355X execution path and the loop termination rule.
"""
from dataclasses import dataclass, field, fields

from . import _abi


def _coerce(cls, value):
    """engine_config.py:38-53: accept a dict-like override for a sub-config."""
    if isinstance(value, cls):
        return value
    try:
        items = dict(value)
    except Exception:
        return cls()
    valid = {f.name for f in fields(cls)}
    return cls(**{k: v for k, v in items.items() if k in valid})


@dataclass(frozen=True)
class NewtonRaphsonConfig:
    """engine_config.py:102-127."""

    max_iters: int = 16
    backtrack_min_iter: int = 2
    atol: float = 1e-3

    def __post_init__(self):
        if self.max_iters < 1:
            raise ValueError(f"nr.max_iters must be >= 1, got {self.max_iters}")
        if self.backtrack_min_iter < 0:
            raise ValueError(f"nr.backtrack_min_iter must be >= 0, got {self.backtrack_min_iter}")
        if self.backtrack_min_iter >= self.max_iters:
            raise ValueError(
                "nr.backtrack_min_iter must be < nr.max_iters, "
                f"got {self.backtrack_min_iter} and {self.max_iters}"
            )
        if self.atol < 0:
            raise ValueError(f"nr.atol must be >= 0, got {self.atol}")


@dataclass(frozen=True)
class LinearSolverConfig:
    """engine_config.py:130-159."""

    max_iters: int = 16
    tol: float = 1e-5
    atol: float = 1e-5
    preconditioner_type: str = "jacobi"
    regularization: float = 1e-6

    def __post_init__(self):
        if self.max_iters < 1:
            raise ValueError(f"linear.max_iters must be >= 1, got {self.max_iters}")
        if self.tol < 0:
            raise ValueError(f"linear.tol must be >= 0, got {self.tol}")
        if self.atol < 0:
            raise ValueError(f"linear.atol must be >= 0, got {self.atol}")
        if self.regularization < 0:
            raise ValueError(f"linear.regularization must be >= 0, got {self.regularization}")
        if self.preconditioner_type not in ("jacobi", "per_body_pair"):
            raise ValueError(
                "linear.preconditioner_type must be 'jacobi' | 'per_body_pair', "
                f"got {self.preconditioner_type!r}"
            )


@dataclass(frozen=True)
class ComplianceConfig:
    """engine_config.py:162-185."""

    joint: float = 1e-5
    contact: float = 1e-4
    friction: float = 1e-6
    contact_fb_smooth_eps_sq: float = 1e-8

    def __post_init__(self):
        for name in ("joint", "contact", "friction"):
            v = getattr(self, name)
            if v < 0:
                raise ValueError(f"compliance.{name} must be >= 0, got {v}")


@dataclass(frozen=True)
class RelaxationConfig:
    """engine_config.py:188-287."""

    friction: str = "cone"
    contact: str = "complementarity"
    gyro: bool = True
    friction_load_gate: float = 1e-6

    _FRICTION_MODES = ("cone", "bilateral", "bilateral_patch")
    _CONTACT_MODES = ("complementarity", "equality")

    def __post_init__(self):
        if self.friction not in self._FRICTION_MODES:
            raise ValueError(
                f"relaxation.friction must be one of {self._FRICTION_MODES}, got {self.friction!r}"
            )
        if self.friction_load_gate < 0:
            raise ValueError(
                f"relaxation.friction_load_gate must be >= 0, got {self.friction_load_gate}"
            )
        if self.contact not in self._CONTACT_MODES:
            raise ValueError(
                f"relaxation.contact must be one of {self._CONTACT_MODES}, got {self.contact!r}"
            )

    @property
    def friction_mode(self) -> int:
        return self._FRICTION_MODES.index(self.friction)

    @property
    def contact_mode(self) -> int:
        return self._CONTACT_MODES.index(self.contact)

    @property
    def gyro_scale(self) -> float:
        return 1.0 if self.gyro else 0.0

    @property
    def is_default(self) -> bool:
        return (
            self.friction == "cone"
            and self.contact == "complementarity"
            and self.gyro
            and self.friction_load_gate == 1e-6
        )


@dataclass(frozen=True)
class LinesearchConfig:
    """engine_config.py:290-333 (merit line search over a fixed candidate set; fused path)."""

    enabled: bool = False
    min_step: float = 1e-6
    conservative_step_count: int = 32
    conservative_upper_bound: float = 0.05
    optimistic_step_count: int = 32
    optimistic_window: float = 0.2

    @property
    def num_steps(self) -> int:
        return self.conservative_step_count + self.optimistic_step_count

    def __post_init__(self):
        if not self.enabled:
            return  # engine_config.py:306-308: no validation when off
        if self.conservative_step_count < 1:
            raise ValueError(f"linesearch.conservative_step_count must be >= 1, got {self.conservative_step_count}")
        if self.optimistic_step_count < 1:
            raise ValueError(f"linesearch.optimistic_step_count must be >= 1, got {self.optimistic_step_count}")
        if self.min_step < 0:
            raise ValueError(f"linesearch.min_step must be >= 0, got {self.min_step}")
        if self.conservative_upper_bound <= self.min_step:
            raise ValueError("linesearch.conservative_upper_bound must be > min_step, "
                             f"got {self.conservative_upper_bound} and {self.min_step}")
        if self.optimistic_window < 0:
            raise ValueError(f"linesearch.optimistic_window must be >= 0, got {self.optimistic_window}")


@dataclass(frozen=True)
class WarmStartConfig:
    """engine_config.py:336-377.

    `enabled` only has an effect together with `seed_iterate`: the reference copies
    lambda (zero at step start) over the friction-lag buffer at the top of every Newton
    iteration (base_engine.py:441), so the warm starter's writes to that buffer in
    load_data (base_engine.py:322-335) never reach the solve unless the iterate is seeded.
    """

    enabled: bool = True
    cold_gravity: bool = True
    cold_impact: bool = True
    cold_friction_v_threshold: float = 0.1
    method: str = "position_match"
    seed_iterate: bool = False

    def __post_init__(self):
        if self.method not in ("position_match", "pair_aggregate"):
            raise ValueError(
                "WarmStartConfig.method must be 'position_match' or 'pair_aggregate', "
                f"got {self.method!r}"
            )


REDUCTION_POLICIES = {"none": 0, "top_k": 1, "fps": 2, "cluster": 3}


@dataclass(frozen=True)
class ContactReductionConfig:
    """collision/config.py:15-56.  The reference's "hull" policy is declared there but
    has no reducer (collision/__init__.py builds none for it); it raises here."""

    policy: str = "none"
    max_per_pair: int = 4
    cluster_normal_dot_thresh: float = 0.996
    cluster_pos_thresh: float = 5e-3

    def __post_init__(self):
        if self.policy not in REDUCTION_POLICIES:
            raise NotImplementedError(f"contact reduction policy {self.policy!r} is not built "
                                      f"(built: {', '.join(REDUCTION_POLICIES)})")
        if self.max_per_pair < 1:
            raise ValueError(f"max_per_pair must be >= 1, got {self.max_per_pair}")
        if not (-1.0 <= self.cluster_normal_dot_thresh <= 1.0):
            raise ValueError(f"cluster_normal_dot_thresh must be in [-1, 1], got {self.cluster_normal_dot_thresh}")
        if self.cluster_pos_thresh < 0.0:
            raise ValueError(f"cluster_pos_thresh must be >= 0, got {self.cluster_pos_thresh}")


@dataclass(frozen=True)
class ContactsConfig:
    """engine_config.py:380-400: contact slots per world and the per-pair reduction."""

    max_per_world: int = 128
    reduction: ContactReductionConfig = field(default_factory=ContactReductionConfig)

    def __post_init__(self):
        if self.max_per_world < 1:
            raise ValueError(f"contacts.max_per_world must be >= 1, got {self.max_per_world}")


@dataclass(frozen=True)
class AdjointConfig:
    """engine_config.py:403-447."""

    pose_vjp: bool = True
    friction_linearization: str = "true"
    soft_blending: bool = True
    soft_blending_temperature: float = 0.05
    regularization: float = 0.0
    gradient_normalization: bool = False
    # step of the central differences that form w^T dR/dq+ and w^T dR/dq- on the GPU
    # (the reference differentiates its residual kernels on a Warp tape instead).  Their
    # truncation is second order with a large constant where contacts sit near the kink
    # of the FB function: against steps of 1e-6 in FP64, 5e-4 leaves at most 4e-4 of the
    # gradient's scale on skid-steer robots among obstacles and tracked robots on a
    # heightfield (2e-3 left up to 5e-3), while FP32 rounding stays below 1e-4 of scale.
    pose_fd_eps: float = 5e-4

    def __post_init__(self):
        if self.friction_linearization not in ("true", "frozen"):
            raise ValueError(
                "adjoint.friction_linearization must be 'true' or 'frozen', "
                f"got {self.friction_linearization!r}"
            )
        if self.soft_blending_temperature <= 0:
            raise ValueError(
                "adjoint.soft_blending_temperature must be > 0, "
                f"got {self.soft_blending_temperature}"
            )
        if self.regularization < 0:
            raise ValueError(f"adjoint.regularization must be >= 0, got {self.regularization}")
        if not self.pose_fd_eps > 0:
            raise ValueError(f"adjoint.pose_fd_eps must be > 0, got {self.pose_fd_eps}")


@dataclass(frozen=True)
class ExecutionConfig:
    """MI355X execution choices (no reference counterpart).

    path:        "fused" - one persistent launch, one wavefront per world;
                 "plain" - one kernel per reference kernel family, data in HBM.
    termination: "per_world" - each world leaves the Newton / PCR loops on its own
                 criterion; "global" - the reference's batch-global stop flags
                 (base_engine.py:42-80, pcr_solver.py:84-117).
    pair_small_worlds: fused path - worlds of at most 32 constraint rows run two per
                 wavefront (32 lanes each) and solve PCR on the assembled Schur matrix;
                 False keeps one world per wavefront and the matrix-free operator.
    large_world_slots: fused path, Jacobi - layout of the large worlds: 2 spreads a world
                 of more than 128 constraint rows over 2, 4 or 8 wavefronts with two row
                 slots per lane (the shortest solve of one world); 4, the wide layout,
                 runs 65-512 rows on two wavefronts with two to four slots per lane and
                 513-1024 on four (more of them resident at once: the faster layout for
                 batches of many large worlds).
    """

    path: str = "fused"
    termination: str = "per_world"
    pair_small_worlds: bool = True
    large_world_slots: int = 2

    def __post_init__(self):
        if self.path not in ("fused", "plain"):
            raise ValueError(f"execution.path must be 'fused' or 'plain', got {self.path!r}")
        if self.termination not in ("per_world", "global"):
            raise ValueError(
                f"execution.termination must be 'per_world' or 'global', got {self.termination!r}"
            )
        if self.large_world_slots not in (2, 4):
            raise ValueError(f"execution.large_world_slots must be 2 or 4, got {self.large_world_slots!r}")


_SUBS = (
    ("nr", NewtonRaphsonConfig),
    ("linear", LinearSolverConfig),
    ("compliance", ComplianceConfig),
    ("relaxation", RelaxationConfig),
    ("linesearch", LinesearchConfig),
    ("warm_start", WarmStartConfig),
    ("contacts", ContactsConfig),
    ("adjoint", AdjointConfig),
    ("execution", ExecutionConfig),
)


@dataclass(frozen=True)
class OstrichEngineConfig:
    """engine_config.py:493-577 (solver sub-configs + `execution`)."""

    differentiable: bool = False
    nr: NewtonRaphsonConfig = field(default_factory=NewtonRaphsonConfig)
    linear: LinearSolverConfig = field(default_factory=LinearSolverConfig)
    compliance: ComplianceConfig = field(default_factory=ComplianceConfig)
    relaxation: RelaxationConfig = field(default_factory=RelaxationConfig)
    linesearch: LinesearchConfig = field(default_factory=LinesearchConfig)
    warm_start: WarmStartConfig = field(default_factory=WarmStartConfig)
    contacts: ContactsConfig = field(default_factory=ContactsConfig)
    adjoint: AdjointConfig = field(default_factory=AdjointConfig)
    execution: ExecutionConfig = field(default_factory=ExecutionConfig)

    def __post_init__(self):
        for name, cls in _SUBS:
            cur = getattr(self, name)
            coerced = _coerce(cls, cur)
            if coerced is not cur:
                object.__setattr__(self, name, coerced)

    def create_engine(self, model, sim_steps=None, logging_config=None, differentiable_simulation=None):
        """engine_config.py:527-557."""
        from .engine import OstrichEngine

        diff = self.differentiable if differentiable_simulation is None else differentiable_simulation
        if diff and not self.relaxation.is_default:
            raise NotImplementedError(
                "relaxed rungs have no adjoint: run them forward-only (differentiable=False)"
            )
        return OstrichEngine(model=model, sim_steps=sim_steps, config=self,
                             differentiable_simulation=diff)

    def to_abi(self) -> _abi.OfbConfig:
        c = _abi.OfbConfig()
        c.nr_max_iters = self.nr.max_iters
        c.nr_backtrack_min_iter = self.nr.backtrack_min_iter
        c.nr_atol = self.nr.atol
        c.lin_max_iters = self.linear.max_iters
        c.lin_tol = self.linear.tol
        c.lin_atol = self.linear.atol
        c.preconditioner = (
            _abi.OFB_PRECOND_PER_BODY_PAIR
            if self.linear.preconditioner_type == "per_body_pair"
            else _abi.OFB_PRECOND_JACOBI
        )
        c.regularization = self.linear.regularization
        c.compliance_joint = self.compliance.joint
        c.compliance_contact = self.compliance.contact
        c.compliance_friction = self.compliance.friction
        c.contact_fb_eps_sq = self.compliance.contact_fb_smooth_eps_sq
        c.contact_fb_alpha = 1.0
        c.gyro_scale = self.relaxation.gyro_scale
        c.friction_load_gate = self.relaxation.friction_load_gate
        c.friction_mode = self.relaxation.friction_mode
        c.contact_mode = self.relaxation.contact_mode
        c.max_contacts_per_world = self.contacts.max_per_world
        c.termination = (
            _abi.OFB_TERM_GLOBAL if self.execution.termination == "global" else _abi.OFB_TERM_PER_WORLD
        )
        c.seed_iterate = 1 if self.warm_start.seed_iterate else 0
        c.rigid_contact_max = 0
        c.adjoint_regularization = self.adjoint.regularization
        c.adjoint_pose_vjp = 1 if self.adjoint.pose_vjp else 0
        c.adjoint_pose_eps = self.adjoint.pose_fd_eps
        c.warm_start_method = (_abi.OFB_WS_PAIR_AGGREGATE if self.warm_start.method == "pair_aggregate"
                               else _abi.OFB_WS_POSITION_MATCH)
        c.adjoint_friction_frozen = 1 if self.adjoint.friction_linearization == "frozen" else 0
        c.adjoint_soft_blending = 1 if self.adjoint.soft_blending else 0
        c.adjoint_soft_temperature = self.adjoint.soft_blending_temperature
        c.warm_start = 1 if self.warm_start.enabled else 0
        c.warm_start_cold_gravity = 1 if self.warm_start.cold_gravity else 0
        c.warm_start_cold_impact = 1 if self.warm_start.cold_impact else 0
        c.warm_start_cold_friction_v = self.warm_start.cold_friction_v_threshold
        ls = self.linesearch
        c.linesearch = 1 if ls.enabled else 0
        c.linesearch_min_step = ls.min_step
        c.linesearch_conservative_count = ls.conservative_step_count
        c.linesearch_conservative_upper = ls.conservative_upper_bound
        c.linesearch_optimistic_count = ls.optimistic_step_count
        c.linesearch_optimistic_window = ls.optimistic_window
        c.pair_small_worlds = 1 if self.execution.pair_small_worlds else 0
        c.large_world_slots = int(self.execution.large_world_slots)
        red = self.contacts.reduction
        c.contact_reduction = REDUCTION_POLICIES[red.policy]
        c.contact_max_per_pair = red.max_per_pair
        c.cluster_normal_dot = red.cluster_normal_dot_thresh
        c.cluster_pos = red.cluster_pos_thresh
        return c

    @property
    def path_id(self) -> int:
        return _abi.OFB_PATH_PLAIN if self.execution.path == "plain" else _abi.OFB_PATH_FUSED
