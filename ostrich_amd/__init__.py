"""MI355X (gfx950) Fisher-Burmeister Newton contact solver with Ostrich's solver entry point."""
from .config import (AdjointConfig, ComplianceConfig, ContactsConfig, ExecutionConfig, LinearSolverConfig,
                     LinesearchConfig, NewtonRaphsonConfig, OstrichEngineConfig, RelaxationConfig, WarmStartConfig)
from .containers import Contacts, Control, State
from .engine import OstrichEngine

__all__ = [
    "OstrichEngine", "OstrichEngineConfig", "NewtonRaphsonConfig", "LinearSolverConfig", "ComplianceConfig",
    "RelaxationConfig", "LinesearchConfig", "WarmStartConfig", "ContactsConfig", "AdjointConfig",
    "ExecutionConfig", "State", "Control", "Contacts",
]
