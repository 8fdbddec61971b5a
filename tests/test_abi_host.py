"""Host-side checks that need no GPU: the C-ABI library loads and exports every entry
point include/ostrich_fb.h declares; the ctypes structs match the header; config
validation mirrors core/engine_config.py."""
import ctypes
import os
import re

import pytest

from ostrich_amd import _abi, _lib
from ostrich_amd.config import (ExecutionConfig, LinearSolverConfig, NewtonRaphsonConfig, OstrichEngineConfig,
                                RelaxationConfig)

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ostrich_fb.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ofb_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)


def _struct_fields(name):
    text = open(HEADER).read()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return re.findall(r"(\w+)(?:\[\d+\])?;", body)


@pytest.mark.parametrize("cname,pyc", [("ofb_model", _abi.OfbModel), ("ofb_config", _abi.OfbConfig),
                                       ("ofb_contacts", _abi.OfbContacts), ("ofb_state_in", _abi.OfbStateIn),
                                       ("ofb_state_out", _abi.OfbStateOut), ("ofb_control", _abi.OfbControl),
                                       ("ofb_adjoint_in", _abi.OfbAdjointIn),
                                       ("ofb_adjoint_out", _abi.OfbAdjointOut),
                                       ("ofb_collider_desc", _abi.OfbColliderDesc),
                                       ("ofb_contact_buffers", _abi.OfbContactBuffers)])
def test_ctypes_structs_match_header(cname, pyc):
    assert _struct_fields(cname) == [f[0] for f in pyc._fields_]


def test_config_validation_mirrors_reference():
    with pytest.raises(ValueError):
        NewtonRaphsonConfig(max_iters=2, backtrack_min_iter=2)
    with pytest.raises(ValueError):
        LinearSolverConfig(preconditioner_type="ilu")
    with pytest.raises(ValueError):
        RelaxationConfig(friction="viscous")
    with pytest.raises(ValueError):
        ExecutionConfig(path="cuda")
    cfg = OstrichEngineConfig(nr={"max_iters": 7, "backtrack_min_iter": 3})
    assert cfg.nr.max_iters == 7 and cfg.to_abi().nr_backtrack_min_iter == 3
    with pytest.raises(NotImplementedError):
        OstrichEngineConfig(relaxation=RelaxationConfig(friction="bilateral")).create_engine(
            model=None, differentiable_simulation=True)
    # the reference's legacy adjoint and warm-start modes reach the ABI
    fz = OstrichEngineConfig(adjoint={"friction_linearization": "frozen", "soft_blending": False}).to_abi()
    assert fz.adjoint_friction_frozen == 1 and fz.adjoint_soft_blending == 0
    assert OstrichEngineConfig(warm_start={"method": "pair_aggregate"}).to_abi().warm_start_method == \
        _abi.OFB_WS_PAIR_AGGREGATE
    # the large-world layout (no reference counterpart) reaches the ABI; 2 or 4 only
    assert OstrichEngineConfig(execution={"large_world_slots": 4}).to_abi().large_world_slots == 4
    assert OstrichEngineConfig().to_abi().large_world_slots == 2
    with pytest.raises(ValueError):
        ExecutionConfig(large_world_slots=3)


def test_engine_requires_gpu_on_host():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    from ostrich_amd import OstrichEngine, scenes
    model, _, _ = scenes.box_stack(1, 1)
    with pytest.raises(RuntimeError, match="HIP device"):
        OstrichEngine(model)
