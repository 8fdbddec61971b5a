"""A full fused step recorded into a HIP graph, and the launch-shape autotuner.

The graph is replayed several times with the program's own eager work and
synchronisations between replays; every replay must equal, bit for bit, the same step
run eagerly on a second solver (same kernels, same inputs).  The autotuner may only
change the launch shape: with pairing fixed, a tuned solver's results equal an untuned
solver's bit for bit.
"""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.config import WarmStartConfig
from ostrich_amd.containers import Contacts, Control, State
from tests.helpers import config

pytestmark = pytest.mark.gpu


def _setup(worlds, device):
    from ostrich_amd import OstrichEngine
    model, state, control = scenes.skid_steer(worlds, seed=1)
    cfg = config(K=128, warm_start=WarmStartConfig(enabled=True, seed_iterate=True))
    col = collide.PrimitiveCollider(model, device=device, gap=0.02, max_contacts_per_world=128)
    mk = lambda: OstrichEngine(model, config=cfg, device=device)
    return model, State.from_numpy(state, device), Control.from_numpy(control, device), col, mk


def test_captured_step_replays_match_eager(gpu_device):
    model, s0, ctl, col, mk = _setup(128, gpu_device)
    eager, graphed = mk(), mk()
    a_in, a_out = s0.clone(), s0.clone()
    b_in, b_out = s0.clone(), s0.clone()
    b_con = Contacts.from_dict(col.collide(b_in.body_q), gpu_device)
    g = graphed.capture(b_in, b_out, ctl, b_con, 0.05)
    for k in range(4):
        con = Contacts.from_dict(col.collide(a_in.body_q), gpu_device)
        eager.step(a_in, a_out, ctl, con, 0.05)
        b_con.copy_(con)
        g.replay()
        torch.cuda.synchronize()
        print(f"[graph] replay {k}: nr mean {graphed.stats()['nr_iters'].float().mean().item():.2f}", flush=True)
        assert torch.equal(a_out.body_q, b_out.body_q) and torch.equal(a_out.body_qd, b_out.body_qd), k
        assert torch.equal(eager.constraint_impulses(), graphed.constraint_impulses()), k
        for key in ("nr_iters", "pcr_iters", "flags"):
            assert torch.equal(eager.stats()[key], graphed.stats()[key]), (k, key)
        # the program's own eager work between replays
        junk = torch.randn(1 << 16, device=gpu_device).sum()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        a_in.copy_(a_out)
        b_in.copy_(b_out)
    with pytest.raises(RuntimeError, match="not timed"):
        graphed.last_solve_ms()


def test_autotune_changes_only_the_launch_shape(gpu_device):
    model, s0, ctl, col, mk = _setup(256, gpu_device)
    tuned, plain = mk(), mk()
    con = Contacts.from_dict(col.collide(s0.body_q), gpu_device)
    choice, times = tuned.autotune(s0, ctl, con, 0.05, pairing=False, layouts=(2,), repeats=2)
    print(f"[autotune] {choice} from {times}", flush=True)
    assert choice in times and tuned.fused_launch() + (tuned.fused_layout(),) == choice
    assert all(np.isfinite(v) and v > 0 for v in times.values())
    out_t, out_p = s0.clone(), s0.clone()
    tuned.step(s0, out_t, ctl, con, 0.05)
    plain.step(s0, out_p, ctl, con, 0.05)
    torch.cuda.synchronize()
    assert torch.equal(out_t.body_q, out_p.body_q) and torch.equal(out_t.body_qd, out_p.body_qd)
    assert tuned.last_kernel_ms() > 0.0


@pytest.mark.parametrize("waves", [8, 16])
def test_wide_workgroup_requests_stay_within_the_kernel_bound(gpu_device, waves):
    """Worlds small enough that 8 or 16 wavefronts' worth would fit the 64 KB of LDS per
    workgroup (3-cube box stacks): the launches still keep to the kernels' 256-thread
    workgroup bound, and the step equals the default 4-wavefront setting bit for bit."""
    from ostrich_amd import OstrichEngine
    model, state, control = scenes.box_stack(16, 3, seed=0)
    cfg = config(K=16)
    col = collide.PrimitiveCollider(model, device=gpu_device, gap=0.02, max_contacts_per_world=16)
    s0, ctl = State.from_numpy(state, gpu_device), Control.from_numpy(control, gpu_device)
    con = Contacts.from_dict(col.collide(s0.body_q), gpu_device)
    outs = []
    for w in (4, waves):
        eng = OstrichEngine(model, config=cfg, device=gpu_device)
        eng.set_fused_launch(w, True)
        out = s0.clone()
        eng.step(s0, out, ctl, con, 0.05)
        torch.cuda.synchronize()
        outs.append((out, eng.constraint_impulses()))
    print(f"[launch] {waves} waves requested on box stacks", flush=True)
    assert torch.equal(outs[0][0].body_q, outs[1][0].body_q) and torch.equal(outs[0][0].body_qd, outs[1][0].body_qd)
    assert torch.equal(outs[0][1], outs[1][1])


def test_captured_step_in_the_wide_layout_matches_eager(gpu_device):
    """The wide layout of large worlds (four row slots per lane on 2 or 4 wavefronts)
    recorded into a graph: tracked robots with grid-cell contacts within 0.1 m (worlds of
    up to ~700 rows), replays equal eager steps bit for bit."""
    import dataclasses
    from ostrich_amd import OstrichEngine
    model, state, control = scenes.tracked_flipper(32, seed=2)
    base = config(K=256, warm_start=WarmStartConfig(enabled=True, seed_iterate=True))
    cfg = dataclasses.replace(base, execution=dataclasses.replace(base.execution, large_world_slots=4))
    col = collide.HipCollider(model, device=gpu_device, gap=0.1, max_contacts_per_world=256, hf_contacts="cells")
    eager, graphed = OstrichEngine(model, config=cfg, device=gpu_device), OstrichEngine(model, config=cfg,
                                                                                           device=gpu_device)
    assert eager.fused_layout() == 4 and graphed.fused_layout() == 4
    s0, ctl = State.from_numpy(state, gpu_device), Control.from_numpy(control, gpu_device)
    a_in, a_out = s0.clone(), s0.clone()
    b_in, b_out = s0.clone(), s0.clone()
    b_con = col.collide(b_in.body_q).clone()
    g = graphed.capture(b_in, b_out, ctl, b_con, 0.05)
    for k in range(12):
        con = col.collide(a_in.body_q)
        eager.step(a_in, a_out, ctl, con, 0.05)
        b_con.copy_(con)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(a_out.body_q, b_out.body_q) and torch.equal(a_out.body_qd, b_out.body_qd), k
        assert torch.equal(eager.constraint_impulses(), graphed.constraint_impulses()), k
        a_in.copy_(a_out)
        b_in.copy_(b_out)
    cnt = graphed.packed_contacts()["count"]
    rows = (graphed.num_constraints - 3 * 256) + 3 * cnt
    print(f"[graph wide] 12 replays, rows {int(rows.min())}..{int(rows.max())}, "
          f"above 256: {int((rows > 256).sum())}", flush=True)
    assert int((rows > 256).sum()) > 0


def test_captured_collide_step_advance_runs_consecutive_steps(gpu_device):
    """collide + step + state advance recorded in one graph: N replays equal N eager
    simulation steps (same kernels) bit for bit, with no host work between replays."""
    model, s0, ctl, _, mk = _setup(128, gpu_device)
    col_e = collide.HipCollider(model, device=gpu_device, max_contacts_per_world=128)
    col_g = collide.HipCollider(model, device=gpu_device, max_contacts_per_world=128)
    eager, graphed = mk(), mk()
    a_in, a_out = s0.clone(), s0.clone()
    b_in, b_out = s0.clone(), s0.clone()
    g = graphed.capture(b_in, b_out, ctl, None, 0.05, collider=col_g, advance=True)
    for k in range(6):
        eager.step(a_in, a_out, ctl, col_e.collide(a_in.body_q), 0.05)
        a_in.body_q.copy_(a_out.body_q)
        a_in.body_qd.copy_(a_out.body_qd)
        g.replay()
    torch.cuda.synchronize()
    print(f"[graph] 6 collide+step replays, {int(col_g.contacts.rigid_contact_count.item())} contacts", flush=True)
    assert torch.equal(a_in.body_q, b_in.body_q) and torch.equal(a_in.body_qd, b_in.body_qd)
    assert torch.equal(eager.constraint_impulses(), graphed.constraint_impulses())
    assert torch.equal(col_e.contacts.rigid_contact_count, col_g.contacts.rigid_contact_count)
