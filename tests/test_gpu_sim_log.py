"""Device step logs (ofb_log_*) and the per-Newton-iteration history of the fused step.

* Every logged slot equals, bit for bit, the value the engine reports for that step, and
  captures past the capacity write nothing (core/history_group.py:29-31).
* The iteration history is consistent with the step's own statistics: ||h||^2 of the
  picked iteration equals the reported merit, the per-iteration PCR counts sum to the
  step's PCR count, and exactly nr_iters entries per world are filled.
* A graph holding collide + step + log capture + state advance, replayed N times, logs
  the same as N eager steps.
"""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.containers import Control, State
from ostrich_amd.sim_log import DatasetLog, SimulationLog, load_log
from tests.helpers import config

pytestmark = pytest.mark.gpu
DT = 0.05


def _setup(device, W=64):
    from ostrich_amd import OstrichEngine
    model, state, control = scenes.skid_steer(W, seed=1)
    eng = OstrichEngine(model, config=config(K=128), device=device, differentiable_simulation=True)
    col = collide.HipCollider(model, device=device, max_contacts_per_world=128)
    return model, eng, col, State.from_numpy(state, device), Control.from_numpy(control, device)


def test_simulation_log_matches_the_engine(gpu_device, tmp_path):
    model, eng, col, s0, ctl = _setup(gpu_device)
    s_in, s_out = s0.clone(), s0.clone()
    log = SimulationLog(3, eng, s_in, s_out, ctl, DT)
    ref = []
    for t in range(5):
        con = col.collide(s_in.body_q)
        eng.step(s_in, s_out, ctl, con, DT)
        log.capture_step()
        st = eng.stats()
        ref.append(dict(body_pose=s_out.body_q.clone(), body_pose_prev=s_in.body_q.clone(),
                        _constr_force=eng.constraint_impulses(), iter_count=st["nr_iters"],
                        res_norm_sq=st["merit"], candidates_best_idx=st["best_iter"],
                        pcr_iter_count=st["pcr_iters"], flags=st["flags"]))
        s_in.body_q.copy_(s_out.body_q)
        s_in.body_qd.copy_(s_out.body_qd)
    torch.cuda.synchronize()
    print(f"[sim log] 5 steps into 3 slots, step index {int(log.step_index.item())}", flush=True)
    assert int(log.step_index.item()) == 5
    buf = log.log.buffers
    for t in range(3):
        for k, v in ref[t].items():
            assert torch.equal(buf[k][t], v), (t, k)
        hm, hp = buf["candidates_res_norm_sq"][t], buf["pcr_history_iter_count"][t]
        nr, best = ref[t]["iter_count"].long(), ref[t]["candidates_best_idx"].long()
        ran = ~torch.isnan(hm)
        assert torch.equal(ran.sum(0), nr)
        assert torch.equal((hp >= 0).sum(0), nr)
        assert torch.equal(torch.where(hp >= 0, hp, 0).sum(0).to(torch.int32), ref[t]["pcr_iter_count"])
        picked = hm.gather(0, best[None, :])[0]
        assert torch.equal(picked, ref[t]["res_norm_sq"])
    p = tmp_path / "sim.npz"
    log.save(p)
    out = load_log(p)
    assert out["attrs"]["num_steps"] == 3
    np.testing.assert_array_equal(out["data"]["body_pose"], buf["body_pose"].cpu().numpy())
    assert out["dims"]["num_worlds"] == model.num_worlds
    log.close()


def test_logs_in_a_captured_graph_match_eager_steps(gpu_device):
    model, eng_e, col_e, s0, ctl = _setup(gpu_device)
    _, eng_g, col_g, _, _ = _setup(gpu_device)
    a_in, a_out = s0.clone(), s0.clone()
    b_in, b_out = s0.clone(), s0.clone()
    log_e = SimulationLog(6, eng_e, a_in, a_out, ctl, DT)
    log_g = SimulationLog(6, eng_g, b_in, b_out, ctl, DT)
    dlog_e = DatasetLog(6, eng_e, a_in, a_out, ctl, col_e.contacts, DT)
    dlog_g = DatasetLog(6, eng_g, b_in, b_out, ctl, col_g.contacts, DT)
    s = torch.cuda.Stream(device=gpu_device)
    s.wait_stream(torch.cuda.current_stream(gpu_device))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        con = col_g.collide(b_in.body_q)
        eng_g.step(b_in, b_out, ctl, con, DT)
        log_g.capture_step()
        dlog_g.capture_step()
        b_in.body_q.copy_(b_out.body_q)
        b_in.body_qd.copy_(b_out.body_qd)
    torch.cuda.current_stream(gpu_device).wait_stream(s)
    for t in range(6):
        con = col_e.collide(a_in.body_q)
        eng_e.step(a_in, a_out, ctl, con, DT)
        log_e.capture_step()
        dlog_e.capture_step()
        a_in.body_q.copy_(a_out.body_q)
        a_in.body_qd.copy_(a_out.body_qd)
        g.replay()
    torch.cuda.synchronize()
    print(f"[sim log] graph vs eager, 6 steps, {int(col_g.contacts.rigid_contact_count.item())} contacts", flush=True)
    for a, b in ((log_e, log_g), (dlog_e, dlog_g)):
        assert int(a.step_index.item()) == int(b.step_index.item()) == 6
        for k in a.log.buffers:
            x, y = a.log.buffers[k], b.log.buffers[k]
            assert torch.equal(torch.nan_to_num(x, nan=-7.0) if x.is_floating_point() else x,
                               torch.nan_to_num(y, nan=-7.0) if y.is_floating_point() else y), k
    log_e.close()
    log_g.close()


def test_adjoint_log_matches_the_backward(gpu_device):
    """AdjointLog through CheckpointedRollout.backward: slot t holds step t's adjoint,
    equal bit for bit to the same backward run again without the log (deterministic)."""
    from ostrich_amd.differentiable import CheckpointedRollout
    from ostrich_amd.sim_log import AdjointLog
    model, eng, col, s0, ctl = _setup(gpu_device, W=32)
    collide_fn = lambda q: col.collide(q).clone()
    N = 6
    controls = [ctl] * N
    ro = CheckpointedRollout(eng, collide_fn, DT, checkpoint_every=3)
    fin = ro.forward(s0, controls)
    gq = torch.ones_like(fin.body_q) * 0.1
    gu = torch.zeros_like(fin.body_qd)
    log = AdjointLog(N + 2, eng, DT)
    g1 = ro.backward(gq, gu, adjoint_log=log)
    # the same backward again, recording each step's outputs by hand
    seen = {}
    inner = eng.step_backward

    def rec(*a, **kw):
        out = inner(*a, **kw)
        seen[len(seen)] = (out, a[7].clone(), a[8].clone())
        return out
    eng.step_backward = rec
    g2 = ro.backward(gq, gu)
    eng.step_backward = inner
    torch.cuda.synchronize()
    print(f"[adjoint log] {N} steps, slots {N + 2}", flush=True)
    assert torch.equal(g1["body_q"], g2["body_q"])
    buf = log.log.buffers
    for j, (out, gq_in, gu_in) in seen.items():
        t = N - 1 - j  # the backward runs from the last step
        assert torch.equal(buf["body_pose_prev_grad"][t], out["body_q"].reshape(buf["body_pose_prev_grad"][t].shape))
        assert torch.equal(buf["ext_force_grad"][t], out["body_f"].reshape(buf["ext_force_grad"][t].shape))
        assert torch.equal(buf["body_pose_grad"][t], gq_in.reshape(buf["body_pose_grad"][t].shape))
        assert torch.equal(buf["body_vel_grad"][t], gu_in.reshape(buf["body_vel_grad"][t].shape))
    assert torch.count_nonzero(buf["body_pose_grad"][N:]) == 0  # slots past the rollout untouched
