"""Host side of the step logs: the on-disk layout (data/, model/, dims/ groups and the
num_steps / dt attributes of logging/sim_logger.py:113-218, as keys of an .npz archive)
round-trips through save_log / load_log."""
import numpy as np

from ostrich_amd.sim_log import load_log, save_log


def test_log_file_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    data = dict(body_pose=rng.normal(size=(4, 6, 7)).astype(np.float32),
                iter_count=rng.integers(0, 20, size=(4, 3)).astype(np.int32),
                candidates_res_norm_sq=np.where(rng.random((4, 20, 3)) < 0.5, np.nan, 1.0).astype(np.float32))
    model = dict(joint_type=np.array([4, 1, 1], np.int32), body_mass=np.ones(6, np.float32))
    dims = dict(num_worlds=3, bodies_per_world=2, newton_max_iters=20)
    p = tmp_path / "sim.npz"
    save_log(p, 4, 0.05, data, model, dims)
    out = load_log(p)
    assert out["attrs"]["num_steps"] == 4 and out["attrs"]["dt"] == 0.05
    for k, v in data.items():
        np.testing.assert_array_equal(out["data"][k], v)
        assert out["data"][k].dtype == v.dtype
    for k, v in model.items():
        np.testing.assert_array_equal(out["model"][k], v)
    assert out["dims"] == dims
