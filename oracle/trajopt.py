"""Test infrastructure: numpy restatement of the trajectory Adam step
(optim/trajectory_adam.py:6-53 and TrajectoryAdam.step, :105-131), the checker of the
gfx950 kernel ofb_trajectory_adam_step.  Arithmetic in float32, in the kernel's order.
"""
import numpy as np


def adam_step(params, grad, m, v, t, lr=0.05, beta1=0.9, beta2=0.999, eps=1e-7, clip=100.0, dof_offset=0,
              num_dofs=None):
    """One step on params [steps, worlds, dofs] (float32, updated in place with m and v)."""
    f = np.float32
    nd = params.shape[2] - dof_offset if num_dofs is None else num_dofs
    sl = (slice(None), slice(None), slice(dof_offset, dof_offset + nd))
    g = grad[sl].astype(np.float32)
    if clip > 0:
        g = np.clip(g, f(-clip), f(clip))
    b1, b2 = f(beta1), f(beta2)
    mn = b1 * m[sl] + (f(1) - b1) * g
    vn = b2 * v[sl] + (f(1) - b2) * g * g
    m[sl], v[sl] = mn, vn
    bc1 = f(1) - f(np.power(f(beta1), f(t + 1)))
    bc2 = f(1) - f(np.power(f(beta2), f(t + 1)))
    params[sl] = params[sl] - f(lr) * (mn / bc1) / (np.sqrt(vn / bc2) + f(eps))
    return params
