"""The learned per-contact warm start (host side).

`WarmStartNet` owns a gfx950 network (include/ostrich_fb.h, ofb_mlp_*): 12 -> 64 -> 64 -> 3
ReLU in BF16 with FP32 accumulation on MFMA, trained with Adam by the library's own
forward/backward kernels.  The learned warm start has no counterpart in the reference
(its learning/warm_start_net.py:15-48 is a per-world network); the solver uses this one
through `OstrichEngine.set_warm_start_net`.

Features of a contact (all from the start of the step, dimensionless; the solver's
feature kernel computes them on the device):
  0     gap / (|g| dt^2)
  1-3   contact normal
  4-5   relative tangential velocity in the contact basis / (|g| dt)
  6-8   previous-step impulse matched by the warm starter (n, t1, t2) / (m_eff |g|)
  9     mu
  10    dt / 0.05
  11    body-mass ratio m0 / (m0 + m1) (a static side counts as infinitely heavy)
Targets / outputs: (lambda_n, lambda_t1, lambda_t2) / (m_eff |g|).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib

NUM_FEATURES = 12
NUM_OUTPUTS = 3


def he_init(num_params, seed=0):
    """He-normal weights, zero biases and zero padding (the device parameter layout)."""
    IN, INP, H, OUTP = 12, 16, 64, 16
    rng = np.random.default_rng(seed)
    p = np.zeros(num_params, np.float32)
    o = 0
    W1 = np.zeros((INP, H), np.float32)
    W1[:IN] = rng.normal(0, np.sqrt(2 / IN), (IN, H))
    p[o:o + INP * H] = W1.reshape(-1); o += INP * H + H
    p[o:o + H * H] = rng.normal(0, np.sqrt(2 / H), (H, H)).reshape(-1); o += H * H + H
    W3 = np.zeros((H, OUTP), np.float32)
    W3[:, :NUM_OUTPUTS] = rng.normal(0, np.sqrt(1 / H), (H, NUM_OUTPUTS))
    p[o:o + H * OUTP] = W3.reshape(-1)
    return p


class WarmStartNet:
    def __init__(self, device="cuda:0", seed=0, params=None):
        self.lib = _lib.load()
        self.device = torch.device(device)
        h = C.c_void_p()
        _lib.check(self.lib.ofb_mlp_create(C.byref(h)), "ofb_mlp_create")
        self._h = h
        self.num_params = int(self.lib.ofb_mlp_num_params())
        self.set_params(torch.as_tensor(he_init(self.num_params, seed) if params is None else params))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
                self.lib.ofb_mlp_destroy(h)
            except Exception:
                pass
            self._h = None

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _dev(self, t, n=None):
        t = torch.as_tensor(t).to(self.device, torch.float32).contiguous()
        if n is not None and t.numel() != n:
            raise ValueError(f"expected {n} values, got {t.numel()}")
        return t

    def set_params(self, params):
        p = self._dev(params, self.num_params)
        _lib.check(self.lib.ofb_mlp_set_params(self._h, C.c_void_p(p.data_ptr()), self._stream()),
                   "ofb_mlp_set_params")
        self._keep = p

    def params(self):
        out = torch.zeros(self.num_params, device=self.device)
        _lib.check(self.lib.ofb_mlp_get_params(self._h, C.c_void_p(out.data_ptr()), self._stream()),
                   "ofb_mlp_get_params")
        return out

    def forward(self, x):
        x = self._dev(x).reshape(-1, NUM_FEATURES)
        y = torch.empty(x.shape[0], NUM_OUTPUTS, device=self.device)
        _lib.check(self.lib.ofb_mlp_forward(self._h, C.c_void_p(x.data_ptr()), x.shape[0],
                                            C.c_void_p(y.data_ptr()), self._stream()), "ofb_mlp_forward")
        return y

    def gradient(self, x, target):
        x = self._dev(x).reshape(-1, NUM_FEATURES)
        t = self._dev(target).reshape(-1, NUM_OUTPUTS)
        g = torch.empty(self.num_params, device=self.device)
        loss = torch.empty(1, device=self.device)
        _lib.check(self.lib.ofb_mlp_gradient(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                             x.shape[0], C.c_void_p(g.data_ptr()), C.c_void_p(loss.data_ptr()),
                                             self._stream()), "ofb_mlp_gradient")
        return g, loss

    def train_step(self, x, target, lr):
        x = self._dev(x).reshape(-1, NUM_FEATURES)
        t = self._dev(target).reshape(-1, NUM_OUTPUTS)
        loss = torch.empty(1, device=self.device)
        _lib.check(self.lib.ofb_mlp_train_step(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                               x.shape[0], C.c_float(lr), C.c_void_p(loss.data_ptr()),
                                               self._stream()), "ofb_mlp_train_step")
        return loss

    def fit(self, x, target, epochs=10, batch=4096, lr=1e-3, seed=0, log=None):
        """Adam over shuffled mini-batches; returns the per-epoch mean loss."""
        x = self._dev(x).reshape(-1, NUM_FEATURES)
        t = self._dev(target).reshape(-1, NUM_OUTPUTS)
        n = x.shape[0]
        gen = torch.Generator(device=self.device)
        gen.manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen, device=self.device)
            losses = []
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                losses.append(self.train_step(x[idx], t[idx], lr))
            hist.append(torch.cat(losses).mean().item())
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e}")
        return hist
