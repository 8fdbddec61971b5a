"""ofb_mlp_gradient_noalloc (the entry point that allocates nothing) against
ofb_mlp_gradient, through the C ABI: two calls in a row, the second with fewer rows,
give the same gradient and loss bits; more rows than the scratch holds are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(lib, fn, net, x, t, n, grad, loss):
    ptr = lambda a: C.c_void_p(a.data_ptr())
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return getattr(lib, fn)(net, ptr(x), ptr(t), n, ptr(grad), ptr(loss), stream)


def test_noalloc_gradient_matches_allocating_one(gpu_device):
    from ostrich_amd import _lib
    lib = _lib.load()
    lib.ofb_mlp_gradient_noalloc.argtypes = [C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 3
    lib.ofb_mlp_gradient_noalloc.restype = C.c_int
    net = C.c_void_p()
    _lib.check(lib.ofb_mlp_create(C.byref(net)), "ofb_mlp_create")
    try:
        npar = lib.ofb_mlp_num_params()
        rng = np.random.default_rng(7)
        params = torch.as_tensor(rng.normal(scale=0.2, size=npar).astype(np.float32), device=gpu_device)
        _lib.check(lib.ofb_mlp_set_params(net, C.c_void_p(params.data_ptr()),
                                          C.c_void_p(torch.cuda.current_stream().cuda_stream)), "set_params")
        for n in (5000, 700):  # second call smaller than the first
            x = torch.as_tensor(rng.normal(size=(n, 12)).astype(np.float32), device=gpu_device)
            t = torch.as_tensor(rng.normal(size=(n, 3)).astype(np.float32), device=gpu_device)
            g_ref, l_ref = torch.zeros(npar, device=gpu_device), torch.zeros(1, device=gpu_device)
            g_new, l_new = torch.full((npar,), float("nan"), device=gpu_device), torch.zeros(1, device=gpu_device)
            _lib.check(_call(lib, "ofb_mlp_gradient", net, x, t, n, g_ref, l_ref), "ofb_mlp_gradient")
            _lib.check(_call(lib, "ofb_mlp_gradient_noalloc", net, x, t, n, g_new, l_new), "ofb_mlp_gradient_noalloc")
            torch.cuda.synchronize()
            print(f"[no-alloc] n={n}: loss {l_ref.item():.6f} vs {l_new.item():.6f}, max |dg| "
                  f"{(g_ref - g_new).abs().max().item():.3e}", flush=True)
            assert torch.equal(g_ref, g_new)
            assert torch.equal(l_ref, l_new)
            assert g_ref.abs().max() > 0
        # more rows than the scratch holds: refused before any launch
        n = 8193
        x = torch.zeros(n, 12, device=gpu_device)
        t = torch.zeros(n, 3, device=gpu_device)
        g, l = torch.zeros(npar, device=gpu_device), torch.zeros(1, device=gpu_device)
        assert _call(lib, "ofb_mlp_gradient_noalloc", net, x, t, n, g, l) == -2
        assert b"at most 8192 rows" in lib.ofb_last_error()
    finally:
        torch.cuda.synchronize()
        lib.ofb_mlp_destroy(net)
