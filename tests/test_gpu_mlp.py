"""MLP surrogate kernels: forward/backward against the FP64 CPU oracle, and training."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _mlp(ctx, sizes, seed=1):
    from csim_amd._lib import check, lib
    p = C.c_void_p()
    check(lib().csx_mlp_create(ctx.handle, *sizes, seed, C.byref(p)))
    return p


def _params(p, sizes):
    from csim_amd._lib import check, lib
    n = lib().csx_mlp_param_count(p)
    flat = np.zeros(n)
    check(lib().csx_mlp_get_params(p, flat.ctypes.data_as(C.c_void_p)))
    nin, h1, h2, no = sizes
    out, o = [], 0
    for shp in ((nin, h1), (h1,), (h1, h2), (h2,), (h2, no), (no,)):
        k = int(np.prod(shp))
        out.append(flat[o:o + k].reshape(shp))
        o += k
    return flat, out


@pytest.mark.parametrize("nb", [96, 1000])       # 1000: the weight gradients split K (7 ranges, a short last one)
@pytest.mark.parametrize("fp32", [True, False])
def test_mlp_gradients_vs_fp64_oracle(ctx, fp32, nb, monkeypatch):
    import torch
    from csim_amd._lib import check, lib
    monkeypatch.setenv("CSX_MLP_FP32", "1" if fp32 else "0")
    sizes = (218, 512, 512, 32)
    p = _mlp(ctx, sizes, seed=5)
    flat, P = _params(p, sizes)
    rng = np.random.default_rng(2)
    x = rng.random((nb, 218)).astype(np.float32)
    t = (rng.standard_normal((nb, 32)) * 0.5 - 6).astype(np.float32)
    dx = torch.from_numpy(x).cuda()
    dt = torch.from_numpy(t).cuda()
    loss = C.c_double()
    g = np.zeros(flat.size)
    check(lib().csx_mlp_loss_grad(p, C.c_void_p(dx.data_ptr()), C.c_void_p(dt.data_ptr()), nb, C.byref(loss),
                                  g.ctypes.data_as(C.c_void_p)))
    if not fp32:
        # the oracle evaluates the same BF16-rounded weights
        P = [np.asarray(torch.from_numpy(q.astype(np.float32)).to(torch.bfloat16).to(torch.float64)) for q in P]
    lref, gref = oracle.mlp_loss_grad(x.astype(np.float64), t.astype(np.float64), P)
    gflat = np.concatenate([v.ravel() for v in gref])
    rel = np.linalg.norm(g - gflat) / np.linalg.norm(gflat)
    print(f"fp32={fp32} batch {nb}: loss {loss.value:.6e} vs {lref:.6e}; gradient rel err {rel:.3e}")
    tol = 1e-5 if fp32 else 3e-2
    assert abs(loss.value - lref) <= tol * abs(lref)
    assert rel < tol
    lib().csx_mlp_destroy(p)


def test_mlp_training_reduces_loss(ctx):
    import torch
    from csim_amd._lib import check, lib
    sizes = (218, 512, 512, 32)
    p = _mlp(ctx, sizes, seed=9)
    rng = np.random.default_rng(4)
    n = 2048
    x = rng.random((n, 218)).astype(np.float32)
    Wt = rng.standard_normal((218, 32)).astype(np.float32) * 0.1
    t = np.tanh(x @ Wt).astype(np.float32) - 5.0
    dx = torch.from_numpy(x).cuda()
    dt = torch.from_numpy(t).cuda()
    loss = C.c_double()
    first = None
    for it in range(300):
        check(lib().csx_mlp_train_step(p, C.c_void_p(dx.data_ptr()), C.c_void_p(dt.data_ptr()), n, 1e-3, C.byref(loss)))
        if first is None:
            first = loss.value
    print(f"loss {first:.4e} -> {loss.value:.4e}")
    assert loss.value < 0.05 * first
    y = torch.empty((n, 32), dtype=torch.float32, device="cuda")
    check(lib().csx_mlp_forward(p, C.c_void_p(dx.data_ptr()), n, C.c_void_p(y.data_ptr())))
    assert float(((y.cpu().numpy() - t) ** 2).mean()) < 0.05 * first
    lib().csx_mlp_destroy(p)
