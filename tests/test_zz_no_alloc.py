"""The allocation-free MFT entry point (csx_mft_noalloc) against csx_mft, through the
C ABI: two calls in a row, the second on a smaller input, in every precision."""
import ctypes as C

import numpy as np
import pytest

from csim_amd import ops
from csim_amd._lib import CsxError, PRECISIONS, check, lib

pytestmark = pytest.mark.gpu


def _rand(rng, n):
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))


def _sym(n, d):
    return (n, -(n * d - d) / 2.0, d)


def _noalloc(ctx, E, gi, go, lfs, prec):
    import torch
    batch = E.shape[0]
    out = torch.empty((batch, go[0], go[0]), dtype=torch.complex128, device=E.device)
    lf = np.ascontiguousarray(lfs, dtype=np.float64)
    check(lib().csx_mft_noalloc(ctx.handle, ops._grid(*gi), ops._grid(*gi), ops._grid(*go), ops._grid(*go),
                                lf.ctypes.data_as(C.c_void_p), batch, ops._ptr(E), ops._ptr(out), PRECISIONS[prec]))
    return out


def _cases(seed):
    rng = np.random.default_rng(seed)
    out = []
    for nin, nout, batch in ((512, 128, 3), (256, 64, 2)):      # the second call smaller
        dx, lf = 1e-5, 650e-9
        gi, go = _sym(nin, dx), _sym(nout, lf / (nin * dx) / 2)
        E = np.stack([_rand(rng, nin) for _ in range(batch)])
        out.append((E, gi, go, [lf * (1 + 0.03 * b) for b in range(batch)]))
    return out


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-6), ("bf16x3", 1e-5), ("bf16x6", 1e-7)])
def test_noalloc_mft_matches_mft(ctx, prec, tol):
    """On the tuned path csx_mft may run another tile variant (or split K), so the two
    agree to the precision's rounding; FP64 to 1e-13 relative RMS."""
    for E, gi, go, lfs in _cases(5):
        Ed = ops.to_device(E)
        ref = ops.to_host(ops.mft(ctx, Ed, gi, gi, go, go, lfs, prec))
        got = ops.to_host(_noalloc(ctx, Ed, gi, go, lfs, prec))
        err = float(np.sqrt(np.mean(np.abs(got - ref) ** 2) / np.mean(np.abs(ref) ** 2)))
        print(f"{prec} {gi[0]}->{go[0]} x{E.shape[0]}: noalloc vs mft rel rms {err:.3e}")
        assert err < tol


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "bf16x6"])
def test_noalloc_mft_bit_exact_on_one_variant(ctx, monkeypatch, prec):
    """With one tile variant for both, the allocation-free call gives the same bits."""
    monkeypatch.setenv("CSX_GEMM_VARIANT", "0")
    for E, gi, go, lfs in _cases(9):
        Ed = ops.to_device(E)
        ref = ops.to_host(ops.mft(ctx, Ed, gi, gi, go, go, lfs, prec))
        got = ops.to_host(_noalloc(ctx, Ed, gi, go, lfs, prec))
        assert np.array_equal(got, ref)


def test_noalloc_mft_refuses_beyond_its_scratch(ctx):
    """A call whose work buffers exceed the scratch array is refused before any GPU
    work, with a message naming the limit."""
    cap = int(lib().csx_mft_noalloc_capacity())
    assert cap == 128 << 20
    n, m, batch = 4096, 1024, 8
    need = ops.mft_workspace_size(_sym(n, 1e-5), _sym(n, 1e-5), _sym(m, 1e-4), _sym(m, 1e-4), batch, "fp64")
    assert need > cap
    E = ops.to_device(np.zeros((1, 8, 8), dtype=np.complex128))
    out = ops.to_device(np.zeros((1, 8, 8), dtype=np.complex128))
    lf = np.full(batch, 650e-9)
    with pytest.raises(CsxError) as ei:
        check(lib().csx_mft_noalloc(ctx.handle, ops._grid(*_sym(n, 1e-5)), ops._grid(*_sym(n, 1e-5)),
                                    ops._grid(*_sym(m, 1e-4)), ops._grid(*_sym(m, 1e-4)),
                                    lf.ctypes.data_as(C.c_void_p), batch, ops._ptr(E), ops._ptr(out),
                                    PRECISIONS["fp64"]))
    assert "scratch" in str(ei.value)
