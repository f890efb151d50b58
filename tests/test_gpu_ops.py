"""GPU parity of the operator API against the FP64 oracle (csim restatement).

Tolerances: FP64 paths <= 1e-12 relative field RMS (the plain-path gate);
mixed paths report their error level (the contrast gate is in the chain tests).
"""
import numpy as np
import pytest

import oracle
from csim_amd import ops

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean(np.abs(a - b) ** 2)) / np.sqrt(np.mean(np.abs(b) ** 2)))


def rand_field(rng, rows, cols):
    return rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))


def sym(n, d):
    return (n, -(n * d - d) / 2.0, d)


@pytest.mark.parametrize("nin,nout", [(16, 16), (64, 32), (96, 40), (256, 64)])
def test_mft_fp64_small_vs_czt(ctx, nin, nout):
    rng = np.random.default_rng(nin + nout)
    E = rand_field(rng, nin, nin)
    dx, lf = 1e-5, 650e-9 * 0.7
    dxo = lf / (nin * dx) / 3.0
    gi, go = sym(nin, dx), sym(nout, dxo)
    ref = oracle.zoomfft(E, (gi[1], gi[1], dx, dx), (nout, go[1], go[1], dxo, dxo), lf)
    out = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf))[0]
    assert rel_rms(out, ref) < 1e-12


def test_mft_fp64_rectangular_vs_direct(ctx):
    rng = np.random.default_rng(7)
    rows, cols, orows, ocols = 80, 48, 36, 20
    E = rand_field(rng, rows, cols)
    lf = -1.3e-6
    gix, giy = (cols, -2.2e-4, 1e-5), (rows, -4e-4, 1e-5)
    gox, goy = (ocols, -1.1e-3, 1.1e-4), (orows, 0.5e-3, -0.7e-4)
    ref = oracle.mft_direct(E, (gix[1], giy[1], gix[2], giy[2]), orows, ocols, (gox[1], goy[1], gox[2], goy[2]), lf)
    out = ops.to_host(ops.mft(ctx, ops.to_device(E), gix, giy, gox, goy, lf))[0]
    assert rel_rms(out, ref) < 1e-12


def test_mft_fp64_1024_to_256_batched(ctx):
    rng = np.random.default_rng(11)
    N, M, B = 1024, 256, 3
    dx = 1e-5
    lams = np.array([600e-9, 630e-9, 660e-9])
    lf = lams * 2 * 5e-3 * 10.0
    Es = [rand_field(rng, N, N) for _ in range(B)]
    gi = sym(N, dx)
    outs = []
    go_list = []
    dxo = 10.0 * 630e-9 / 8
    go = sym(M, dxo)
    out = ops.to_host(ops.mft(ctx, ops.to_device(np.stack(Es)), gi, gi, go, go, lf))
    for b in range(B):
        ref = oracle.zoomfft(Es[b], (gi[1], gi[1], dx, dx), (M, go[1], go[1], dxo, dxo), lf[b])
        e = rel_rms(out[b], ref)
        print(f"field {b}: rel rms vs Bluestein CZT {e:.3e}")
        assert e < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 5e-5), ("bf16x6", 2e-7)])
def test_mft_mixed_error_level(ctx, prec, tol):
    rng = np.random.default_rng(3)
    N, M = 256, 64
    E = rand_field(rng, N, N)
    dx, lf = 1e-5, 650e-9
    dxo = lf / (N * dx) / 2
    gi, go = sym(N, dx), sym(M, dxo)
    ref = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, "fp64"))[0]
    out = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, prec))[0]
    e = rel_rms(out, ref)
    print(f"{prec}: rel rms vs fp64 {e:.3e}")
    assert e < tol


VARIANTS = {"bf16x3": 17, "bf16x6": 10}
MFMA32 = {"bf16x3": {7, 8, 9, 10}, "bf16x6": {4, 5, 6, 9}}   # variants on the 32x32x16 MFMA (gemm.hip)


PLANES_SIZES = {"bf16x3": [(1024, 256, 2), (200, 72, 3), (96, 136, 2)],
                # three-term planes are read in windows of 32 along K: sizes divisible by 32,
                # ragged against the 64- and 128-wide tiles
                "bf16x6": [(1024, 256, 2), (224, 96, 3), (96, 160, 2)]}


@pytest.mark.parametrize("prec,tol", [("bf16x3", 5e-5), ("bf16x6", 2e-7)])
@pytest.mark.parametrize("case", [0, 1, 2])
def test_mft_planes_tile_variants(ctx, monkeypatch, prec, tol, case):
    """Every tile variant of the split-BF16 kernels stays at the split-BF16 error level
    against FP64 on ragged sizes.  Variants on the same MFMA shape as variant 0 take the
    same products in the same order per K slab (equal to 1e-6); the 32x32x16 ones sum
    the k values of a slab in another order inside the MFMA (equal to tol / 10)."""
    nin, nout, batch = PLANES_SIZES[prec][case]
    rng = np.random.default_rng(nin + 7 * nout)
    E = np.stack([rand_field(rng, nin, nin) for _ in range(batch)])
    dx, lf = 1e-5, 650e-9
    dxo = lf / (nin * dx) / 2
    gi, go = sym(nin, dx), sym(nout, dxo)
    ref = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, "fp64"))
    outs = []
    for v in range(VARIANTS[prec]):
        monkeypatch.setenv("CSX_GEMM_VARIANT", str(v))
        outs.append(ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, prec)))
    monkeypatch.delenv("CSX_GEMM_VARIANT")
    for v, o in enumerate(outs):
        for b in range(batch):
            e, d = rel_rms(o[b], ref[b]), rel_rms(o[b], outs[0][b])
            print(f"{prec} {nin}->{nout} variant {v} field {b}: vs fp64 {e:.3e}, vs variant 0 {d:.3e}")
            assert e < tol
            assert d < (tol / 10 if v in MFMA32[prec] else 1e-6)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3", "bf16x6"])
def test_mft_caller_workspace(ctx, monkeypatch, prec):
    """csx_mft with a caller workspace (d_work of csx_mft_workspace_size bytes, every
    buffer carved from it) gives the fields of the call that allocates, bit for bit on
    the same tile variant."""
    import torch
    rng = np.random.default_rng(17)
    nin, nout, batch = 256, 64, 3
    E = np.stack([rand_field(rng, nin, nin) for _ in range(batch)])
    dx, lf = 1e-5, 650e-9
    gi, go = sym(nin, dx), sym(nout, lf / (nin * dx) / 2)
    lfs = [lf, lf * 1.05, lf * 0.95]
    size = ops.mft_workspace_size(gi, gi, go, go, batch, prec)
    work = torch.empty(size, dtype=torch.uint8, device="cuda")
    monkeypatch.setenv("CSX_GEMM_VARIANT", "0")
    ref = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lfs, prec))
    got = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lfs, prec, work=work))
    monkeypatch.delenv("CSX_GEMM_VARIANT")
    print(f"{prec}: workspace {size} bytes, max |diff| {np.max(np.abs(got - ref)):.3e}")
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("ksplit", [2, 4])
@pytest.mark.parametrize("nin,nout,batch", [(1024, 128, 2), (200, 72, 3)])
def test_mft_planes_split_k(ctx, monkeypatch, ksplit, nin, nout, batch):
    """Split-K contractions (FP32 sums per K range, added in order by the finishing
    launch) stay at the split-BF16 error level against FP64 and agree with the
    unsplit ones to a tenth of it (the FP32 sums round differently), for a 64x64, a
    128x128 and a direct-to-LDS tile variant, on ragged K (200 = 3 slabs + 8)."""
    rng = np.random.default_rng(nin + 11 * nout)
    E = np.stack([rand_field(rng, nin, nin) for _ in range(batch)])
    dx, lf = 1e-5, 650e-9
    dxo = lf / (nin * dx) / 2
    gi, go = sym(nin, dx), sym(nout, dxo)
    ref = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, "fp64"))
    monkeypatch.setenv("CSX_GEMM_VARIANT", "0")
    base = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, "bf16x3"))
    monkeypatch.setenv("CSX_GEMM_KSPLIT", str(ksplit))
    for v in (0, 2, 11):
        monkeypatch.setenv("CSX_GEMM_VARIANT", str(v))
        out = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, "bf16x3"))
        for b in range(batch):
            e, d = rel_rms(out[b], ref[b]), rel_rms(out[b], base[b])
            print(f"split {ksplit} {nin}->{nout} variant {v} field {b}: vs fp64 {e:.3e}, vs unsplit {d:.3e}")
            assert e < 5e-5 and d < 5e-6
    monkeypatch.delenv("CSX_GEMM_KSPLIT")
    monkeypatch.delenv("CSX_GEMM_VARIANT")


SPLIT_SIZES = {"fp64": [(1024, 128, 2), (200, 72, 3)], "fp32": [(1024, 128, 2), (200, 72, 3)],
               # three-term planes: sizes divisible by 32
               "bf16x6": [(1024, 128, 2), (224, 96, 3)]}
SPLIT_TOL = {"fp64": (1e-12, 1e-14), "fp32": (2e-6, 2e-6), "bf16x6": (2e-7, 2e-8)}


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x6"])
@pytest.mark.parametrize("ksplit", [3, 16])
@pytest.mark.parametrize("case", [0, 1])
def test_mft_split_k_fp64_workspace(ctx, monkeypatch, prec, ksplit, case):
    """Split-K contractions whose splits leave FP64 sums (FP64, FP32 and three-term
    planes: the finishing launch adds them in order and, for three-term plane outputs,
    puts each window of 32 columns on its fixed-point grid) meet their precision's level
    against the Bluestein restatement (FP64) or the FP64 path, and agree with the
    unsplit launch to a hundredth of it (FP64) or a tenth (three-term planes; FP32, whose
    unsplit sums round in FP32 over the whole K, to its own level), on K split into
    ragged ranges (1024 / 3, 200 / 16 below one slab per split) for the first tile
    variant and a 128-wide one."""
    nin, nout, batch = SPLIT_SIZES[prec][case]
    tol, tol_split = SPLIT_TOL[prec]
    rng = np.random.default_rng(nin + 13 * nout + ksplit)
    E = np.stack([rand_field(rng, nin, nin) for _ in range(batch)])
    dx, lf = 1e-5, 650e-9
    dxo = lf / (nin * dx) / 2
    gi, go = sym(nin, dx), sym(nout, dxo)
    if prec == "fp64":
        ref = np.stack([oracle.zoomfft(E[b], (gi[1], gi[1], dx, dx), (nout, go[1], go[1], dxo, dxo), lf)
                        for b in range(batch)])
    else:
        ref = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, "fp64"))
    variants = {"fp64": (0, 4), "fp32": (0,), "bf16x6": (0, 1)}[prec]
    for v in variants:
        monkeypatch.setenv("CSX_GEMM_VARIANT", str(v))
        monkeypatch.delenv("CSX_GEMM_KSPLIT", raising=False)
        base = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, prec))
        monkeypatch.setenv("CSX_GEMM_KSPLIT", str(ksplit))
        out = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, prec))
        for b in range(batch):
            e, d = rel_rms(out[b], ref[b]), rel_rms(out[b], base[b])
            print(f"{prec} split {ksplit} {nin}->{nout} variant {v} field {b}: vs reference {e:.3e}, "
                  f"vs unsplit {d:.3e}")
            assert e < tol and d < tol_split
    monkeypatch.delenv("CSX_GEMM_KSPLIT")
    monkeypatch.delenv("CSX_GEMM_VARIANT")


@pytest.mark.parametrize("nin,nout,batch", [(1024, 256, 2), (200, 72, 3), (96, 136, 2)])
def test_mft_fp64_tile_variants(ctx, monkeypatch, nin, nout, batch):
    """Every FP64 tile variant (single- or double-buffered, 64 to 128 wide) meets the
    plain-path gate against the Bluestein restatement on ragged sizes."""
    rng = np.random.default_rng(nin + 3 * nout)
    E = np.stack([rand_field(rng, nin, nin) for _ in range(batch)])
    dx, lf = 1e-5, 650e-9
    dxo = lf / (nin * dx) / 2
    gi, go = sym(nin, dx), sym(nout, dxo)
    refs = [oracle.zoomfft(E[b], (gi[1], gi[1], dx, dx), (nout, go[1], go[1], dxo, dxo), lf) for b in range(batch)]
    for v in range(5):
        monkeypatch.setenv("CSX_GEMM_VARIANT", str(v))
        out = ops.to_host(ops.mft(ctx, ops.to_device(E), gi, gi, go, go, lf, "fp64"))
        for b in range(batch):
            e = rel_rms(out[b], refs[b])
            print(f"fp64 {nin}->{nout} variant {v} field {b}: vs CZT {e:.3e}")
            assert e < 1e-12
    monkeypatch.delenv("CSX_GEMM_VARIANT")


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_fft2_vs_numpy(ctx, n):
    # every length the size-specialised kernels cover; batches of 3 (below n = 64 a
    # row-pass block spans several fields)
    rng = np.random.default_rng(n)
    batch = 3 if n <= 1024 else 1
    E = np.stack([rand_field(rng, n, n) for _ in range(batch)])
    d = ops.to_device(E)
    out = ops.to_host(ops.fft2(ctx, d, -1))
    for b in range(batch):
        assert rel_rms(out[b], np.fft.fft2(E[b])) < 1e-14, (n, b)
    back = ops.to_host(ops.fft2(ctx, d, +1))
    for b in range(batch):
        assert rel_rms(back[b], E[b]) < 1e-14, (n, b)


@pytest.mark.parametrize("n,batch", [(16, 3), (128, 2), (512, 2), (2048, 1), (4096, 1)])
def test_fresnel_vs_oracle(ctx, n, batch):
    rng = np.random.default_rng(5)
    E = np.stack([rand_field(rng, n, n) for _ in range(batch)])
    lams = [650e-9, 800e-9, 550e-9][:batch]
    z = 1.0
    ap = n * 10e-6 / 2
    out = ops.to_host(ops.fresnel_as(ctx, ops.to_device(E), lams, ap, z))
    for b in range(batch):
        e = rel_rms(out[b], oracle.fresnel_as(E[b], lams[b], ap, z))
        print(f"fresnel n={n} field {b}: {e:.3e}")
        assert e < 1e-12


def test_mirror_phasor_accuracy(ctx):
    # The phasor kernels' sine and cosine (Cody-Waite reduction by pi/2, Taylor series)
    # against numpy's on phases up to 1.2e7 rad (the angular-spectrum phases of a 1 m
    # step reach that) and near the quadrant boundaries: unit field, E = exp(i a h).
    rng = np.random.default_rng(21)
    n = 256
    lam, sign = 633e-9, -1.0
    a = -sign * 2 * 2 * np.pi / lam
    ph = np.concatenate([rng.uniform(-1.2e7, 1.2e7, n * n // 2),
                         (np.arange(n * n // 2) - n * n // 4) * (np.pi / 4) + rng.uniform(-1e-9, 1e-9, n * n // 2)])
    h = (ph / a).reshape(n, n)
    ph = a * h                                   # the phase the kernel forms, same rounding
    d = ops.to_device(np.ones((n, n), dtype=np.complex128))
    ops.apply_mirror(ctx, d, ops.real_to_device(h), [lam], sign=sign)
    out = ops.to_host(d)[0]
    err = np.max(np.abs(out - (np.cos(ph) + 1j * np.sin(ph))))
    print(f"phasor max abs error vs numpy: {err:.3e}")
    assert err < 3e-16


def test_downsample_vs_oracle(ctx):
    rng = np.random.default_rng(9)
    n, f, dx = 256, 2, 1e-5
    E = rand_field(rng, n, n)
    ref = oracle.downsample(E, f, dx)
    out = ops.to_host(ops.downsample(ctx, ops.to_device(E), f, dx))[0]
    assert rel_rms(out, ref) < 1e-14


def test_mask_mirror_vs_oracle(ctx):
    rng = np.random.default_rng(13)
    n = 128
    E = rand_field(rng, n, n)
    m = rng.random((n, n))
    h = rng.standard_normal((n, n)) * 1e-7
    lam = 633e-9
    ref = oracle.mirror_apply(oracle.mask_apply(E, m), h, lam, sign=-1.0)
    d = ops.to_device(E)
    ops.apply_mask(ctx, d, ops.real_to_device(m))
    ops.apply_mirror(ctx, d, ops.real_to_device(h), [lam], sign=-1.0)
    assert rel_rms(ops.to_host(d)[0], ref) < 1e-14


def test_hex_mask_vs_oracle(ctx):
    import torch
    S, nsp = 48, 8
    nsub, hexnum, nhex = ops.hex_rasterize(1, 4, 1.0, 512, 0.9, 0.02, S, nsp, 1.0 / 40, 0.5)
    rng = np.random.default_rng(17)
    ncand = 2
    sags = rng.random((ncand, nhex)) * 0.5e-6
    lams = np.array([600e-9, 650e-9])
    d_nsub = torch.from_numpy(np.ascontiguousarray(np.transpose(nsub, (2, 1, 0)))).cuda()
    d_hex = torch.from_numpy(np.ascontiguousarray(np.transpose(hexnum, (2, 1, 0)))).cuda()
    d_sags = torch.from_numpy(sags).cuda()
    out = ops.hex_masks(ctx, d_nsub, d_hex, S, nsp, d_sags, nhex, ncand, lams)
    got = ops.to_host(out)
    for c in range(ncand):
        for li, lam in enumerate(lams):
            ref = oracle.hex_mask(nsub, hexnum, nsp, sags[c], lam)
            assert rel_rms(got[c * len(lams) + li], ref) < 1e-14


def test_intensity_region_mean(ctx):
    import torch
    rng = np.random.default_rng(21)
    n, per, groups = 32, 3, 2
    Es = np.stack([rand_field(rng, n, n) for _ in range(per * groups)])
    I = ops.intensity(ctx, ops.to_device(Es), per)
    got = np.transpose(I.cpu().numpy(), (0, 2, 1))
    for g in range(groups):
        ref = np.zeros((n, n), order="F")
        for f in range(per):
            oracle.intensity_accumulate(Es[g * per + f], ref)
        assert np.allclose(got[g], ref, rtol=1e-15, atol=0)
    idx = oracle.region_pixels(n, 1.0, 2.0, radii=(1.0, 5.0), angles=(-180, 180))
    means = ops.region_mean(ctx, I, torch.from_numpy(idx).cuda(), 7.0).cpu().numpy()
    for g in range(groups):
        ref = np.asfortranarray(got[g]).ravel(order="F")[idx].mean() / 7.0
        assert abs(means[g] - ref) <= 1e-14 * abs(ref)


def test_hex_resp_objective_known_answer_and_batch(ctx):
    """The reference's GPU path (its zgemm, optObjective.cpp:46) on MFMA: the shipped
    response matrix and DE optimum give the logged 1.35012e-10."""
    import ctypes as C
    import os
    import torch
    from csim_amd._lib import check, lib
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mono_kat.npz"))
    M = fx["M"]                                           # 636 x 500
    rng = np.random.default_rng(23)
    ncand = 37
    X = np.concatenate([fx["best"][None, :], rng.uniform(-5e-8, 5e-8, (ncand - 1, 499))])   # candidate-major
    d_M = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()    # column-major 636 x 500
    d_X = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    d_o = torch.zeros(ncand, dtype=torch.float64, device="cuda")
    lam = np.array([float(fx["wavelength"])])
    check(lib().csx_hex_resp_objective(ctx.handle, C.c_void_p(d_M.data_ptr()), 636, 500, 1,
                                       lam.ctypes.data_as(C.c_void_p), 1.0, float(fx["calib"]),
                                       C.c_void_p(d_X.data_ptr()), 499, ncand, 0, C.c_void_p(d_o.data_ptr()), 0))
    got = d_o.cpu().numpy()
    ref = oracle.hex_resp_objective(M[:, :, None], lam, 1.0, float(fx["calib"]), X.T)
    assert abs(got[0] - 1.3501189541e-10) < 1e-19
    assert np.max(np.abs(got - ref) / ref) < 1e-12
