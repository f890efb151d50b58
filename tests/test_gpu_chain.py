"""GPU element chain (csim scripts) against the oracle chain on the same inputs.

Gates: plain FP64 GPU path <= 1e-12 relative field RMS against the CPU
restatement of csim (Bluestein chirp-z zoomFft); mixed precision <= 1e-11
absolute contrast (dark-hole mean, 3-10 lambda/D) against the FP64 GPU path,
per-pixel maximum reported.
"""
import numpy as np
import pytest

from csim_amd import workloads as W
from csim_amd.chain import Chain
from oracle.chain import OracleChain

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean(np.abs(a - b) ** 2)) / np.sqrt(np.mean(np.abs(b) ** 2)))


def build(ctx, w, precision="fp64"):
    ch = Chain(ctx, w.script)
    W.apply(ch, w)
    ch.set_precision(precision)
    return ch


def test_workload0_fields_and_contrast(ctx):
    w = W.workload0()
    ch = build(ctx, w)
    oc = OracleChain(w.script, w.arrays)
    ch.propagate(calibrating=True)
    e_cal = rel_rms(ch.field(), oc.field(0, 0, calibrating=True))
    ch.propagate(calibrating=False)
    e_full = rel_rms(ch.field(), oc.field(0, 0, calibrating=False))
    print(f"workload0 calibration field {e_cal:.3e}, occulted field {e_full:.3e}")
    assert e_cal < 1e-12 and e_full < 1e-12
    c, rep = ch.contrast()
    c_ref, img_ref, peak_ref = oc.contrast()
    print(f"contrast gpu {c[0]:.12e} oracle {c_ref:.12e} peak {rep['calib_peak']:.6e}")
    assert abs(rep["calib_peak"] - peak_ref) <= 1e-12 * peak_ref
    assert abs(c[0] - c_ref) <= 1e-12 * c_ref
    img = ch.normalized_image()
    assert np.max(np.abs(img - img_ref)) <= 1e-12 * img_ref.max()


@pytest.mark.parametrize("pad", [1.0, 2.0])
def test_zoomfft_propagator_vs_oracle(ctx, pad):
    """zoomFftPropagator pair (dir 1 onto csim's offset FPM grid, an occulting mask
    there, dir -1 back) between pupil and Lyot stop: fields and contrast against the
    oracle chain, FP64."""
    w = W.workload_zoomprop(n=256, pad=pad)
    ch = build(ctx, w)
    oc = OracleChain(w.script, w.arrays)
    ch.propagate(calibrating=False)
    e = rel_rms(ch.field(), oc.field(0, 0))
    c, _ = ch.contrast()
    c_ref, _, _ = oc.contrast()
    print(f"zoomFftPropagator pad {pad}: field {e:.3e}, contrast {c[0]:.10e} vs {c_ref:.10e}")
    assert e < 1e-12
    assert abs(c[0] - c_ref) <= 1e-12 * c_ref


def test_workload2_reduced_fp64(ctx):
    w = W.workload2(n_lambda=3, offsets=(0.0, 2.0, 6.0))
    ch = build(ctx, w)
    oc = OracleChain(w.script, w.arrays, sags=w.meta["sags"])
    ch.propagate(False)
    worst = 0.0
    for s in range(3):
        for li in range(3):
            worst = max(worst, rel_rms(ch.field(0, s, li), oc.field(s, li)))
    print(f"workload2 (3 lambda x 3 sources) worst field rel rms {worst:.3e}")
    assert worst < 1e-12
    c, _ = ch.contrast()
    c_ref, _, _ = oc.contrast()
    assert abs(c[0] - c_ref) <= 1e-12 * c_ref


def test_contrast_curve_vs_oracle(ctx):
    """contrastCurve::make_contrast_curve on a 3-wavelength, 2-source chain (the
    middle wavelength sets lambda/D): sample radii equal, annulus means to 1e-12."""
    w = W.workload2(n_lambda=3, offsets=(0.0, 2.0))
    ch = build(ctx, w)
    pos, avg = ch.contrast_curve()
    oc = OracleChain(w.script, w.arrays, sags=w.meta["sags"])
    pos_ref, avg_ref = oc.contrast_curve()
    assert pos.shape == pos_ref.shape and np.allclose(pos, pos_ref, rtol=1e-15, atol=0)
    assert np.array_equal(np.isnan(avg), np.isnan(avg_ref))
    ok = ~np.isnan(avg_ref)
    err = float(np.max(np.abs(avg[ok] - avg_ref[ok]) / np.abs(avg_ref[ok])))
    print(f"contrast curve: {pos.size} rows, {ok.sum()} non-empty, max rel err {err:.3e}")
    assert err < 1e-12


def _dark_hole_errors(ctx, w, precision):
    ref = build(ctx, w, "fp64")
    c64, _ = ref.contrast()
    img64 = ref.normalized_image()
    mix = build(ctx, w, precision)
    cm, _ = mix.contrast()
    imgm = mix.normalized_image()
    inf = ref.info()
    n, d = inf["out_rows"], inf["out_pixel_size"]
    sel = OracleChain(w.script, w.arrays, sags=w.meta.get("sags")).region_index((n, -(n * d - d) / 2, d))
    return c64[0], abs(cm[0] - c64[0]), float(np.max(np.abs(imgm - img64)[sel]))


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16x6"])
def test_mixed_contrast_gate_on_dark_hole(ctx, precision):
    """1e-11 absolute contrast gate (3-10 lambda/D dark-hole mean) on the 9-wavelength
    occulter chain (mean C ~ 3e-6), every mixed precision; the per-pixel maximum is
    reported."""
    w = W.workload2_dark()
    c, dmean, dmax = _dark_hole_errors(ctx, w, precision)
    print(f"{precision}: dark hole mean C {c:.4e}; |delta mean| {dmean:.3e}; max |delta C| per pixel {dmax:.3e}")
    assert dmean <= 1e-11


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16x6"])
def test_mixed_contrast_gate_on_bench_workload(ctx, precision):
    """The bench's own workload: 9 wavelengths x 3 sources (0, 2, 6 lambda/D), the
    random 217-hex FPM, 27 fields.  The 6 lambda/D source passes outside the FPM and
    its core lies in the 3-10 lambda/D region (mean C ~ 7e-3), so the 1e-11 absolute
    gate asks for a relative error of ~1.4e-9.  bf16x6 (leading parts on a
    fixed-point grid, their products summed exactly) meets it; fp32 and bf16x3 round
    every sum or operand at 2^-24 / 2^-17 and stay at a relative error below 1e-6."""
    w = W.workload2()
    c, dmean, dmax = _dark_hole_errors(ctx, w, precision)
    print(f"{precision}: region mean C {c:.4e}; |delta mean| {dmean:.3e} (relative {dmean / c:.3e}); "
          f"max |delta C| per pixel {dmax:.3e}")
    assert dmean / c < 1e-6
    if precision == "bf16x6":
        assert dmean <= 1e-11


@pytest.mark.parametrize("kind", ["doubleFraunhoferInterpFPM", "fpmPupToLyot"])
def test_candidate_batch_matches_single(ctx, kind):
    """A candidate batch (the pupil-plane fields carried once and expanded by the FPM
    element) gives every candidate's contrast as a chain run with that candidate alone."""
    w = W.workload2(n_lambda=2, offsets=(0.0,)) if kind != "fpmPupToLyot" else W.workload_p2l("hex", offsets=(0.0, 2.0))
    ch = build(ctx, w)
    n_hex = w.meta["n_hex"]
    t = W.hex_thickness(w.meta["seed"], 7, n_hex, n_cand=3)
    sags = W.thickness_to_sag(t)
    ch.set_candidates("fpm", sags)
    cb, _ = ch.contrast()
    for k in range(3):
        single = build(ctx, w)
        single.set_opt_data("fpm", "sags", sags[k])
        cs, _ = single.contrast()
        assert abs(cb[k] - cs[0]) <= 1e-13 * cs[0]


def test_opt_data_roundtrip(ctx):
    w = W.workload2(n_lambda=1, offsets=(0.0,))
    ch = build(ctx, w)
    v = np.linspace(0, 1e-7, w.meta["n_hex"])
    ch.set_opt_data("fpm", "sags", v)
    assert np.array_equal(ch.get_opt_data("fpm", "sags"), v)
    ch.set_opt_data("lyot", "position", [2.5])
    assert ch.get_opt_data("lyot", "position")[0] == 2.5


def test_workload1_fresnel_chain(ctx):
    w = W.workload1()
    ch = build(ctx, w)
    oc = OracleChain(w.script, w.arrays)
    ch.propagate(False)
    e = rel_rms(ch.field(), oc.field(0, 0))
    print(f"workload1 (2048^2 PIAA + Fresnel + downsample + MFT) field rel rms {e:.3e}")
    assert e < 1e-12


@pytest.mark.parametrize("mask", ["hex", "cmc", "binary"])
def test_fpm_pup_to_lyot_vs_oracle(ctx, mask):
    """fpmPupToLyot (2n x 2n FFT convolution, FP64) with each mask kind against the
    oracle's restatement of csim: calibration and occulted fields, contrast."""
    w = W.workload_p2l(mask)
    ch = build(ctx, w)
    arrays = dict(w.arrays)
    for (elem, key, sl), a in w.complex_arrays.items():
        arrays[(elem, key)] = a
    oc = OracleChain(w.script, arrays, sags=w.meta.get("sags"))
    if w.meta.get("sags") is not None:
        ch.set_opt_data("fpm", "sags", w.meta["sags"])
    nl, ns = len(w.meta["lambdas"]), 2
    for cal in (True, False):
        ch.propagate(calibrating=cal)
        worst = 0.0
        for s in range(ns):
            for li in range(nl):
                worst = max(worst, rel_rms(ch.field(0, s, li), oc.field(s, li, calibrating=cal)))
        print(f"fpmPupToLyot {mask} calibrating={cal}: worst field rel rms {worst:.3e}")
        assert worst < 1e-12
    c, rep = ch.contrast()
    c_ref, _, peak_ref = oc.contrast()
    print(f"contrast gpu {c[0]:.12e} oracle {c_ref:.12e}")
    assert abs(rep["calib_peak"] - peak_ref) <= 1e-12 * peak_ref
    assert abs(c[0] - c_ref) <= 1e-11 * abs(c_ref)


def test_graph_replay_matches_eager(ctx):
    """Graph replays of a compiled chain give the eager result on every replay, with
    eager work and synchronisations between replays; sags updated in place reach
    the replayed stages."""
    import torch
    w = W.workload2(n_lambda=3, offsets=(0.0, 2.0))
    eager = build(ctx, w, "bf16x3")
    eager.propagate(False)
    ref = [eager.field(0, s, l) for s in range(2) for l in range(3)]
    g = build(ctx, w, "bf16x3")
    g.set_graph(True)
    for rep in range(4):
        g.propagate(False)          # rep 0 eager, rep 1 captures, later replays
        got = [g.field(0, s, l) for s in range(2) for l in range(3)]
        worst = max(float(np.max(np.abs(a - b))) for a, b in zip(got, ref))
        print(f"replay {rep}: max |diff| {worst:.3e}")
        assert worst == 0.0
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        assert torch.isfinite(x.sum())
    sags2 = w.meta["sags"][::-1].copy()
    g.set_opt_data("fpm", "sags", sags2)
    eager.set_opt_data("fpm", "sags", sags2)
    g.propagate(False)
    eager.propagate(False)
    for s in range(2):
        for l in range(3):
            assert np.array_equal(g.field(0, s, l), eager.field(0, s, l))


def test_linear_optimizer_jacobian_and_step(ctx):
    """linearOptimizer on the batched chain: the finite-difference Jacobian of the
    region fields (one candidate batch) equals the derivative of the response-matrix
    model E(s) = sum_k M_k exp(i 4 pi s_k / lambda) + M_out; the optimiser's steps
    lower the dark-hole contrast from a random start."""
    import torch
    from csim_amd.linear import LinearConfig, LinearOptimizer
    w = W.workload2(n=256, n_lambda=2, offsets=(0.0,), rings=3)
    ch = build(ctx, w)
    nhex = w.meta["n_hex"]
    rng = np.random.default_rng(41)
    s0 = rng.random(nhex) * 0.2e-6
    ch.set_opt_data("fpm", "sags", s0)
    M, _ = ch.hex_response("fpm")                                 # (nslice, nhex + 1, npix)
    lams = np.asarray(w.meta["lambdas"])
    opt = LinearOptimizer(ch, "fpm", LinearConfig(jacobian_dx=1e-12))
    c0, J, e = opt.jacobian(s0)
    t = torch.from_numpy(np.exp(1j * 4 * np.pi * s0[None, :] / lams[:, None])).cuda()       # (nl, nhex)
    dE = M[:, :nhex, :] * (t * (1j * 4 * np.pi / torch.from_numpy(lams).cuda()[:, None]))[:, :, None]
    Jref = torch.cat([dE.real, dE.imag], dim=2).permute(0, 2, 1).reshape(-1, nhex)
    err = float(torch.linalg.norm(J - Jref) / torch.linalg.norm(Jref))
    Eref = (M[:, :nhex, :] * t[:, :, None]).sum(dim=1) + M[:, nhex, :]
    eref = -torch.cat([Eref.real, Eref.imag], dim=1).reshape(-1)
    e_err = float(torch.linalg.norm(e - eref) / torch.linalg.norm(eref))
    print(f"jacobian vs response derivative {err:.2e}, objective vector {e_err:.2e}")
    assert err < 1e-4 and e_err < 1e-11
    scale = float(torch.linalg.matrix_norm(J, ord=2))
    cfg = LinearConfig(jacobian_dx=1e-12, mu_scan=(0.01 * scale, 0.3 * scale, 0.01 * scale),
                       global_bounds=(0.0, 1e-6), max_iterations=3, contrast_rel_tolerance=0.0, stop_contrast=0.0)
    res = LinearOptimizer(ch, "fpm", cfg).run(s0)
    print([(h[0], f"{h[2]:.3e}") for h in res.history])
    assert res.best_value < 0.5 * res.history[0][2]


def _de_batch_chains(ctx, precision, n_cand=32, n_lambda=9, seed=31):
    """The DE-generation batch with uneven pupils: n_cand candidates x n_lambda
    wavelengths on a 512^2 and on a 1024^2 pupil chain (radii halved on the small one)."""
    rng = np.random.default_rng(seed)
    chains, sags = [], []
    for n in (512, 1024):
        w = W.workload3(n=n, n_lambda=n_lambda)
        s = W.thickness_to_sag(rng.random((n_cand, w.meta["n_hex"])) * 1e-6)
        ch = build(ctx, w, precision)
        ch.set_candidates("fpm", s)
        chains.append(ch)
        sags.append((w, s))
    return chains, sags


@pytest.mark.parametrize("precision,tol", [("bf16x3", 2e-5), ("bf16x6", 1e-8), ("fp64", 1e-12)])
def test_chain_group_uneven_pupils_matches_chains(ctx, precision, tol):
    """The full DE generation of uneven pupils: 32 candidates x 9 wavelengths on a 512^2
    and 32 x 9 on a 1024^2 pupil, 576 fields, in one group whose every contraction
    stage is one grouped launch over both sizes.  Its contrasts equal those of the two
    chains run alone, each launch on the tile variant the tuner chose for it (grouped
    and single launches may sum K in another order: the tolerance is the precision's)."""
    from csim_amd.chain import ChainGroup
    chains, _ = _de_batch_chains(ctx, precision)
    alone = [ch.contrast()[0] for ch in chains]
    g = ChainGroup(chains)
    got = g.contrast()
    ref = np.concatenate(alone)
    err = float(np.max(np.abs(got - ref) / ref))
    print(f"{precision}: grouped MFT stages {g.merged_launches()}, {got.size} candidates, "
          f"contrast {ref.min():.3e}..{ref.max():.3e}, max rel diff grouped vs alone {err:.3e}")
    # pupil->FPM, FPM->Lyot (the shared unmasked field, then each candidate's hex window,
    # by Babinet), Lyot->image: two GEMM launches each
    assert g.merged_launches() == 4
    assert err < tol
    g.close()
    for ch in chains:
        ch.close()


def test_de_batch_mixed_gate_against_fp64(ctx):
    """The same 576-field generation in bf16x6 against FP64: every candidate's
    3-10 lambda/D region contrast within the 1e-11 absolute gate."""
    from csim_amd.chain import ChainGroup
    res = {}
    for precision in ("fp64", "bf16x6"):
        chains, _ = _de_batch_chains(ctx, precision)
        g = ChainGroup(chains)
        res[precision] = g.contrast()
        g.close()
        for ch in chains:
            ch.close()
    d = np.abs(res["bf16x6"] - res["fp64"])
    print(f"bf16x6 vs fp64 over {d.size} candidates: contrast {res['fp64'].min():.3e}..{res['fp64'].max():.3e}, "
          f"max |delta C| {d.max():.3e}")
    assert d.max() <= 1e-11


def test_de_batch_candidates_vs_oracle(ctx):
    """One candidate of each pupil size of the DE batch (9 wavelengths) against the
    oracle chain in FP64: fields to 1e-12 relative RMS, contrasts to 1e-12."""
    chains, sags = _de_batch_chains(ctx, "fp64", n_cand=4)
    for ch, (w, s) in zip(chains, sags):
        c, _ = ch.contrast()
        oc = OracleChain(w.script, w.arrays, sags=w.meta["sags"])
        k = 2
        worst = max(rel_rms(ch.field(k, 0, li), oc.field(0, li, cand_sags=s[k])) for li in range(9))
        c_ref, _, _ = oc.contrast(cand_sags=s[k])
        print(f"{w.meta['n']}^2 candidate {k}: worst field rel rms {worst:.3e}, contrast {c[k]:.10e} vs {c_ref:.10e}")
        assert worst < 1e-12
        assert abs(c[k] - c_ref) <= 1e-12 * c_ref
        ch.close()


@pytest.mark.parametrize("kind", ["doubleFraunhoferInterpFPM", "fpmPupToLyot"])
def test_hex_response_matrix_reproduces_propagation(ctx, kind):
    """makeHexCFpmResponse as one batch: the response matrix M and the reference's
    linear objective |M [exp(i 4 pi s / lambda); 1]|^2 (csx_hex_resp_objective) give
    the region contrast of full FP64 propagations of random sag vectors."""
    import ctypes as C
    import torch
    from csim_amd._lib import check, lib
    if kind == "fpmPupToLyot":
        w = W.workload_p2l("hex")
    else:
        w = W.workload2(n=256, n_lambda=2, offsets=(0.0, 2.0))
    ch = build(ctx, w, "fp64")
    M, calib = ch.hex_response("fpm")
    nsl, ncol, npix = M.shape
    nhex = ncol - 1
    rng = np.random.default_rng(41)
    ncand = 5
    sags = W.thickness_to_sag(rng.random((ncand, nhex)) * 1e-6)
    ch.set_candidates("fpm", sags)
    c, rep = ch.contrast()
    assert abs(rep["calib_peak"] - calib) <= 1e-14 * calib
    nsrc = nsl // len(ch.lambdas())
    lam = np.ascontiguousarray(np.tile(ch.lambdas(), nsrc))
    d_X = torch.from_numpy(np.ascontiguousarray(sags)).cuda()
    d_o = torch.zeros(ncand, dtype=torch.float64, device="cuda")
    check(lib().csx_hex_resp_objective(ctx.handle, C.c_void_p(M.data_ptr()), npix, ncol, nsl,
                                       lam.ctypes.data_as(C.c_void_p), 1.0, float(calib),
                                       C.c_void_p(d_X.data_ptr()), nhex, ncand, 0, C.c_void_p(d_o.data_ptr()), 0))
    lin = d_o.cpu().numpy()
    print(kind, "propagated", c, "linear", lin)
    assert np.max(np.abs(lin - c) / c) < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_zero_pattern_skipping_is_exact(ctx, precision, monkeypatch):
    """The pupil and Lyot masks zero whole rows and blocks of the fields the MFTs
    contract; the contractions skip those K slabs and the Lyot-zeroed output blocks.
    On one tile variant the fields equal the dense contractions' exactly, and the
    counted (needed) flops drop."""
    monkeypatch.setenv("CSX_GEMM_VARIANT", "0")
    w = W.workload2(n_lambda=3, offsets=(0.0, 2.0))
    runs = {}
    for skip in ("1", "0"):
        monkeypatch.setenv("CSX_ZERO_SKIP", skip)
        ch = build(ctx, w, precision)
        ch.set_profiling(True)
        ch.propagate(False)
        prof = ch.profile()
        flops = sum(v["flops_total"] for k, v in prof.items() if k.startswith("gemm"))
        fields = [ch.field(0, s, li) for s in range(2) for li in range(3)]
        runs[skip] = (flops, fields)
        ch.close()
    (f_skip, e_skip), (f_dense, e_dense) = runs["1"], runs["0"]
    print(f"{precision}: gemm flops with zero skipping {f_skip:.4e}, dense {f_dense:.4e} "
          f"({f_skip / f_dense:.3f})")
    assert 0.5 < f_skip / f_dense < 0.97
    for a, b in zip(e_skip, e_dense):
        assert np.array_equal(a, b)


def test_point_sources_folded_into_the_first_mft(ctx, monkeypatch):
    """Several point sources before the first MFT: the generation pass writes one unit
    source per wavelength and the MFT applies each source's flux and tilt through its
    DFT kernels (shifted output grids).  Sources with flux != 1 and a y tilt (their
    first contraction no longer shared) against the oracle chain, against the path
    that generates every source, and the mixed path against FP64."""
    import re
    w = W.workload2(n_lambda=3, offsets=(0.0, 2.0))
    script = re.sub(r"(\tpointSource:[^\n]*\n)+",
                    "\tpointSource:1.0, 0.0, 0.0\n\tpointSource:0.7, 3.5, 1.25\n"
                    "\tpointSource:1.3, -2.0, 0.0\n", w.script)
    assert script.count("pointSource:") == 3

    def run(precision, fold):
        monkeypatch.setenv("CSX_SOURCE_FOLD", fold)
        ch = Chain(ctx, script)
        W.apply(ch, w)
        ch.set_precision(precision)
        ch.set_profiling(True)
        ch.propagate(False)
        flops = sum(v["flops_total"] for k, v in ch.profile().items() if k.startswith("gemm"))
        fields = [ch.field(0, s, li) for s in range(3) for li in range(3)]
        c, _ = ch.contrast()
        ch.close()
        return flops, fields, c[0]

    f_fold, e_fold, c_fold = run("fp64", "1")
    f_gen, e_gen, c_gen = run("fp64", "0")
    oc = OracleChain(script, w.arrays, sags=w.meta["sags"])
    worst_gen = max(rel_rms(a, b) for a, b in zip(e_fold, e_gen))
    worst_orc = max(rel_rms(e_fold[3 * s + li], oc.field(s, li)) for s in range(3) for li in range(3))
    c_ref, _, _ = oc.contrast()
    print(f"folded sources: vs generated {worst_gen:.3e}, vs oracle {worst_orc:.3e}, "
          f"contrast {c_fold:.10e} / {c_gen:.10e} / {c_ref:.10e}, gemm flops {f_fold:.4e} vs {f_gen:.4e}")
    assert worst_gen < 1e-13 and worst_orc < 1e-12
    assert abs(c_fold - c_ref) <= 1e-12 * c_ref
    assert f_fold < 0.9 * f_gen            # the two untilted-in-y sources share stage 1
    _, e_mix, _ = run("bf16x3", "1")
    worst_mix = max(rel_rms(a, b) for a, b in zip(e_mix, e_fold))
    print(f"bf16x3 folded vs fp64 {worst_mix:.3e}")
    assert worst_mix < 3e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3", "bf16x6"])
def test_candidate_batch_babinet_split(ctx, precision, monkeypatch):
    """A candidate batch through doubleFraunhoferInterpFPM: the hex mask is 1 outside its
    window, so the return MFT of the shared unmasked FPM field runs once per (source,
    wavelength) and each candidate adds the window's (M - 1) E.  Fields and contrasts
    against the direct per-candidate path (CSX_BABINET=0), and the flops it saves."""
    w = W.workload2(n_lambda=3, offsets=(0.0, 2.0))
    ncand = 5
    sags = W.thickness_to_sag(W.hex_thickness(w.meta["seed"], 11, w.meta["n_hex"], n_cand=ncand))
    res = {}
    for bab in ("1", "0"):
        monkeypatch.setenv("CSX_BABINET", bab)
        ch = build(ctx, w, precision)
        ch.set_candidates("fpm", sags)
        ch.set_profiling(True)
        c, _ = ch.contrast()
        flops = sum(v["flops_total"] for k, v in ch.profile().items() if k.startswith("gemm"))
        fields = [ch.field(k, s, li) for k in range(ncand) for s in range(2) for li in range(3)]
        res[bab] = (np.asarray(c), flops, fields)
        ch.close()
    (c1, f1, e1), (c0, f0, e0) = res["1"], res["0"]
    worst = max(rel_rms(a, b) for a, b in zip(e1, e0))
    dc = float(np.max(np.abs(c1 - c0) / c0))
    print(f"{precision} babinet vs direct: fields {worst:.3e}, contrast rel {dc:.3e}, "
          f"gemm flops {f1:.4e} vs {f0:.4e} ({f1 / f0:.3f})")
    if precision == "fp64":
        assert worst < 1e-12 and dc < 1e-12
    else:
        assert worst < 3e-5 and dc < 1e-5
    assert f1 < 0.75 * f0


@pytest.mark.parametrize("precision,tol", [("bf16x3", (5e-6, 1e-5)), ("bf16x6", (1e-9, 1e-10)),
                                           ("fp64", (1e-13, 1e-13))])
def test_split_k_on_every_stage(ctx, monkeypatch, precision, tol):
    """Split K forced on every contraction of a chain, the FPM -> Lyot stage included
    (its Lyot-zeroed blocks that the next contraction never reads get no sums and no
    stores): fields and contrast equal the unsplit ones to the rounding of the split
    sums (FP32 for two-term planes, FP64 otherwise)."""
    w = W.workload2(n_lambda=3, offsets=(0.0, 2.0))
    monkeypatch.setenv("CSX_GEMM_VARIANT", "0")
    res = []
    for split in (None, "2"):
        if split:
            monkeypatch.setenv("CSX_GEMM_KSPLIT", split)
        ch = build(ctx, w, precision)
        c, _ = ch.contrast()
        res.append((c[0], [ch.field(0, s, li) for s in range(2) for li in range(3)]))
        ch.close()
    monkeypatch.delenv("CSX_GEMM_KSPLIT")
    (c0, f0), (c1, f1) = res
    worst = max(rel_rms(a, b) for a, b in zip(f1, f0))
    print(f"{precision} split K on every stage: fields {worst:.3e}, contrast {c1:.10e} vs {c0:.10e}")
    assert worst < tol[0] and abs(c1 - c0) <= tol[1] * c0


def _with_cube(w):
    """The workload's arrays with the FPM mask cube stacked (rows x cols x slices) for the oracle."""
    arrays = dict(w.arrays)
    sl = sorted(k[2] for k in w.complex_arrays if k[:2] == ("fpm", "mask"))
    if sl:
        arrays[("fpm", "mask")] = np.stack([w.complex_arrays[("fpm", "mask", k)] for k in sl], axis=2)
    return arrays


def _native_vs_oracle(ctx, w, sags=None):
    ch = build(ctx, w)
    oc = OracleChain(w.script, _with_cube(w), sags=sags)
    ch.propagate(False)
    ns, nl = ch.info()["n_sources"], len(w.meta["lambdas"])
    worst = max(rel_rms(ch.field(0, s, li), oc.field(s, li)) for s in range(ns) for li in range(nl))
    c, _ = ch.contrast()
    c_ref, _, _ = oc.contrast()
    ch.close()
    return worst, c[0], c_ref


@pytest.mark.parametrize("n,pad,mask_index", [(256, 0.25, None), (256, 1.0, 1), (1024, 0.25, None)])
def test_native_double_fraunhofer_fpm_vs_oracle(ctx, n, pad, mask_index):
    """doubleFraunhoferFPM on csim's own geometry (no additive keys): lambda*f = 3.52
    lambda, FPM pixel dxprime = 2 (R / (dx N zoom)) lambda 80 on the offset grid of pad N
    pixels, the complex mask cube bound as maskFilename and centred, slice maskIndex or,
    unset, the latch to slice 0 for every wavelength (doubleFraunhoferFPM.cpp:45-97).
    Fields and contrast against the oracle chain, FP64."""
    w = W.workload_native("cube", n=n, pad=pad, n_lambda=3 if n == 256 else 2,
                          offsets=(0.0, 2.0) if n == 256 else (0.0,), mask_index=mask_index)
    worst, c, c_ref = _native_vs_oracle(ctx, w)
    print(f"doubleFraunhoferFPM {n}^2 pad {pad} maskIndex {mask_index}: field {worst:.3e}, "
          f"contrast {c:.10e} vs {c_ref:.10e}")
    assert worst < 1e-12 and abs(c - c_ref) <= 1e-12 * c_ref


@pytest.mark.parametrize("n,pad", [(256, 0.25), (256, 1.0), (1024, 0.25)])
def test_native_interp_fpm_vs_oracle(ctx, n, pad):
    """doubleFraunhoferInterpFPM on the hex mask's own pixel grid: no pixelScale key, so
    complexHexMaskFPM derives it from lambdaRef and fpmFRatio (complexHexMaskFPM.cpp:
    197-203), and the FPM grid is pad N pixels of it with lambda*f = 3.52 lambda
    (doubleFraunhoferInterpFPM.cpp:86-125).  FP64 against the oracle chain."""
    w = W.workload_native("hex", n=n, pad=pad, n_lambda=2, offsets=(0.0, 2.0) if n == 256 else (0.0,))
    worst, c, c_ref = _native_vs_oracle(ctx, w, sags=w.meta["sags"])
    print(f"doubleFraunhoferInterpFPM {n}^2 pad {pad}: field {worst:.3e}, contrast {c:.10e} vs {c_ref:.10e}")
    assert worst < 1e-12 and abs(c - c_ref) <= 1e-12 * c_ref


def test_fresnel_pad_factor_chain_vs_oracle(ctx):
    """fresnelPropagator padFactor:2 (fresnelPropagator.cpp:45-54): the field embedded in
    a 2N zero pad, apRad * padFactor, cropped back; the PIAA chain at 512^2 (1024^2 FFTs)
    against the oracle, FP64."""
    w = W.workload1(n=512, fresnel_pad=2)
    ch = build(ctx, w)
    oc = OracleChain(w.script, w.arrays)
    ch.propagate(False)
    e = rel_rms(ch.field(), oc.field(0, 0))
    w1 = W.workload1(n=512)
    ch1 = build(ctx, w1)
    ch1.propagate(False)
    d = rel_rms(ch1.field(), ch.field())
    print(f"fresnel padFactor 2 chain: field rel rms vs oracle {e:.3e} (differs from unpadded by {d:.3e})")
    assert e < 1e-12 and d > 1e-6
    ch.close()
    ch1.close()


def test_fresnel_chain_off_axis_sources_vs_oracle(ctx):
    """Point sources generated in the Fresnel step's first pass (efield.cpp:179-206: the
    tip/tilt phase at every pixel, then the pupil mask and both mirrors before the first
    transform): an on-axis and two off-axis sources of the PIAA chain at 512^2 against
    the oracle, FP64."""
    w = W.workload1(n=512, sources=[(1.0, 0.0, 0.0), (0.5, 1.5, -2.0), (2.0, -3.25, 0.0)])
    ch = build(ctx, w)
    oc = OracleChain(w.script, w.arrays)
    ch.propagate(False)
    worst = max(rel_rms(ch.field(0, s, 0), oc.field(s, 0)) for s in range(3))
    print(f"fresnel chain, 3 point sources: worst field rel rms {worst:.3e}")
    assert worst < 1e-12
    ch.close()


def test_fresnel_reading_its_field_vs_oracle(ctx):
    """A Fresnel step after a zoomFftPropagator: its first pass reads the field (prefetched
    a column ahead) and applies the occulting mask before it, its last pass applies the
    mirror after it (map loaded before the transform); 512^2 against the oracle, FP64."""
    w = W.workload_fresnel_mid(512)
    ch = build(ctx, w)
    oc = OracleChain(w.script, w.arrays)
    ch.propagate(False)
    e = rel_rms(ch.field(), oc.field(0, 0))
    c, _ = ch.contrast()
    c_ref, _, _ = oc.contrast()
    print(f"Fresnel step reading its field: field rel rms {e:.3e}, contrast {c[0]:.12e} vs {c_ref:.12e}")
    assert e < 1e-12 and abs(c[0] - c_ref) <= 1e-12 * abs(c_ref)
    ch.close()


def test_complex_masks_vs_oracle(ctx):
    """complexMask elements (complexMask.cpp:24-67): centred complex windows in the pupil
    plane (the general pointwise pass), before a Fresnel step (its first pass, which
    reads the field) and after it (its last pass, before a mirror); 512^2 against the
    oracle, FP64, calibration and occulted."""
    w = W.workload_complex_masks(512)
    ch = build(ctx, w)
    arrays = dict(w.arrays)
    for (elem, key, sl), a in w.complex_arrays.items():
        arrays[(elem, key)] = a
    oc = OracleChain(w.script, arrays)
    for cal in (True, False):
        ch.propagate(calibrating=cal)
        e = rel_rms(ch.field(), oc.field(0, 0, calibrating=cal))
        print(f"complexMask chain, calibrating={cal}: field rel rms {e:.3e}")
        assert e < 1e-12
    c, _ = ch.contrast()
    c_ref, _, _ = oc.contrast()
    assert abs(c[0] - c_ref) <= 1e-12 * abs(c_ref)
    ch.close()


def test_c4_1261_hex_fp64_vs_oracle(ctx):
    """The 20-ring standard layout (1261 hexes) of the surrogate's scale-up through the
    chain at a 256^2 pupil, 2 wavelengths: FP64 fields and contrast against the oracle."""
    w = W.workload2(n=256, rings=20, n_lambda=2, offsets=(0.0,))
    assert w.meta["n_hex"] == 1261
    ch = build(ctx, w)
    oc = OracleChain(w.script, w.arrays, sags=w.meta["sags"])
    ch.propagate(False)
    worst = max(rel_rms(ch.field(0, 0, li), oc.field(0, li)) for li in range(2))
    c, _ = ch.contrast()
    c_ref, _, _ = oc.contrast()
    print(f"1261 hexes, 256^2: field {worst:.3e}, contrast {c[0]:.10e} vs {c_ref:.10e}")
    assert worst < 1e-12 and abs(c[0] - c_ref) <= 1e-12 * c_ref
    ch.close()


@pytest.mark.parametrize("precision", ["bf16x3", "bf16x6"])
def test_c4_1261_hex_mixed_vs_fp64(ctx, precision):
    """The 1261-hex layout at a 1024^2 pupil, 9 wavelengths, a batch of 8 candidates:
    mixed precision against FP64 (3-10 lambda/D region contrasts; bf16x6 within the
    1e-11 absolute gate, bf16x3 to 1e-6 relative)."""
    w = W.workload2(n=1024, rings=20, n_lambda=9, offsets=(0.0,))
    sags = W.thickness_to_sag(W.hex_thickness(w.meta["seed"], 21, w.meta["n_hex"], n_cand=8))
    res = {}
    for p in ("fp64", precision):
        ch = build(ctx, w, p)
        ch.set_candidates("fpm", sags)
        res[p], _ = ch.contrast()
        ch.close()
    d = np.abs(res[precision] - res["fp64"])
    print(f"1261 hexes, 1024^2, {precision}: contrast {res['fp64'].min():.3e}..{res['fp64'].max():.3e}, "
          f"max |delta C| {d.max():.3e}, max rel {np.max(d / res['fp64']):.3e}")
    assert np.max(d / res["fp64"]) < 1e-6
    if precision == "bf16x6":
        assert d.max() <= 1e-11
