"""csim's script driver on the GPU (csx_run_script): a whole script with the tool
blocks csim's main loop runs (src/csim.cpp:54-108), checked against the chain API
and the oracle; the response objective in every precision against the FP64
oracle on the reference's shipped response; the differential-evolution
optimiser over the response objective and over full propagations."""
import os

import numpy as np
import pytest

import oracle
from csim_amd import fits, script
from csim_amd import workloads as W
from csim_amd.chain import Chain

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REGION = "\tregionContrast\n\treferenceLambda:6.3e-07\n\tradii:3, 10\n\tanglesInDegrees:-180, 180\n"


def small_hex_workload():
    return W.workload2(n=256, n_sub_pix=4, n_lambda=2, offsets=(0.0,), rings=3)


def script_with_files(w, d):
    """The workload's script with its arrays and sags in FITS files (csim's way of binding
    them): filename keys for the pupil, aberration and Lyot elements, sagFilename for the
    hex mask."""
    text = w.script
    for (elem, key), a in w.arrays.items():
        p = os.path.join(d, f"{elem}.fits")
        fits.write_fits(p, a)
        text = text.replace(f"\t\tname:{elem}\n", f"\t\tname:{elem}\n\t\t{key}:{p}\n", 1)
    p = os.path.join(d, "sags.fits")
    fits.write_fits(p, np.asarray(w.meta["sags"], dtype=np.float64)[:, None])
    text = text.replace("\t\t\tname:hexmask\n", f"\t\t\tname:hexmask\n\t\t\tsagFilename:{p}\n", 1)
    return text


def test_script_run_tools_match_the_chain_api(ctx, tmp_path):
    """One script through csx_run_script: regionContrast (its file and PSF files),
    makeContrastCurve, makeHexCFpmResponse (a responseData directory) and a short
    differentialEvolutionOptimizer over that response.  The region contrast and curve
    equal the chain API's on the same inputs bound in memory; the saved response equals
    csx_chain_hex_response's matrix; the DE's best vector (bestOptVector.fits) gives its
    logged best value (optBestVal.txt) under the oracle's response objective."""
    w = small_hex_workload()
    d = str(tmp_path)
    text = script_with_files(w, d)
    text += f"start\n\tmakeContrastCurve\n\tfilename:cc.txt\nend\n"
    text += f"start\n{REGION}\tfilename:region.txt\nend\n"
    text += (f"start\n\tmakeHexCFpmResponse\n\tdataToOptimize:fpm, sags\n\toutputDirectory:{d}/resp\n"
             f"\tstart\n{REGION}\tend\nend\n")
    text += (f"start\n\tdifferentialEvolutionOptimizer\n\tstart\n\t\toptHexRespObjective\n"
             f"\t\tresponseDirectory:{d}/resp\n\t\tobjectiveType:mean\n\tend\n"
             "\tstartStrategy:deLocalToBest1\n\tbaseStrategy:eitherOr\n\tendStrategy:deLocalToBest1\n"
             "\treinitStrategy:eitherOr\n\tpopulationSize:32\n\tinitPopulationMean:zero\n\tinitSpread:1.0\n"
             "\treinitCompression:0.1\n\tglobalBounds:-5e-7, 5e-7\n\tinitBounds:-1e-7, 1e-7\n"
             "\tinitType:uniformRandom\n\tmaxIterations:40\n\tfWeight:0.5\n\tcrossoverProb:1.0\n"
             f"\toutputInterval:10\n\toutputDirectory:{d}/opt\n\tsaveState:0\nend\n")
    log = script.run_script(ctx, text, workdir=d)
    print(log)
    # the same chain through the API, arrays bound in memory
    ch = Chain(ctx, w.script)
    W.apply(ch, w)
    c_api, _ = ch.contrast(radii=(3, 10), angles=(-180, 180), reference_lambda=6.3e-7)
    lines = open(os.path.join(d, "region.txt")).read().splitlines()
    assert lines[0] == "-180.000000, 180.000000" and lines[1] == "3.000000, 10.000000"
    assert abs(float(lines[3]) - c_api[0]) <= 1e-6 * c_api[0]          # written %e
    img = fits.read_fits(os.path.join(d, "normalized_PSF.fits"))
    assert img.shape == (128, 128) and np.isclose(img.max(), ch.normalized_image().max(), rtol=1e-14)
    pos, avg = ch.contrast_curve()
    rows = np.loadtxt(os.path.join(d, "cc.txt"))
    assert rows.shape == (len(pos), 2)
    assert np.allclose(rows[:, 0], pos, atol=1e-6) and np.allclose(rows[:, 1], avg, rtol=1e-5)
    M_api, calib_api = ch.hex_response("fpm", radii=(3, 10), reference_lambda=6.3e-7)
    M_api = M_api.cpu().numpy()                                       # (slices, ncol, npix)
    M, calib, lam, sizes = script.response_load(os.path.join(d, "resp"))
    assert sizes["cols"] == w.meta["n_hex"] + 1 and sizes["slices"] == 2 and sizes["nsrc"] == 1
    assert calib == pytest.approx(calib_api, rel=1e-14) and np.allclose(lam, w.meta["lambdas"])
    for s in range(2):
        ref = M_api[s].T                                              # npix x ncol
        assert np.max(np.abs(M[0, s] - ref)) <= 1e-13 * np.max(np.abs(ref))
    best = np.asarray(fits.read_fits(os.path.join(d, "opt", "bestOptVector.fits"))).ravel()
    logged = [float(ln.split()[1]) for ln in open(os.path.join(d, "opt", "optBestVal.txt")).read().splitlines()]
    Mo = np.transpose(M[0], (1, 2, 0))                                # npix x ncol x slices
    v = oracle.hex_resp_objective(Mo, lam, 1.0, calib, best[:, None])[0]
    v0 = oracle.hex_resp_objective(Mo, lam, 1.0, calib, np.zeros((len(best), 1)))[0]
    print(f"DE over the response: zero sags {v0:.4e}, best {v:.4e} (logged {logged[-1]:.4e})")
    assert abs(v - logged[-1]) <= 1e-5 * v                             # optBestVal.txt is %g
    assert v < 0.5 * v0


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5), ("bf16x3", 5e-4), ("bf16x6", 2e-6)])
def test_response_objective_every_precision(ctx, prec, tol):
    """csx_hex_resp_objective (the reference's zgemm, optObjective.cpp:46) in every
    precision on the shipped mono response (636 pixels x 500 columns; K padded to 512 for
    the split-BF16 planes) against the FP64 oracle, the logged optimum among the
    candidates.  The objective is the mean |M x|^2 of fields that largely cancel, so the
    relative error grows with the operands' rounding: measured 1.4e-4 (bf16x3, 16
    significant bits), 7.3e-6 (FP32) and 4.1e-7 (bf16x6); at the optimum (contrast
    1.35e-10) every precision stays within 1e-13 absolute contrast."""
    import ctypes as C
    import torch
    from csim_amd._lib import PRECISIONS, check, lib
    fx = np.load(os.path.join(GOLDEN, "mono_kat.npz"))
    M = fx["M"]
    rng = np.random.default_rng(29)
    X = np.concatenate([fx["best"][None, :], rng.uniform(-5e-8, 5e-8, (99, 499))])
    d_M = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
    d_X = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    d_o = torch.zeros(100, dtype=torch.float64, device="cuda")
    lam = np.array([float(fx["wavelength"])])
    check(lib().csx_hex_resp_objective(ctx.handle, C.c_void_p(d_M.data_ptr()), 636, 500, 1,
                                       lam.ctypes.data_as(C.c_void_p), 1.0, float(fx["calib"]),
                                       C.c_void_p(d_X.data_ptr()), 499, 100, 0, C.c_void_p(d_o.data_ptr()),
                                       PRECISIONS[prec]))
    got = d_o.cpu().numpy()
    ref = oracle.hex_resp_objective(M[:, :, None], lam, 1.0, float(fx["calib"]), X.T)
    err = np.max(np.abs(got - ref) / ref)
    print(f"{prec}: max relative error {err:.3e}, optimum {got[0]:.10e} (abs error {abs(got[0] - ref[0]):.2e})")
    assert err < tol
    assert abs(got[0] - ref[0]) <= 1e-13


def test_de_over_full_propagation(ctx, tmp_path):
    """differentialEvolutionOptimizer with the optChainObjective block: each generation is
    one candidate batch through the chain (every candidate's sags, full propagation, region
    contrast).  The best vector, propagated alone, gives the logged best value, and the
    best is below the zero-sag start."""
    w = small_hex_workload()
    d = str(tmp_path)
    text = script_with_files(w, d)
    text += (f"start\n\tdifferentialEvolutionOptimizer\n\tstart\n\t\toptChainObjective\n\t\tcomponent:fpm\n"
             f"\t\tstart\n{REGION}\t\tend\n\tend\n"
             "\tstartStrategy:deRand1\n\tpopulationSize:16\n\tinitPopulationMean:zero\n\tinitSpread:1.0\n"
             "\tglobalBounds:-3e-7, 3e-7\n\tinitBounds:-1e-7, 1e-7\n\tmaxIterations:6\n\tfWeight:0.5\n"
             f"\toutputInterval:2\n\toutputDirectory:{d}/opt\n\tsaveState:0\nend\n")
    log = script.run_script(ctx, text, workdir=d)
    print(log)
    best = np.asarray(fits.read_fits(os.path.join(d, "opt", "bestOptVector.fits"))).ravel()
    logged = float(open(os.path.join(d, "opt", "optBestVal.txt")).read().splitlines()[-1].split()[1])
    ch = Chain(ctx, w.script)
    W.apply(ch, w)
    ch.set_opt_data("fpm", "sags", best)
    c_best, _ = ch.contrast(radii=(3, 10), angles=(-180, 180), reference_lambda=6.3e-7)
    ch.set_opt_data("fpm", "sags", np.zeros_like(best))
    c_zero, _ = ch.contrast(radii=(3, 10), angles=(-180, 180), reference_lambda=6.3e-7)
    print(f"DE over propagations: zero sags {c_zero[0]:.4e}, best {c_best[0]:.4e} (logged {logged:.4e})")
    assert abs(c_best[0] - logged) <= 1e-5 * logged
    assert c_best[0] < c_zero[0]


def test_linear_optimizer_script_block_matches_the_python_driver(ctx, tmp_path):
    """The linearOptimizer block in C++ (Jacobian batch, normal equations on the device,
    Cholesky solves over the mu scan, the scan's contrasts as one batch) against the
    Python driver of the same algorithm (csim_amd.linear, torch solves) on the same chain
    and start: the same contrast after every step and the same best vector."""
    from csim_amd.linear import LinearConfig, LinearOptimizer
    w = W.workload2(n=256, n_lambda=2, offsets=(0.0,), rings=3)
    nhex = w.meta["n_hex"]
    s0 = np.random.default_rng(41).random(nhex) * 0.2e-6
    w.meta["sags"] = s0
    ch = Chain(ctx, w.script)
    W.apply(ch, w)
    opt = LinearOptimizer(ch, "fpm", LinearConfig(jacobian_dx=1e-12))
    import torch
    _, J, _ = opt.jacobian(s0)
    scale = float(torch.linalg.matrix_norm(J, ord=2))
    mu = (0.01 * scale, 0.3 * scale, 0.01 * scale)
    hi = float(np.float32(1e-6))                      # globalBounds is read through float
    cfg = LinearConfig(jacobian_dx=1e-12, mu_scan=mu, global_bounds=(0.0, hi), max_iterations=3,
                       contrast_rel_tolerance=0.0, stop_contrast=0.0)
    ch.set_opt_data("fpm", "sags", s0)
    ref = LinearOptimizer(ch, "fpm", cfg).run(s0)
    d = str(tmp_path)
    text = script_with_files(w, d)
    text += ("start\n\tlinearOptimizer\n\tdataToOptimize:fpm, sags\n\tjacobianDx:1e-12\n\tglobalBounds:0, 1e-6\n"
             f"\tmaxIterations:3\n\tstopContrast:0\n\tcontrastRelTolerance:0\n\tmuScan:{mu[0]!r}, {mu[1]!r}, {mu[2]!r}\n"
             f"\toutputDirectory:{d}/lin\nend\n")
    log = script.run_script(ctx, text, workdir=d)
    print(log)
    hist = np.loadtxt(os.path.join(d, "lin", "optValHistory.txt"))[:, 1]
    refc = np.array([h[2] for h in ref.history])
    print("C++", hist, "python", refc)
    assert hist.shape == refc.shape
    assert np.allclose(hist, refc, rtol=1e-5)               # optValHistory.txt is %g
    best = np.asarray(fits.read_fits(os.path.join(d, "lin", "bestOptVector.fits"))).ravel()
    assert np.max(np.abs(best - ref.best_vector)) <= 1e-6 * np.max(np.abs(ref.best_vector))
    assert refc.min() < 0.5 * refc[0]
