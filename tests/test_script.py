"""Script driver and tools of libcsx.so on the CPU (no device): the device-free
script check on csim's own example scripts, the responseData directory format
(src/data/responseData.cpp:175-273) against an independent FITS parse and the
reference's shipped response, and the differential-evolution optimiser
(differentialEvolutionOptimizer.cpp:216-330) over Python objectives."""
import glob
import os

import numpy as np
import pytest

from csim_amd import script

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = "/root/reference"


def fits_data(path):
    """(header dict, float64 data in file order) of a BITPIX -64 FITS file, parsed per the
    standard (80-byte cards, 2880-byte blocks, big-endian)."""
    raw = open(path, "rb").read()
    hdr, off = {}, 0
    while True:
        card = raw[off:off + 80].decode("ascii")
        off += 80
        if card[:8].strip() == "END":
            break
        if card[8:10] == "= ":
            hdr[card[:8].strip()] = card[10:].split("/")[0].strip()
    off = (off + 2879) // 2880 * 2880
    dims = [int(hdr[f"NAXIS{i + 1}"]) for i in range(int(hdr["NAXIS"]))]
    n = int(np.prod(dims))
    return hdr, np.frombuffer(raw[off:off + 8 * n], dtype=">f8").astype(np.float64), dims


# --------------------------------------------------------------- parsing ---
# element lists of two example scripts, read off their coronagraph blocks
EXPECTED_ELEMENTS = {
    "donutPiaaHexFPMExample.txt": ["mask", "mirror", "fpmPupToLyot", "mask", "fraunhoferFocal"],
    "zoomFftPIAACMCDoubleFrIntMask.txt": ["mask", "fresnelPropagator", "mirror", "fresnelPropagator", "mirror",
                                          "mask", "doubleFraunhoferInterpFPM", "mask", "fresnelPropagator", "mask",
                                          "mask"],
}
# element types the reference scripts name that this build leaves out (deformable mirror,
# the debugging zoomDft propagator, an element type that no longer exists in csim)
OUT_OF_SCOPE = {"zoomDftPropagator", "deformableMirror", "propagator"}


def reference_scripts():
    if not os.path.isdir(REF):
        return []
    out = []
    for f in sorted(glob.glob(os.path.join(REF, "*.txt")) + glob.glob(os.path.join(REF, "*", "*", "*", "*.txt"))):
        if "start" in open(f).read():
            out.append(f)
    return out


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's example scripts are not present")
def test_reference_example_scripts_parse():
    """Every csim example script parses with no unknown key, its elements built by the
    engine (files named, not read) and only out-of-scope element types unknown.
    test_input.txt names a single file for an amp,ph mask pair, which csim's own parser
    cannot split either: it is the one expected failure."""
    scripts = reference_scripts()
    assert len(scripts) >= 15
    for f in scripts:
        r = script.parse_script(open(f).read())
        name = os.path.basename(f)
        if name == "test_input.txt":
            assert not r["ok"] and "first,second" in r["error"]
            continue
        assert r["ok"], (name, r["error"])
        assert r["unknown_keys"] == [], (name, r["unknown_keys"])
        assert set(r["unknown_elements"]) <= OUT_OF_SCOPE, (name, r["unknown_elements"])
        if name in EXPECTED_ELEMENTS:
            assert [e["type"] for e in r["elements"]] == EXPECTED_ELEMENTS[name]
        if "coronagraph" in r["top_level"]:
            assert r["elements"] or r["unknown_elements"]


def test_parse_reports_structure_unknown_keys_and_files():
    text = """
start
  telescope
  primaryDiameter: 2.4
  primaryFRatio: 10
end
start
  efield
  setSize: 64, 64, 1
  wavelength: 650e-9
  beamRadiusPhysical: 1.2
  pixelScale: 0.04
  colour: blue
end
start
  coronagraph
  start
    mask
    name: pupil
    filename: pupil.fits
    sparkle: 3
  end
  start
    fraunhoferFocal
    samplesPerflD: 4
    FOVflD: 10
  end
end
start
  makeContrastCurve
  filename: cc.txt
end
start
  differentialEvolutionOptimizer
  start
    optHexRespObjective
    responseDirectory: resp
  end
  populationSize: 20
end
"""
    r = script.parse_script(text)
    assert r["ok"], r["error"]
    assert r["top_level"] == ["telescope", "efield", "coronagraph", "makeContrastCurve",
                              "differentialEvolutionOptimizer"]
    assert [e["type"] for e in r["elements"]] == ["mask", "fraunhoferFocal"]
    assert [e["name"] for e in r["elements"]] == ["pupil", "fraunhoferFocal"]
    assert sorted(r["unknown_keys"]) == ["efield:colour", "mask:sparkle"]
    assert r["files"] == ["pupil.fits"]
    de = r["blocks"][4]
    assert de["keys"] == ["populationSize"]
    assert de["blocks"][0]["type"] == "optHexRespObjective" and de["blocks"][0]["keys"] == ["responseDirectory"]


# ---------------------------------------------------------- responseData ---
def test_response_directory_roundtrip_and_layout(tmp_path):
    """csx_response_save / csx_response_load: the files' layout read independently
    (NAXIS1 = response columns, NAXIS2 = pixels, NAXIS3 = wavelengths; each slice
    transposed, as csim's save_cube) and the round trip bit for bit."""
    rng = np.random.default_rng(5)
    nsp, slices, rows, cols = 2, 3, 7, 5
    M = rng.standard_normal((nsp, slices, rows, cols)) + 1j * rng.standard_normal((nsp, slices, rows, cols))
    lam = np.array([6.175e-07, 6.5e-07, 6.825e-07])
    d = str(tmp_path / "resp")
    script.response_save(d, M, 229329.604278, lam)
    assert open(os.path.join(d, "responseParams.txt")).read() == \
        "nRows = 7, nCols = 5, nSlices = 3, nPolarizations = 1, nSources = 2\n"
    pp = open(os.path.join(d, "problemParams.txt")).read().splitlines()
    assert pp[0] == "calibMaxIntensity: 229329.604278"
    assert [float(x) for x in pp[1].split(":")[1].split()] == list(lam)
    for s in range(nsp):
        for part, f in (("re", np.real), ("im", np.imag)):
            hdr, data, dims = fits_data(os.path.join(d, f"responseData_s{s}_p0_{part}.fits"))
            assert dims == [cols, rows, slices]
            assert np.array_equal(data.reshape(slices, rows, cols), f(M[s]))
    M2, calib, lam2, sizes = script.response_load(d)
    assert sizes == {"rows": rows, "cols": cols, "slices": slices, "npol": 1, "nsrc": nsp}
    assert np.array_equal(M2, M) and calib == 229329.604278 and np.array_equal(lam2, lam)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "donutPiaaFullCircleExample")),
                    reason="the reference's shipped response is not present")
def test_response_load_reads_the_reference_response():
    """The reference's mono response directory loads to the matrix, calibration peak and
    wavelength of the golden fixture (written by tests/golden/make_golden.py)."""
    fx = np.load(os.path.join(GOLDEN, "mono_kat.npz"))
    M, calib, lam, sizes = script.response_load(os.path.join(REF, "donutPiaaFullCircleExample/mono_2_6/responseData"))
    assert sizes == {"rows": 636, "cols": 500, "slices": 1, "npol": 1, "nsrc": 1}
    assert np.array_equal(M[0, 0], fx["M"])
    assert calib == float(fx["calib"]) and lam[0] == float(fx["wavelength"])


# -------------------------------------------------------------------- DE ---
DE_BLOCK = """
start
  differentialEvolutionOptimizer
  startStrategy:{start}
  baseStrategy:eitherOr
  endStrategy:deLocalToBest1
  reinitStrategy:eitherOr
  populationSize:{pop}
  initPopulationMean:zero
  initSpread:1.0
  globalBounds:-1, 1
  initBounds:-1, 1
  initType:uniformRandom
  maxIterations:{iters}
  stopObjectiveValue:1e-14
  reinitChangefracThreshold:{thr}
  maxSmallChange:10
  maxRetries:20
  fWeight:0.5
  crossoverProb:{cross}
  outputInterval:10
  outputDirectory:{out}
  saveState:0
end
"""


@pytest.mark.parametrize("start", ["deLocalToBest1", "deRand1", "eitherOr"])
def test_de_converges_and_writes_csim_files(tmp_path, start):
    """The C++ optimiser over a shifted quadratic, with csim's restart rule (a best value
    changing by less than 1e-3 at 10 reports in a row restarts the population around the
    global best, compressed 100x): the optimum is found, and csim's output files are
    written (optBestVal.txt lines "iteration value" decreasing to the best,
    bestOptVector.fits the best vector)."""
    D = 8
    target = np.linspace(-0.6, 0.7, D)

    def f(pop):
        return np.sum((pop - target[:, None]) ** 2, axis=0)

    out = str(tmp_path / "opt")
    block = DE_BLOCK.format(start=start, pop=40, iters=600, cross=1.0, thr=1e-3, out=out)
    block = block.replace("saveState:0", "saveState:0\n  reinitCompression:0.01\n  seed:1")
    best, vec, it = script.de_optimize(block, D, f)
    assert best < 1e-10 and np.allclose(vec, target, atol=1e-4)
    lines = [ln.split() for ln in open(os.path.join(out, "optBestVal.txt")).read().splitlines()]
    vals = [float(v) for _, v in lines]
    assert vals == sorted(vals, reverse=True) and abs(vals[-1] - best) <= 1e-5 * best
    hdr, data, dims = fits_data(os.path.join(out, "bestOptVector.fits"))
    assert dims == [D] and np.array_equal(data, vec)
    assert os.path.getsize(os.path.join(out, "optValHistory.txt")) > 0


def test_de_trial_vectors_stay_inside_the_bounds(tmp_path):
    """With no restart (a zero change threshold), every evaluated population (the initial
    one and the bounced-back trial vectors of enforce_boundaries, :395-434) lies inside
    the global bounds even when the optimum sits on the bound; a restart population is
    not bounded, as in csim."""
    D = 6
    target = np.array([1.0, -1.0, 0.99, -0.5, 0.0, 1.0])
    seen = []

    def f(pop):
        seen.append(pop.copy())
        return np.sum((pop - target[:, None]) ** 2, axis=0)

    block = DE_BLOCK.format(start="deRand1", pop=30, iters=100, cross=0.9, thr=0, out=str(tmp_path / "opt"))
    script.de_optimize(block, D, f)
    assert all(p.shape == (D, 30) for p in seen) and len(seen) == 102
    allp = np.concatenate(seen, axis=1)
    assert allp.min() >= -1.0 and allp.max() <= 1.0
    assert allp.max() > 0.95 and allp.min() < -0.95          # the bounds were reached


def test_de_without_crossover_keeps_the_initial_population(tmp_path):
    """crossoverProb 0: every trial vector is its parent (do_crossover, :360-369), so the
    best value never moves from the initial population's."""
    D = 4
    calls = []

    def f(pop):
        calls.append(pop.copy())
        return np.sum(pop ** 2, axis=0)

    out = str(tmp_path / "opt")
    # initBounds wide: the bounce-back window (mean + initBounds, :395-434) keeps every member
    block = DE_BLOCK.format(start="deRand1", pop=10, iters=5, cross=0.0, thr=1e-4, out=out)
    block = block.replace("initBounds:-1, 1", "initBounds:-10, 10").replace("initSpread:1.0", "initSpread:0.1")
    best, vec, it = script.de_optimize(block, D, f)
    assert np.array_equal(calls[1], calls[0])
    assert best == np.min(np.sum(calls[0] ** 2, axis=0))


def test_de_objective_error_stops_the_run(tmp_path):
    def f(pop):
        raise ValueError("objective failed")

    with pytest.raises(ValueError):
        script.de_optimize(DE_BLOCK.format(start="deRand1", pop=10, iters=5, cross=1.0, thr=1e-4, out=str(tmp_path)), 3, f)
