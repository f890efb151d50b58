"""Pins of the FP64 oracle (CPU restatement of csim) before it is trusted.

Each check is independent of the oracle's own formulas: library routines
(scipy.signal.czt, numpy.fft, numpy.interp), brute-force sums, closed forms
(delta input, plane wave, Parseval, Babinet), the reference's known-answer
fixture (tests/golden/mono_kat.npz) and counts the reference's files fix.
A planted mistake (flipped sign, dropped factor) must fail the pin.
"""
import math
import os

import numpy as np
import pytest
from scipy.signal import czt as scipy_czt
from scipy.special import j1

import oracle
from oracle.chain import OracleChain

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_rms(a, b):
    return float(np.sqrt(np.mean(np.abs(a - b) ** 2)) / np.sqrt(np.mean(np.abs(b) ** 2)))


def brute_czt(x, k, w, a):
    n = np.arange(len(x))
    q = np.arange(k)
    return (x[None, :] * np.exp(-1j * n[None, :] * (a - w * q[:, None]))).sum(1)


# ----------------------------------------------------------------- czt ---
@pytest.mark.parametrize("m,k", [(5, 3), (37, 23), (64, 64), (100, 31)])
def test_czt_vs_brute_force_and_scipy(m, k):
    rng = np.random.default_rng(m * 100 + k)
    x = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    w, a = -0.0713, 0.42
    o = oracle.czt1d_cols(x, k, w, a)[:, 0]
    assert rel_rms(o, brute_czt(x, k, w, a)) < 1e-13
    # csim's (w, a) is scipy's W = e^{i w}, A = e^{i a} (derivation: out(q) = sum x(n) A^-n W^(nq))
    s = scipy_czt(x, m=k, w=np.exp(1j * w), a=np.exp(1j * a))
    assert rel_rms(o, s) < 1e-12


def test_czt_nfft_rule():
    # nfft = 2^ceil(log2(m+k-1)), src/lib/csim_fft.cpp:464
    assert oracle.czt_nfft(1024, 256) == 2048
    assert oracle.czt_nfft(256, 1024) == 2048
    assert oracle.czt_nfft(1024, 128) == 2048
    assert oracle.czt_nfft(3, 2) == 4
    assert oracle.czt_nfft(5, 4) == 8


# ---------------------------------------------------------- zoomFft/MFT ---
def sym(n, d):
    return -(n * d - d) / 2.0


def test_zoomfft_delta_input_closed_form():
    """A one-pixel delta at (m0, n0) maps to s * exp(-2 pi i (x_n0 u + y_m0 v) / lf)."""
    N, M = 32, 24
    dx, lf = 1e-5, 7.1e-6
    dxo = lf / (N * dx) / 2.5
    E = np.zeros((N, N), dtype=complex)
    m0, n0 = 5, 21
    E[m0, n0] = 1.0
    x0, xo0 = sym(N, dx), sym(M, dxo)
    out = oracle.zoomfft(E, (x0, x0, dx, dx), (M, xo0, xo0, dxo, dxo), lf)
    xi, yi = x0 + n0 * dx, x0 + m0 * dx
    u = xo0 + np.arange(M) * dxo
    closed = (dx * dx / (1j * lf)) * np.exp(-2j * np.pi * (xi * u[None, :] + yi * u[:, None]) / lf)
    assert rel_rms(out, closed) < 1e-13
    # planted mistake: the inverse sign (lf -> -lf) must fail this pin
    wrong = oracle.zoomfft(E, (x0, x0, dx, dx), (M, xo0, xo0, dxo, dxo), -lf)
    assert rel_rms(wrong, closed) > 0.5


def test_zoomfft_equals_gemm_form():
    rng = np.random.default_rng(1)
    N, M = 1024, 256
    E = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    dx, lf = 1e-5, 650e-9 * 0.1
    dxo = lf / (N * dx) / 8
    g_in = (sym(N, dx), sym(N, dx), dx, dx)
    a = oracle.zoomfft(E, g_in, (M, sym(M, dxo), sym(M, dxo), dxo, dxo), lf)
    b = oracle.mft_direct(E, g_in, M, M, (sym(M, dxo), sym(M, dxo), dxo, dxo), lf)
    assert rel_rms(a, b) < 1e-13


def test_mft_airy_pattern():
    """Uniform disk of diameter D -> 2 J1(pi D u / lf) / (pi D u / lf) profile (coarse, sampling-limited)."""
    N = 512
    dx = 1.0 / N * 2.0
    x = sym(N, dx) + np.arange(N) * dx
    r = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2)
    D = 1.0
    E = (r < D / 2).astype(complex)
    lf = 1.0
    M, du = 64, 0.125
    out = oracle.mft_direct(E, (x[0], x[0], dx, dx), M, M, (sym(M, du), sym(M, du), du, du), lf)
    u = sym(M, du) + np.arange(M) * du
    rho = np.sqrt(u[None, :] ** 2 + u[:, None] ** 2)
    z = np.pi * D * rho / lf
    airy = np.where(z > 0, 2 * j1(z) / np.where(z > 0, z, 1), 1.0)
    prof = np.abs(out) / np.abs(out).max()
    assert np.max(np.abs(prof - np.abs(airy))) < 0.01


def test_mft_gaussian_to_gaussian():
    N = 256
    dx = 4.0 / N
    x = sym(N, dx) + np.arange(N) * dx
    s = 0.3
    E = np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / (2 * s * s)).astype(complex)
    M, du = 48, 0.08
    out = oracle.zoomfft(E, (x[0], x[0], dx, dx), (M, sym(M, du), sym(M, du), du, du), 1.0)
    u = sym(M, du) + np.arange(M) * du
    # continuous FT of exp(-r^2/2s^2) is 2 pi s^2 exp(-2 pi^2 s^2 k^2); times 1/(i lf)
    ref = (2 * np.pi * s * s) * np.exp(-2 * np.pi ** 2 * s * s * (u[None, :] ** 2 + u[:, None] ** 2)) / 1j
    assert rel_rms(out, ref) < 1e-10


# ------------------------------------------------------------ fft/shift ---
@pytest.mark.parametrize("shape", [(8, 8), (16, 32), (5, 7), (64, 64)])
def test_fft_and_shift_vs_numpy(shape):
    rng = np.random.default_rng(shape[0] * shape[1])
    E = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert rel_rms(oracle.fft2(E), np.fft.fft2(E)) < 1e-14
    assert rel_rms(oracle.ifft2(E), np.fft.ifft2(E)) < 1e-14
    # csim shifts by floor(n/2): numpy's fftshift (not ifftshift) also for odd sizes
    assert np.array_equal(oracle.fft_shift(E), np.fft.fftshift(E))


# -------------------------------------------------------------- Fresnel ---
def test_fresnel_plane_wave_energy_and_reversibility():
    n, lam, z = 256, 650e-9, 0.7
    ap = n * 10e-6 / 2
    E = np.full((n, n), 1.0 + 0j)
    out = oracle.fresnel_as(E, lam, ap, z)
    assert np.allclose(out, np.exp(1j * 2 * np.pi / lam * z), rtol=0, atol=1e-12)
    rng = np.random.default_rng(3)
    E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    F = oracle.fresnel_as(E, lam, ap, z)
    assert abs(np.sum(np.abs(F) ** 2) / np.sum(np.abs(E) ** 2) - 1) < 1e-13   # |H| = 1 (Parseval)
    back = oracle.fresnel_as(F, lam, ap, -z)
    assert rel_rms(back, E) < 1e-12


def test_fresnel_vs_numpy_restatement():
    n, lam, z = 64, 633e-9, 0.3
    ap = n * 20e-6 / 2
    rng = np.random.default_rng(4)
    E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    p = np.fft.fftfreq(n) * n
    H = p[None, :] ** 2 + p[:, None] ** 2
    ref = np.fft.ifft2(np.fft.fft2(E) * np.exp(1j * 2 * np.pi / lam * z * np.sqrt(1 - 0.25 * (lam / ap) ** 2 * H)))
    assert rel_rms(oracle.fresnel_as(E, lam, ap, z), ref) < 1e-12


# ---------------------------------------------------------- resampling ---
def test_interp1_vs_numpy_with_zero_extrapolation():
    rng = np.random.default_rng(5)
    xg = np.sort(rng.random(50)) * 10
    yg = rng.standard_normal(50)
    xi = np.sort(rng.random(80) * 12 - 1)
    ref = np.interp(xi, xg, yg, left=0.0, right=0.0)
    ref[(xi < xg[0]) | (xi > xg[-1])] = 0.0
    assert np.max(np.abs(oracle.interp1_linear(xg, yg, xi) - ref)) < 1e-13


def test_downsample_factor2_is_box_mean():
    rng = np.random.default_rng(6)
    n = 64
    E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    out = oracle.downsample(E, 2, 1e-5)
    box = 0.25 * (E[0::2, 0::2] + E[1::2, 0::2] + E[0::2, 1::2] + E[1::2, 1::2])
    # weights come from physical coordinates (~1e-4 m): 0.5 to ~1e-14, as in csim
    assert np.max(np.abs(out - box)) < 5e-14


# ------------------------------------------------------------ pointwise ---
def test_mask_and_mirror_closed_forms():
    rng = np.random.default_rng(7)
    E = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    m = rng.random((16, 16))
    assert np.allclose(oracle.mask_apply(E, m), E * (1 + 1j) * m, rtol=1e-15, atol=0)   # mask.cpp (1+i) m
    h = rng.standard_normal((16, 16)) * 1e-7
    lam = 600e-9
    assert np.allclose(oracle.mirror_apply(E, h, lam, 1.0), E * np.exp(-4j * np.pi * h / lam), rtol=1e-14)
    assert np.allclose(oracle.mirror_apply(E, h, lam, -1.0), E * np.exp(4j * np.pi * h / lam), rtol=1e-14)


def test_point_source_tilt_offsets_the_psf():
    """tipX = t * Lx / (2 D) moves the focal-plane peak by t lambda/D (efield.cpp:199-201)."""
    n, pix = 256, 1e-5
    D = 200 * pix
    t = 3.0
    E = oracle.point_source(n, n, pix, 1.0, t * n * pix / (2 * D), 0.0)
    x = sym(n, pix) + np.arange(n) * pix
    r = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2)
    E = E * (r < D / 2)
    lf = 1.0
    du = lf / D / 4   # 4 px per lambda/D
    M = 64
    out = oracle.mft_direct(E, (x[0], x[0], pix, pix), M, M, (sym(M, du), sym(M, du), du, du), lf)
    u = sym(M, du) + np.arange(M) * du
    i, j = np.unravel_index(np.argmax(np.abs(out)), out.shape)
    # the grid samples lambda/D at (k - 31.5)/4: the peak lands within half a pixel
    assert abs(u[j] * D / lf - t) <= 0.125 + 1e-9 and abs(u[i] * D / lf) <= 0.125 + 1e-9


# ------------------------------------------------------------------ hex ---
def test_hex_counts_match_reference_files():
    fx = np.load(os.path.join(GOLDEN, "mono_kat.npz"))
    hx, _, _ = oracle.hex_centers(20, "csim")
    assert hx.size == 499 == fx["best"].size            # bestOptVector.fits length
    assert hx.size + 1 == fx["M"].shape[1]              # + Babinet column (responseParams nCols = 500)
    hx32, _, _ = oracle.hex_centers(32, "csim")
    assert hx32.size == int(fx["wfirst_nhex"]) == 1237  # wfirst_sags.fits length
    for rings, n in ((8, 217), (20, 1261), (0, 1), (1, 7)):
        assert oracle.hex_centers(rings, "standard")[0].size == n


def test_hex_raster_invariants():
    hx, hy, _ = oracle.hex_centers(4, "standard")
    lr = 4 * math.sqrt(3) + 1
    design = oracle.hex_design_array(hx, hy, lr, 512, 0.9, 0.02)
    S, nsp = 40, 8
    nsub, hexnum = oracle.hex_interp_data(design, S, nsp, 1.0 / 30, 0.5, 0.9)
    assert np.all(nsub.sum(axis=2) == nsp * nsp)         # every pixel accounts for nSubPix^2 samples
    assert set(np.unique(hexnum)) <= set(range(-1, hx.size))
    m0 = oracle.hex_mask(nsub, hexnum, nsp, np.zeros(hx.size), 650e-9)
    assert np.allclose(m0, 1.0, rtol=0, atol=1e-15)        # zero sag -> unit transmission
    # a uniform sag s on every hex: pixels fully inside hexes get exp(i 4 pi s / lambda)
    s = 0.1e-6
    m1 = oracle.hex_mask(nsub, hexnum, nsp, np.full(hx.size, s), 650e-9)
    full = (nsub[:, :, 0] == nsp * nsp) & (hexnum[:, :, 0] >= 0)
    assert full.any()
    assert np.allclose(m1[full], np.exp(1j * 4 * np.pi * s / 650e-9), atol=1e-14)
    # phase linear in sag / lambda
    m2 = oracle.hex_mask(nsub, hexnum, nsp, np.full(hx.size, s), 325e-9)
    assert np.allclose(m2[full], m1[full] ** 2, atol=1e-13)


def _interior_pixel_hexes(hx, hy, lr, A, scale, gap, S, nsp, ps, fpm_radius):
    """For every mask pixel whose sub-samples all land on design cells inside one hex's
    inscribed circle, that hex (else -1).  Geometry of complexHexMaskFPM.cpp, stated
    independently of the rasteriser: hex h is centred at design column 0.5 A + hx s
    and row 0.5 A + hy s with s = 0.5 A / lr * fpmScaleFactor (make_hex_array,
    :317-366), its inner radius is R = s (1 - hexGap) sqrt(3) / 2, and a mask pixel
    (row i, column j) samples design row trunc((0.5 + x xc1) A + 0.5) at
    x = (i - S/2 - 0.5 + (0.5 + a) / nSubPix) pixelScale and design column likewise
    from j (make_interpolation_data, :372-469).  A hexagon contains its inscribed
    circle, so these pixels are wholly one hex whatever the hexagon test."""
    s = 0.5 * A / lr * scale
    R = s * (1 - gap) * math.sqrt(3) / 2
    Xc, Yc = 0.5 * A + hx * s, 0.5 * A + hy * s
    xc1 = 0.5 / fpm_radius * scale
    sub = (np.arange(S)[:, None] - S / 2 - 0.5 + (0.5 + np.arange(nsp)[None, :]) / nsp) * ps
    idx = np.trunc((0.5 + sub * xc1) * A + 0.5)
    lo, hi = idx.min(axis=1), idx.max(axis=1)
    out = np.full((S, S), -1)
    for i in range(S):
        for j in range(S):
            if lo[i] < 0 or hi[i] >= A or lo[j] < 0 or hi[j] >= A:
                continue
            yc, xc = 0.5 * (lo[i] + hi[i]), 0.5 * (lo[j] + hi[j])
            h = int(np.argmin((Xc - xc) ** 2 + (Yc - yc) ** 2))
            far = max(math.hypot(c - Xc[h], r - Yc[h]) for r in (lo[i], hi[i]) for c in (lo[j], hi[j]))
            if far < R - 1e-9:
                out[i, j] = h
    return out, R, 1.0 / (xc1 * A)


@pytest.mark.parametrize("layout,rings", [("csim", 5), ("standard", 3)])
def test_hex_raster_geometry_pinned(layout, rings):
    """The oracle's hex rasterisation against the geometry of complexHexMaskFPM.cpp.
    (1) Pixels wholly inside one hex carry that hex's number with all nSubPix^2
    samples: the hex lattice is not symmetric under x <-> y, so a transposed raster
    fails (planted below).  (2) Each hex of the design array holds the area of a
    hexagon of inner radius R = s (1 - hexGap) sqrt(3) / 2 in design cells, which a
    hexStepPix or hexGap off by 10 % fails.  (3) On a design without gaps, each
    interior hex (six neighbours; csim's slot 0 also takes the off-design samples of
    a pixel, which only edge hexes meet) holds that area in samples of the mask."""
    hx, hy, _ = oracle.hex_centers(rings, layout)
    lr = rings if layout == "csim" else rings * math.sqrt(3) + 1
    A, scale, gap, S, nsp, ps, rad = 600, 0.85, 0.03, 56, 6, 1.0 / 24, 1.0
    design = oracle.hex_design_array(hx, hy, lr, A, scale, gap)
    nsub, hexnum = oracle.hex_interp_data(design, S, nsp, ps, rad, scale)
    want, R, cell = _interior_pixel_hexes(hx, hy, lr, A, scale, gap, S, nsp, ps, rad)
    inside = want >= 0
    assert inside.sum() > S * S // 4 and len(np.unique(want[inside])) >= 0.5 * hx.size
    assert np.array_equal(hexnum[..., 0][inside], want[inside])
    assert np.all(nsub[..., 0][inside] == nsp * nsp) and np.all(hexnum[..., 1][inside] == -1)
    assert np.mean(hexnum.transpose(1, 0, 2)[..., 0][inside] != want[inside]) > 0.3
    # (2) design cells per hex (hexes within the design array)
    s = 0.5 * A / lr * scale
    Xc, Yc = 0.5 * A + hx * s, 0.5 * A + hy * s
    ext = 2 * R / math.sqrt(3) + 2
    on_design = (np.abs(Xc - 0.5 * A) + ext < 0.5 * A) & (np.abs(Yc - 0.5 * A) + ext < 0.5 * A)
    cells = np.bincount(design[design >= 0].ravel(), minlength=hx.size)[on_design]
    area = 2 * math.sqrt(3) * R ** 2
    assert on_design.sum() >= 7
    assert np.max(np.abs(cells / area - 1)) < 1.0 / R
    for wrong in (scale * 1.1, scale / 1.1):
        bad = oracle.hex_design_array(hx, hy, lr, A, wrong, gap)
        cb = np.bincount(bad[bad >= 0].ravel(), minlength=hx.size)[on_design]
        assert np.max(np.abs(cb / area - 1)) > 0.1
    # (3) samples per interior hex on a design without gaps
    d0 = oracle.hex_design_array(hx, hy, lr, A, scale, 0.0)
    n0, h0 = oracle.hex_interp_data(d0, S, nsp, ps, rad, scale)
    counts = np.zeros(hx.size)
    for k in range(3):
        ok = h0[..., k] >= 0
        np.add.at(counts, h0[..., k][ok], n0[..., k][ok])
    R0 = s * math.sqrt(3) / 2
    dist = np.hypot(hx[:, None] - hx[None, :], hy[:, None] - hy[None, :])
    interior = (np.sum(np.abs(dist - math.sqrt(3)) < 1e-9, axis=1) == 6)
    half = min(S / 2 * ps / cell, 0.5 * A)
    interior &= (np.abs(Xc - 0.5 * A) + ext < half) & (np.abs(Yc - 0.5 * A) + ext < half)
    assert interior.sum() >= 1
    expect = 2 * math.sqrt(3) * R0 ** 2 * (cell / (ps / nsp)) ** 2
    rel = np.abs(counts[interior] / expect - 1)
    print(f"{layout}: {inside.sum()} interior pixels; design area error {np.max(np.abs(cells / area - 1)):.4f} "
          f"over {on_design.sum()} hexes; sample area error {rel.max():.4f} over {interior.sum()} interior hexes")
    assert rel.max() < 2.0 / R0


# --------------------------------------------------------- known answer ---
def test_linear_objective_known_answer():
    """optHexRespObjective on csim's shipped response matrix and DE optimum reproduces
    the logged best value 1.35012e-10 (optBestVal.txt:24938)."""
    fx = np.load(os.path.join(GOLDEN, "mono_kat.npz"))
    M = fx["M"][:, :, None]
    X = fx["best"].reshape(-1, 1)
    obj = oracle.hex_resp_objective(M, [float(fx["wavelength"])], 1.0, float(fx["calib"]), X)
    assert abs(obj[0] - float(fx["logged_value"])) <= 0.5e-15 * 1e10 * 1e-10 + 5e-16    # logged to 6 digits
    assert abs(obj[0] - 1.3501189541e-10) < 1e-19
    zero = oracle.hex_resp_objective(M, [float(fx["wavelength"])], 1.0, float(fx["calib"]), np.zeros((499, 1)))
    assert abs(zero[0] - 2.590e-4) < 5e-7
    # planted mistake: dropping the factor 2 in the phase (2 pi instead of 4 pi) breaks the pin
    wrong = np.mean(np.abs(fx["M"] @ np.exp(2j * np.pi * np.concatenate([fx["best"], [0.0]]) / 6.5e-7)) ** 2) / float(fx["calib"])
    assert abs(wrong - float(fx["logged_value"])) > 1e-12


# ---------------------------------------------------------- Babinet FPM ---
def test_fpm_pup_to_lyot_babinet_identity_and_numpy():
    rng = np.random.default_rng(8)
    n = 32
    E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    S = 16
    ones_minus_1 = np.zeros((S, S), dtype=complex)
    out = oracle.fpm_pup_to_lyot(E, 1e-5, ones_minus_1, 2e-6, 1e-3, 1.0, 1.0, True)
    assert np.allclose(out, E, rtol=0, atol=1e-15)          # mask == 1 -> field unchanged
    mk = rng.standard_normal((S, S)) * 0.1 + 1j * rng.standard_normal((S, S)) * 0.1
    dx, mdx, lf = 1e-5, 2e-6, 1e-3
    out = oracle.fpm_pup_to_lyot(E, dx, mk, mdx, lf, -1.0, -1.0, True)
    # numpy restatement of fpmPupToLyot.cpp:24-167
    n2 = 2 * n
    xm = sym(S, mdx) + np.arange(S) * mdx
    xp = sym(n2, dx) + np.arange(n2) * dx
    A = np.exp(-2j * np.pi * xp[:, None] * xm[None, :] / lf)
    hat = -1.0 * (mdx * mdx / (1j * lf)) * (A @ mk @ A.T)
    F = np.fft.fft2(np.fft.fftshift(hat))
    P = np.zeros((n2, n2), dtype=complex)
    P[n // 2:n // 2 + n, n // 2:n // 2 + n] = E
    res = np.fft.ifft2(F * np.fft.fft2(P)) * dx * dx / (-1j * -1.0 * lf) + P
    assert rel_rms(out, res[n // 2:n // 2 + n, n // 2:n // 2 + n]) < 1e-12


def test_fpm_pup_to_lyot_equals_double_mft():
    """The 2n-padded circular convolution never wraps on the centre n x n, so it equals
    a linear convolution with maskHat sampled half a pixel off the pupil grid:
      E' = E + s z c  B (K . (A E A^T)) B^T,  A = exp(+2 pi i u x / lf),
      B = exp(-2 pi i u (x + dx/2) / lf),  c = du^2 / (i lf),  s = dx^2 / (-i fs lf).
    A wrong pad offset, a missing fft_shift or a flipped sign fails it (the
    correction term, not the whole field, is compared)."""
    rng = np.random.default_rng(12)
    n, S = 32, 12
    dx, du, lf = 1.0, 0.5, 5.0          # a few cycles across both grids, O(1) correction
    E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    K = rng.standard_normal((S, S)) * 0.2 + 1j * rng.standard_normal((S, S)) * 0.2
    x = sym(n, dx) + np.arange(n) * dx
    u = sym(S, du) + np.arange(S) * du
    A = np.exp(2j * np.pi * u[:, None] * x[None, :] / lf)
    B = np.exp(-2j * np.pi * (x[:, None] + dx / 2) * u[None, :] / lf)
    for zs, fs, bab in ((-1.0, -1.0, True), (1.0, 1.0, False)):
        out = oracle.fpm_pup_to_lyot(E, dx, K, du, lf, zs, fs, bab)
        corr = dx * dx / (-1j * fs * lf) * zs * (du * du / (1j * lf)) * (B @ (K * (A @ E @ A.T)) @ B.T)
        base = E if bab else 0.0
        assert np.sqrt(np.mean(np.abs(corr) ** 2)) > 0.1 * np.sqrt(np.mean(np.abs(E) ** 2))
        assert rel_rms(out - base, corr) < 1e-12
    # planted mistake: the kernel sampled on the pupil grid itself (no half-pixel shift)
    B0 = np.exp(-2j * np.pi * x[:, None] * u[None, :] / lf)
    bad = dx * dx / (1j * lf) * -1.0 * (du * du / (1j * lf)) * (B0 @ (K * (A @ E @ A.T)) @ B0.T)
    assert rel_rms(oracle.fpm_pup_to_lyot(E, dx, K, du, lf, -1.0, -1.0, True) - E, bad) > 1e-3


# -------------------------------------------------------------- region ---
def test_region_pixels_annulus():
    n, d, loD = 64, 0.5, 2.0
    idx = oracle.region_pixels(n, d, loD, radii=(3, 10), angles=(-180, 180))
    x = sym(n, d) + np.arange(n) * d
    XX, YY = np.meshgrid(x, x)
    RR = np.sqrt(XX ** 2 + YY ** 2) / loD
    sel = (RR >= 3) & (RR <= 10)
    assert set(idx.tolist()) == set(np.flatnonzero(sel.T.ravel()).tolist())   # column-major indices
    half = oracle.region_pixels(n, d, loD, radii=(3, 10), angles=(0, 180))
    assert 0 < half.size < idx.size


def test_annulus_curve_pins():
    """csim's contrast-curve averaging (strict annuli of two sample spacings around each
    sample radius): a constant image gives the constant; a one-pixel image is seen only
    by the rows whose centre lies within one spacing of its radius; a disk image gives
    exactly 1 inside and 0 outside; rows start at the third sample."""
    from oracle.chain import annulus_curve
    n, d, loD = 128, 0.37, 1.3
    x = sym(n, d) + np.arange(n) * d
    m = n - 2
    smp = np.linspace(0.0, x.max() / loD, m)
    step = smp[1]
    pos, avg = annulus_curve(np.full((n, n), 2.5), x, loD)
    assert pos.size == m - 3 and pos[0] == smp[2] and pos[-1] == smp[m - 2]
    assert np.all(np.abs(avg - 2.5) <= 1e-15 * 2.5)
    I = np.zeros((n, n))
    I[90, 101] = 1.0                                   # row y index 90, column x index 101
    r = np.hypot(x[101], x[90]) / loD
    pos, avg = annulus_curve(I, x, loD)
    hit = avg > 0
    assert np.array_equal(hit, np.abs(pos - r) < step)
    assert 1 <= hit.sum() <= 2
    R0 = 12.0
    RR = np.hypot(x[None, :], x[:, None]) / loD
    pos, avg = annulus_curve((RR < R0).astype(float), x, loD)
    inside = pos + step <= R0
    outside = pos - step >= R0
    assert inside.sum() > 5 and outside.sum() > 5
    assert np.all(avg[inside] == 1.0) and np.all(avg[outside] == 0.0)


# ------------------------------------------------------------------ MLP ---
def test_mlp_gradients_vs_finite_differences():
    rng = np.random.default_rng(9)
    nin, h1, h2, nout, nb = 7, 6, 5, 3, 4
    P = [rng.standard_normal((nin, h1)) * 0.5, rng.standard_normal(h1) * 0.1, rng.standard_normal((h1, h2)) * 0.5,
         rng.standard_normal(h2) * 0.1, rng.standard_normal((h2, nout)) * 0.5, rng.standard_normal(nout) * 0.1]
    x = rng.standard_normal((nb, nin))
    t = rng.standard_normal((nb, nout))
    loss, g = oracle.mlp_loss_grad(x, t, P)
    eps = 1e-6
    for pi, p in enumerate(P):
        flat = p.reshape(-1)
        for k in range(0, flat.size, max(1, flat.size // 5)):
            old = flat[k]
            flat[k] = old + eps
            lp, _ = oracle.mlp_loss_grad(x, t, P)
            flat[k] = old - eps
            lm, _ = oracle.mlp_loss_grad(x, t, P)
            flat[k] = old
            assert abs((lp - lm) / (2 * eps) - g[pi].reshape(-1)[k]) < 1e-7


# --------------------------------------------------------- oracle chain ---
def test_oracle_chain_small_workload_runs_and_normalises():
    from csim_amd import workloads as W
    w = W.workload0(n=128, r_ap=62, r_gauss=50, r_lyot=56)
    oc = OracleChain(w.script, w.arrays)
    c, img, peak = oc.contrast()
    assert 0 < c < 1e-3 and peak > 0
    cal = oc.intensity(True) / peak
    assert abs(cal.max() - 1.0) < 1e-15


@pytest.mark.parametrize("rows,cols,mr,mc", [(8, 8, 3, 3), (9, 8, 4, 5), (7, 10, 2, 7), (6, 5, 6, 5)])
def test_complex_mask_centring(rows, cols, mr, mc):
    """The centred window of complexMask and of the FPM mask cubes (complexMask.cpp:33-43,
    doubleFraunhoferFPM.cpp:84-97): cube element (r, c) multiplies field element
    (rows/2 - mr/2 + r, cols/2 - mc/2 + c) with integer halves, nothing else changes.
    Non-square shapes with odd and even sizes: a transposed or rounded-up offset fails."""
    rng = np.random.default_rng(rows * 100 + cols)
    E = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
    K = np.ones((mr, mc), dtype=np.complex128)
    r, c = mr - 1, mc // 2
    K[r, c] = 2.0 - 0.5j
    out = oracle.center_mask_multiply(E, K)
    expected = E.copy()
    expected[rows // 2 - mr // 2 + r, cols // 2 - mc // 2 + c] *= 2.0 - 0.5j
    assert np.array_equal(out, expected)

