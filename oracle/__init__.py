"""FP64 CPU oracle for csim's E-field propagation path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package: it is the parity
checker and the timed CPU baseline, never the thing measured or shipped.

Arrays follow csim's mathematical convention: a field ``E`` of ``rows x cols``
is a numpy array of shape ``(rows, cols)`` (row index = y, column index = x).
The C library stores them column-major exactly like Armadillo; the wrappers
convert with ``np.asfortranarray``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_d = C.c_double
_i = C.c_int
_p = C.c_void_p


def _declare(L):
    sig = {
        "orc_fft1d": (None, [_p, _i, _i]),
        "orc_fft2d": (None, [_p, _i, _i, _i]),
        "orc_fft2_csim": (None, [_p, _i, _i]),
        "orc_ifft2_csim": (None, [_p, _i, _i]),
        "orc_fft_shift": (None, [_p, _p, _i, _i]),
        "orc_czt_nfft": (_i, [_i, _i]),
        "orc_czt1d_cols": (None, [_p, _i, _i, _i, _d, _d, _p]),
        "orc_czt2d": (None, [_p, _i, _i, _i, _d, _d, _d, _d, _p]),
        "orc_zoomfft": (None, [_p, _i, _i, _d, _d, _d, _d, _i, _d, _d, _d, _d, _d, _p]),
        "orc_mft_direct": (None, [_p, _i, _i, _d, _d, _d, _d, _i, _i, _d, _d, _d, _d, _d, _p]),
        "orc_fresnel_as": (None, [_p, _i, _d, _d, _d]),
        "orc_grid_symmetric": (None, [_i, _d, _d, _p]),
        "orc_grid_offset": (None, [_i, _d, _p]),
        "orc_interp1_linear": (None, [_p, _p, _i, _p, _p, _i, _d]),
        "orc_interp2": (None, [_p, _i, _i, _p, _i, _i, _p, _p, _p, _p]),
        "orc_downsample": (_i, [_p, _i, _i, _i, _d, _d, _d, _d, _p]),
        "orc_mask_apply": (None, [_p, _p, _i]),
        "orc_mirror_apply": (None, [_p, _p, _i, _d, _d]),
        "orc_point_source": (None, [_p, _i, _i, _d, _d, _d, _d]),
        "orc_hex_centers_csim": (_i, [_i, _d, _p, _p, _p, _i]),
        "orc_hex_centers_standard": (_i, [_i, _d, _p, _p, _p, _i]),
        "orc_hex_design_array": (None, [_p, _p, _i, _d, _i, _d, _d, _p]),
        "orc_hex_interp_data": (None, [_p, _i, _i, _i, _d, _d, _d, _p, _p]),
        "orc_hex_mask": (None, [_p, _p, _i, _i, _p, _i, _d, _d, _i, _p]),
        "orc_center_mask_multiply": (None, [_p, _i, _i, _p, _i, _i]),
        "orc_fpm_pup_to_lyot": (None, [_p, _i, _d, _p, _i, _d, _d, _d, _d, _i]),
        "orc_intensity_accumulate": (None, [_p, _i, _p]),
        "orc_region_pixels": (_i, [_i, _d, _d, _d, _d, _d, _d, _d, _d, _p]),
        "orc_hex_resp_objective": (None, [_p, _i, _i, _i, _p, _d, _d, _p, _i, _i, _i, _p]),
        "orc_mlp_forward": (None, [_p, _i, _i, _p, _p, _i, _p, _p, _i, _p, _p, _i,
                                   _p, _p, _p, _p, _p]),
        "orc_mlp_loss_grad": (_d, [_p, _p, _i, _i, _p, _p, _i, _p, _p, _i, _p, _p, _i,
                                   _p, _p, _p, _p, _p, _p]),
        "orc_num_threads": (_i, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _cf(a, dtype=np.complex128):
    """Column-major contiguous copy (csim memory order)."""
    return np.array(a, dtype=dtype, order="F", copy=True)


def num_threads() -> int:
    return lib().orc_num_threads()


# ------------------------------------------------------------------ FFT ---
def fft2(E):
    a = _cf(E)
    lib().orc_fft2_csim(_ptr(a), a.shape[0], a.shape[1])
    return a


def ifft2(E):
    a = _cf(E)
    lib().orc_ifft2_csim(_ptr(a), a.shape[0], a.shape[1])
    return a


def fft1d(x, sign=-1):
    a = np.array(x, dtype=np.complex128, copy=True)
    lib().orc_fft1d(_ptr(a), a.size, sign)
    return a


def fft_shift(E):
    a = _cf(E)
    out = np.empty_like(a, order="F")
    lib().orc_fft_shift(_ptr(a), _ptr(out), a.shape[0], a.shape[1])
    return out


# ------------------------------------------------------------- chirp-z ---
def czt_nfft(m, k):
    return lib().orc_czt_nfft(m, k)


def czt1d_cols(x, k, w, a):
    x = _cf(x if np.ndim(x) == 2 else np.asarray(x).reshape(-1, 1))
    out = np.empty((k, x.shape[1]), dtype=np.complex128, order="F")
    lib().orc_czt1d_cols(_ptr(x), x.shape[0], x.shape[1], k, w, a, _ptr(out))
    return out


def czt2d(E, k, wx, ax, wy, ay):
    a = _cf(E)
    out = np.empty((k, k), dtype=np.complex128, order="F")
    lib().orc_czt2d(_ptr(a), a.shape[0], a.shape[1], k, wx, ax, wy, ay, _ptr(out))
    return out


def zoomfft(E, in_geom, out_geom, lambda_f):
    """csim zoomFft.  in_geom = (x0, y0, dx, dy); out_geom = (n, x0, y0, dx, dy)."""
    a = _cf(E)
    n = int(out_geom[0])
    out = np.empty((n, n), dtype=np.complex128, order="F")
    lib().orc_zoomfft(_ptr(a), a.shape[0], a.shape[1], *map(float, in_geom), n,
                      *map(float, out_geom[1:]), float(lambda_f), _ptr(out))
    return out


def mft_direct(E, in_geom, out_rows, out_cols, out_geom, lambda_f):
    """Closed-form MFT.  in_geom = (x0, y0, dx, dy); out_geom = (x0, y0, dx, dy)."""
    a = _cf(E)
    out = np.empty((out_rows, out_cols), dtype=np.complex128, order="F")
    lib().orc_mft_direct(_ptr(a), a.shape[0], a.shape[1], *map(float, in_geom),
                         out_rows, out_cols, *map(float, out_geom), float(lambda_f), _ptr(out))
    return out


# ------------------------------------------------------------- Fresnel ---
def fresnel_as(E, lam, ap_rad, z):
    a = _cf(E)
    assert a.shape[0] == a.shape[1]
    lib().orc_fresnel_as(_ptr(a), a.shape[0], lam, ap_rad, z)
    return a


# ------------------------------------------------------------ geometry ---
def grid_symmetric(n, d, origin=0.0):
    x = np.empty(n)
    lib().orc_grid_symmetric(n, d, origin, _ptr(x))
    return x


def grid_offset(n, d):
    x = np.empty(n)
    lib().orc_grid_offset(n, d, _ptr(x))
    return x


def interp1_linear(xg, yg, xi, extrap=0.0):
    xg = np.ascontiguousarray(xg, dtype=np.float64)
    yg = np.ascontiguousarray(yg, dtype=np.float64)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    yi = np.empty_like(xi)
    lib().orc_interp1_linear(_ptr(xg), _ptr(yg), xg.size, _ptr(xi), _ptr(yi), xi.size, extrap)
    return yi


def interp2(M, in_x, in_y, out_x, out_y):
    a = _cf(M, np.float64)
    out = np.empty((len(out_y), len(out_x)), order="F")
    ix, iy, ox, oy = (np.ascontiguousarray(v, dtype=np.float64) for v in (in_x, in_y, out_x, out_y))
    lib().orc_interp2(_ptr(a), a.shape[0], a.shape[1], _ptr(out), out.shape[0], out.shape[1],
                      _ptr(ix), _ptr(iy), _ptr(ox), _ptr(oy))
    return out


def downsample(E, factor, dx, dy=None, ox=0.0, oy=0.0):
    a = _cf(E)
    out = np.empty((a.shape[0] // factor, a.shape[1] // factor), dtype=np.complex128, order="F")
    lib().orc_downsample(_ptr(a), a.shape[0], a.shape[1], int(factor), dx,
                         dx if dy is None else dy, ox, oy, _ptr(out))
    return out


# ----------------------------------------------------------- pointwise ---
def mask_apply(E, m):
    a = _cf(E)
    mm = _cf(m, np.float64)
    lib().orc_mask_apply(_ptr(a), _ptr(mm), a.size)
    return a


def mirror_apply(E, h, lam, sign=1.0):
    a = _cf(E)
    hh = _cf(h, np.float64)
    lib().orc_mirror_apply(_ptr(a), _ptr(hh), a.size, sign, lam)
    return a


def point_source(rows, cols, pixel_scale, flux, tip_x, tilt_y):
    out = np.empty((rows, cols), dtype=np.complex128, order="F")
    lib().orc_point_source(_ptr(out), rows, cols, pixel_scale, flux, tip_x, tilt_y)
    return out


def center_mask_multiply(F, m):
    a = _cf(F)
    mm = _cf(m)
    lib().orc_center_mask_multiply(_ptr(a), a.shape[0], a.shape[1], _ptr(mm), mm.shape[0], mm.shape[1])
    return a


# ----------------------------------------------------------------- hex ---
def hex_centers(n_rings, layout="csim", hex_step=1.0, max_hex=20000):
    hx = np.zeros(max_hex)
    hy = np.zeros(max_hex)
    ring = np.zeros(max_hex, dtype=np.int32)
    f = lib().orc_hex_centers_csim if layout == "csim" else lib().orc_hex_centers_standard
    n = f(int(n_rings), float(hex_step), _ptr(hx), _ptr(hy), _ptr(ring), max_hex)
    return hx[:n].copy(), hy[:n].copy(), ring[:n].copy()


def hex_design_array(hx, hy, layout_radius, array_size, scale_factor, hex_gap):
    hx = np.ascontiguousarray(hx, dtype=np.float64)
    hy = np.ascontiguousarray(hy, dtype=np.float64)
    d = np.empty((array_size, array_size), dtype=np.int32, order="F")
    lib().orc_hex_design_array(_ptr(hx), _ptr(hy), hx.size, float(layout_radius), int(array_size),
                               float(scale_factor), float(hex_gap), _ptr(d))
    return d


def hex_interp_data(design, mask_size, n_sub_pix, pixel_scale, fpm_radius, scale_factor):
    design = np.asfortranarray(design, dtype=np.int32)
    S = int(mask_size)
    nsub = np.empty((S, S, 3), dtype=np.int32, order="F")
    hexnum = np.empty((S, S, 3), dtype=np.int32, order="F")
    lib().orc_hex_interp_data(_ptr(design), design.shape[0], S, int(n_sub_pix), float(pixel_scale),
                              float(fpm_radius), float(scale_factor), _ptr(nsub), _ptr(hexnum))
    return nsub, hexnum


def hex_mask(nsub, hexnum, n_sub_pix, sags, lam, phase_sign=1.0, use_only=-1):
    nsub = np.asfortranarray(nsub, dtype=np.int32)
    hexnum = np.asfortranarray(hexnum, dtype=np.int32)
    sags = np.ascontiguousarray(sags, dtype=np.float64)
    S = nsub.shape[0]
    out = np.empty((S, S), dtype=np.complex128, order="F")
    lib().orc_hex_mask(_ptr(nsub), _ptr(hexnum), S, int(n_sub_pix), _ptr(sags), sags.size,
                       float(lam), float(phase_sign), int(use_only), _ptr(out))
    return out


def fpm_pup_to_lyot(E, dx, mask_minus1, mask_dx, lambda_f, zoom_fft_sign, f_ratio_sign, babinet):
    a = _cf(E)
    m = _cf(mask_minus1)
    lib().orc_fpm_pup_to_lyot(_ptr(a), a.shape[0], dx, _ptr(m), m.shape[0], mask_dx,
                              lambda_f, zoom_fft_sign, f_ratio_sign, int(babinet))
    return a


# ------------------------------------------------------------ contrast ---
def intensity_accumulate(E, I):
    a = _cf(E)
    assert I.flags["F_CONTIGUOUS"] and I.dtype == np.float64
    lib().orc_intensity_accumulate(_ptr(a), a.size, _ptr(I))
    return I


def region_pixels(n, d, loD, xlim=(-1e16, 1e16), radii=(1.0, 2.0), angles=(0.0, 30.0)):
    idx = np.empty(n * n, dtype=np.int32)
    cnt = lib().orc_region_pixels(n, d, loD, xlim[0], xlim[1], radii[0], radii[1],
                                  angles[0], angles[1], _ptr(idx))
    return idx[:cnt].copy()


def hex_resp_objective(M, lambdas, phase_sign, calib, X, objective="mean"):
    """M: (npix, ncol, nslices) complex; X: (nopt, ncand)."""
    M = np.asfortranarray(M, dtype=np.complex128)
    X = np.asfortranarray(X, dtype=np.float64)
    lam = np.ascontiguousarray(lambdas, dtype=np.float64)
    obj = np.empty(X.shape[1])
    lib().orc_hex_resp_objective(_ptr(M), M.shape[0], M.shape[1], M.shape[2], _ptr(lam),
                                 phase_sign, calib, _ptr(X), X.shape[0], X.shape[1],
                                 1 if objective == "max" else 0, _ptr(obj))
    return obj


# ----------------------------------------------------------------- MLP ---
def mlp_forward(x, params):
    W1, b1, W2, b2, W3, b3 = (np.ascontiguousarray(p, dtype=np.float64) for p in params)
    x = np.ascontiguousarray(x, dtype=np.float64)
    nb, nin = x.shape
    nh1, nh2, nout = W1.shape[1], W2.shape[1], W3.shape[1]
    h1 = np.empty((nb, nh1)); a1 = np.empty((nb, nh1))
    h2 = np.empty((nb, nh2)); a2 = np.empty((nb, nh2)); y = np.empty((nb, nout))
    lib().orc_mlp_forward(_ptr(x), nb, nin, _ptr(W1), _ptr(b1), nh1, _ptr(W2), _ptr(b2), nh2,
                          _ptr(W3), _ptr(b3), nout, _ptr(h1), _ptr(a1), _ptr(h2), _ptr(a2), _ptr(y))
    return y


def mlp_loss_grad(x, t, params):
    W1, b1, W2, b2, W3, b3 = (np.ascontiguousarray(p, dtype=np.float64) for p in params)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    nb, nin = x.shape
    nh1, nh2, nout = W1.shape[1], W2.shape[1], W3.shape[1]
    g = [np.empty_like(p) for p in (W1, b1, W2, b2, W3, b3)]
    loss = lib().orc_mlp_loss_grad(_ptr(x), _ptr(t), nb, nin, _ptr(W1), _ptr(b1), nh1,
                                   _ptr(W2), _ptr(b2), nh2, _ptr(W3), _ptr(b3), nout,
                                   *[_ptr(v) for v in g])
    return loss, g
