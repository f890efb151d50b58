"""Operator API on device tensors (torch is only the device-memory plumbing).

Fields follow csim's layout: a ``rows x cols`` field is stored column-major,
which for a torch tensor means shape ``(batch, cols, rows)`` contiguous
(``tensor[b, c, r]`` = E(r, c)).  Helpers ``to_device`` / ``to_host`` convert
from / to numpy arrays of shape ``(rows, cols)`` (or ``(batch, rows, cols)``).
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from ._lib import Grid, PRECISIONS, check, lib


def _torch():
    import torch
    return torch


_live_contexts = weakref.WeakSet()


CTX_DEFAULT_STREAM = 1     # csx.h CSX_CTX_DEFAULT_STREAM


class Context:
    """One HIP device and stream (csx_ctx).  Objects created on a context
    (chains, surrogates) register with it and are closed before it."""

    def __init__(self, device: int = 0):
        # The library's work is ordered with torch's on torch's current stream of
        # the device (tensors torch produces asynchronously are then complete
        # when a library kernel reads them, and the reverse).
        torch = _torch()
        with torch.cuda.device(device):
            ts = torch.cuda.current_stream().cuda_stream
        p = C.c_void_p()
        check(lib().csx_ctx_create(device, CTX_DEFAULT_STREAM if ts == 0 else 0, C.byref(p)))
        if ts != 0:
            check(lib().csx_ctx_set_stream(p, C.c_void_p(ts)))
        self._p = p
        self.device = device
        self._children = weakref.WeakSet()
        _live_contexts.add(self)

    @property
    def handle(self):
        if not self._p:
            raise RuntimeError("context is closed")
        return self._p

    def adopt(self, obj):
        self._children.add(obj)

    def synchronize(self):
        check(lib().csx_ctx_synchronize(self._p))

    def close(self):
        for c in list(self._children):
            c.close()
        if self._p:
            lib().csx_ctx_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@atexit.register
def _close_all():
    # release device objects while the HIP runtime is still up
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


def to_device(E, device="cuda"):
    """numpy (rows, cols) or (batch, rows, cols) complex -> torch (batch, cols, rows) complex128."""
    torch = _torch()
    a = np.asarray(E, dtype=np.complex128)
    if a.ndim == 2:
        a = a[None]
    t = np.ascontiguousarray(np.transpose(a, (0, 2, 1)))
    return torch.from_numpy(t).to(device)


def to_host(t):
    """torch (batch, cols, rows) -> numpy (batch, rows, cols)."""
    a = t.detach().cpu().numpy()
    return np.transpose(a, (0, 2, 1)).copy()


def real_to_device(m, device="cuda"):
    torch = _torch()
    a = np.asarray(m, dtype=np.float64)
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(device)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _grid(n, x0, dx):
    return Grid(int(n), float(x0), float(dx))


def mft_workspace_size(in_x, in_y, out_x, out_y, batch, precision="fp64"):
    return int(lib().csx_mft_workspace_size(_grid(*in_x), _grid(*in_y), _grid(*out_x), _grid(*out_y),
                                            int(batch), PRECISIONS[precision]))


def mft(ctx: Context, E, in_x, in_y, out_x, out_y, lambda_f, precision="fp64", work=None):
    """csim zoomFft in closed form on a batch.  Grids are (n, x0, dx) tuples.  work:
    None (the library allocates) or a device tensor of at least mft_workspace_size
    bytes that holds every work buffer of the call."""
    torch = _torch()
    assert E.dtype == torch.complex128 and E.is_contiguous()
    batch = E.shape[0]
    assert E.shape[1] == in_x[0] and E.shape[2] == in_y[0]
    lf = np.ascontiguousarray(np.broadcast_to(np.asarray(lambda_f, dtype=np.float64), (batch,)))
    out = torch.empty((batch, out_x[0], out_y[0]), dtype=torch.complex128, device=E.device)
    if work is not None:
        assert work.is_cuda and work.numel() * work.element_size() >= mft_workspace_size(in_x, in_y, out_x, out_y,
                                                                                          batch, precision)
    check(lib().csx_mft(ctx.handle, _grid(*in_x), _grid(*in_y), _grid(*out_x), _grid(*out_y),
                        lf.ctypes.data_as(C.c_void_p), batch, _ptr(E), _ptr(out),
                        PRECISIONS[precision], None if work is None else _ptr(work)))
    return out


def fresnel_as(ctx: Context, E, lambdas, ap_rad, z):
    n = E.shape[1]
    lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lambdas, dtype=np.float64), (E.shape[0],)))
    check(lib().csx_fresnel_as(ctx.handle, _ptr(E), n, E.shape[0], lam.ctypes.data_as(C.c_void_p),
                               float(ap_rad), float(z)))
    return E


def fft2(ctx: Context, E, direction=-1):
    check(lib().csx_fft2(ctx.handle, _ptr(E), E.shape[1], E.shape[0], int(direction)))
    return E


def downsample(ctx: Context, E, factor, pixel_size):
    torch = _torch()
    b, cols, rows = E.shape
    out = torch.empty((b, cols // factor, rows // factor), dtype=torch.complex128, device=E.device)
    check(lib().csx_downsample(ctx.handle, _ptr(E), rows, cols, b, int(factor), float(pixel_size), _ptr(out)))
    return out


def apply_mask(ctx: Context, E, m_dev):
    check(lib().csx_apply_mask(ctx.handle, _ptr(E), E.shape[1] * E.shape[2], E.shape[0], _ptr(m_dev)))
    return E


def apply_mirror(ctx: Context, E, h_dev, lambdas, sign=1.0):
    lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lambdas, dtype=np.float64), (E.shape[0],)))
    check(lib().csx_apply_mirror(ctx.handle, _ptr(E), E.shape[1] * E.shape[2], E.shape[0], _ptr(h_dev),
                                 float(sign), lam.ctypes.data_as(C.c_void_p)))
    return E


def hex_layout(layout, n_rings, hex_step=1.0):
    n = check(lib().csx_hex_layout(int(layout), int(n_rings), float(hex_step), None, None, 0))
    hx = np.zeros(n)
    hy = np.zeros(n)
    check(lib().csx_hex_layout(int(layout), int(n_rings), float(hex_step),
                               hx.ctypes.data_as(C.c_void_p), hy.ctypes.data_as(C.c_void_p), n))
    return hx, hy


def hex_rasterize(layout, n_rings, hex_step, array_size, scale_factor, hex_gap, mask_size, n_sub_pix,
                  pixel_scale, fpm_radius):
    """Returns (nsub, hexnum, nhex); arrays are (S, S, 3) in csim's index order."""
    S = int(mask_size)
    nsub = np.zeros((3, S, S), dtype=np.int32)     # memory = column-major S x S x 3
    hexnum = np.zeros((3, S, S), dtype=np.int32)
    nhex = check(lib().csx_hex_rasterize(int(layout), int(n_rings), float(hex_step), int(array_size),
                                         float(scale_factor), float(hex_gap), S, int(n_sub_pix),
                                         float(pixel_scale), float(fpm_radius),
                                         nsub.ctypes.data_as(C.c_void_p), hexnum.ctypes.data_as(C.c_void_p)))
    return np.transpose(nsub, (2, 1, 0)), np.transpose(hexnum, (2, 1, 0)), nhex


def hex_masks(ctx: Context, nsub_dev, hexnum_dev, S, n_sub_pix, sags_dev, n_hex, n_cand, lambdas, phase_sign=1.0):
    torch = _torch()
    lam = np.ascontiguousarray(np.asarray(lambdas, dtype=np.float64))
    out = torch.empty((n_cand * lam.size, S, S), dtype=torch.complex128, device=sags_dev.device)
    check(lib().csx_hex_mask(ctx.handle, _ptr(nsub_dev), _ptr(hexnum_dev), S, int(n_sub_pix), _ptr(sags_dev),
                             int(n_hex), int(n_cand), lam.ctypes.data_as(C.c_void_p), lam.size,
                             float(phase_sign), _ptr(out)))
    return out


def intensity(ctx: Context, E, fields_per_group):
    torch = _torch()
    b = E.shape[0]
    npix = E.shape[1] * E.shape[2]
    out = torch.empty((b // fields_per_group, E.shape[1], E.shape[2]), dtype=torch.float64, device=E.device)
    check(lib().csx_intensity(ctx.handle, _ptr(E), npix, b // fields_per_group, fields_per_group, _ptr(out)))
    return out


def region_mean(ctx: Context, I, idx_dev, peak):
    torch = _torch()
    g = I.shape[0]
    out = torch.empty((g,), dtype=torch.float64, device=I.device)
    check(lib().csx_region_mean(ctx.handle, _ptr(I), I.shape[1] * I.shape[2], g, _ptr(idx_dev),
                                idx_dev.numel(), float(peak), _ptr(out)))
    return out


def philox_fill(seed, stream, n, normal=False):
    out = np.empty(int(n))
    check(lib().csx_philox_fill(int(seed), int(stream), out.ctypes.data_as(C.c_void_p), out.size, int(normal)))
    return out


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().csx_philox4x32_10(c.ctypes.data_as(C.c_void_p), k.ctypes.data_as(C.c_void_p),
                            out.ctypes.data_as(C.c_void_p))
    return out
