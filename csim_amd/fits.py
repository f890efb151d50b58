"""FITS matrices in csim's convention, read and written by ``libcsx.so``.

``load_mat`` / ``save_mat`` follow ``src/lib/csim_fits.cpp``: a matrix of ``rows x
cols`` is stored with NAXIS1 = cols and NAXIS2 = rows, its rows one after another
(``save_mat`` writes the transposed column-major matrix, :480-493), BITPIX -64,
big-endian.  Complex matrices are file pairs named from the root before ``.fits``:
``_re``/``_im`` (``load_mat``, :451-477, and ``save_mat`` with "reIm") or
``_am``/``_ph`` (``save_mat`` with "amPh", :495-528).  Arrays here are numpy
``(rows, cols)`` (``(rows, cols, slices)`` for cubes) in csim's index order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


def _root(path: str) -> str:
    i = path.find(".fits")
    return path[:i] if i > 0 else path


def read_fits(path: str) -> np.ndarray:
    """A real FITS image as (rows, cols) or (rows, cols, slices) float64."""
    r, c, s = C.c_int(), C.c_int(), C.c_int()
    check(lib().csx_fits_read(path.encode(), None, 0, C.byref(r), C.byref(c), C.byref(s)))
    out = np.empty(r.value * c.value * s.value, dtype=np.float64)
    check(lib().csx_fits_read(path.encode(), out.ctypes.data_as(C.c_void_p), out.size,
                              C.byref(r), C.byref(c), C.byref(s)))
    m = out.reshape((s.value, c.value, r.value)).transpose(2, 1, 0)
    return np.ascontiguousarray(m[:, :, 0] if s.value == 1 else m)


def write_fits(path: str, m) -> None:
    """A real 2-D matrix as a BITPIX -64 FITS image (save_mat)."""
    a = np.asarray(m, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    if a.ndim != 2:
        raise ValueError("write_fits: a matrix is 2-D")
    f = np.asfortranarray(a)
    check(lib().csx_fits_write(path.encode(), f.ctypes.data_as(C.c_void_p), a.shape[0], a.shape[1]))


def load_mat(path: str, complex_pair: bool = False) -> np.ndarray:
    """load_mat: a real matrix from ``path``, or with ``complex_pair`` the complex
    matrix from ``<root>_re.fits`` and ``<root>_im.fits``."""
    if not complex_pair:
        return read_fits(path)
    root = _root(path)
    re, im = read_fits(root + "_re.fits"), read_fits(root + "_im.fits")
    if re.shape != im.shape:
        raise ValueError(f"load_mat: {root}_re.fits and _im.fits differ in shape")
    return re + 1j * im


def save_mat(path: str, m, save_type: str = "reIm") -> None:
    """save_mat: a real matrix to ``path``; a complex one to the ``_re``/``_im``
    ("reIm") or ``_am``/``_ph`` ("amPh") pair next to it."""
    a = np.asarray(m)
    if not np.iscomplexobj(a):
        write_fits(path, a)
        return
    root = _root(path)
    if save_type == "reIm":
        write_fits(root + "_re.fits", a.real)
        write_fits(root + "_im.fits", a.imag)
    elif save_type == "amPh":
        write_fits(root + "_am.fits", np.abs(a))
        write_fits(root + "_ph.fits", np.angle(a))
    else:
        raise ValueError(f"save_mat: save type {save_type!r} is neither 'reIm' nor 'amPh'")
