"""Element-chain API: csim input scripts run on the GPU (Python host mirror).

Mirrors the reference's script-driven flow (``src/csim.cpp:54-108``): a script
with ``telescope``, ``efield``, ``coronagraph`` (the element list,
``src/coronagraph/coronagraph.cpp:67-155``) and ``regionContrast`` blocks.
FITS ``filename:`` keys are honoured; ``set_array`` binds the same data from
memory.  Arrays are numpy ``(rows, cols)`` in csim's index order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import ChainInfo, PRECISIONS, Region, Report, check, lib
from .ops import Context


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _colmajor(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


class Chain:
    def __init__(self, ctx: Context, script: str):
        self.ctx = ctx
        p = C.c_void_p()
        check(lib().csx_chain_from_script(ctx.handle, script.encode(), C.byref(p)))
        self._p = p
        ctx.adopt(self)

    def close(self):
        if self._p:
            lib().csx_chain_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- setup --
    def set_array(self, element: str, key: str, a):
        self._info = None
        a = _colmajor(a, np.float64)
        if a.ndim == 1:
            a = a.reshape(-1, 1, order="F")
        check(lib().csx_chain_set_array(self._p, element.encode(), key.encode(), _ptr(a), a.shape[0], a.shape[1]))

    def set_complex_array(self, element: str, key: str, a, slice_index: int = 0):
        self._info = None
        a = _colmajor(a, np.complex128)
        check(lib().csx_chain_set_complex_array(self._p, element.encode(), key.encode(), _ptr(a),
                                                a.shape[0], a.shape[1], int(slice_index)))

    def set_opt_data(self, element: str, name: str, v):
        self._info = None
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        check(lib().csx_chain_set_opt_data(self._p, element.encode(), name.encode(), _ptr(v), v.size))

    def get_opt_data(self, element: str, name: str):
        n = C.c_size_t(0)
        check(lib().csx_chain_get_opt_data(self._p, element.encode(), name.encode(), None, C.byref(n)))
        v = np.zeros(n.value)
        check(lib().csx_chain_get_opt_data(self._p, element.encode(), name.encode(), _ptr(v), C.byref(n)))
        return v

    def set_candidates(self, element: str, sags):
        """sags: (n_cand, n_hex) array, or None to return to the single sag set."""
        self._info = None
        if sags is None:
            check(lib().csx_chain_set_candidates(self._p, element.encode(), None, 0, 0))
            return
        s = np.ascontiguousarray(sags, dtype=np.float64)
        check(lib().csx_chain_set_candidates(self._p, element.encode(), _ptr(s), s.shape[1], s.shape[0]))

    def set_precision(self, precision: str):
        self._info = None
        check(lib().csx_chain_set_precision(self._p, PRECISIONS[precision]))

    def info(self) -> dict:
        """Geometry and batch shape; cached until a setter changes the chain."""
        if getattr(self, "_info", None) is None:
            i = ChainInfo()
            check(lib().csx_chain_get_info(self._p, C.byref(i)))
            self._info = {f: getattr(i, f) for f, _ in ChainInfo._fields_}
        return dict(self._info)

    def lambdas(self):
        n = check(lib().csx_chain_get_lambdas(self._p, None, 0))
        out = np.zeros(n)
        check(lib().csx_chain_get_lambdas(self._p, _ptr(out), n))
        return out

    # ------------------------------------------------------------ execute --
    def propagate(self, calibrating: bool = False):
        check(lib().csx_chain_propagate(self._p, int(bool(calibrating))))

    def field(self, cand: int = 0, source: int = 0, lam: int = 0):
        inf = self.info()
        out = np.zeros((inf["out_cols"], inf["out_rows"]), dtype=np.complex128)
        check(lib().csx_chain_get_field(self._p, cand, source, lam, _ptr(out)))
        return out.T.copy()

    def intensity(self, cand: int = 0):
        inf = self.info()
        out = np.zeros((inf["out_cols"], inf["out_rows"]))
        check(lib().csx_chain_get_intensity(self._p, cand, _ptr(out)))
        return out.T.copy()

    def contrast(self, radii=None, angles=(-180.0, 180.0), xlim=(-1e16, 1e16), reference_lambda=-1.0):
        """Region-mean contrast per candidate (regionContrast::compute_contrast)."""
        n = self.info()["n_candidates"]
        out = np.zeros(n)
        rep = Report()
        if radii is None:
            check(lib().csx_chain_contrast(self._p, None, _ptr(out), C.byref(rep)))
        else:
            r = Region(float(radii[0]), float(radii[1]), float(angles[0]), float(angles[1]),
                       float(xlim[0]), float(xlim[1]), float(reference_lambda))
            check(lib().csx_chain_contrast(self._p, C.byref(r), _ptr(out), C.byref(rep)))
        return out, {f: getattr(rep, f) for f, _ in Report._fields_}

    def field_profiles(self, bins, nbins):
        """Sum of |E|^2 per bin for every field of the last propagation -> (n_fields, nbins).
        bins: (out_rows, out_cols) int array, -1 outside every bin."""
        b = np.asfortranarray(np.asarray(bins, dtype=np.int32))
        inf = self.info()
        nf = inf["n_candidates"] * inf["n_sources"] * inf["n_lambdas"]
        out = np.zeros((nf, nbins))
        check(lib().csx_chain_field_profiles(self._p, _ptr(b), int(nbins), _ptr(out)))
        return out

    def hex_response(self, element: str, radii=None, angles=(-180.0, 180.0), xlim=(-1e16, 1e16),
                     reference_lambda=-1.0):
        """Response matrix of the hex FPM (makeHexCFpmResponse): device complex128
        tensor (n_slices, n_hex + 1, n_pix) (each slice column-major npix x ncol),
        and the calibration peak."""
        import torch
        r = None
        if radii is not None:
            r = C.byref(Region(float(radii[0]), float(radii[1]), float(angles[0]), float(angles[1]),
                               float(xlim[0]), float(xlim[1]), float(reference_lambda)))
        npix, ncol, nsl, calib = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(lib().csx_chain_hex_response(self._p, element.encode(), r, None, C.byref(npix), C.byref(ncol),
                                           C.byref(nsl), C.byref(calib)))
        M = torch.empty((nsl.value, ncol.value, npix.value), dtype=torch.complex128, device="cuda")
        check(lib().csx_chain_hex_response(self._p, element.encode(), r, C.c_void_p(M.data_ptr()), C.byref(npix),
                                           C.byref(ncol), C.byref(nsl), C.byref(calib)))
        self._info = None
        return M, calib.value

    def region_fields(self, radii=None, angles=(-180.0, 180.0), xlim=(-1e16, 1e16), reference_lambda=-1.0):
        """Fields of every candidate at the region pixels after the last full
        propagation: device complex128 tensor (n_slices, n_candidates, n_pix)."""
        import torch
        r = None
        if radii is not None:
            r = C.byref(Region(float(radii[0]), float(radii[1]), float(angles[0]), float(angles[1]),
                               float(xlim[0]), float(xlim[1]), float(reference_lambda)))
        npix, nc, nsl = C.c_int(), C.c_int(), C.c_int()
        check(lib().csx_chain_region_fields(self._p, r, None, C.byref(npix), C.byref(nc), C.byref(nsl)))
        F = torch.empty((nsl.value, nc.value, npix.value), dtype=torch.complex128, device="cuda")
        check(lib().csx_chain_region_fields(self._p, r, C.c_void_p(F.data_ptr()), C.byref(npix), C.byref(nc),
                                            C.byref(nsl)))
        return F

    def set_graph(self, on: bool):
        """Replay every propagation after the first as one HIP graph."""
        check(lib().csx_chain_set_graph(self._p, int(bool(on))))

    def set_profiling(self, on: bool):
        check(lib().csx_chain_set_profiling(self._p, int(bool(on))))

    def profile(self) -> dict:
        import json
        n = check(lib().csx_chain_profile_json(self._p, None, 0))
        buf = C.create_string_buffer(n + 1)
        check(lib().csx_chain_profile_json(self._p, buf, n + 1))
        return json.loads(buf.value.decode())

    def normalized_image(self, cand: int = 0):
        inf = self.info()
        out = np.zeros((inf["out_cols"], inf["out_rows"]))
        check(lib().csx_chain_get_normalized_image(self._p, cand, _ptr(out)))
        return out.T.copy()

    def contrast_curve(self, cand: int = 0):
        """(radius in lambda/D, mean normalised intensity) rows of csim's contrast
        curve (contrastCurve::make_contrast_curve) for one candidate."""
        n = check(lib().csx_chain_contrast_curve(self._p, cand, None, None, 0))
        pos = np.zeros(n)
        avg = np.zeros(n)
        check(lib().csx_chain_contrast_curve(self._p, cand, _ptr(pos), _ptr(avg), n))
        return pos, avg


class ChainGroup:
    """Chains with identical element lists (e.g. one script at several pupil sizes)
    propagated together: each contraction stage is one grouped launch."""

    def __init__(self, chains):
        self.chains = list(chains)
        arr = (C.c_void_p * len(self.chains))(*[c._p.value for c in self.chains])
        p = C.c_void_p()
        check(lib().csx_group_create(arr, len(self.chains), C.byref(p)))
        self._p = p
        self.chains[0].ctx.adopt(self)

    def close(self):
        if self._p:
            lib().csx_group_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def merged_launches(self):
        return check(lib().csx_group_merged_launches(self._p))

    def contrast(self, radii=None, angles=(-180.0, 180.0), xlim=(-1e16, 1e16), reference_lambda=-1.0):
        n = check(lib().csx_group_candidates(self._p))
        out = np.zeros(n)
        if radii is None:
            check(lib().csx_group_contrast(self._p, None, _ptr(out)))
        else:
            r = Region(float(radii[0]), float(radii[1]), float(angles[0]), float(angles[1]),
                       float(xlim[0]), float(xlim[1]), float(reference_lambda))
            check(lib().csx_group_contrast(self._p, C.byref(r), _ptr(out)))
        return out
