"""ctypes mirror of the script driver and tool entry points (include/csx.h):
csim's main loop (src/csim.cpp:54-108), the responseData directory format
(src/data/responseData.cpp:175-273) and the differential-evolution optimiser
(src/optimization/differentialEvolution/differentialEvolutionOptimizer.cpp:216-330)
over a Python objective.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from ._lib import check, lib

OBJECTIVE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_double))


def parse_script(text: str) -> dict:
    """Device-free check of a csim input script (csx_script_parse)."""
    b = text.encode()
    n = check(lib().csx_script_parse(b, None, 0))
    buf = C.create_string_buffer(n + 1)
    check(lib().csx_script_parse(b, buf, n + 1))
    return json.loads(buf.value.decode())


def run_script(ctx, text: str, workdir: str | None = None, log_capacity: int = 1 << 20) -> str:
    """Runs the script's blocks in order (csx_run_script); returns the run log."""
    buf = C.create_string_buffer(log_capacity)
    check(lib().csx_run_script(ctx.handle, text.encode(), workdir.encode() if workdir else None, buf, log_capacity))
    return buf.value.decode()


def response_save(directory: str, M, calib: float, lambdas, npol: int = 1):
    """M: (nsrc * npol, slices, rows, cols) complex; written as csim's responseData."""
    M = np.asarray(M, dtype=np.complex128)
    nsp, slices, rows, cols = M.shape
    # each slice column-major rows x cols
    flat = np.ascontiguousarray(np.transpose(M, (0, 1, 3, 2)))
    lam = np.ascontiguousarray(np.asarray(lambdas, dtype=np.float64))
    check(lib().csx_response_save(directory.encode(), flat.ctypes.data, rows, cols, slices, npol, nsp // npol,
                                  float(calib), lam.ctypes.data))


def response_load(directory: str):
    """Returns (M (nsrc * npol, slices, rows, cols) complex, calib, lambdas, sizes dict)."""
    sizes = (C.c_int * 5)()
    calib = C.c_double()
    check(lib().csx_response_load(directory.encode(), None, 0, sizes, C.byref(calib), None, 0))
    rows, cols, slices, npol, nsrc = list(sizes)
    lam = np.zeros(slices)
    flat = np.zeros((nsrc * npol, slices, cols, rows), dtype=np.complex128)
    check(lib().csx_response_load(directory.encode(), flat.ctypes.data, flat.size, sizes, C.byref(calib),
                                  lam.ctypes.data, slices))
    M = np.transpose(flat, (0, 1, 3, 2)).copy()
    return M, calib.value, lam, {"rows": rows, "cols": cols, "slices": slices, "npol": npol, "nsrc": nsrc}


def de_optimize(block_text: str, dim: int, objective):
    """csx_de_optimize over objective(pop (dim, npop) array) -> values (npop,).
    Returns (best value, best vector, iterations)."""
    err = []

    def cb(user, pop, d, n, out):
        try:
            P = np.ctypeslib.as_array(pop, shape=(n * d,)).reshape(n, d).T.copy()
            v = np.asarray(objective(P), dtype=np.float64)
            np.ctypeslib.as_array(out, shape=(n,))[:] = v
            return 0
        except Exception as e:       # reported after the run
            err.append(e)
            return -1

    fn = OBJECTIVE_FN(cb)
    best = np.zeros(dim)
    val = C.c_double()
    it = C.c_int()
    rc = lib().csx_de_optimize(block_text.encode(), dim, C.cast(fn, C.c_void_p), None, best.ctypes.data,
                               C.byref(val), C.byref(it))
    if err:
        raise err[0]
    check(rc)
    return val.value, best, it.value
