"""Writes the golden fixtures that come from the reference's own data files.

Reads (as data, never executing anything) from /root/reference:
  donutPiaaFullCircleExample/mono_2_6/responseData/responseData_s0_p0_{re,im}.fits
      csim response matrix M (636 pixels x 500 columns, FP64 big-endian FITS)
  donutPiaaFullCircleExample/mono_2_6/responseData/problemParams.txt
      calibMaxIntensity and the wavelength
  donutPiaaFullCircleExample/mono_2_6/optTest/bestOptVector.fits
      the DE optimum (499 hex sags)
  donutPiaaFullCircleExample/mono_2_6/optTest/optBestVal.txt (last line)
      the objective value csim logged for that optimum
  wfirst_sags.fits
      a 32-ring csim hex-sag vector (its length is the hex count)
and writes tests/golden/mono_kat.npz.  Run in the build container:
    python tests/golden/make_golden.py
"""
import os

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mono_kat.npz")


def read_fits(path):
    raw = open(path, "rb").read()
    hdr, off = {}, 0
    while True:
        block = raw[off:off + 2880]
        off += 2880
        done = False
        for i in range(36):
            card = block[i * 80:(i + 1) * 80].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= ":
                hdr[key] = card[10:].split("/")[0].strip()
        if done:
            break
    dims = [int(hdr[f"NAXIS{i + 1}"]) for i in range(int(hdr["NAXIS"]))]
    dt = {-64: ">f8", -32: ">f4"}[int(hdr["BITPIX"])]
    n = int(np.prod(dims))
    data = np.frombuffer(raw[off:off + n * np.dtype(dt).itemsize], dtype=dt).astype(np.float64)
    return data.reshape(dims[::-1]), dims


def main():
    base = os.path.join(REF, "donutPiaaFullCircleExample", "mono_2_6")
    re_, dims = read_fits(os.path.join(base, "responseData", "responseData_s0_p0_re.fits"))
    im_, _ = read_fits(os.path.join(base, "responseData", "responseData_s0_p0_im.fits"))
    # FITS NAXIS1 = 500 columns, NAXIS2 = 636 rows; numpy order (1, 636, 500) is the matrix
    M = (re_ + 1j * im_).reshape(dims[1], dims[0])
    best, _ = read_fits(os.path.join(base, "optTest", "bestOptVector.fits"))
    params = open(os.path.join(base, "responseData", "problemParams.txt")).read().split()
    calib = float(params[params.index("calibMaxIntensity:") + 1])
    lam = float(params[params.index("wavelengths:") + 1])
    last = open(os.path.join(base, "optTest", "optBestVal.txt")).read().strip().splitlines()[-1].split()
    wf, wfd = read_fits(os.path.join(REF, "wfirst_sags.fits"))
    np.savez_compressed(OUT, M=M, best=best.ravel(), calib=calib, wavelength=lam,
                        logged_iteration=int(last[0]), logged_value=float(last[1]),
                        wfirst_nhex=int(np.asarray(wf).size))
    print(f"wrote {OUT}: M {M.shape}, best {best.size}, calib {calib}, lambda {lam}, "
          f"logged {last}, wfirst hexes {np.asarray(wf).size}")


if __name__ == "__main__":
    main()
