"""Oracle element chain: csim's script-driven propagation restated on the CPU.

TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).  Runs the same
script text and in-memory arrays as the GPU chain, element by element, with the
FP64 C oracle (Bluestein chirp-z zoomFft exactly as ``src/lib/csim_fft.cpp``).
Geometry is restated here from the reference files, independently of the
product code:

  parser          src/lib/csim_parser.cpp:60-190
  telescope       src/telescope/telescope.cpp:21-60 (loD = mag * F * D * lambda / D)
  efield          src/data/efield.cpp:179-206, 263-430
  mask            src/elements/mask.cpp:23-58            E *= (1+i) m
  mirror          src/elements/mirror.cpp:23-72          E *= exp(-sign 4 pi i h / lambda)
  fresnel         src/elements/fresnelPropagator.cpp:23-61
  downsample      src/elements/downsample.cpp:23-96
  doubleFraunhofer src/elements/doubleFraunhoferFPM.cpp:23-110, doubleFraunhoferInterpFPM.cpp
  hex FPM         src/elements/complexHexMaskFPM.cpp
  fpmPupToLyot    src/elements/fpmPupToLyot.cpp:24-167, 240-496 (three mask kinds)
  fraunhoferFocal src/elements/fraunhoferFocal.cpp:23-72
  zoomFftPropagator src/elements/zoomFftPropagator.cpp:23-91
  region contrast src/tools/regionContrast.cpp:89-205
  contrast curve  src/tools/contrastCurve.cpp:68-158
The additive keys (fpmSamplesPerflD, fpmFOVflD, occulterRadiusFld, hexLayout)
place the FPM on fraunhoferFocal's grid convention: symmetric grid of pixel
loD(referenceLambda)/samplesPerflD, lambda*f = lambda * 2 R * F#.
"""
from __future__ import annotations

import math

import numpy as np

from . import (center_mask_multiply, downsample, fpm_pup_to_lyot, fresnel_as, grid_symmetric, hex_centers,
               hex_design_array, hex_interp_data, hex_mask, interp2, mask_apply, mirror_apply, point_source, zoomfft)


# ------------------------------------------------------- contrast curve ---
def annulus_curve(I, x, loD, pixel_sampling=1):
    """The averaging of contrastCurve::make_contrast_curve
    (src/tools/contrastCurve.cpp:124-139) on a square image I over grid x (both
    axes): n = (len(x) - 2) / pixelSampling sample radii s = linspace(0, max(x)/loD, n);
    bin i (1 <= i <= n - 2) at s[i] is the mean of I over s[i-1] < r/loD < s[i+1];
    the file gets bins i >= 2."""
    RR = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2) / loD
    n = int((len(x) - 2) / pixel_sampling)
    smp = np.linspace(0.0, (x / loD).max(), n)
    pos, avg = [], []
    for i in range(1, n - 1):
        sel = (RR > smp[i - 1]) & (RR < smp[i + 1])
        avg.append(I[sel].mean() if sel.any() else np.nan)
        pos.append(smp[i])
    return np.array(pos[1:]), np.array(avg[1:])


# ------------------------------------------------------------- parsing ---
def parse_lines(text):
    cmds = []
    for line in text.splitlines():
        line = line.lstrip(" \t").rstrip("\r")
        if not line or line[0] == "#":
            continue
        if ":" in line:
            k, a = line.split(":", 1)
        else:
            k, a = line, ""
        cmds.append((k, a))
    return cmds


def blocks(cmds):
    out, loose, depth = [], [], 0
    for k, a in cmds:
        if k == "start":
            if depth == 0:
                out.append([])
            else:
                out[-1].append((k, a))
            depth += 1
        elif k == "end":
            depth -= 1
            if depth > 0:
                out[-1].append((k, a))
        elif depth == 0:
            loose.append((k, a))
        else:
            out[-1].append((k, a))
    if loose:
        out.insert(0, loose)
    return out


def _f(a):
    return float(a.strip())


P2L_MASKS = ("fpmCMCForPupToLyot", "fpmIntHexCMCForPupToLyot", "fpmBinaryForPupToLyot")


class Model:
    """The parsed script: telescope, efield, element list, region."""

    def __init__(self, text, arrays=None):
        self.arrays = dict(arrays or {})
        self.elements = []
        self.region = dict(r1=1.0, r2=2.0, a1=0.0, a2=30.0, x1=-1e16, x2=1e16, ref=-1.0)
        bl = blocks(parse_lines(text))
        bl.sort(key=lambda b: 0 if b and b[0][0] == "telescope" else 1)
        for b in bl:
            t = b[0][0]
            kv = b
            if t == "telescope":
                d = {k: a for k, a in kv}
                self.D = _f(d["primaryDiameter"])
                self.F = _f(d.get("primaryFRatio", d.get("primaryfRatio")))
                self.mag = _f(d.get("magnification", "1"))
            elif t == "efield":
                self.sources = []
                for k, a in kv:
                    if k == "setSize":
                        r, c, n = (int(v) for v in a.split(","))
                        self.rows, self.cols = r, c
                    elif k == "wavelength":
                        self.lambdas = [_f(a)]
                    elif k == "wavelengthList":
                        self.lambdas = [float(v) for v in a.split()]
                    elif k == "pointSource":
                        fl, tx, ty = (float(v) for v in a.split(","))
                        self.sources.append((math.sqrt(fl), tx * self.mag, ty * self.mag))
                    elif k == "beamRadiusPhysical":
                        self.R = _f(a)
                    elif k == "pixelScale":
                        self.pix = _f(a)
                    elif k == "arraySizePixels":
                        self.array_pix = int(a)
                self.beam_pix = math.ceil(2 * self.R / self.pix)
                self.ap_rad = self.R * self.array_pix / self.beam_pix
            elif t == "coronagraph":
                for sb in blocks(kv):
                    if not sb or sb[0][0] == "coronagraph":
                        continue
                    self.elements.append(self._element(sb))
            elif t == "regionContrast":
                for k, a in kv:
                    if k == "referenceLambda":
                        self.region["ref"] = _f(a)
                    elif k in ("radii", "anglesInDegrees", "xlimit"):
                        v = [float(np.float32(float(s))) for s in a.split(",")]   # parsed as float
                        key = {"radii": ("r1", "r2"), "anglesInDegrees": ("a1", "a2"), "xlimit": ("x1", "x2")}[k]
                        self.region[key[0]], self.region[key[1]] = v
        self.nsrc = max(1, len(self.sources))

    @staticmethod
    def _element(sb):
        e = {"type": sb[0][0], "name": sb[0][0], "position": 0.0, "disable": False, "hex": None,
             "mask_type": None, "mask": {}}
        for sub in blocks(sb):
            if sub and sub[0][0] == "complexHexMaskFPM":
                e["hex"] = {k: a.strip() for k, a in sub[1:]}
            elif sub and sub[0][0] in P2L_MASKS:
                # fpmPupToLyot mask sub-block (fpmPupToLyot.cpp:172-195), possibly
                # holding a complexHexMaskFPM block
                e["mask_type"] = sub[0][0]
                for msub in blocks(sub):
                    if msub and msub[0][0] == "complexHexMaskFPM":
                        e["hex"] = {k: a.strip() for k, a in msub[1:]}
                    else:
                        e["mask"].update({k: a.strip() for k, a in msub if k != sub[0][0]})
            else:
                for k, a in sub:
                    a = a.strip()
                    if k == "name":
                        e["name"] = a
                    elif k == "position":
                        e["position"] = float(a)
                    elif k == "disableForCalibration":
                        e["disable"] = a == "on"
                    elif k != e["type"]:
                        e[k] = a
        return e

    def loD(self, lam):
        return self.mag * self.F * self.D * lam / self.D


# ------------------------------------------------------------ geometry ---
def sym_grid(n, d):
    return (n, -(n * d - d) / 2.0, d)


def resample_to(m, in_pix, n_in, g_out):
    """mask/mirror data defined on the initial grid, resampled when sizes differ."""
    if m.shape[0] == g_out[0]:
        return m
    xin = grid_symmetric(n_in, in_pix)
    xout = g_out[1] + np.arange(g_out[0]) * g_out[2]
    return interp2(m, xin, xin, xout, xout)


class HexGeometry:
    def __init__(self, h, model=None, zoom=1.0):
        self.layout = h.get("hexLayout", "csim")
        self.rings = int(h["maskNRings"])
        self.step = float(h.get("hexStep", 1))
        self.A = int(h["fpmArraySize"])
        self.scale = float(h["fpmScaleFactor"])
        self.gap = float(h["hexGap"])
        self.nsp = int(h["nSubPix"])
        self.S = int(h["fpmMaskSize"])
        # init_mask (complexHexMaskFPM.cpp:193-234): pixelScale, else the physical mask
        # size over the mask size, else 2 (R / (dx S zoom)) lambdaRef fpmFRatio with the
        # parent element's zoomFactor
        if "pixelScale" in h:
            self.pixel = float(h["pixelScale"])
        elif "fpmMaskPhysicalSize" in h:
            self.pixel = float(h["fpmMaskPhysicalSize"]) / self.S
        else:
            self.pixel = (2 * (model.R / (model.pix * self.S * zoom)) * float(h["lambdaRef"]) *
                          float(h["fpmFRatio"]))
        self.radius = float(h["fpmRadius"])
        self.phase_sign = float(h.get("phaseSign", 1))
        hx, hy, _ = hex_centers(self.rings, self.layout, self.step)
        self.nhex = hx.size
        lr = self.rings if self.layout == "csim" else self.rings * math.sqrt(3.0) * self.step + self.step
        design = hex_design_array(hx, hy, lr, self.A, self.scale, self.gap)
        self.nsub, self.hexnum = hex_interp_data(design, self.S, self.nsp, self.pixel, self.radius, self.scale)

    def mask(self, sags, lam):
        return hex_mask(self.nsub, self.hexnum, self.nsp, sags, lam, self.phase_sign)


class OracleChain:
    def __init__(self, text, arrays, sags=None):
        self.m = Model(text, arrays)
        self.sags = sags
        self.hex = {}
        for e in self.m.elements:
            if e["hex"] is not None:
                self.hex[e["name"]] = HexGeometry(e["hex"], self.m, float(e.get("zoomFactor", 1.0)))

    def _fpm_geom(self, e, lam, g):
        m = self.m
        if float(e.get("fpmSamplesPerflD", 0)) > 0:
            spl, fov = float(e["fpmSamplesPerflD"]), float(e["fpmFOVflD"])
            ref = float(e.get("referenceLambda", m.lambdas[0]))
            return sym_grid(math.ceil(fov * spl), m.loD(ref) / spl), lam * 2 * m.R * m.F
        pad = float(e.get("padFactor", 1.0))
        zoom = float(e.get("zoomFactor", 1.0))
        n = int(pad * g[0])
        if e["hex"] is not None:
            d = self.hex[e["name"]].pixel
        else:
            d = 2 * (m.R / (m.pix * g[0] * zoom)) * lam * 80
        return (n, (0 - n / 2.0) * d, d), lam * 3.52

    def field(self, src, li, calibrating=False, cand_sags=None):
        m = self.m
        lam = m.lambdas[li]
        if m.sources:
            fl, tx, ty = m.sources[src]
        else:
            fl, tx, ty = 1.0, 0.0, 0.0
        E = point_source(m.rows, m.cols, m.pix, fl, tx, ty)
        g = sym_grid(m.cols, m.pix)
        els = m.elements
        for i, e in enumerate(els):
            t = e["type"]
            if calibrating and e["disable"] and t != "fpmPupToLyot":
                continue
            if t == "mask":
                E = mask_apply(E, resample_to(m.arrays[(e["name"], "filename")], m.pix, m.cols, g))
            elif t == "mirror":
                h = resample_to(m.arrays[(e["name"], "filename")], m.pix, m.cols, g)
                E = mirror_apply(E, h, lam, float(e.get("sign", 1)))
            elif t == "complexMask":
                # complexMask::execute (src/elements/complexMask.cpp:24-55): the field's
                # centred window (row0 = rows/2 - maskRows/2, likewise columns) times the
                # maskIndex slice of the complex cube amp exp(-i ph) (:58-67)
                K = np.asarray(m.arrays[(e["name"], "mask")], dtype=np.complex128)
                if K.ndim == 3:
                    K = K[..., int(e.get("maskIndex", 0))]
                E = center_mask_multiply(E, K)
            elif t == "fresnelPropagator":
                z = float(e.get("propagationSign", 1)) * abs(els[i + 1]["position"] - els[i - 1]["position"])
                pad = int(float(e.get("padFactor", 1)))
                if pad == 1:
                    E = fresnel_as(E, lam, m.ap_rad, z)
                else:
                    # fresnelPropagator.cpp:45-54: any other padFactor embeds the field in a
                    # 2N x 2N zero pad (rows N/2 .. 3N/2 - 1) with apRad * padFactor
                    r, c = E.shape
                    P = np.zeros((2 * r, 2 * c), dtype=np.complex128)
                    P[r // 2:r // 2 + r, c // 2:c // 2 + c] = E
                    P = fresnel_as(P, lam, pad * m.ap_rad, z)
                    E = P[r // 2:r // 2 + r, c // 2:c // 2 + c].copy()
            elif t == "downsample":
                f = int(e["sampleFactor"])
                E = downsample(E, f, g[2])
                g = sym_grid(g[0] // f, g[2] * f)
            elif t in ("doubleFraunhoferFPM", "doubleFraunhoferInterpFPM"):
                gf, lf = self._fpm_geom(e, lam, g)
                Ef = zoomfft(E, (g[1], g[1], g[2], g[2]), (gf[0], gf[1], gf[1], gf[2], gf[2]), lf)
                if e["hex"] is not None:
                    hg = self.hex[e["name"]]
                    sags = cand_sags if cand_sags is not None else self.sags
                    Ef = center_mask_multiply(Ef, hg.mask(sags, lam))
                elif float(e.get("occulterRadiusFld", -1)) >= 0:
                    ref = float(e.get("referenceLambda", m.lambdas[0]))
                    x = gf[1] + np.arange(gf[0]) * gf[2]
                    rr = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2) / m.loD(ref)
                    Ef = Ef * np.where(rr < float(e["occulterRadiusFld"]), 0.0, 1.0)
                elif (e["name"], "mask") in m.arrays:
                    # doubleFraunhoferFPM.cpp:84-97: the complex mask cube, centred; an unset
                    # maskIndex latches to the first slice executed (slice 0) for good
                    K = np.asarray(m.arrays[(e["name"], "mask")], dtype=np.complex128)
                    if K.ndim == 3:
                        K = K[..., max(int(e.get("maskIndex", -1)), 0)]
                    Ef = center_mask_multiply(Ef, K)
                E = zoomfft(Ef, (gf[1], gf[1], gf[2], gf[2]), (g[0], g[1], g[1], g[2], g[2]), -lf)
            elif t == "fpmPupToLyot":
                E = self._pup_to_lyot(e, E, g, lam, calibrating, cand_sags)
            elif t == "zoomFftPropagator":
                # zoomFftPropagator::execute (src/elements/zoomFftPropagator.cpp:23-91):
                # offset output grid (i - n/2) d; lambda*f from the first wavelength
                d, pad = int(e.get("dir", 1)), float(e.get("padFactor", 1.0))
                lam0 = m.lambdas[0]
                if d == 1:
                    n = int(pad * g[0])
                    dxp = 2 * (m.R / (m.pix * g[0] * float(e.get("zoomFactor", 1.0)))) * lam0 * 80
                    lf = lam0 * 3.52
                else:
                    n, dxp, lf = int(g[0] / pad), m.pix, -lam0 * 3.52
                go = (n, (0 - n / 2.0) * dxp, dxp)
                E = zoomfft(E, (g[1], g[1], g[2], g[2]), (go[0], go[1], go[1], go[2], go[2]), lf)
                g = go
            elif t == "fraunhoferFocal":
                ref = float(e["referenceLambda"])
                spl, fov = float(e["samplesPerflD"]), float(e["FOVflD"])
                go = sym_grid(math.ceil(fov * spl), m.loD(ref) / spl)
                E = zoomfft(E, (g[1], g[1], g[2], g[2]), (go[0], go[1], go[1], go[2], go[2]), lam * 2 * m.R * m.F)
                g = go
            else:
                raise NotImplementedError(t)
        self.out_grid = g
        return np.asarray(E)

    def _pup_to_lyot(self, e, E, g, lam, calibrating, cand_sags):
        """fpmPupToLyot::execute (src/elements/fpmPupToLyot.cpp:24-167) with the mask
        of its sub-block: K = fliplr(hex mask) - 1, K = cube - 1 (Babinet, identity
        when disabled for calibration) or the opaque mask itself."""
        m = self.m
        mt, mk = e["mask_type"], e["mask"]
        frs = float(e.get("fRatioSign", 1))
        zs = float(e.get("zoomFftSign", 1))
        if mt != "fpmBinaryForPupToLyot":
            if calibrating and e["disable"]:
                return E                                  # K = exp(0) - 1 = 0
            lf = frs * lam * float(e["focalRatio"]) * 2 * m.R
            if mt == "fpmIntHexCMCForPupToLyot":
                hg = self.hex[e["name"]]
                sags = cand_sags if cand_sags is not None else self.sags
                K = hg.mask(sags, lam)[:, ::-1] - 1.0
                mdx = hg.pixel
            else:
                idx = int(mk.get("maskIndex", -1))
                K = np.asarray(m.arrays[(e["name"], "mask")], dtype=np.complex128)
                if K.ndim == 3:
                    K = K[..., max(idx, 0)]
                K = K - 1.0
                mdx = float(mk["fpmMatScale"])
            return fpm_pup_to_lyot(E, g[2], K, mdx, lf, zs, frs, True)
        key = "calibFilename" if (calibrating and e["disable"]) else "maskFilename"
        K = np.asarray(m.arrays[(e["name"], key)], dtype=np.complex128)
        mdx = 2 * float(mk["outerRadiusFld"]) * m.loD(float(mk["referenceLambda"])) / K.shape[0]
        lf = frs * lam * m.F * m.D
        return fpm_pup_to_lyot(E, g[2], K, mdx, lf, zs, frs, False)

    def intensity(self, calibrating=False, cand_sags=None):
        I = None
        for s in range(self.m.nsrc):
            for li in range(len(self.m.lambdas)):
                E = self.field(s, li, calibrating, cand_sags)
                I = np.abs(E) ** 2 if I is None else I + np.abs(E) ** 2
        return I

    def region_index(self, g=None):
        r = self.m.region
        g = g or self.out_grid
        x = g[1] + np.arange(g[0]) * g[2]
        XX = np.broadcast_to(x[None, :], (g[0], g[0]))
        YY = np.broadcast_to(x[:, None], (g[0], g[0]))
        ref = r["ref"] if r["ref"] != -1.0 else float(np.mean(self.m.lambdas))
        loD = self.m.loD(ref)
        RR = np.sqrt(XX ** 2 + YY ** 2)
        TT = np.arctan2(YY, XX)
        sel = ((XX / loD >= r["x1"]) & (XX / loD <= r["x2"]) & (RR / loD >= r["r1"]) & (RR / loD <= r["r2"])
               & (TT > r["a1"] * np.pi / 180) & (TT < r["a2"] * np.pi / 180))
        return sel

    def contrast_curve(self, cand_sags=None, pixel_sampling=1):
        """contrastCurve::make_contrast_curve (src/tools/contrastCurve.cpp:68-158): the
        source/wavelength intensity sum over the calibration peak, averaged over the
        annuli sample[i-1] < r/loD < sample[i+1] (strict) with
        sample = linspace(0, max(pixelX)/loD, (nX - 2)/pixelSampling) and loD at
        lambdaData[(int) round(n_slices/2)]; returns the rows the reference writes
        to its file (i >= 2)."""
        m = self.m
        loD = m.loD(m.lambdas[len(m.lambdas) // 2])
        peak = self.intensity(True).max()
        I = self.intensity(False, cand_sags) / peak
        g = self.out_grid
        x = g[1] + np.arange(g[0]) * g[2]
        return annulus_curve(I, x, loD, pixel_sampling)

    def contrast(self, cand_sags=None):
        Ic = self.intensity(True)
        peak = Ic.max()
        I = self.intensity(False, cand_sags)
        sel = self.region_index()
        return float(I[sel].mean() / peak), I / peak, peak
