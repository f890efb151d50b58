"""Synthetic workloads of the propagation path as csim input scripts.

Each builder returns a ``Workload``: the script text (csim grammar and keys,
plus the additive geometry keys ``fpmSamplesPerflD``/``fpmFOVflD``/
``occulterRadiusFld``/``hexLayout``), the in-memory arrays that stand in for
csim's FITS inputs, and metadata.  Random inputs come from Philox4x32-10
(``csx_philox_fill``) with seed ``20261203 + workload id`` so the GPU chain,
the CPU oracle and the bench see identical data.

Workloads (sizes and physics of the headline configurations):
  0  1024^2 pupil, mono 650 nm, occulting disk 3 lambda/D, Lyot 0.9 D, 128^2 image
  1  2048^2 field, two PIAA sag mirrors, 1 m angular-spectrum step, downsample x2, then 0
  2  1024^2 pupil, 9 wavelengths 600-660 nm x 3 sources (0, 2, 6 lambda/D), 217-hex FPM
  3  DE batch: 9 wavelengths, on-axis, hex FPM; pupil 1024^2 or 512^2 (radii halved)
  p2l  the FPM as a pupil-plane convolution (csim's fpmPupToLyot) with each of its
       three masks: hex (fpmIntHexCMCForPupToLyot), complex cube (fpmCMCForPupToLyot),
       opaque disk (fpmBinaryForPupToLyot)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

BASE_SEED = 20261203
PIX = 10e-6            # pupil pixel (m) of the 1024^2 pupils
F_RATIO = 10.0         # telescope F#
LAMBDA0 = 650e-9


@dataclass
class Workload:
    wid: int
    name: str
    script: str
    arrays: dict = field(default_factory=dict)          # (element, key) -> real array
    complex_arrays: dict = field(default_factory=dict)  # (element, key, slice) -> complex array
    meta: dict = field(default_factory=dict)


def philox(seed, stream, n, normal=False):
    from .ops import philox_fill
    return philox_fill(seed, stream, n, normal)


def _radius_px(n):
    c = (np.arange(n) - (n - 1) / 2.0)
    return np.sqrt(c[None, :] ** 2 + c[:, None] ** 2), c


def pupil_amplitude(n, r_ap, r_gauss):
    r, _ = _radius_px(n)
    return (r < r_ap) * np.exp(-(r / r_gauss) ** 2)


def circle(n, radius):
    r, _ = _radius_px(n)
    return (r < radius).astype(np.float64)


def aberration_sag(seed, n, sigma_rad=0.01, lam0=LAMBDA0):
    """White phase noise phi ~ N(0, sigma) as a mirror sag h = -phi lam0 / (4 pi)."""
    phi = philox(seed, 0, n * n, normal=True).reshape(n, n) * sigma_rad
    return -phi * lam0 / (4 * math.pi)


def noll_nm(jmax):
    out, j, n = {}, 1, 0
    while j <= jmax:
        for mabs in range(n % 2, n + 1, 2):
            if mabs == 0:
                out[j] = (n, 0)
                j += 1
            else:
                for jj in (j, j + 1):
                    out[jj] = (n, mabs if jj % 2 == 0 else -mabs)
                j += 2
        n += 1
    return out


def zernike(j, rho, theta):
    n, m = noll_nm(j)[j]
    ma = abs(m)
    R = np.zeros_like(rho)
    for k in range((n - ma) // 2 + 1):
        c = (-1) ** k * math.factorial(n - k) / (
            math.factorial(k) * math.factorial((n + ma) // 2 - k) * math.factorial((n - ma) // 2 - k))
        R = R + c * rho ** (n - 2 * k)
    if m == 0:
        return math.sqrt(n + 1) * R
    norm = math.sqrt(2 * (n + 1))
    return norm * R * (np.cos(ma * theta) if m > 0 else np.sin(ma * theta))


def piaa_sag(seed, stream, n, r_unit_px, sigma=50e-9, radial=0.5e-6):
    """sum_{j=4..36} c_j Z_j + radial (rho^2 - rho^4/2), rho clipped to 1."""
    coef = philox(seed, stream, 33, normal=True) * sigma
    r, c = _radius_px(n)
    rho = np.minimum(r / r_unit_px, 1.0)
    theta = np.arctan2(c[:, None] + 0 * c[None, :], c[None, :] + 0 * c[:, None])
    h = np.zeros((n, n))
    for i, j in enumerate(range(4, 37)):
        h += coef[i] * zernike(j, rho, theta)
    return h + radial * (rho ** 2 - rho ** 4 / 2), coef


def hex_thickness(seed, stream, n_hex, n_cand=1, tmax=1e-6):
    return philox(seed, stream, n_hex * n_cand).reshape(n_cand, n_hex) * tmax


def thickness_to_sag(t, n_index=1.46):
    """tau = exp(i 2 pi (n-1) t / lambda) == csim's exp(i 4 pi sag / lambda) with sag = (n-1) t / 2."""
    return (n_index - 1.0) * np.asarray(t) / 2.0


# ---------------------------------------------------------------- scripts --
def _telescope(D):
    return f"""start
\ttelescope
\tprimaryDiameter:{D!r}
\tprimaryFRatio:{F_RATIO!r}
\tmagnification:1
end
"""


def _efield(n, pix, beam_r, lambdas, sources):
    lines = ["start", "\tefield", f"\tsetSize:{n}, {n}, {len(lambdas)}"]
    if len(lambdas) == 1:
        lines.append(f"\twavelength:{lambdas[0]!r}")
    else:
        lines.append("\twavelengthList:" + " ".join(repr(float(l)) for l in lambdas))
    for (flux, tx, ty) in sources:
        lines.append(f"\tpointSource:{flux!r}, {tx!r}, {ty!r}")
    lines += [f"\tbeamRadiusPhysical:{beam_r!r}", f"\tpixelScale:{pix!r}", f"\tarraySizePixels:{n}", "end", ""]
    return "\n".join(lines)


def _fpm_block(kind, lam_ref, occulter=None, hex_opts=None):
    lines = ["\tstart", f"\t\t{kind}", "\t\tname:fpm", "\t\tdisableForCalibration:on",
             "\t\tfpmSamplesPerflD:8", "\t\tfpmFOVflD:32", f"\t\treferenceLambda:{lam_ref!r}"]
    if occulter is not None:
        lines.append(f"\t\tocculterRadiusFld:{occulter!r}")
    if hex_opts:
        lines += ["\t\tstart", "\t\t\tcomplexHexMaskFPM", "\t\t\tname:hexmask"]
        lines += [f"\t\t\t{k}:{v}" for k, v in hex_opts.items()]
        lines.append("\t\tend")
    lines.append("\tend")
    return "\n".join(lines)


def _tail(lam_ref):
    return f"""\tstart
\t\tmask
\t\tname:lyot
\tend
\tstart
\t\tfraunhoferFocal
\t\tname:camera
\t\treferenceLambda:{lam_ref!r}
\t\tsamplesPerflD:2.56
\t\tFOVflD:50
\tend
end
start
\tregionContrast
\treferenceLambda:{lam_ref!r}
\tradii:3, 10
\tanglesInDegrees:-180, 180
end
"""


def hex_options(lam_ref=630e-9, n_sub_pix=16, layout="standard", rings=8, mask_size=64):
    loD = F_RATIO * lam_ref
    return {"hexLayout": layout, "maskNRings": rings, "hexStep": 1, "fpmArraySize": 1024,
            "fpmScaleFactor": 1.0, "hexGap": 0.0, "nSubPix": n_sub_pix, "fpmMaskSize": mask_size,
            "pixelScale": repr(loD / 8), "fpmRadius": repr(3.0 * loD), "phaseSign": 1}


def workload0(n=1024, r_ap=500, r_gauss=400, r_lyot=450):
    seed = BASE_SEED + 0
    pix = PIX * 1024 / n
    beam_r = r_ap * pix
    script = (_telescope(2 * beam_r) + _efield(n, pix, beam_r, [LAMBDA0], [(1.0, 0.0, 0.0)]) +
              "start\n\tcoronagraph\n\tstart\n\t\tmask\n\t\tname:pupil\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:aberration\n\tend\n" +
              _fpm_block("doubleFraunhoferFPM", LAMBDA0, occulter=3.0) + "\n" + _tail(LAMBDA0))
    arrays = {("pupil", "filename"): pupil_amplitude(n, r_ap, r_gauss),
              ("aberration", "filename"): aberration_sag(seed, n),
              ("lyot", "filename"): circle(n, r_lyot)}
    return Workload(0, "mono 1024^2 pupil, occulter 3 l/D, 128^2 image", script, arrays,
                    meta={"seed": seed, "n": n, "fields": 1, "lambdas": [LAMBDA0]})


def workload_zoomprop(n=256, pad=1.0, zoom=1.0):
    """pupil -> zoomFftPropagator (dir 1) -> occulting mask on its grid ->
    zoomFftPropagator (dir -1) -> Lyot stop -> 128^2-class image, one wavelength."""
    seed = BASE_SEED + 11
    pix = PIX * 1024 / n
    r_ap = 500 * n // 1024
    beam_r = r_ap * pix
    nf = int(pad * n)
    script = (_telescope(2 * beam_r) + _efield(n, pix, beam_r, [LAMBDA0], [(1.0, 0.0, 0.0)]) +
              "start\n\tcoronagraph\n\tstart\n\t\tmask\n\t\tname:pupil\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:aberration\n\tend\n"
              f"\tstart\n\t\tzoomFftPropagator\n\t\tname:to fpm\n\t\tdir:1\n\t\tpadFactor:{pad!r}\n"
              f"\t\tzoomFactor:{zoom!r}\n\tend\n"
              "\tstart\n\t\tmask\n\t\tname:occulter\n\tend\n"
              f"\tstart\n\t\tzoomFftPropagator\n\t\tname:to lyot\n\t\tdir:-1\n\t\tpadFactor:{pad!r}\n\tend\n" +
              _tail(LAMBDA0))
    arrays = {("pupil", "filename"): pupil_amplitude(n, r_ap, 400 * n / 1024),
              ("aberration", "filename"): aberration_sag(seed, n),
              ("occulter", "filename"): 1.0 - circle(nf, 0.06 * nf),
              ("lyot", "filename"): circle(n, 450 * n / 1024)}
    return Workload(11, f"zoomFftPropagator pair around an occulter, {n}^2 pupil", script, arrays,
                    meta={"seed": seed, "n": n, "fields": 1, "lambdas": [LAMBDA0]})


def workload_fresnel_mid(n=512):
    """pupil, aberration -> zoomFftPropagator (dir 1) -> occulting mask -> 0.5 m
    angular-spectrum step -> mirror -> zoomFftPropagator (dir -1) -> Lyot stop -> image:
    a Fresnel step that reads its input field (no point-source generation before it),
    with a mask applied as its first pass reads and a mirror as its last pass writes."""
    seed = BASE_SEED + 13
    pix = PIX * 1024 / n
    r_ap = 500 * n // 1024
    beam_r = r_ap * pix
    script = (_telescope(2 * beam_r) + _efield(n, pix, beam_r, [LAMBDA0], [(1.0, 0.0, 0.0)]) +
              "start\n\tcoronagraph\n\tstart\n\t\tmask\n\t\tname:pupil\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:aberration\n\tend\n"
              "\tstart\n\t\tzoomFftPropagator\n\t\tname:to fpm\n\t\tdir:1\n\t\tpadFactor:1.0\n"
              "\t\tzoomFactor:1.0\n\tend\n"
              "\tstart\n\t\tmask\n\t\tname:occulter\n\t\tposition:0\n\tend\n"
              "\tstart\n\t\tfresnelPropagator\n\t\tname:gap\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:late\n\t\tposition:0.5\n\tend\n"
              "\tstart\n\t\tzoomFftPropagator\n\t\tname:to lyot\n\t\tdir:-1\n\t\tpadFactor:1.0\n\tend\n" +
              _tail(LAMBDA0))
    arrays = {("pupil", "filename"): pupil_amplitude(n, r_ap, 400 * n / 1024),
              ("aberration", "filename"): aberration_sag(seed, n),
              ("occulter", "filename"): 1.0 - circle(n, 0.06 * n),
              ("late", "filename"): aberration_sag(seed + 1, n),
              ("lyot", "filename"): circle(n, 450 * n / 1024)}
    return Workload(13, f"Fresnel step between two zoomFftPropagators, {n}^2", script, arrays,
                    meta={"seed": seed, "n": n, "fields": 1, "lambdas": [LAMBDA0]})


def workload_complex_masks(n=512):
    """complexMask elements (src/elements/complexMask.cpp): a centred complex factor in
    the pupil plane (before the first MFT: the general pointwise pass), one before and
    one after a 0.5 m Fresnel step (applied by its first and last passes), and a mirror
    after the second; the window sizes differ from the field's so the centring counts."""
    seed = BASE_SEED + 14
    pix = PIX * 1024 / n
    r_ap = 500 * n // 1024
    beam_r = r_ap * pix
    rng = np.random.Generator(np.random.Philox(seed))
    script = (_telescope(2 * beam_r) + _efield(n, pix, beam_r, [LAMBDA0], [(1.0, 0.0, 0.0)]) +
              "start\n\tcoronagraph\n\tstart\n\t\tmask\n\t\tname:pupil\n\tend\n"
              "\tstart\n\t\tcomplexMask\n\t\tname:apod\n\tend\n"
              "\tstart\n\t\tzoomFftPropagator\n\t\tname:to fpm\n\t\tdir:1\n\t\tpadFactor:1.0\n"
              "\t\tzoomFactor:1.0\n\tend\n"
              "\tstart\n\t\tcomplexMask\n\t\tname:cm1\n\t\tposition:0\n\tend\n"
              "\tstart\n\t\tfresnelPropagator\n\t\tname:gap\n\tend\n"
              "\tstart\n\t\tcomplexMask\n\t\tname:cm2\n\t\tposition:0.5\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:late\n\tend\n"
              "\tstart\n\t\tzoomFftPropagator\n\t\tname:to lyot\n\t\tdir:-1\n\t\tpadFactor:1.0\n\tend\n" +
              _tail(LAMBDA0))

    def cmask(m):
        yy, xx = np.mgrid[0:m, 0:m] / m
        amp = 0.8 + 0.2 * np.cos(2 * np.pi * (xx + 0.5 * yy)) * rng.uniform(0.5, 1.0)
        ph = rng.uniform(-1, 1) * np.sin(2 * np.pi * xx * 1.5) + rng.uniform(-1, 1) * np.cos(2 * np.pi * yy * 2.5)
        return amp * np.exp(-1j * ph)
    arrays = {("pupil", "filename"): pupil_amplitude(n, r_ap, 400 * n / 1024),
              ("late", "filename"): aberration_sag(seed + 1, n),
              ("lyot", "filename"): circle(n, 450 * n / 1024)}
    complex_arrays = {("apod", "mask", 0): cmask(int(0.8 * n) + 1),
                      ("cm1", "mask", 0): cmask(int(0.75 * n)),
                      ("cm2", "mask", 0): cmask(int(0.6 * n) + 1)}
    return Workload(14, f"complexMask elements around a Fresnel step, {n}^2", script, arrays, complex_arrays,
                    meta={"seed": seed, "n": n, "fields": 1, "lambdas": [LAMBDA0]})


def workload1(n=2048, fresnel_pad=1, sources=((1.0, 0.0, 0.0),)):
    """2048^2 field, PIAA mirrors around a 1 m Fresnel step, downsampled to 1024^2.
    fresnel_pad 2: the step on csim's 2n zero pad (fresnelPropagator padFactor).
    sources: point sources (flux, tipX, tiltY), one on axis by default."""
    seed = BASE_SEED + 1
    pix = PIX
    r_ap = 1000 * n // 2048                  # 1000 px at 2048
    beam_r = r_ap * pix
    h1, c1 = piaa_sag(seed, 2, n, r_ap, radial=0.5e-6)
    h2, c2 = piaa_sag(seed, 3, n, r_ap, radial=-0.5e-6)
    script = (_telescope(2 * beam_r) + _efield(n, pix, beam_r, [LAMBDA0], list(sources)) +
              "start\n\tcoronagraph\n"
              "\tstart\n\t\tmask\n\t\tname:pupil\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:aberration\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:piaa1\n\t\tposition:0\n\tend\n"
              "\tstart\n\t\tfresnelPropagator\n\t\tname:piaa gap\n" +
              (f"\t\tpadFactor:{fresnel_pad}\n" if fresnel_pad != 1 else "") + "\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:piaa2\n\t\tposition:1.0\n\t\tsign:-1\n\tend\n"
              "\tstart\n\t\tdownsample\n\t\tname:ds\n\t\tsampleFactor:2\n\tend\n" +
              _fpm_block("doubleFraunhoferFPM", LAMBDA0, occulter=3.0) + "\n" + _tail(LAMBDA0))
    arrays = {("pupil", "filename"): pupil_amplitude(n, r_ap, 800 * n / 2048),
              ("aberration", "filename"): aberration_sag(seed, n),
              ("piaa1", "filename"): h1, ("piaa2", "filename"): h2,
              ("lyot", "filename"): circle(n, 900 * n / 2048)}
    return Workload(1, "2048^2 PIAA + 1 m Fresnel + downsample + MFT chain", script, arrays,
                    meta={"seed": seed, "n": n, "fields": len(sources), "lambdas": [LAMBDA0], "zernike": (c1, c2)})


def poly_lambdas(n=9):
    return [600e-9 + 7.5e-9 * i for i in range(n)]


def workload2(n=1024, n_sub_pix=16, n_lambda=9, offsets=(0.0, 2.0, 6.0), layout="standard", rings=8):
    """Batched polychromatic x 3 sources through a 217-hex phase FPM."""
    seed = BASE_SEED + 2
    pix = PIX * 1024 / n
    r_ap = 500 * n // 1024
    beam_r = r_ap * pix
    Lx_over_D = n / (2.0 * r_ap)
    sources = [(1.0, t * Lx_over_D / 2.0, 0.0) for t in offsets]   # tipX = t Lx / (2 D)
    lams = poly_lambdas(n_lambda)
    lam_ref = 630e-9
    hopt = hex_options(lam_ref, n_sub_pix, layout, rings)
    script = (_telescope(2 * beam_r) + _efield(n, pix, beam_r, lams, sources) +
              "start\n\tcoronagraph\n\tstart\n\t\tmask\n\t\tname:pupil\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:aberration\n\tend\n" +
              _fpm_block("doubleFraunhoferInterpFPM", lam_ref, hex_opts=hopt) + "\n" + _tail(lam_ref))
    n_hex = 1 + 3 * rings * (rings + 1) if layout == "standard" else None
    t = hex_thickness(seed, 1, n_hex)[0] if n_hex else None
    arrays = {("pupil", "filename"): pupil_amplitude(n, r_ap, 400 * n / 1024),
              ("aberration", "filename"): aberration_sag(BASE_SEED + 0, n),
              ("lyot", "filename"): circle(n, 450 * n / 1024)}
    return Workload(2, f"{n_lambda} lambda x {len(offsets)} sources, {n_hex}-hex FPM, {n}^2 pupil",
                    script, arrays,
                    meta={"seed": seed, "n": n, "fields": n_lambda * len(offsets), "lambdas": lams,
                          "thickness": t, "sags": thickness_to_sag(t) if t is not None else None,
                          "n_hex": n_hex, "hex": hopt})


def workload2_dark(n=1024, n_lambda=9):
    """Workload 2's pupil, wavelengths and image with the on-axis source only and the
    3 lambda/D occulting disk in place of the random hex mask: a real dark hole
    (mean contrast ~1e-6), on which the 1e-11 absolute-contrast gate is evaluated."""
    seed = BASE_SEED + 2
    pix = PIX * 1024 / n
    r_ap = 500 * n // 1024
    beam_r = r_ap * pix
    lams = poly_lambdas(n_lambda)
    lam_ref = 630e-9
    script = (_telescope(2 * beam_r) + _efield(n, pix, beam_r, lams, [(1.0, 0.0, 0.0)]) +
              "start\n\tcoronagraph\n\tstart\n\t\tmask\n\t\tname:pupil\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:aberration\n\tend\n" +
              _fpm_block("doubleFraunhoferFPM", lam_ref, occulter=3.0) + "\n" + _tail(lam_ref))
    arrays = {("pupil", "filename"): pupil_amplitude(n, r_ap, 400 * n / 1024),
              ("aberration", "filename"): aberration_sag(BASE_SEED + 0, n),
              ("lyot", "filename"): circle(n, 450 * n / 1024)}
    return Workload(2, f"{n_lambda} lambda on-axis, occulter 3 l/D, {n}^2 pupil (dark hole)", script, arrays,
                    meta={"seed": seed, "n": n, "fields": n_lambda, "lambdas": lams})


def workload3(n=1024, n_sub_pix=16, n_lambda=9):
    """One pupil size of the DE batch (on-axis, hex FPM); 512^2 halves every radius."""
    w = workload2(n=n, n_sub_pix=n_sub_pix, n_lambda=n_lambda, offsets=(0.0,))
    w.wid = 3
    w.name = f"DE batch member: {n_lambda} lambda, {w.meta['n_hex']}-hex FPM, {n}^2 pupil"
    return w


def workload_native(kind="cube", n=256, pad=0.25, n_lambda=3, offsets=(0.0, 2.0), mask_index=None):
    """csim's own FPM geometry, none of the additive keys.
    cube: doubleFraunhoferFPM (doubleFraunhoferFPM.cpp:45-97): lambda*f = 3.52 lambda, FPM
      pixel dxprime = 2 (R / (dx N zoomFactor)) lambda 80 on the offset grid (i - pad N / 2)
      dxprime, a two-slice complex mask cube (maskFilename amp,ph; bound from memory as
      "mask") placed at the centre, slice maskIndex (unset: the latch to slice 0).
    hex: doubleFraunhoferInterpFPM (doubleFraunhoferInterpFPM.cpp:86-125): the FPM grid
      pixel is the hex mask's, which complexHexMaskFPM derives from lambdaRef and
      fpmFRatio (complexHexMaskFPM.cpp:197-203); lambda*f = 3.52 lambda."""
    seed = BASE_SEED + 12
    pix = PIX * 1024 / n
    r_ap = 500 * n // 1024
    beam_r = r_ap * pix
    Lx_over_D = n / (2.0 * r_ap)
    sources = [(1.0, t * Lx_over_D / 2.0, 0.0) for t in offsets]
    lams = poly_lambdas(n_lambda)
    lam_ref = 630e-9
    lines = ["\tstart", "\t\tdoubleFraunhoferFPM" if kind == "cube" else "\t\tdoubleFraunhoferInterpFPM",
             "\t\tname:fpm", "\t\tdisableForCalibration:on", f"\t\tpadFactor:{pad!r}", "\t\tzoomFactor:1"]
    complex_arrays, meta = {}, {"seed": seed, "n": n, "fields": n_lambda * len(offsets), "lambdas": lams}
    if kind == "cube":
        if mask_index is not None:
            lines.append(f"\t\tmaskIndex:{mask_index}")
        S = max(8, int(n * pad) // 4)
        x = np.arange(S) - (S - 1) / 2.0
        spot = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2) < S / 5.0
        for sl in range(2):
            amp = np.where(spot, 0.05 * sl, 0.6 + 0.4 * philox(seed, 2 + 2 * sl, S * S).reshape(S, S))
            ph = 2 * math.pi * philox(seed, 3 + 2 * sl, S * S).reshape(S, S)
            complex_arrays[("fpm", "mask", sl)] = amp * np.exp(-1j * ph)   # initMask: amp exp(-i ph)
    elif kind == "hex":
        S, rings = 32, 3
        lamf_over_D = 3.52 * lam_ref / (2 * beam_r)
        fratio = 0.25 * lamf_over_D / (2 * (r_ap / S) * lam_ref)      # hex pixel ~ lambda f / (4 D)
        pixel = 2 * (r_ap / S) * lam_ref * fratio
        lines += ["\t\tstart", "\t\t\tcomplexHexMaskFPM", "\t\t\tname:hexmask", "\t\t\thexLayout:standard",
                  f"\t\t\tmaskNRings:{rings}", "\t\t\thexStep:1", "\t\t\tfpmArraySize:512",
                  "\t\t\tfpmScaleFactor:1.0", "\t\t\thexGap:0.02", "\t\t\tnSubPix:4", f"\t\t\tfpmMaskSize:{S}",
                  f"\t\t\tlambdaRef:{lam_ref!r}", f"\t\t\tfpmFRatio:{fratio!r}",
                  f"\t\t\tfpmRadius:{S / 2 * pixel!r}", "\t\t\tphaseSign:1", "\t\tend"]
        n_hex = 1 + 3 * rings * (rings + 1)
        t = hex_thickness(seed, 1, n_hex)[0]
        meta.update(n_hex=n_hex, thickness=t, sags=thickness_to_sag(t))
    else:
        raise ValueError(kind)
    lines.append("\tend")
    script = (_telescope(2 * beam_r) + _efield(n, pix, beam_r, lams, sources) +
              "start\n\tcoronagraph\n\tstart\n\t\tmask\n\t\tname:pupil\n\tend\n"
              "\tstart\n\t\tmirror\n\t\tname:aberration\n\tend\n" +
              "\n".join(lines) + "\n" + _tail(lam_ref))
    arrays = {("pupil", "filename"): pupil_amplitude(n, r_ap, 400 * n / 1024),
              ("aberration", "filename"): aberration_sag(seed, n),
              ("lyot", "filename"): circle(n, 450 * n / 1024)}
    return Workload(12, f"native {kind} FPM geometry, pad {pad}, {n}^2 pupil", script, arrays, complex_arrays, meta)


def apply(chain, w: Workload):
    for (elem, key), a in w.arrays.items():
        chain.set_array(elem, key, a)
    for (elem, key, sl), a in w.complex_arrays.items():
        chain.set_complex_array(elem, key, a, sl)
    if w.meta.get("sags") is not None:
        chain.set_opt_data("fpm", "sags", w.meta["sags"])


def workload_p2l(mask="hex", n=256, n_lambda=2, offsets=(0.0, 2.0), rings=3, mask_size=64, n_sub_pix=4):
    """pupil -> fpmPupToLyot (2n x 2n convolution) -> Lyot stop -> 128^2-class image."""
    seed = BASE_SEED + 10
    pix = PIX * 1024 / n
    r_ap = 500 * n // 1024
    beam_r = r_ap * pix
    Lx_over_D = n / (2.0 * r_ap)
    sources = [(1.0, t * Lx_over_D / 2.0, 0.0) for t in offsets]
    lams = poly_lambdas(n_lambda)
    lam_ref = 630e-9
    loD = F_RATIO * lam_ref
    lines = ["	start", "		fpmPupToLyot", "		name:fpm", "		disableForCalibration:on",
             f"		focalRatio:{F_RATIO!r}", "		fRatioSign:-1", "		zoomFftSign:-1"]
    arrays = {("pupil", "filename"): pupil_amplitude(n, r_ap, 400 * n / 1024),
              ("lyot", "filename"): circle(n, 450 * n / 1024)}
    complex_arrays = {}
    meta = {"seed": seed, "n": n, "fields": n_lambda * len(offsets), "lambdas": lams}
    if mask == "hex":
        hopt = hex_options(lam_ref, n_sub_pix, "standard", rings, mask_size=mask_size)
        lines += ["		start", "			fpmIntHexCMCForPupToLyot", "			zoomFactor:1", "			start",
                  "				complexHexMaskFPM", "				name:hexmask"]
        lines += [f"				{k}:{v}" for k, v in hopt.items()]
        lines += ["			end", "		end"]
        n_hex = 1 + 3 * rings * (rings + 1)
        t = hex_thickness(seed, 1, n_hex)[0]
        meta.update(n_hex=n_hex, thickness=t, sags=thickness_to_sag(t), hex=hopt)
    elif mask == "cmc":
        S = 48
        amp = 0.5 + 0.5 * philox(seed, 2, S * S).reshape(S, S)
        ph = 2 * math.pi * philox(seed, 3, S * S).reshape(S, S)
        complex_arrays[("fpm", "mask", 0)] = amp * np.exp(-1j * ph)
        lines += ["		start", "			fpmCMCForPupToLyot", f"			fpmMatScale:{loD / 6!r}", "		end"]
    elif mask == "binary":
        S = 64
        outer = 4.0
        x = (np.arange(S) - (S - 1) / 2.0) * (2 * outer / S)
        rr = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2)
        arrays[("fpm", "maskFilename")] = (rr < 2.5).astype(np.float64)
        arrays[("fpm", "calibFilename")] = np.ones((S, S))
        lines += ["		start", "			fpmBinaryForPupToLyot", f"			referenceLambda:{lam_ref!r}",
                  "			innerRadiusFld:0", f"			outerRadiusFld:{outer!r}", "		end"]
    else:
        raise ValueError(mask)
    lines.append("	end")
    script = (_telescope(2 * beam_r) + _efield(n, pix, beam_r, lams, sources) +
              "start\n\tcoronagraph\n\tstart\n\t\tmask\n\t\tname:pupil\n\tend\n" +
              "\n".join(lines) + "\n" + _tail(lam_ref))
    return Workload(10, f"fpmPupToLyot ({mask}) {n}^2 pupil, {n_lambda} lambda x {len(offsets)} sources",
                    script, arrays, complex_arrays, meta)
