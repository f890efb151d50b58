// Pointwise, FFT, resampling, hex-mask and reduction kernels (gfx950).
#pragma once
#include "common.hpp"

namespace csx {

// E[f][p] *= (1+i) m[p]            (csim mask, src/elements/mask.cpp:32-46)
int k_apply_mask(double2* E, long npix, int nfield, const double* m, hipStream_t s);
// E[f][p] *= exp(-sign 4 pi i h[p] / lambda[f])   (mirror.cpp:66-72); lambdas on device
int k_apply_mirror(double2* E, long npix, int nfield, const double* h, double sign,
                   const double* lambdas_dev, hipStream_t s);
// E[f][p] *= P[f / per][p]   complex factor shared by `per` consecutive fields
int k_apply_cfactor(double2* E, long npix, int nfield, const double2* P, int per, hipStream_t s);
// Point sources (efield.cpp:179-206): E[f] = flux_f exp(i 2 pi (2 x tipX_f/Lx + 2 y tiltY_f/Ly))
int k_point_sources(double2* E, int rows, int cols, int nfield, const double* params_dev,
                    double pixel_scale, hipStream_t s);

// Fused pointwise pass over a batch: optional point-source generation followed by
// up to kMaxPwOps element factors, one read and one write of every field value.
enum PwKind : int { PW_SOURCE = 0, PW_MASK_1PI = 1, PW_MIRROR = 2, PW_CFACTOR = 3 };
constexpr int kMaxPwOps = 6;
struct PwOp { int kind; const void* p; double sign; int per; };   // per: fields sharing a PW_CFACTOR plane
enum PwOut : int { PW_OUT_C128 = 0, PW_OUT_C64 = 1, PW_OUT_PLANES = 2 };
struct PwChain {
  PwOp op[kMaxPwOps];
  int nops;
  int rows, cols;
  double pixel_scale;               // PW_SOURCE grid
  const double* src_params;         // PW_SOURCE: flux, tipX, tiltY per field
  const double* lambdas;            // PW_MIRROR: per field
  int out_kind;                     // PwOut: in place, or to `out` at a working precision
  void* out;                        // complex64 fields or BF16 planes (field f at f * 2 terms npix)
  int terms;
  // > 0: pixel-major generation pass.  Requires op[0] = PW_SOURCE, no PW_CFACTOR, an
  // even pixel count, nl <= kMaxPwLam, lambdas[f] = lambdas[f % nl] and source
  // parameters that depend on f / nl only (the chain's field order).
  int nl;
};
constexpr int kMaxPwLam = 16;
int pointwise_chain(double2* E, int nfield, const PwChain& pc, hipStream_t s);

// 2-D FFT, csim semantics: dir -1 forward unnormalised, +1 inverse / n^2.
int fft2_batched(double2* E, int n, int nfield, int dir, hipStream_t s);
// Angular-spectrum Fresnel step (src/lib/csim_fft.cpp:371-392); lambdas on device.
// pre / post (optional): pointwise chains on the n x n fields applied as the first pass
// reads them and as the last pass writes them (the elements around the step, fused).
int fresnel_batched(double2* E, int n, int nfield, const double* lambdas_dev, double ap_rad,
                    double z, hipStream_t s, const PwChain* pre = nullptr, const PwChain* post = nullptr);

// Separable linear resampling with precomputed (index, weight) tables, extrap 0.
struct InterpTable { int lo, hi; double w; };
int resample_batched(const double2* in, int rows, int cols, int nfield, double2* out, int orows,
                     int ocols, const InterpTable* ty_dev, const InterpTable* tx_dev, hipStream_t s);

// Hex FPM masks for (candidate, lambda) pairs (complexHexMaskFPM.cpp:44-99).
// use_only: -1 every hex, h one hex (h == nhex: only the area outside the hexes),
// kUseOnlyPerCandidate: candidate c uses only hex c (the response batch).
constexpr int kUseOnlyPerCandidate = -2;
int hex_masks(const int32_t* nsub, const int32_t* hexnum, int S, int nsubpix, const double* sags,
              int nhex, int ncand, const double* lambdas_dev, int nlam, double phase_sign,
              int use_only, double2* out, hipStream_t s);

// I[g][p] = sum_{f < per} |E[g*per + f][p]|^2
int intensity_groups(const double2* E, long npix, int ngroups, int per, double* I, hipStream_t s);
// out[g] = max_p I[g][p]
int max_groups(const double* I, long npix, int ngroups, double* out, hipStream_t s);
// out[g] = mean_{idx} I[g][idx] / peak_dev[0] (peak on device) or / peak
int region_mean_groups(const double* I, long npix, int ngroups, const int32_t* idx, int nidx,
                       const double* peak_dev, double peak, double* out, hipStream_t s);

// out[b] = mean of I over pixels (column-major ny x nx grid) with s[b] < r / loD < s[b + 2],
// divided by peak_dev[0]; b < nbins (samples_dev holds nbins + 2 values)
int annulus_means(const double* I, int nx, int ny, double x0, double dx, double y0, double dy, double loD,
                  const double* samples_dev, int nbins, const double* peak_dev, double* out, hipStream_t s);

// out[f][b] = sum_{p : bin[p] == b} |E_f[p]|^2 over nfield fields (bins < 0 skipped)
int bin_sums(const double2* E, long npix, int nfield, const int32_t* bin, int nbins, double* out, hipStream_t s);

// Centred copy: dst (drows x dcols) <- src (srows x scols) placed at the centre
// offsets csim uses (rows/2 - srows/2); zero elsewhere (pad) or crop.
int embed_center(const double2* src, int srows, int scols, double2* dst, int drows, int dcols,
                 int nfield, int crop, hipStream_t s);

// Convolution-form FPM (fpmPupToLyot.cpp:24-167), square even fields, 2n padding.
// out[k] = factor (src[k] - (k >= minus1_from)) over n masks of rows x cols (columns mirrored if fliplr)
int fpm_prep(const double2* src, long src_stride, int rows, int cols, int n, int fliplr, int minus1_from,
             double factor, double2* out, hipStream_t s);
// P[f] = H[cand(f) * nl + l(f)] Pin[f % nin] (-1)^(r+c), cand = f / per_cand (clamped),
// l = f % nl; Pin == P (with nin == nfield) multiplies in place
int spectrum_mul(double2* P, const double2* Pin, int nin, int n2, int nfield, const double2* H, int per_cand, int nl,
                 int nmask_cand, hipStream_t s);
// E[f] = scale[f % nl] * centre(P[f]) + (f >= babinet_from ? Ein[f % nin] : 0); Ein == E
// (with nin == nfield) adds in place
int crop_add(double2* E, const double2* Ein, int nin, int n, const double2* P, int n2, int nfield, int nl,
             const double2* scale_dev, int babinet_from, hipStream_t s);
// M[s * nl + l][c][j] = E[(c * nslice + s * nl + l)][idx[j]], nslice = nsrc * nl
// Babinet window of a candidate batch (doubleFraunhofer(Interp)FPM with the hex mask
// transparent outside its S x S window): W[f] = (M[cand, lambda] - 1) * E[f_in] on the
// window at (row0, col0) of the FPM grid (rows x ...), f = (cand * ns + s) * nl + l,
// f_in = s * nl + l, masks of candidate min(cand, ncm - 1).  kind: 0 complex128,
// 1 complex64, 2 `terms` BF16 planes (input field stride e_npix, output S * S).
int babinet_window(const void* E, int kind, int terms, long e_npix, int rows, const double2* masks, int S,
                   int row0, int col0, int nfield, int nl, int ns, int ncm, void* W, hipStream_t s);
int gather_response(const double2* E, long npix_field, const int32_t* idx, int nidx, int ncol, int nslice,
                    double2* M, hipStream_t s);

}  // namespace csx
