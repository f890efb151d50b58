/*
 * csx_oracle.h -- FP64 CPU restatement of csim's E-field propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product in csim_amd/ and the timed CPU baseline of bench.py.  Product code
 * never links, loads or calls it.
 *
 * Conventions (identical to csim / Armadillo):
 *   - matrices are column-major: element (r, c) of an R x C matrix sits at
 *     index c*R + r (row index = y, column index = x);
 *   - complex numbers are interleaved (re, im) doubles (C99 double _Complex).
 * Every function cites the reference file:line it restates.
 */
#ifndef CSX_ORACLE_H
#define CSX_ORACLE_H
#include <complex.h>
#include <stdint.h>

typedef double _Complex ocx;

/* ---- FFT (FFTW_FORWARD / FFTW_BACKWARD semantics, src/lib/csim_fft.cpp:98-291) ---- */
void orc_fft1d(ocx* x, int n, int sign);                    /* in place, unnormalised */
void orc_fft2d(ocx* a, int rows, int cols, int sign);       /* in place, unnormalised */
void orc_fft2_csim(ocx* a, int rows, int cols);             /* forward, unnormalised */
void orc_ifft2_csim(ocx* a, int rows, int cols);            /* backward, / (rows*cols) */
void orc_fft_shift(const ocx* in, ocx* out, int rows, int cols);
void orc_fft_shift_real(const double* in, double* out, int rows, int cols);

/* ---- chirp-z (Bluestein) and zoomFft (src/lib/csim_fft.cpp:422-826) ---- */
int  orc_czt_nfft(int m, int k);
void orc_czt1d_cols(const ocx* in, int m, int ncols, int k, double w, double a, ocx* out);
void orc_czt2d(const ocx* in, int rows, int cols, int k,
               double wx, double ax, double wy, double ay, ocx* out);
/* zoomFft::init(arrayGeom, arrayGeom, lambdaF) + execute: out is nOut x nOut */
void orc_zoomfft(const ocx* in, int rows, int cols,
                 double inX0, double inY0, double inDx, double inDy,
                 int nOut, double outX0, double outY0, double outDx, double outDy,
                 double lambdaF, ocx* out);
/* Closed form of the same transform as two dense DFT products (rectangular out). */
void orc_mft_direct(const ocx* in, int rows, int cols,
                    double inX0, double inY0, double inDx, double inDy,
                    int outRows, int outCols, double outX0, double outY0,
                    double outDx, double outDy, double lambdaF, ocx* out);

/* ---- angular-spectrum Fresnel (src/lib/csim_fft.cpp:371-392) ---- */
void orc_fresnel_as(ocx* E, int n, double lambda, double apRad, double z);

/* ---- geometry / resampling (src/lib/csim_geom.cpp:113-266) ---- */
void orc_grid_symmetric(int n, double pixelSize, double origin, double* x);
void orc_grid_offset(int n, double pixelSize, double* x);
void orc_interp1_linear(const double* xg, const double* yg, int ng,
                        const double* xi, double* yi, int ni, double extrap);
void orc_interp2(const double* in, int rows, int cols, double* out, int orows, int ocols,
                 const double* inX, const double* inY, const double* outX, const double* outY);
int  orc_downsample(const ocx* in, int rows, int cols, int factor,
                    double pixelSizeX, double pixelSizeY, double originX, double originY,
                    ocx* out);

/* ---- pointwise elements ---- */
void orc_mask_apply(ocx* E, const double* m, int n);                      /* mask.cpp:32-46 */
void orc_mirror_apply(ocx* E, const double* h, int n, double sign, double lambda); /* mirror.cpp:66-72 */
void orc_point_source(ocx* E, int rows, int cols, double pixelScale,
                      double flux, double tipX, double tiltY);            /* efield.cpp:179-206 */

/* ---- hex-cell FPM (src/elements/complexHexMaskFPM.cpp) ---- */
int  orc_hex_centers_csim(int maskNRings, double hexStep, double* hx, double* hy,
                          int* ring, int maxHex);
int  orc_hex_centers_standard(int nRings, double hexStep, double* hx, double* hy,
                              int* ring, int maxHex);
void orc_hex_design_array(const double* hx, const double* hy, int nHex, double layoutRadius,
                          int fpmArraySize, double fpmScaleFactor, double hexGap,
                          int32_t* design);
void orc_hex_interp_data(const int32_t* design, int fpmArraySize, int fpmMaskSize,
                         int nSubPix, double pixelScale, double fpmRadius,
                         double fpmScaleFactor, int32_t* nsub, int32_t* hexnum);
void orc_hex_mask(const int32_t* nsub, const int32_t* hexnum, int S, int nSubPix,
                  const double* sags, int nHex, double lambda, double phaseSign,
                  int useOnlyThisHex, ocx* mask);

/* ---- doubleFraunhoferFPM centred mask multiply (doubleFraunhoferFPM.cpp:84-97) ---- */
void orc_center_mask_multiply(ocx* F, int rows, int cols, const ocx* m, int mrows, int mcols);

/* ---- convolution-form FPM (src/elements/fpmPupToLyot.cpp:24-167) ---- */
void orc_fpm_pup_to_lyot(ocx* E, int n, double pixelSize, const ocx* maskMinus1, int S,
                         double maskPixelSize, double lambdaF, double zoomFftSign,
                         double fRatioSign, int babinet);

/* ---- contrast (src/tools/regionContrast.cpp:89-205) ---- */
void orc_intensity_accumulate(const ocx* E, int n, double* Iacc);
int  orc_region_pixels(int n, double pixelSize, double loD, double xlim1, double xlim2,
                       double r1, double r2, double a1deg, double a2deg, int32_t* idx);

/* ---- linear hex-response objective (src/optimization/optObjective.cpp:24-69) ---- */
void orc_hex_resp_objective(const ocx* M, int npix, int ncol, int nslices,
                            const double* lambdas, double phaseSign, double calib,
                            const double* X, int nopt, int ncand, int objMax, double* obj);

/* ---- small dense MLP in FP64 (gradient oracle for the surrogate) ---- */
void orc_mlp_forward(const double* x, int nb, int nin, const double* W1, const double* b1,
                     int nh1, const double* W2, const double* b2, int nh2,
                     const double* W3, const double* b3, int nout,
                     double* h1, double* a1, double* h2, double* a2, double* y);
double orc_mlp_loss_grad(const double* x, const double* t, int nb, int nin,
                         const double* W1, const double* b1, int nh1,
                         const double* W2, const double* b2, int nh2,
                         const double* W3, const double* b3, int nout,
                         double* gW1, double* gb1, double* gW2, double* gb2,
                         double* gW3, double* gb3);

int orc_num_threads(void);
#endif
