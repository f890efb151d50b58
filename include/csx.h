/*
 * csx.h -- C ABI of the MI355X-native csim E-field propagation path.
 *
 * Two levels:
 *   1. Element/script API: a csim input script (same grammar and keys) builds a
 *      chain of elements that runs on one GPU over a batch of fields
 *      (candidate x source x wavelength).  Host buffers, csim layout.
 *   2. Operator API: the hot kernels on device pointers (HIP memory) and a HIP
 *      stream, for callers that keep fields resident in HBM.
 *
 * Conventions
 *   - Complex values are interleaved (re, im) doubles / floats.
 *   - A field of rows x cols is stored column-major exactly like csim's
 *     arma::cx_mat: element (r, c) at index c*rows + r.  Batches of fields are
 *     contiguous, field-major.
 *   - Every function returns CSX_OK (0) or a negative CSX_ERR_* code; the
 *     message of the last failure on the calling thread is csx_last_error().
 *     Nothing aborts and no C++ exception crosses this boundary.
 *   - A context is used by one host thread at a time; distinct contexts are
 *     independent.
 */
#ifndef CSX_H
#define CSX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSX_OK 0
#define CSX_ERR_ARG (-1)
#define CSX_ERR_HIP (-2)
#define CSX_ERR_PARSE (-3)
#define CSX_ERR_STATE (-4)
#define CSX_ERR_NOMEM (-5)
#define CSX_ERR_UNSUPPORTED (-6)
#define CSX_ERR_NUMERIC (-7)

/* Arithmetic of the matrix-Fourier-transform contractions.  Phase arguments
 * of every DFT kernel are always formed in FP64. */
typedef enum {
  CSX_FP64 = 0,     /* FP64 MFMA, the plain path */
  CSX_FP32 = 1,     /* FP32 MFMA */
  CSX_BF16X3 = 2,   /* 2-term split BF16, 3 products, FP32 accumulation */
  CSX_BF16X6 = 3    /* 3-term split BF16, 6 products, FP32 accumulation, FP64 across K blocks */
} csx_precision;

typedef struct csx_ctx csx_ctx;
typedef struct csx_chain csx_chain;

/* ---------------------------------------------------------------- context */
/* Replaces csim's global state set up in src/csim.cpp:28-50 (init_fft_lib,
 * globalTelescope/globalCoronagraph singletons) with an explicit context bound
 * to one HIP device.  flags: 0 = the context creates its own non-blocking
 * stream; CSX_CTX_DEFAULT_STREAM = all work goes to the device's default (null)
 * stream, ordered with a framework that allocates and computes there. */
#define CSX_CTX_DEFAULT_STREAM 1u
int csx_ctx_create(int device, unsigned flags, csx_ctx** out);
void csx_ctx_destroy(csx_ctx* ctx);
/* Stream for all subsequent work of this context (NULL = the context's own).
 * Device buffers produced on another stream must be complete before use. */
int csx_ctx_set_stream(csx_ctx* ctx, void* hip_stream);
int csx_ctx_synchronize(csx_ctx* ctx);
const char* csx_last_error(void);
const char* csx_version(void);
/* Contraction tile tuner (no reference counterpart; csim's nearest relative is FFTW's
 * wisdom file, src/lib/csim_fft.cpp:20-36): the (tile variant, K split) of every
 * grouped launch is timed once per problem signature and kept in a JSON file
 * (CSX_TUNE_CACHE, default tune_cache.json beside the library) that later processes
 * read, so every run launches the same kernels.  JSON summary of this process's
 * lookups {"file", "kernels", "from_file", "timed", "choices": {signature: [variant,
 * ksplit, source]}}; returns its length (buf NULL: only the length). */
int csx_tune_report(char* buf, size_t cap);

/* ------------------------------------------------------ element / script API */
/* Parse a csim input script (src/lib/csim_parser.cpp:60-190 grammar; blocks
 * telescope, efield, coronagraph, regionContrast as in src/csim.cpp:54-108)
 * and build the element chain (src/coronagraph/coronagraph.cpp:50-168).
 * FITS `filename:` keys are read from disk; csx_chain_set_array binds the same
 * data from memory instead. */
int csx_chain_from_script(csx_ctx* ctx, const char* script_text, csx_chain** out);
void csx_chain_destroy(csx_chain* chain);

/* Bind a real array (rows x cols, column-major) to an element key that
 * csim reads from a FITS file: mask "filename", mirror "filename",
 * complexHexMaskFPM "sagFilename" (rows = nHex, cols = 1).  Additive API. */
int csx_chain_set_array(csx_chain* chain, const char* element_name, const char* key,
                        const double* data, int rows, int cols);
/* Bind a complex array (rows x cols, interleaved, column-major): efield "field"
 * (one source x lambda slice, replaces csim's efield filename: re,im) and
 * doubleFraunhoferFPM "mask" (replaces maskFilename: amp,ph). */
int csx_chain_set_complex_array(csx_chain* chain, const char* element_name, const char* key,
                                const double* data, int rows, int cols, int slice);

/* celem::get/set_optimization_data (src/elements/celem.hpp:39-40,
 * src/coronagraph/coronagraph.cpp:253-268) with the intended virtual dispatch:
 * "position" on any element; "sags", "useOnlyThisHex" on complexHexMaskFPM. */
int csx_chain_set_opt_data(csx_chain* chain, const char* element_name, const char* data_name,
                           const double* v, size_t n);
int csx_chain_get_opt_data(csx_chain* chain, const char* element_name, const char* data_name,
                           double* v, size_t* n);

/* Candidate batch axis: n_cand sag vectors (n_hex each, candidate-major) for the
 * hex FPM named element_name (doubleFraunhoferInterpFPM, or fpmPupToLyot with an
 * fpmIntHexCMCForPupToLyot mask).  Every propagation then runs
 * n_cand x n_sources x n_lambda fields. n_cand = 0 returns to the single sag set. */
int csx_chain_set_candidates(csx_chain* chain, const char* element_name,
                             const double* sags, int n_hex, int n_cand);

int csx_chain_set_precision(csx_chain* chain, csx_precision p);

typedef struct {
  int n_sources, n_lambdas, n_candidates;
  int in_rows, in_cols;       /* efield size */
  int out_rows, out_cols;     /* field size after the last element */
  double out_pixel_size;      /* physical pixel of the output plane */
  double lambda_over_d;       /* telescope loD at the region's reference lambda */
} csx_chain_info;
int csx_chain_get_info(csx_chain* chain, csx_chain_info* info);
int csx_chain_get_lambdas(csx_chain* chain, double* lambdas, int n);

/* coronagraph::execute (src/coronagraph/coronagraph.cpp:175-200) on every field
 * of the batch, with the calibration switch (coronagraph.hpp:64-65). */
int csx_chain_propagate(csx_chain* chain, int calibrating);
/* Output field of (candidate, source, lambda index), out_rows x out_cols complex. */
int csx_chain_get_field(csx_chain* chain, int cand, int source, int lambda_index, double* out);
/* I = sum_s sum_lambda |E|^2 for one candidate (src/tools/regionContrast.cpp:124-143). */
int csx_chain_get_intensity(csx_chain* chain, int cand, double* out);

typedef struct {
  double calib_peak;          /* max of the calibration intensity */
  int n_region_pixels;
  double seconds_calibration; /* wall time of the calibration propagation (0 if cached) */
  double seconds_propagation; /* wall time of the full propagation and the region reduction */
} csx_report;
/* regionContrast::compute_contrast (src/tools/regionContrast.cpp:145-205): runs
 * the calibration chain (cached until the chain changes), the full chain, and
 * the region mean of I / max(I_cal) per candidate. radii/angles/xlim as the
 * regionContrast block; pass NULL region to use the script's block. */
typedef struct {
  double radius1, radius2;    /* lambda/D, inclusive */
  double angle1, angle2;      /* degrees, strict */
  double xlim1, xlim2;        /* lambda/D, inclusive */
  double reference_lambda;    /* <= 0: mean of the chain's lambdas */
} csx_region;
int csx_chain_contrast(csx_chain* chain, const csx_region* region, double* contrast_per_cand,
                       csx_report* rep);
/* Normalised image I / max(I_cal) of one candidate (after csx_chain_contrast). */
int csx_chain_get_normalized_image(csx_chain* chain, int cand, double* out);

/* Contrast curve of one candidate (contrastCurve::make_contrast_curve,
 * src/tools/contrastCurve.cpp:68-158): calibration and full propagation, the
 * source/wavelength intensity sum over the calibration peak, averaged over
 * overlapping annuli.  With n = out_cols - 2 sample radii s_i = linspace(0,
 * max x / loD, n) (loD at the middle wavelength, lambdas[n_lambdas / 2]), row k
 * (k = 0 .. n - 4) is pos[k] = s[k + 2] and avg[k] = mean of I / peak over the
 * pixels with s[k + 1] < r / loD < s[k + 3] (NaN if none): the rows csim writes
 * to its curve file.  pos = avg = NULL returns the row count without running;
 * otherwise returns the row count, or CSX_ERR_ARG when capacity is smaller. */
int csx_chain_contrast_curve(csx_chain* chain, int cand, double* pos, double* avg, int capacity);

/* Per-field binned intensity of the last propagation: out[f][b] = sum of |E_f|^2
 * over the output pixels p with bin_of_pixel[p] == b (column-major pixel order,
 * bins < 0 ignored, nbins <= 64), for every field f = (cand * n_sources + s) *
 * n_lambdas + l.  Annulus profiles of this kind are the surrogate's targets. */
int csx_chain_field_profiles(csx_chain* chain, const int32_t* bin_of_pixel, int nbins, double* out);

/* Measurement: with profiling on, every contraction launch (and its input
 * conversion) is bracketed by HIP events on the chain's stream; the JSON
 * report gives per kernel {launches, ms_total, flops_total, bytes_total}
 * (algorithmic flops, 8 per complex MAC; algorithmic HBM bytes).  Enabling
 * resets the statistics.  Returns the JSON length. */
int csx_chain_set_profiling(csx_chain* chain, int enable);
int csx_chain_profile_json(csx_chain* chain, char* buf, size_t cap);

/* Graph replay (no reference counterpart; csim runs elements one by one on the
 * CPU): with graphs on, the second and later propagations of a compiled chain
 * replay one HIP graph holding every stage (captured on a private stream,
 * launched on the chain's stream).  Device data the stages read (sags,
 * candidate batches, bound arrays updated in place) are read at replay time;
 * recompilation discards the graphs.  Ignored while profiling is on. */
int csx_chain_set_graph(csx_chain* chain, int enable);

/* Hex-FPM response matrix: makeHexCFpmResponse::compute_response
 * (src/tools/makeHexCFpmResponse.cpp:91-190) as one batch of n_hex + 1
 * candidates.  d_M (device, complex128): nslices = n_sources * n_lambdas slices
 * (slice s * n_lambdas + l), each npix x ncol column-major, ncol = n_hex + 1;
 * column h < n_hex is the region field of hex h alone, column n_hex the area
 * outside the hexes (with the Babinet term for fpmPupToLyot), so that
 * csx_hex_resp_objective(d_M, ...) gives the region contrast of any sag vector.
 * d_M NULL: only the sizes.  The element's hex selection and the chain's
 * candidate batch are restored afterwards.  calib_peak: the calibration peak. */
int csx_chain_hex_response(csx_chain* chain, const char* element_name, const csx_region* region, void* d_M,
                           int* npix, int* ncol, int* nslices, double* calib_peak);

/* FITS matrices in csim's convention (src/lib/csim_fits.cpp): csx_fits_read
 * reads BITPIX 8/16/32/-32/-64 images (NAXIS1 = columns, NAXIS2 = rows, NAXIS3 =
 * slices) into a column-major rows x cols x slices array (out NULL: only the
 * sizes); csx_fits_write writes a column-major rows x cols matrix as save_mat does
 * (:480-493: BITPIX -64, the transposed matrix, big-endian). */
int csx_fits_read(const char* path, double* out, size_t capacity, int* rows, int* cols, int* slices);
int csx_fits_write(const char* path, const double* colmajor, int rows, int cols);

/* Region-pixel fields of every candidate after the last full propagation: the
 * fields linearOptimizer::compute_jacobian and compute_objective_vector read
 * (src/optimization/linearOptimizer.cpp:420-505).  d_out (device, complex128):
 * nslices = n_sources * n_lambdas slices (slice s * n_lambdas + l), each npix x
 * ncand column-major (column c = candidate c), region pixels in
 * regionContrast::get_region_pixels order.  d_out NULL: only the sizes. */
int csx_chain_region_fields(csx_chain* chain, const csx_region* region, void* d_out, int* npix, int* ncand,
                            int* nslices);

/* Chain group (no reference counterpart; csim runs one configuration at a time):
 * several chains with identical element lists, e.g. the same script at 512^2 and
 * 1024^2 pupils, propagated together.  Every contraction stage of all chains is
 * one grouped launch over uneven problem sizes; other stages run chain by chain.
 * The chains must share the context and precision and outlive the group; a
 * chain recompiled since the last call is picked up. */
typedef struct csx_group csx_group;
int csx_group_create(csx_chain* const* chains, int n, csx_group** out);
void csx_group_destroy(csx_group* group);
int csx_group_candidates(csx_group* group);          /* sum of the chains' candidate counts */
/* Region contrast (regionContrast::compute_contrast) of every candidate of every
 * chain, chain by chain; each chain's calibration is its own (cached). */
int csx_group_contrast(csx_group* group, const csx_region* region, double* out);
int csx_group_merged_launches(csx_group* group);     /* grouped MFT stages per propagation (2 GEMM launches each) */

/* ------------------------------------------------------ operator API (device) */
/* Uniform 1-D grid: x_i = x0 + i*dx, n points. */
typedef struct { int n; double x0; double dx; } csx_grid;

/* Matrix Fourier transform = csim zoomFft (src/lib/csim_fft.cpp:670-826) in its
 * closed form  E_out = dx_in dy_in / (i lf) * A_y * E_in * B_x^T  with
 * A_y[r][m] = exp(-2 pi i y_out[r] y_in[m] / lf),  B_x[c][n] likewise in x,
 * for a batch of fields.  lambda_f[b] is the signed lambda*focal length of field
 * b.  d_in: batch x in_y.n x in_x.n complex128; d_out: batch x out_y.n x out_x.n.
 * Fields are column-major (rows = y).  d_work: NULL (the library allocates its
 * buffers and times tile variants on first use of a shape) or a device buffer of at
 * least csx_mft_workspace_size bytes, 256-byte aligned, which then holds every work
 * buffer of the call (nothing allocated; no tile timing: the persisted tuner choice
 * or the 64 x 64 tile).  A d_work buffer may be reused once the call has returned. */
size_t csx_mft_workspace_size(csx_grid in_x, csx_grid in_y, csx_grid out_x, csx_grid out_y,
                              int batch, csx_precision p);
int csx_mft(csx_ctx* ctx, csx_grid in_x, csx_grid in_y, csx_grid out_x, csx_grid out_y,
            const double* lambda_f, int batch, const void* d_in, void* d_out,
            csx_precision p, void* d_work);

/* Allocation-free csx_mft (no counterpart in the reference): for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  The same contraction as csx_mft, with every work buffer in
 * one 128 MiB scratch array that each GPU holds for the life of the process; calls
 * whose csx_mft_workspace_size exceeds it fail with CSX_ERR_ARG.  No tile timing: the
 * persisted tuner choice, else the 64 x 64 tile.  Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (two threads or streams) race; calls in
 * sequence do not (each returns after its work is done). */
int csx_mft_noalloc(csx_ctx* ctx, csx_grid in_x, csx_grid in_y, csx_grid out_x, csx_grid out_y,
                    const double* lambda_f, int batch, const void* d_in, void* d_out, csx_precision p);
/* Bytes of the scratch array csx_mft_noalloc works in. */
size_t csx_mft_noalloc_capacity(void);

/* Angular-spectrum Fresnel step, fresnelPropagateAS::execute
 * (src/lib/csim_fft.cpp:371-392): per field b, E <- IFFT2(FFT2(E) . H_b),
 * H = exp(i 2pi/lambda z sqrt(1 - (lambda/apRad)^2/4 (p^2+q^2))).  n x n, n a
 * power of two <= 4096; d_field complex128, in place. */
int csx_fresnel_as(csx_ctx* ctx, void* d_field, int n, int batch, const double* lambdas,
                   double ap_rad, double z);
/* 2-D FFT with csim/FFTW semantics (src/lib/csim_fft.cpp:98-228): dir -1 forward
 * unnormalised, +1 inverse divided by n*n.  In place, n power of two. */
int csx_fft2(csx_ctx* ctx, void* d_field, int n, int batch, int dir);

/* downsample::execute (src/elements/downsample.cpp:23-96) for an integer factor
 * on csim's symmetric grids: separable linear interpolation, extrapolation 0. */
int csx_downsample(csx_ctx* ctx, const void* d_in, int rows, int cols, int batch, int factor,
                   double pixel_size, void* d_out);

/* Pointwise elements on a batch of fields (nfield = batch, npix each):
 * mask.cpp:32-46  E *= (1+i) m   (m real, npix)
 * mirror.cpp:66-72 E *= exp(-sign 4 pi i h / lambda_b)  (h real, npix) */
int csx_apply_mask(csx_ctx* ctx, void* d_field, int npix, int batch, const double* d_mask);
int csx_apply_mirror(csx_ctx* ctx, void* d_field, int npix, int batch, const double* d_sag,
                     double sign, const double* lambdas);

/* Hex-cell FPM per-lambda complex mask, complexHexMaskFPM::make_complex_intpolated_mask
 * (src/elements/complexHexMaskFPM.cpp:44-99): for each (candidate c, lambda l)
 * mask = sum_s nsub_s exp(i phaseSign 4 pi sag_c[hex_s] / lambda_l) / nSubPix^2.
 * d_nsub/d_hexnum: S x S x 3 int32 (column-major); d_sags: n_cand x n_hex;
 * d_out: n_cand x n_lambda x S x S complex128. */
int csx_hex_mask(csx_ctx* ctx, const int32_t* d_nsub, const int32_t* d_hexnum, int S, int n_sub_pix,
                 const double* d_sags, int n_hex, int n_cand, const double* lambdas, int n_lambda,
                 double phase_sign, void* d_out);
/* Host-side hex geometry (src/elements/complexHexMaskFPM.cpp:236-469):
 * layout 0 = csim radial cut-off, 1 = standard rings.  Returns the hex count
 * (and writes nothing) when the arrays are NULL. */
int csx_hex_layout(int layout, int n_rings, double hex_step, double* hx, double* hy, int max_hex);
int csx_hex_rasterize(int layout, int n_rings, double hex_step, int fpm_array_size,
                      double fpm_scale_factor, double hex_gap, int fpm_mask_size, int n_sub_pix,
                      double pixel_scale, double fpm_radius, int32_t* nsub, int32_t* hexnum);

/* I[g] = sum over the group's fields of |E|^2 (fields_per_group consecutive
 * fields each); regionContrast::compute_intensity (src/tools/regionContrast.cpp:124-143). */
int csx_intensity(csx_ctx* ctx, const void* d_fields, int npix, int n_groups,
                  int fields_per_group, double* d_intensity);
/* Mean of I[g] / peak over a pixel index list, regionContrast.cpp:195. */
int csx_region_mean(csx_ctx* ctx, const double* d_intensity, int npix, int n_groups,
                    const int32_t* d_idx, int n_idx, double peak, double* d_out);

/* optHexRespObjective::execute (src/optimization/optObjective.cpp:24-69): the
 * response-matrix objective as a batched complex GEMM on MFMA.  d_M: nslices x
 * (ncol x npix) complex128 with M[s] column-major npix x ncol, ncol = nopt+1;
 * d_X: ncand x nopt; d_obj: ncand.  objective 0 = mean, 1 = max.
 * Every precision: FP64, FP32, or split-BF16 (bf16x3 / bf16x6: both operands split
 * into BF16 planes with K = ncol padded by zeros to a multiple of 32). */
int csx_hex_resp_objective(csx_ctx* ctx, const void* d_M, int npix, int ncol, int nslices,
                           const double* lambdas, double phase_sign, double calib,
                           const double* d_X, int nopt, int ncand, int objective,
                           double* d_obj, csx_precision p);

/* ------------------------------------------------ script driver and tools */
/* Device-free check of a csim input script (no context, no device, no file read):
 * a JSON report {"ok", "error", "blocks": [{"type", "keys", "blocks"}...],
 * "elements": [{"type", "name"}...] (the coronagraph as the engine builds it),
 * "unknown_keys", "unknown_elements", "files" (every FITS file the script names),
 * "top_level"}.  The block grammar of src/lib/csim_parser.cpp:60-190; keys as the
 * element set() functions take them (unknown ones listed, as csim prints and goes
 * on).  Returns the report length (written up to cap - 1 bytes, NUL-terminated). */
int csx_script_parse(const char* script_text, char* json, size_t cap);

/* csim's main loop (src/csim.cpp:54-108): the telescope, efield, coronagraph and
 * regionContrast blocks define the chain; execute, regionContrast
 * (regionContrast.cpp:145-205: calibration_PSF.fits, normalized_PSF.fits, its
 * filename), makeContrastCurve (contrastCurve.cpp:68-158), makeHexCFpmResponse
 * (makeHexCFpmResponse.cpp:91-190: a responseData directory) and
 * differentialEvolutionOptimizer (differentialEvolutionOptimizer.cpp:216-330,
 * with an optHexRespObjective sub-block or the additive optChainObjective sub-block
 * {component, regionContrast} that evaluates every candidate by full propagation)
 * run in script order; break stops.  PSF, region and curve files go to workdir
 * (NULL: the current directory); the tools' own directories are taken as given.
 * log receives the run's messages.  Returns the log length or a negative error. */
int csx_run_script(csx_ctx* ctx, const char* script_text, const char* workdir, char* log, size_t cap);

/* responseData directory (src/data/responseData.cpp:175-273): responseParams.txt,
 * responseData_s<s>_p<p>_{re,im}.fits (each slice transposed, NAXIS1 = columns)
 * and problemParams.txt.  M: interleaved complex, nsrc x npol cubes (source-major),
 * each slices x (cols x rows) with every slice column-major rows x cols. */
int csx_response_save(const char* dir, const double* M, int rows, int cols, int slices, int npol, int nsrc,
                      double calib_max_intensity, const double* lambdas);
/* sizes[5] = rows, cols, slices, npol, nsrc; M NULL: only the sizes, calibration
 * peak and wavelengths (lambdas NULL: not returned). */
int csx_response_load(const char* dir, double* M, size_t capacity_complex, int* sizes,
                      double* calib_max_intensity, double* lambdas, int lambda_capacity);

/* differentialEvolutionOptimizer::optimize (differentialEvolutionOptimizer.cpp:
 * 216-330, strategies deStrategy.cpp:23-158) over a caller's objective: the
 * differentialEvolutionOptimizer block's keys (:66-197; the additive key seed
 * selects the Philox stream), every generation evaluated as one call
 * f(user, population (dim x npop column-major), dim, npop, values); a negative
 * return stops the run.  Writes optBestVal.txt, optValHistory.txt,
 * bestOptVector.fits (and the population files when saveState) to
 * outputDirectory. */
typedef int (*csx_objective_fn)(void* user, const double* population, int dim, int npop, double* values);
int csx_de_optimize(const char* de_block_text, int dim, csx_objective_fn f, void* user,
                    double* best_vector, double* best_value, int* iterations);

/* Philox4x32-10 (Salmon et al., SC'11) counter-based generator used for every
 * synthetic input: fills n doubles, uniform [0,1) (normal = 0) or standard
 * normal via Box-Muller in FP64 (normal = 1), stream (seed, stream_id). */
int csx_philox_fill(uint64_t seed, uint64_t stream_id, double* out, size_t n, int normal);
void csx_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ------------------------------------------------------------ MLP surrogate */
/* 218 -> 512 -> 512 -> 32 GELU perceptron trained by the library's own HIP
 * kernels (BF16 weights, FP32 accumulation, FP32 master weights, Adam).  No
 * counterpart in the reference. */
/* Surrogate input embedding (no reference counterpart): for every candidate c
 * (d_thickness: ncand x nhex, metres) and wavelength l, one float row
 * [cos phi_h, sin phi_h, (lambda_l - lam_center) / lam_scale], h < nhex,
 * phi_h = 2 pi (n_index - 1) t_h / lambda_l; d_out: (ncand nlam) x (2 nhex + 1). */
int csx_hex_phasor_features(csx_ctx* ctx, const double* d_thickness, int ncand, int nhex,
                            const double* lambdas, int nlam, double n_index, double lam_center,
                            double lam_scale, float* d_out);

typedef struct csx_mlp csx_mlp;
int csx_mlp_create(csx_ctx* ctx, int n_in, int n_h1, int n_h2, int n_out, uint64_t seed,
                   csx_mlp** out);
void csx_mlp_destroy(csx_mlp* mlp);
/* Host <-> device parameter copies, order W1 (n_in x n_h1, row-major), b1, W2, b2, W3, b3.
 * Setting parameters also restarts the Adam state (moments and step count). */
int csx_mlp_get_params(csx_mlp* mlp, double* flat);
int csx_mlp_set_params(csx_mlp* mlp, const double* flat);
size_t csx_mlp_param_count(csx_mlp* mlp);
/* Forward on n rows of device x (n x n_in float32) -> device y (n x n_out float32). */
int csx_mlp_forward(csx_mlp* mlp, const float* d_x, int n, float* d_y);
/* Loss sum (y-t)^2/(n n_out) and its gradient (host, same order as params). */
int csx_mlp_loss_grad(csx_mlp* mlp, const float* d_x, const float* d_t, int n, double* loss,
                      double* grad_flat);
/* One Adam step on a device minibatch; returns the loss before the step (loss
 * NULL: no read-back, the step is only enqueued on the context's stream). */
int csx_mlp_train_step(csx_mlp* mlp, const float* d_x, const float* d_t, int n, double lr,
                       double* loss);

#ifdef __cplusplus
}
#endif
#endif /* CSX_H */
