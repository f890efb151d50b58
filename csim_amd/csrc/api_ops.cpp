// C ABI: context, operator API (device pointers), Philox generator.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "hostgeom.hpp"
#include "mft.hpp"
#include "ops.hpp"

namespace csx {
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int fail(int code, const std::string& m) { g_err = m; return code; }
}  // namespace csx

using namespace csx;

#define CSX_GUARD_BEGIN try {
#define CSX_GUARD_END                                                  \
  }                                                                    \
  catch (const std::bad_alloc&) { return fail(CSX_ERR_NOMEM, "host allocation failed"); } \
  catch (const std::exception& e) { return fail(CSX_ERR_STATE, e.what()); } \
  catch (...) { return fail(CSX_ERR_STATE, "unknown exception"); }

static int use_device(csx_ctx* ctx) {
  if (!ctx) return fail(CSX_ERR_ARG, "null context");
  CSX_HIP(hipSetDevice(ctx->device));
  return CSX_OK;
}

// Scratch of the allocation-free matrix Fourier transform (csx_mft_noalloc): every GPU
// the library runs on holds one copy for the life of the process.  Host code uses the
// per-device address hipGetSymbolAddress gives (scratch_ptr), never the symbol itself.
// Nothing in it has to start at zero: each buffer carved from it is written before it
// is read within the call.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;   // by device index

// Resolves this device's copy of `scratch` (the device is current).
static int resolve_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0) return fail(CSX_ERR_ARG, "bad device");
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return CSX_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(CSX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return CSX_OK;
}
static void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

extern "C" {

const char* csx_last_error(void) { return g_err.c_str(); }
const char* csx_version(void) { return "csx 0.1 (gfx950)"; }

int csx_tune_report(char* buf, size_t cap) {
  CSX_GUARD_BEGIN
  const std::string j = tune_report();
  if (buf && cap) {
    const size_t n = std::min(cap - 1, j.size());
    std::memcpy(buf, j.data(), n);
    buf[n] = 0;
  }
  return (int)j.size();
  CSX_GUARD_END
}

int csx_ctx_create(int device, unsigned flags, csx_ctx** out) {
  if (!out) return fail(CSX_ERR_ARG, "null out");
  if (flags & ~CSX_CTX_DEFAULT_STREAM) return fail(CSX_ERR_ARG, "csx_ctx_create: unknown flags");
  *out = nullptr;
  int n = 0;
  CSX_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(CSX_ERR_ARG, "device index out of range");
  CSX_HIP(hipSetDevice(device));
  csx_ctx* c = new (std::nothrow) csx_ctx();
  if (!c) return fail(CSX_ERR_NOMEM, "context allocation failed");
  c->device = device;
  if (int rc = resolve_scratch(device); rc != CSX_OK) { delete c; return rc; }
  if (flags & CSX_CTX_DEFAULT_STREAM) {
    c->stream = nullptr;   // the null stream
    c->own_stream = false;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail(CSX_ERR_HIP, hipGetErrorString(e)); }
    c->own_stream = true;
  }
  *out = c;
  return CSX_OK;
}

void csx_ctx_destroy(csx_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int csx_ctx_set_stream(csx_ctx* ctx, void* s) {
  CSX_TRY(use_device(ctx));
  if (ctx->own_stream && ctx->stream) { CSX_HIP(hipStreamSynchronize(ctx->stream)); CSX_HIP(hipStreamDestroy(ctx->stream)); }
  if (s) { ctx->stream = (hipStream_t)s; ctx->own_stream = false; }
  else { CSX_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)); ctx->own_stream = true; }
  return CSX_OK;
}

int csx_ctx_synchronize(csx_ctx* ctx) {
  CSX_TRY(use_device(ctx));
  CSX_HIP(hipStreamSynchronize(ctx->stream));
  return CSX_OK;
}

// ------------------------------------------------------------------- MFT ---
static int mft_fields(csx_grid in_x, csx_grid in_y, csx_grid out_x, csx_grid out_y, const double* lambda_f,
                      int batch, const void* d_in, void* d_out, int in_kind, size_t ein, std::vector<MftField>& F) {
  if (batch <= 0 || !lambda_f || !d_in || !d_out) return fail(CSX_ERR_ARG, "csx_mft: bad arguments");
  size_t nin = (size_t)in_x.n * in_y.n, nout = (size_t)out_x.n * out_y.n;
  F.resize(batch);
  for (int b = 0; b < batch; b++) {
    MftField& f = F[b];
    f.in = static_cast<const char*>(d_in) + b * nin * ein;
    f.in_kind = in_kind;
    f.out = static_cast<char*>(d_out) + b * nout * sizeof(double2);
    f.out_kind = FK_C128;
    f.in_x = in_x; f.in_y = in_y; f.out_x = out_x; f.out_y = out_y;
    f.lambda_f = lambda_f[b];
  }
  return CSX_OK;
}

// Bytes of the working-precision copy of the input (none for FP64).
static size_t mft_conv_bytes(csx_grid in_x, csx_grid in_y, int batch, csx_precision p) {
  if (p == CSX_FP64 || batch <= 0) return 0;
  const size_t e = p == CSX_FP32 ? sizeof(float2) : (size_t)4 * terms_of(p);
  return Arena::round((size_t)batch * in_x.n * in_y.n * e);
}

size_t csx_mft_workspace_size(csx_grid in_x, csx_grid in_y, csx_grid out_x, csx_grid out_y, int batch,
                              csx_precision p) {
  return mft_conv_bytes(in_x, in_y, batch, p) + MftPlan::arena_bytes(in_x.n, in_y.n, out_x.n, out_y.n, batch, p);
}

// csx_mft over device memory from `arena` (work buffers and plan, nothing allocated)
// or, with arena NULL, from the context's scratch and the plan's own buffers.
static int mft_run(csx_ctx* ctx, csx_grid in_x, csx_grid in_y, csx_grid out_x, csx_grid out_y,
                   const double* lambda_f, int batch, const void* d_in, void* d_out, csx_precision p, Arena* arena) {
  if (batch <= 0 || !lambda_f || !d_in || !d_out) return fail(CSX_ERR_ARG, "csx_mft: empty batch or null pointer");
  const int T = terms_of(p);
  const void* src = d_in;
  int kind = FK_C128;
  size_t ein = sizeof(double2);
  size_t n = (size_t)batch * in_x.n * in_y.n;
  DevBuf conv;
  conv.arena = arena;
  DevBuf& work = arena ? conv : ctx->work;
  if (p == CSX_FP32) {
    CSX_TRY(work.reserve(n * sizeof(float2)));
    CSX_TRY(convert_c128_to_c64(d_in, work.p, n, ctx->stream));
    src = work.p; kind = FK_C64; ein = sizeof(float2);
  } else if (T) {
    CSX_TRY(work.reserve(n * 4 * T));
    CSX_TRY(split_c128_to_planes(d_in, work.p, (long)in_x.n * in_y.n, batch, T, ctx->stream));
    src = work.p; kind = FK_PLANES; ein = 4 * T;
  }
  std::vector<MftField> F;
  CSX_TRY(mft_fields(in_x, in_y, out_x, out_y, lambda_f, batch, src, d_out, kind, ein, F));
  MftPlan plan;
  CSX_TRY(plan.build(F, p, ctx->stream, arena));
  CSX_TRY(plan.run(ctx->stream));
  CSX_HIP(hipStreamSynchronize(ctx->stream));   // plan buffers are released (or free for reuse) on return
  return CSX_OK;
}

int csx_mft(csx_ctx* ctx, csx_grid in_x, csx_grid in_y, csx_grid out_x, csx_grid out_y, const double* lambda_f,
            int batch, const void* d_in, void* d_out, csx_precision p, void* d_work) {
  CSX_GUARD_BEGIN
  CSX_TRY(use_device(ctx));
  if (!d_work) return mft_run(ctx, in_x, in_y, out_x, out_y, lambda_f, batch, d_in, d_out, p, nullptr);
  Arena a;
  a.base = static_cast<char*>(d_work);
  a.cap = csx_mft_workspace_size(in_x, in_y, out_x, out_y, batch, p);
  return mft_run(ctx, in_x, in_y, out_x, out_y, lambda_f, batch, d_in, d_out, p, &a);
  CSX_GUARD_END
}

int csx_mft_noalloc(csx_ctx* ctx, csx_grid in_x, csx_grid in_y, csx_grid out_x, csx_grid out_y,
                    const double* lambda_f, int batch, const void* d_in, void* d_out, csx_precision p) {
  CSX_GUARD_BEGIN
  CSX_TRY(use_device(ctx));
  const size_t need = csx_mft_workspace_size(in_x, in_y, out_x, out_y, batch, p);
  if (need > kNoAllocScratchBytes)
    return fail(CSX_ERR_ARG, "csx_mft_noalloc: the call needs " + std::to_string(need) +
                                 " bytes of scratch, more than its " + std::to_string(kNoAllocScratchBytes) +
                                 " (use csx_mft with a workspace)");
  CSX_TRY(resolve_scratch(ctx->device));
  Arena a;
  a.base = static_cast<char*>(scratch_ptr(ctx->device));
  a.cap = kNoAllocScratchBytes;
  return mft_run(ctx, in_x, in_y, out_x, out_y, lambda_f, batch, d_in, d_out, p, &a);
  CSX_GUARD_END
}

size_t csx_mft_noalloc_capacity(void) { return kNoAllocScratchBytes; }

// ------------------------------------------------------- FFT / Fresnel ---
static int upload_doubles(csx_ctx* ctx, DevBuf& buf, const double* h, int n) {
  CSX_TRY(buf.reserve(sizeof(double) * std::max(1, n)));
  CSX_HIP(hipMemcpyAsync(buf.p, h, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  return CSX_OK;
}

int csx_fresnel_as(csx_ctx* ctx, void* d_field, int n, int batch, const double* lambdas, double ap_rad, double z) {
  CSX_GUARD_BEGIN
  CSX_TRY(use_device(ctx));
  if (!d_field || !lambdas || batch <= 0 || ap_rad <= 0) return fail(CSX_ERR_ARG, "csx_fresnel_as: bad arguments");
  CSX_TRY(upload_doubles(ctx, ctx->small, lambdas, batch));
  CSX_TRY(fresnel_batched((double2*)d_field, n, batch, ctx->small.as<double>(), ap_rad, z, ctx->stream));
  CSX_HIP(hipStreamSynchronize(ctx->stream));
  return CSX_OK;
  CSX_GUARD_END
}

int csx_fft2(csx_ctx* ctx, void* d_field, int n, int batch, int dir) {
  CSX_GUARD_BEGIN
  CSX_TRY(use_device(ctx));
  if (!d_field || batch <= 0) return fail(CSX_ERR_ARG, "csx_fft2: bad arguments");
  CSX_TRY(fft2_batched((double2*)d_field, n, batch, dir, ctx->stream));
  CSX_HIP(hipStreamSynchronize(ctx->stream));
  return CSX_OK;
  CSX_GUARD_END
}

int csx_downsample(csx_ctx* ctx, const void* d_in, int rows, int cols, int batch, int factor, double pixel_size,
                   void* d_out) {
  CSX_GUARD_BEGIN
  CSX_TRY(use_device(ctx));
  if (!d_in || !d_out || factor < 1 || rows <= 0 || cols <= 0 || batch <= 0)
    return fail(CSX_ERR_ARG, "csx_downsample: bad arguments");
  // csim's interp2 pairs its row and column grids consistently only for square arrays
  if (rows != cols) return fail(CSX_ERR_UNSUPPORTED, "csx_downsample: square fields only");
  int nr = rows / factor, nc = cols / factor;
  auto inX = grid_symmetric(cols, pixel_size), inY = grid_symmetric(rows, pixel_size);
  auto outX = grid_symmetric(nc, factor * pixel_size), outY = grid_symmetric(nr, factor * pixel_size);
  auto tx = interp_table(inX, outX), ty = interp_table(inY, outY);
  CSX_TRY(ctx->small.reserve(sizeof(InterpTable) * (tx.size() + ty.size())));
  InterpTable* dtx = ctx->small.as<InterpTable>();
  InterpTable* dty = dtx + tx.size();
  CSX_HIP(hipMemcpyAsync(dtx, tx.data(), sizeof(InterpTable) * tx.size(), hipMemcpyHostToDevice, ctx->stream));
  CSX_HIP(hipMemcpyAsync(dty, ty.data(), sizeof(InterpTable) * ty.size(), hipMemcpyHostToDevice, ctx->stream));
  CSX_TRY(resample_batched((const double2*)d_in, rows, cols, batch, (double2*)d_out, nr, nc, dty, dtx, ctx->stream));
  CSX_HIP(hipStreamSynchronize(ctx->stream));
  return CSX_OK;
  CSX_GUARD_END
}

int csx_apply_mask(csx_ctx* ctx, void* d_field, int npix, int batch, const double* d_mask) {
  CSX_TRY(use_device(ctx));
  if (!d_field || !d_mask || npix <= 0 || batch <= 0) return fail(CSX_ERR_ARG, "csx_apply_mask: bad arguments");
  CSX_TRY(k_apply_mask((double2*)d_field, npix, batch, d_mask, ctx->stream));
  CSX_HIP(hipStreamSynchronize(ctx->stream));
  return CSX_OK;
}

int csx_apply_mirror(csx_ctx* ctx, void* d_field, int npix, int batch, const double* d_sag, double sign,
                     const double* lambdas) {
  CSX_TRY(use_device(ctx));
  if (!d_field || !d_sag || !lambdas || npix <= 0 || batch <= 0)
    return fail(CSX_ERR_ARG, "csx_apply_mirror: bad arguments");
  CSX_TRY(upload_doubles(ctx, ctx->small, lambdas, batch));
  CSX_TRY(k_apply_mirror((double2*)d_field, npix, batch, d_sag, sign, ctx->small.as<double>(), ctx->stream));
  CSX_HIP(hipStreamSynchronize(ctx->stream));
  return CSX_OK;
}

// ------------------------------------------------------------------ hex ---
int csx_hex_mask(csx_ctx* ctx, const int32_t* d_nsub, const int32_t* d_hexnum, int S, int n_sub_pix,
                 const double* d_sags, int n_hex, int n_cand, const double* lambdas, int n_lambda,
                 double phase_sign, void* d_out) {
  CSX_TRY(use_device(ctx));
  if (!d_nsub || !d_hexnum || !d_sags || !lambdas || !d_out || S <= 0 || n_cand <= 0 || n_lambda <= 0)
    return fail(CSX_ERR_ARG, "csx_hex_mask: bad arguments");
  CSX_TRY(upload_doubles(ctx, ctx->small, lambdas, n_lambda));
  CSX_TRY(hex_masks(d_nsub, d_hexnum, S, n_sub_pix, d_sags, n_hex, n_cand, ctx->small.as<double>(), n_lambda,
                    phase_sign, -1, (double2*)d_out, ctx->stream));
  CSX_HIP(hipStreamSynchronize(ctx->stream));
  return CSX_OK;
}

int csx_hex_layout(int layout, int n_rings, double hex_step, double* hx, double* hy, int max_hex) {
  CSX_GUARD_BEGIN
  if (layout != 0 && layout != 1) return fail(CSX_ERR_ARG, "csx_hex_layout: layout must be 0 or 1");
  std::vector<double> x, y;
  int n = hex_centers(layout, n_rings, hex_step, x, y);
  if (hx && hy) {
    if (max_hex < n) return fail(CSX_ERR_ARG, "csx_hex_layout: output arrays too small");
    std::memcpy(hx, x.data(), sizeof(double) * n);
    std::memcpy(hy, y.data(), sizeof(double) * n);
  }
  return n;
  CSX_GUARD_END
}

int csx_hex_rasterize(int layout, int n_rings, double hex_step, int fpm_array_size, double fpm_scale_factor,
                      double hex_gap, int fpm_mask_size, int n_sub_pix, double pixel_scale, double fpm_radius,
                      int32_t* nsub, int32_t* hexnum) {
  CSX_GUARD_BEGIN
  if (!nsub || !hexnum) return fail(CSX_ERR_ARG, "csx_hex_rasterize: null output");
  HexRaster r;
  CSX_TRY(hex_rasterize(layout, n_rings, hex_step, fpm_array_size, fpm_scale_factor, hex_gap, fpm_mask_size,
                        n_sub_pix, pixel_scale, fpm_radius, r));
  std::memcpy(nsub, r.nsub.data(), sizeof(int32_t) * r.nsub.size());
  std::memcpy(hexnum, r.hexnum.data(), sizeof(int32_t) * r.hexnum.size());
  return r.nhex;
  CSX_GUARD_END
}

// ------------------------------------------------------------- contrast ---
int csx_intensity(csx_ctx* ctx, const void* d_fields, int npix, int n_groups, int fields_per_group,
                  double* d_intensity) {
  CSX_TRY(use_device(ctx));
  if (!d_fields || !d_intensity || npix <= 0 || n_groups <= 0 || fields_per_group <= 0)
    return fail(CSX_ERR_ARG, "csx_intensity: bad arguments");
  CSX_TRY(intensity_groups((const double2*)d_fields, npix, n_groups, fields_per_group, d_intensity, ctx->stream));
  CSX_HIP(hipStreamSynchronize(ctx->stream));
  return CSX_OK;
}

int csx_region_mean(csx_ctx* ctx, const double* d_intensity, int npix, int n_groups, const int32_t* d_idx,
                    int n_idx, double peak, double* d_out) {
  CSX_TRY(use_device(ctx));
  if (!d_intensity || !d_idx || !d_out || n_idx <= 0 || peak == 0)
    return fail(CSX_ERR_ARG, "csx_region_mean: bad arguments");
  CSX_TRY(region_mean_groups(d_intensity, npix, n_groups, d_idx, n_idx, nullptr, peak, d_out, ctx->stream));
  CSX_HIP(hipStreamSynchronize(ctx->stream));
  return CSX_OK;
}

// -------------------------------------------------------------- Philox ---
void csx_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0; k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

int csx_philox_fill(uint64_t seed, uint64_t stream_id, double* out, size_t n, int normal) {
  if (!out) return fail(CSX_ERR_ARG, "csx_philox_fill: null output");
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (size_t i = 0, blk = 0; i < n; blk++) {
    uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)stream_id, (uint32_t)(stream_id >> 32)};
    uint32_t r[4];
    csx_philox4x32_10(ctr, key, r);
    // two 53-bit uniforms in [0, 1)
    double u0 = (double)((((uint64_t)r[0] << 32) | r[1]) >> 11) * 0x1.0p-53;
    double u1 = (double)((((uint64_t)r[2] << 32) | r[3]) >> 11) * 0x1.0p-53;
    if (!normal) {
      out[i++] = u0;
      if (i < n) out[i++] = u1;
    } else {
      double rad = std::sqrt(-2.0 * std::log(1.0 - u0));   // 1 - u0 in (0, 1]
      double th = 2.0 * M_PI * u1;
      out[i++] = rad * std::cos(th);
      if (i < n) out[i++] = rad * std::sin(th);
    }
  }
  return CSX_OK;
}

}  // extern "C"
