// C ABI of the element / script API.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"

using namespace csx;

struct csx_chain {
  Chain impl;
  explicit csx_chain(csx_ctx* c) : impl(c) {}
};

struct csx_group {
  ChainGroup impl;
  explicit csx_group(std::vector<Chain*> c) : impl(std::move(c)) {}
};

#define CSX_GUARD_BEGIN try {
#define CSX_GUARD_END                                                               \
  }                                                                                 \
  catch (const std::bad_alloc&) { return fail(CSX_ERR_NOMEM, "host allocation failed"); } \
  catch (const std::exception& e) { return fail(CSX_ERR_STATE, e.what()); }          \
  catch (...) { return fail(CSX_ERR_STATE, "unknown exception"); }

static int chk(csx_chain* c) {
  if (!c) return fail(CSX_ERR_ARG, "null chain");
  CSX_HIP(hipSetDevice(c->impl.ctx()->device));
  return CSX_OK;
}

extern "C" {

int csx_chain_from_script(csx_ctx* ctx, const char* text, csx_chain** out) {
  CSX_GUARD_BEGIN
  if (!ctx || !text || !out) return fail(CSX_ERR_ARG, "csx_chain_from_script: null argument");
  *out = nullptr;
  CSX_HIP(hipSetDevice(ctx->device));
  csx_chain* c = new csx_chain(ctx);
  int rc = c->impl.load_script(text);
  if (rc != CSX_OK) { delete c; return rc; }
  *out = c;
  return CSX_OK;
  CSX_GUARD_END
}

void csx_chain_destroy(csx_chain* c) {
  if (!c) return;
  (void)hipSetDevice(c->impl.ctx()->device);
  (void)hipStreamSynchronize(c->impl.stream());
  delete c;
}

int csx_chain_set_array(csx_chain* c, const char* e, const char* k, const double* d, int rows, int cols) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!e || !k || !d) return fail(CSX_ERR_ARG, "csx_chain_set_array: null argument");
  return c->impl.set_array(e, k, d, rows, cols);
  CSX_GUARD_END
}

int csx_chain_set_complex_array(csx_chain* c, const char* e, const char* k, const double* d, int rows, int cols,
                                int slice) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!e || !k || !d) return fail(CSX_ERR_ARG, "csx_chain_set_complex_array: null argument");
  return c->impl.set_complex_array(e, k, d, rows, cols, slice);
  CSX_GUARD_END
}

int csx_chain_set_opt_data(csx_chain* c, const char* e, const char* n, const double* v, size_t len) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!e || !n || (!v && len)) return fail(CSX_ERR_ARG, "csx_chain_set_opt_data: null argument");
  return c->impl.set_opt(e, n, std::vector<double>(v, v + len));
  CSX_GUARD_END
}

int csx_chain_get_opt_data(csx_chain* c, const char* e, const char* n, double* v, size_t* len) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!e || !n || !len) return fail(CSX_ERR_ARG, "csx_chain_get_opt_data: null argument");
  std::vector<double> out;
  CSX_TRY(c->impl.get_opt(e, n, out));
  if (v) {
    if (*len < out.size()) { *len = out.size(); return fail(CSX_ERR_ARG, "output buffer too small"); }
    std::memcpy(v, out.data(), sizeof(double) * out.size());
  }
  *len = out.size();
  return CSX_OK;
  CSX_GUARD_END
}

int csx_chain_set_candidates(csx_chain* c, const char* e, const double* sags, int n_hex, int n_cand) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!e) return fail(CSX_ERR_ARG, "csx_chain_set_candidates: null element name");
  return c->impl.set_candidates(e, sags, n_hex, n_cand);
  CSX_GUARD_END
}

int csx_chain_set_precision(csx_chain* c, csx_precision p) {
  CSX_TRY(chk(c));
  if (p < CSX_FP64 || p > CSX_BF16X6) return fail(CSX_ERR_ARG, "unknown precision");
  return c->impl.set_precision(p);
}

int csx_chain_get_info(csx_chain* c, csx_chain_info* info) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!info) return fail(CSX_ERR_ARG, "null info");
  return c->impl.info(info);
  CSX_GUARD_END
}

int csx_chain_get_lambdas(csx_chain* c, double* lambdas, int n) {
  CSX_TRY(chk(c));
  const auto& l = c->impl.efield().lambdas;
  if (!lambdas) return (int)l.size();
  if (n < (int)l.size()) return fail(CSX_ERR_ARG, "output too small");
  std::memcpy(lambdas, l.data(), sizeof(double) * l.size());
  return (int)l.size();
}

int csx_chain_propagate(csx_chain* c, int calibrating) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  return c->impl.propagate(calibrating != 0);
  CSX_GUARD_END
}

int csx_chain_get_field(csx_chain* c, int cand, int source, int lambda_index, double* out) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!out) return fail(CSX_ERR_ARG, "null output");
  return c->impl.get_field(cand, source, lambda_index, out);
  CSX_GUARD_END
}

int csx_chain_get_intensity(csx_chain* c, int cand, double* out) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!out) return fail(CSX_ERR_ARG, "null output");
  return c->impl.get_intensity(cand, out);
  CSX_GUARD_END
}

int csx_chain_contrast(csx_chain* c, const csx_region* r, double* out, csx_report* rep) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!out) return fail(CSX_ERR_ARG, "null output");
  return c->impl.contrast(r, out, rep);
  CSX_GUARD_END
}

int csx_chain_get_normalized_image(csx_chain* c, int cand, double* out) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!out) return fail(CSX_ERR_ARG, "null output");
  return c->impl.get_normalized_image(cand, out);
  CSX_GUARD_END
}

int csx_chain_contrast_curve(csx_chain* c, int cand, double* pos, double* avg, int capacity) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if ((pos == nullptr) != (avg == nullptr)) return fail(CSX_ERR_ARG, "contrast curve: pos and avg both or neither");
  int rows = 0;
  CSX_TRY(c->impl.contrast_curve(cand, pos, avg, capacity, &rows));
  return rows;
  CSX_GUARD_END
}

int csx_chain_field_profiles(csx_chain* c, const int32_t* bin_of_pixel, int nbins, double* out) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!bin_of_pixel || !out) return fail(CSX_ERR_ARG, "null argument");
  return c->impl.field_profiles(bin_of_pixel, nbins, out);
  CSX_GUARD_END
}

int csx_chain_set_profiling(csx_chain* c, int enable) {
  CSX_TRY(chk(c));
  return c->impl.set_profiling(enable != 0);
}

int csx_chain_hex_response(csx_chain* c, const char* element_name, const csx_region* region, void* d_M, int* npix,
                           int* ncol, int* nslices, double* calib_peak) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  if (!element_name) return fail(CSX_ERR_ARG, "null element name");
  return c->impl.hex_response(element_name, region, static_cast<double2*>(d_M), npix, ncol, nslices, calib_peak);
  CSX_GUARD_END
}

int csx_fits_read(const char* path, double* out, size_t capacity, int* rows, int* cols, int* slices) {
  CSX_GUARD_BEGIN
  if (!path || !rows || !cols || !slices) return fail(CSX_ERR_ARG, "csx_fits_read: null argument");
  std::vector<double> m;
  CSX_TRY(fits_matrix(path, m, *rows, *cols, *slices));
  if (out) {
    if (capacity < m.size()) return fail(CSX_ERR_ARG, "csx_fits_read: capacity below rows x cols x slices");
    std::memcpy(out, m.data(), sizeof(double) * m.size());
  }
  return CSX_OK;
  CSX_GUARD_END
}

int csx_fits_write(const char* path, const double* data, int rows, int cols) {
  CSX_GUARD_BEGIN
  if (!path) return fail(CSX_ERR_ARG, "csx_fits_write: null path");
  return write_fits_matrix(path, data, rows, cols);
  CSX_GUARD_END
}

int csx_chain_region_fields(csx_chain* c, const csx_region* region, void* d_out, int* npix, int* ncand,
                            int* nslices) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  return c->impl.region_fields(region, static_cast<double2*>(d_out), npix, ncand, nslices);
  CSX_GUARD_END
}

int csx_group_create(csx_chain* const* chains, int n, csx_group** out) {
  CSX_GUARD_BEGIN
  if (!chains || n <= 0 || !out) return fail(CSX_ERR_ARG, "csx_group_create: bad arguments");
  *out = nullptr;
  std::vector<Chain*> v;
  for (int i = 0; i < n; i++) {
    if (!chains[i]) return fail(CSX_ERR_ARG, "csx_group_create: null chain");
    v.push_back(&chains[i]->impl);
  }
  *out = new csx_group(std::move(v));
  return CSX_OK;
  CSX_GUARD_END
}

void csx_group_destroy(csx_group* g) { delete g; }

int csx_group_candidates(csx_group* g) {
  if (!g) return fail(CSX_ERR_ARG, "null group");
  return g->impl.n_candidates();
}

int csx_group_contrast(csx_group* g, const csx_region* region, double* out) {
  CSX_GUARD_BEGIN
  if (!g || !out) return fail(CSX_ERR_ARG, "csx_group_contrast: null argument");
  return g->impl.contrast(region, out);
  CSX_GUARD_END
}

int csx_group_merged_launches(csx_group* g) {
  if (!g) return fail(CSX_ERR_ARG, "null group");
  return g->impl.merged_launches();
}

int csx_chain_set_graph(csx_chain* c, int enable) {
  CSX_TRY(chk(c));
  return c->impl.set_graph(enable != 0);
}

int csx_chain_profile_json(csx_chain* c, char* buf, size_t cap) {
  CSX_GUARD_BEGIN
  CSX_TRY(chk(c));
  std::string j = c->impl.profile_json();
  if (buf && cap) {
    size_t n = std::min(cap - 1, j.size());
    std::memcpy(buf, j.data(), n);
    buf[n] = 0;
  }
  return (int)j.size();
  CSX_GUARD_END
}

}  // extern "C"
