// Matrix Fourier transform plans (host side).
#include "mft.hpp"
#include "planes.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

#include <dlfcn.h>
#include <mutex>
#include <unistd.h>

namespace csx {

int terms_of(csx_precision p) { return p == CSX_BF16X3 ? 2 : (p == CSX_BF16X6 ? 3 : 0); }

// ------------------------------------------------------------- zero patterns --
std::shared_ptr<const Support> Support::of_mask(const double* m, int ny, int nx) {
  auto s = std::make_shared<Support>();
  s->ny = ny; s->nx = nx;
  s->lo.assign(nx, ny);
  s->hi.assign(nx, -1);
  for (int x = 0; x < nx; x++) {
    const double* c = m + (size_t)x * ny;
    int lo = 0, hi = ny - 1;
    while (lo < ny && c[lo] == 0.0) lo++;
    while (hi >= lo && c[hi] == 0.0) hi--;
    if (lo <= hi) { s->lo[x] = lo; s->hi[x] = hi; }
  }
  return s;
}
bool Support::dense() const {
  for (int x = 0; x < nx; x++)
    if (lo[x] != 0 || hi[x] != ny - 1) return false;
  return true;
}
std::shared_ptr<const Support> Support::intersect(const std::shared_ptr<const Support>& a,
                                                  const std::shared_ptr<const Support>& b) {
  if (!a) return b;
  if (!b) return a;
  if (a->ny != b->ny || a->nx != b->nx) return nullptr;   // unrelated grids: assume nothing
  auto s = std::make_shared<Support>(*a);
  for (int x = 0; x < s->nx; x++) {
    s->lo[x] = std::max(a->lo[x], b->lo[x]);
    s->hi[x] = std::min(a->hi[x], b->hi[x]);
    if (s->lo[x] > s->hi[x]) { s->lo[x] = s->ny; s->hi[x] = -1; }
  }
  return s;
}

// K range tables (GemmProb::kr).  Stage 1 (M = out_y, N = in_x, K = in_y): the
// input rows that may be nonzero in the columns of each N block.  Stage 2 (M =
// out_x, N = out_y, K = in_x): the input columns holding any nonzero; a block whose
// outputs the epilogue mask zeroes takes no K at all.  Empty when nothing is skipped.
static std::vector<int> stage1_ranges(const MftField& F) {
  const Support* s = F.in_sup.get();
  if (!s || s->ny != F.in_y.n || s->nx != F.in_x.n || s->dense()) return {};
  const int M = F.out_y.n, N = F.in_x.n, K = F.in_y.n;
  const int mb = (M + kKrBlock - 1) / kKrBlock, nb = (N + kKrBlock - 1) / kKrBlock;
  std::vector<int> r((size_t)2 * mb * nb);
  for (int j = 0; j < nb; j++) {
    int lo = K, hi = 0;
    for (int x = j * kKrBlock; x < std::min(N, (j + 1) * kKrBlock); x++)
      if (s->lo[x] <= s->hi[x]) { lo = std::min(lo, s->lo[x]); hi = std::max(hi, s->hi[x] + 1); }
    for (int i = 0; i < mb; i++) { r[2 * ((size_t)i * nb + j)] = lo; r[2 * ((size_t)i * nb + j) + 1] = hi; }
  }
  return r;
}
static std::vector<int> stage2_ranges(const MftField& F) {
  const int M = F.out_x.n, N = F.out_y.n, K = F.in_x.n;
  int klo = 0, khi = K;
  const Support* si = F.add ? nullptr : F.in_sup.get();   // with an addend only the mask zeroes a block
  if (si && si->ny == F.in_y.n && si->nx == K) {
    klo = K; khi = 0;
    for (int x = 0; x < K; x++)
      if (si->lo[x] <= si->hi[x]) { klo = std::min(klo, x); khi = x + 1; }
  }
  const Support* so = F.out_sup.get();
  const bool use_out = so && so->ny == N && so->nx == M;
  if (!use_out && klo == 0 && khi == K) return {};
  const int mb = (M + kKrBlock - 1) / kKrBlock, nb = (N + kKrBlock - 1) / kKrBlock;
  // rows the next MFT's first contraction reads in each column band (the union of
  // the band's column supports, and of its pair's: a 128-wide tile covers two bands),
  // widened to whole blocks as its slabs are
  std::vector<int> rlo(mb, N), rhi(mb, 0);
  if (use_out && F.out_read_sup) {
    std::vector<int> blo(mb, N), bhi(mb, 0);
    for (int i = 0; i < mb; i++)
      for (int x = i * kKrBlock; x < std::min(M, (i + 1) * kKrBlock); x++)
        if (so->lo[x] <= so->hi[x]) { blo[i] = std::min(blo[i], so->lo[x]); bhi[i] = std::max(bhi[i], so->hi[x] + 1); }
    for (int i = 0; i < mb; i++) {
      const int p = std::min(i ^ 1, mb - 1);
      rlo[i] = std::min(blo[i], blo[p]) / kKrBlock * kKrBlock;
      rhi[i] = (std::max(bhi[i], bhi[p]) + kKrBlock - 1) / kKrBlock * kKrBlock;
    }
  }
  std::vector<int> r((size_t)2 * mb * nb);
  for (int i = 0; i < mb; i++)
    for (int j = 0; j < nb; j++) {
      bool live = !use_out;
      for (int x = i * kKrBlock; !live && x < std::min(M, (i + 1) * kKrBlock); x++)
        live = so->lo[x] <= so->hi[x] && so->lo[x] < (j + 1) * kKrBlock && so->hi[x] >= j * kKrBlock;
      const bool read = !(use_out && F.out_read_sup) || (j * kKrBlock < rhi[i] && (j + 1) * kKrBlock > rlo[i]);
      r[2 * ((size_t)i * nb + j)] = live ? klo : (read ? K : -1);
      r[2 * ((size_t)i * nb + j) + 1] = live ? khi : 0;
    }
  return r;
}

int launch_gemm(const GemmProb* probs_dev, int n, int tiles, csx_precision p, hipStream_t s, int variant) {
  switch (p) {
    case CSX_FP64: return gemm_f64(probs_dev, n, tiles, s, variant);
    case CSX_FP32: return gemm_f32(probs_dev, n, tiles, s);
    case CSX_BF16X3: return gemm_bf16planes(probs_dev, n, tiles, 2, variant, s);
    case CSX_BF16X6: return gemm_bf16planes(probs_dev, n, tiles, 3, variant, s);
  }
  return fail(CSX_ERR_ARG, "unknown precision");
}

// Tile variants a precision can launch (variant 0: 64 x 64).
static int n_variants(csx_precision p) {
  if (p == CSX_FP64) return f64_variant_count();
  return p == CSX_BF16X3 || p == CSX_BF16X6 ? planes_variant_count(terms_of(p)) : 1;
}

// Forced variant from CSX_GEMM_VARIANT (an index), else -1 (choose by timing).
static int forced_variant(csx_precision p) {
  const char* e = std::getenv("CSX_GEMM_VARIANT");
  if (!e || !*e) return -1;
  int v = std::atoi(e);
  return v >= 0 && v < n_variants(p) ? v : -1;
}

// Autotuner cache: problem-list signature -> chosen (variant, K splits).  The choices
// persist in a JSON file so that every process (and every run on a fresh device of the
// same kind) launches the same kernels: CSX_TUNE_CACHE names it, by default
// tune_cache.json beside libcsx.so.  The file is read on first use; a choice timed in
// this process is added and the file rewritten (through a temporary and a rename).
// Entries of another kernel set (gemm_variant_names digest) are ignored.  Layout:
//   {"format": 1, "kernels": "<digest>",
//    "choices": {
//     "<signature>": [variant, ksplit],
//     ...}}
struct TuneState {
  std::mutex mu;
  bool loaded = false;
  std::string path, digest;
  std::map<std::string, std::pair<int, int>> choices;
  std::map<std::string, int> source;   // signature -> 0 file, 1 timed here (this process's lookups)
};
static TuneState& tune_state() {
  static TuneState t;
  return t;
}
static std::string kernel_digest() {
  unsigned long long h = 1469598103934665603ull;   // FNV-1a, 64 bit
  for (unsigned char c : gemm_variant_names()) h = (h ^ c) * 1099511628211ull;
  char b[32];
  std::snprintf(b, sizeof b, "%016llx", h);
  return b;
}
static std::string default_cache_path() {
  if (const char* e = std::getenv("CSX_TUNE_CACHE")) return e;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&kernel_digest), &info) && info.dli_fname) {
    std::string f = info.dli_fname;
    const size_t k = f.rfind('/');
    return (k == std::string::npos ? std::string(".") : f.substr(0, k)) + "/tune_cache.json";
  }
  return "tune_cache.json";
}
static void load_tune_cache(TuneState& t) {
  if (t.loaded) return;
  t.loaded = true;
  t.path = default_cache_path();
  t.digest = kernel_digest();
  std::FILE* f = std::fopen(t.path.c_str(), "r");
  if (!f) return;
  std::string text;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
  std::fclose(f);
  const std::string kk = "\"kernels\": \"";
  const size_t kp = text.find(kk);
  if (kp == std::string::npos || text.compare(kp + kk.size(), t.digest.size(), t.digest) != 0) return;
  size_t pos = text.find("\"choices\"");
  if (pos == std::string::npos) return;
  pos += 9;
  while (true) {                                      // "<signature>": [v, k] entries
    const size_t q0 = text.find('"', pos);
    if (q0 == std::string::npos) break;
    const size_t q1 = text.find('"', q0 + 1);
    if (q1 == std::string::npos) break;
    int v = -1, k = 0;
    if (std::sscanf(text.c_str() + q1 + 1, ": [%d, %d]", &v, &k) == 2 && v >= 0 && k >= 1)
      t.choices[text.substr(q0 + 1, q1 - q0 - 1)] = {v, k};
    pos = q1 + 1;
  }
}
static void save_tune_cache(const TuneState& t) {
  const std::string tmp = t.path + ".tmp" + std::to_string((long)getpid());
  std::FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return;                                     // read-only location: in memory only
  std::fprintf(f, "{\"format\": 1, \"kernels\": \"%s\",\n \"choices\": {", t.digest.c_str());
  bool first = true;
  for (const auto& kv : t.choices) {
    std::fprintf(f, "%s\n  \"%s\": [%d, %d]", first ? "" : ",", kv.first.c_str(), kv.second.first, kv.second.second);
    first = false;
  }
  std::fprintf(f, "}}\n");
  const bool ok = std::fclose(f) == 0;
  if (!ok || std::rename(tmp.c_str(), t.path.c_str()) != 0) std::remove(tmp.c_str());
}

// precision, then the problems' shapes (and K-range table digests) in launch order,
// runs of equal shapes written once with their count: "3:27*1024x256x256"
static std::string signature(const std::vector<GemmProb>& v, csx_precision p) {
  std::string s = std::to_string((int)p);
  std::string last;
  int run = 0;
  auto flush = [&]() {
    if (run) s += ':' + std::to_string(run) + '*' + last;
  };
  for (const auto& g : v) {
    std::string one = std::to_string(g.M) + 'x' + std::to_string(g.N) + 'x' + std::to_string(g.K);
    if (g.kr) one += 'r' + std::to_string(g.kr_tag);
    if (one == last) { run++; continue; }
    flush();
    last = one;
    run = 1;
  }
  flush();
  return s;
}

// K splits worth trying for a launch on a bm x bn tile whose tiles do not fill the chip
// several times over and whose K is long enough to share: the split that brings the
// launch to about `target` blocks and half of it (empty: no split).  Per precision:
// the most splits, the least K per split, and the blocks a launch should reach (1, 2
// or 4 blocks of a variant fit on each of the 256 CUs).  Three-term plane outputs are
// finished in whole windows of kPlaneWindow columns, so their N must hold whole
// windows.
static std::vector<int> split_candidates(const std::vector<GemmProb>& v, csx_precision p, int bm, int bn) {
  std::vector<int> out;
  int kmax_split = 8, kmin_each = 128, target = 2048;
  if (p == CSX_FP64 || p == CSX_FP32) { kmax_split = 32; kmin_each = 32; target = 1024; }
  if (p == CSX_BF16X6) { kmax_split = 16; kmin_each = 64; target = 1024; }
  long tiles = 0;
  int kmin = 1 << 30;
  for (const auto& g : v) {
    tiles += (long)((g.M + bm - 1) / bm) * ((g.N + bn - 1) / bn);
    kmin = std::min(kmin, g.K);
    if (p == CSX_BF16X6 && g.out_kind == OUT_PLANES && (g.N % kPlaneWindow || g.transC)) return out;
  }
  if (const char* e = std::getenv("CSX_GEMM_KSPLIT")) {
    if (std::atoi(e) > 1) out.push_back(std::atoi(e));
    return out;
  }
  if (tiles >= target || kmin < 2 * kmin_each) return out;
  int k = (int)std::min<long>(kmax_split, (target + tiles - 1) / tiles);
  while (k > 1 && kmin / k < kmin_each) k--;
  if (k > 1) out.push_back(k);
  if (k / 2 > 1) out.push_back(k / 2);
  return out;
}

// Sets each problem's K splits and its workspace slice (ws: splits x M x N complex, float
// for two-term planes, double otherwise).
static size_t split_elem_bytes(csx_precision p) { return p == CSX_BF16X3 ? sizeof(float2) : sizeof(double2); }
static void set_split(std::vector<GemmProb>& v, csx_precision p, int ksplit, void* ws) {
  size_t off = 0;
  for (auto& g : v) {
    g.ksplit = ksplit;
    g.ws = ksplit > 1 ? static_cast<char*>(ws) + off : nullptr;
    g.ws_f64 = p != CSX_BF16X3;
    off += ksplit > 1 ? split_elem_bytes(p) * ksplit * g.M * g.N : 0;
  }
}

// Uploads the problem list planned for `variant` into `dev`; returns the tile count.
static int upload(std::vector<GemmProb>& v, csx_precision p, int variant, DevBuf& dev, hipStream_t s, int* tiles) {
  int tm = kTileM, tn = kTileN;
  if (p == CSX_BF16X3 || p == CSX_BF16X6) CSX_TRY(planes_variant_tile(terms_of(p), variant, &tm, &tn));
  if (p == CSX_FP64) CSX_TRY(f64_variant_tile(variant, &tm, &tn));
  *tiles = plan_tiles(v.data(), (int)v.size(), tm, tn);
  CSX_TRY(dev.reserve(v.size() * sizeof(GemmProb)));
  CSX_HIP(hipMemcpyAsync(dev.p, v.data(), v.size() * sizeof(GemmProb), hipMemcpyHostToDevice, s));
  CSX_HIP(hipStreamSynchronize(s));
  return CSX_OK;
}

static long total_elems(const std::vector<GemmProb>& v) {
  long e = 0;
  for (const auto& g : v) e += (long)g.M * g.N;
  return e;
}

// One contraction launch (plus the split-K finishing launch when the problems are split).
static int run_stage(const GemmProb* dev, const std::vector<GemmProb>& v, int tiles, csx_precision p, hipStream_t s,
                     int variant, int ksplit, long elems) {
  CSX_TRY(launch_gemm(dev, (int)v.size(), tiles, p, s, variant));
  if (ksplit > 1) CSX_TRY(gemm_splitk_finish(dev, (int)v.size(), elems, s));
  return CSX_OK;
}

// Picks the fastest (tile variant, K splits) of one grouped launch by timing: two
// passes over the candidates, each timing 3 back-to-back launches between two HIP
// events on the launch stream after one warm-up launch; a candidate's time is its
// better pass, and a later one must beat the best earlier one by 2 % (choices within
// timing noise stay put).  The launches write the stage's own output buffers (and
// its workspace) only, which the real run overwrites.
// tune = false (a plan built in caller memory): no timing launches, no split-K
// workspace; the persisted choice without its K split, else variant 0.
static int choose_variant(std::vector<GemmProb>& v, csx_precision p, DevBuf& dev, DevBuf& ws, hipStream_t s,
                          int* variant, int* ksplit, int* tiles, bool tune = true) {
  const int nv = n_variants(p);
  std::vector<std::vector<int>> kc(nv);   // K splits worth trying with each variant's tile
  int kmax = 1;
  for (int var = 0; var < nv && tune; var++) {
    int bm = kTileM, bn = kTileN;
    if (p == CSX_BF16X3 || p == CSX_BF16X6) CSX_TRY(planes_variant_tile(terms_of(p), var, &bm, &bn));
    if (p == CSX_FP64) CSX_TRY(f64_variant_tile(var, &bm, &bn));
    kc[var] = split_candidates(v, p, bm, bn);
    for (int k : kc[var]) kmax = std::max(kmax, k);
  }
  const size_t wsz = split_elem_bytes(p);
  if (kmax > 1) CSX_TRY(ws.reserve(wsz * (size_t)kmax * total_elems(v)));
  int forced = forced_variant(p);
  if (nv == 1 || forced >= 0) {
    *variant = nv == 1 ? 0 : forced;
    *ksplit = std::getenv("CSX_GEMM_KSPLIT") && tune && !kc[*variant].empty() ? kc[*variant][0] : 1;
    set_split(v, p, *ksplit, ws.p);
    return upload(v, p, *variant, dev, s, tiles);
  }
  const std::string key = signature(v, p);
  TuneState& ts = tune_state();
  {
    std::lock_guard<std::mutex> lock(ts.mu);
    load_tune_cache(ts);
    auto it = ts.choices.find(key);
    if (it != ts.choices.end() && it->second.first < nv &&
        (!tune || it->second.second == 1 ||
         std::count(kc[it->second.first].begin(), kc[it->second.first].end(), it->second.second))) {
      *variant = it->second.first;
      *ksplit = tune ? it->second.second : 1;
      if (!ts.source.count(key)) ts.source[key] = 0;
    } else {
      *variant = tune ? -1 : 0;
      *ksplit = 1;
    }
  }
  if (*variant >= 0) {
    if (*ksplit > 1) CSX_TRY(ws.reserve(wsz * (size_t)*ksplit * total_elems(v)));
    set_split(v, p, *ksplit, ws.p);
    return upload(v, p, *variant, dev, s, tiles);
  }
  std::vector<std::pair<int, int>> cand;
  for (int var = 0; var < nv; var++) cand.push_back({var, 1});
  for (int var = 0; var < nv; var++)
    for (int k : kc[var]) cand.push_back({var, k});
  const long elems = total_elems(v);
  hipEvent_t e0, e1;
  // timing-only events: no system-scope fence (no cache writeback between the timed launches)
  CSX_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
  CSX_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
  std::vector<float> tbest(cand.size(), 1e30f);
  int rc = CSX_OK;
  for (int pass = 0; pass < 2 && rc == CSX_OK; pass++) {
    for (size_t c = 0; c < cand.size() && rc == CSX_OK; c++) {
      const int var = cand[c].first, ks = cand[c].second;
      int nt = 0;
      set_split(v, p, ks, ws.p);
      if ((rc = upload(v, p, var, dev, s, &nt)) != CSX_OK) break;
      if ((rc = run_stage(dev.as<GemmProb>(), v, nt, p, s, var, ks, elems)) != CSX_OK) break;
      if (hipEventRecord(e0, s) != hipSuccess) { rc = fail(CSX_ERR_HIP, "autotune: event record"); break; }
      for (int r = 0; r < 3 && rc == CSX_OK; r++) rc = run_stage(dev.as<GemmProb>(), v, nt, p, s, var, ks, elems);
      if (rc != CSX_OK) break;
      if (hipEventRecord(e1, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        rc = fail(CSX_ERR_HIP, "autotune: event wait");
        break;
      }
      float t = 0;
      (void)hipEventElapsedTime(&t, e0, e1);
      tbest[c] = std::min(tbest[c], t / 3);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != CSX_OK) return rc;
  size_t best = 0;
  for (size_t c = 1; c < cand.size(); c++)
    if (tbest[c] < tbest[best] * 0.98f) best = c;
  if (std::getenv("CSX_TUNE_LOG")) {
    std::fprintf(stderr, "csx tune %s:", key.c_str());
    for (size_t c = 0; c < cand.size(); c++)
      std::fprintf(stderr, " v%d/k%d %.1fus%s", cand[c].first, cand[c].second, tbest[c] * 1e3, c == best ? "*" : "");
    std::fprintf(stderr, "\n");
  }
  {
    std::lock_guard<std::mutex> lock(ts.mu);
    ts.choices[key] = cand[best];
    ts.source[key] = 1;
    save_tune_cache(ts);
  }
  *variant = cand[best].first;
  *ksplit = cand[best].second;
  set_split(v, p, *ksplit, ws.p);
  return upload(v, p, *variant, dev, s, tiles);
}

// Problems of different shapes (a chain group: 512^2 and 1024^2 pupils) have tiles of
// different cost.  Each XCD works a contiguous range of tiles, so the shape classes
// are spread evenly over the list (each keeps its order): every range gets the same
// mix of work instead of one XCD taking the small problems and finishing early.
static void spread_shapes(std::vector<GemmProb>& v) {
  std::map<std::tuple<int, int, int>, std::vector<int>> cls;
  for (int i = 0; i < (int)v.size(); i++) cls[std::make_tuple(v[i].M, v[i].N, v[i].K)].push_back(i);
  if (cls.size() < 2) return;
  std::vector<std::pair<double, int>> key;
  for (auto& kv : cls) {
    const int n = (int)kv.second.size();
    for (int j = 0; j < n; j++) key.push_back({(j + 0.5) / n, kv.second[j]});
  }
  std::stable_sort(key.begin(), key.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  std::vector<GemmProb> out;
  out.reserve(v.size());
  for (const auto& k : key) out.push_back(v[k.second]);
  v.swap(out);
}

namespace {
struct DftKey {
  int nout, nin;
  double xo0, dxo, xi0, dxi, lf;
  bool operator<(const DftKey& o) const {
    return std::tie(nout, nin, xo0, dxo, xi0, dxi, lf) < std::tie(o.nout, o.nin, o.xo0, o.dxo, o.xi0, o.dxi, o.lf);
  }
};
}  // namespace

size_t MftPlan::arena_bytes(int in_x, int in_y, int out_x, int out_y, int batch, csx_precision p) {
  const int T = terms_of(p);
  const size_t esz = T ? (size_t)4 * T : (p == CSX_FP32 ? sizeof(float2) : sizeof(double2));
  const size_t b = (size_t)std::max(batch, 0);
  return Arena::round(b * ((size_t)out_y * in_y + (size_t)out_x * in_x) * esz) +   // DFT kernels
         Arena::round(2 * b * sizeof(DftSpec)) +                                    // their specs
         Arena::round(b * (size_t)out_y * in_x * esz) +                             // stage-1 outputs
         2 * Arena::round(b * sizeof(GemmProb));                                    // problem arrays
}

int MftPlan::build(const std::vector<MftField>& fields, csx_precision p, hipStream_t s, Arena* arena) {
  prec_ = p;
  for (DevBuf* d : {&dft_, &tbuf_, &probs1_, &probs2_, &specs_, &ws1_, &ws2_, &kr_}) {
    if (arena && d->arena != arena) { d->release(); d->arena = arena; }
  }
  nf_ = (int)fields.size();
  flops_ = flops1_ = flops2_ = bytes1_ = bytes2_ = 0;
  if (nf_ == 0) return CSX_OK;
  const int T = terms_of(p);
  const bool c64 = p == CSX_FP32;
  // bytes per complex element of the working operands (DFT kernels, stage-1 output)
  const size_t esz = T ? (size_t)4 * T : (c64 ? sizeof(float2) : sizeof(double2));
  const int want_in = T ? FK_PLANES : (c64 ? FK_C64 : FK_C128);
  for (const auto& f : fields) {
    if (f.in_x.n <= 0 || f.in_y.n <= 0 || f.out_x.n <= 0 || f.out_y.n <= 0)
      return fail(CSX_ERR_ARG, "mft: empty grid");
    if (f.lambda_f == 0) return fail(CSX_ERR_ARG, "mft: lambda*f is zero");
    if (!f.in || !f.out) return fail(CSX_ERR_ARG, "mft: null field pointer");
    if (f.in_kind != want_in) return fail(CSX_ERR_ARG, "mft: input storage does not match the precision");
    if (T && (f.in_x.n % 8 || f.in_y.n % 8 || f.out_y.n % 8))
      return fail(CSX_ERR_UNSUPPORTED, "mft: split-BF16 precisions need input sizes and output rows divisible by 8");
    // three-term planes are read in whole windows of 32 along every K (planes.hpp)
    if (T == 3 && (f.in_x.n % 32 || f.in_y.n % 32 || f.out_y.n % 32))
      return fail(CSX_ERR_UNSUPPORTED, "mft: bf16x6 needs input sizes and output rows divisible by 32");
  }
  // distinct DFT kernels
  std::map<DftKey, size_t> index;     // key -> element offset
  std::vector<DftKey> keys;
  size_t elems = 0;
  auto add = [&](const csx_grid& o, const csx_grid& i, double lf) {
    DftKey k{o.n, i.n, o.x0, o.dx, i.x0, i.dx, lf};
    auto it = index.find(k);
    if (it != index.end()) return it->second;
    size_t off = elems;
    index[k] = off;
    keys.push_back(k);
    elems += (size_t)o.n * i.n;
    return off;
  };
  // output grids of the DFT kernels: a source tilt shifts them (MftField::tilt_x/y)
  std::vector<csx_grid> ky(nf_), kx(nf_);
  for (int f = 0; f < nf_; f++) {
    ky[f] = fields[f].out_y;
    kx[f] = fields[f].out_x;
    ky[f].x0 -= fields[f].tilt_y * fields[f].lambda_f;
    kx[f].x0 -= fields[f].tilt_x * fields[f].lambda_f;
  }
  std::vector<size_t> offAy(nf_), offBx(nf_);
  for (int f = 0; f < nf_; f++) {
    offAy[f] = add(ky[f], fields[f].in_y, fields[f].lambda_f);
    offBx[f] = add(kx[f], fields[f].in_x, fields[f].lambda_f);
  }
  CSX_TRY(dft_.reserve(elems * esz));
  std::vector<DftSpec> specs;
  int maxe = 0;
  for (const auto& k : keys) {
    DftSpec d;
    d.out = static_cast<char*>(dft_.p) + index[k] * esz;
    d.nout = k.nout; d.nin = k.nin;
    d.xo0 = k.xo0; d.dxo = k.dxo; d.xi0 = k.xi0; d.dxi = k.dxi; d.lf = k.lf;
    d.out_c64 = c64 ? 1 : 0;
    d.terms = T;
    specs.push_back(d);
    maxe = std::max(maxe, k.nout * k.nin);
  }
  CSX_TRY(specs_.reserve(specs.size() * sizeof(DftSpec)));
  CSX_HIP(hipMemcpyAsync(specs_.p, specs.data(), specs.size() * sizeof(DftSpec), hipMemcpyHostToDevice, s));
  CSX_TRY(dft_matrices(specs_.as<DftSpec>(), (int)specs.size(), maxe, s));

  // stage-1 workspace T^T (out_y.n x in_x.n per distinct first contraction): fields
  // with the same input, input grids, output rows and lambda*f (the candidates of a
  // batch before their masks differ) share one T
  std::map<std::tuple<const void*, int, int, double, double, int, double, double, int, double, double, double>, int>
      t_index;
  std::vector<int> tof(nf_);
  std::vector<int> first;                 // a field of each distinct T
  std::vector<size_t> offT;
  size_t telems = 0;
  for (int f = 0; f < nf_; f++) {
    const MftField& F = fields[f];
    auto key = std::make_tuple(F.in, F.in_kind, F.in_x.n, F.in_x.x0, F.in_x.dx, F.in_y.n, F.in_y.x0, F.in_y.dx,
                               ky[f].n, ky[f].x0, ky[f].dx, F.lambda_f);
    auto it = t_index.find(key);
    if (it != t_index.end()) { tof[f] = it->second; continue; }
    tof[f] = (int)first.size();
    t_index[key] = tof[f];
    first.push_back(f);
    offT.push_back(telems);
    telems += (size_t)F.out_y.n * F.in_x.n;
  }
  nt_ = (int)first.size();
  CSX_TRY(tbuf_.reserve(telems * esz));

  std::vector<GemmProb> p1(nt_), p2(nf_);
  for (int u = 0; u < nt_; u++) {
    const MftField& F = fields[first[u]];
    const long npix_in = (long)F.in_x.n * F.in_y.n, nt = (long)F.out_y.n * F.in_x.n;
    GemmProb& a = p1[u];
    std::memset(&a, 0, sizeof(a));
    a.A = static_cast<char*>(dft_.p) + offAy[first[u]] * esz; a.lda = F.in_y.n; a.sA = (long)F.out_y.n * F.in_y.n;
    a.B = F.in; a.ldb = F.in_y.n; a.sB = npix_in;
    a.C = static_cast<char*>(tbuf_.p) + offT[u] * esz; a.ldc = F.in_x.n; a.sC = nt;
    a.M = F.out_y.n; a.N = F.in_x.n; a.K = F.in_y.n;
    a.transC = 0;
    a.out_kind = T ? OUT_PLANES : (c64 ? OUT_C64 : OUT_C128);
    a.terms = T;
    a.alpha_re = 1.0; a.alpha_im = 0.0;
    a.mask_kind = MASK_NONE;
    flops1_ += 8.0 * (double)F.in_x.n * F.in_y.n * F.out_y.n;
    const double ein = F.in_kind == FK_PLANES ? 4.0 * T : (F.in_kind == FK_C64 ? 8.0 : 16.0);
    bytes1_ += ein * F.in_x.n * F.in_y.n + (double)esz * F.out_y.n * F.in_x.n;
  }
  for (int f = 0; f < nf_; f++) {
    const MftField& F = fields[f];
    const long nt = (long)F.out_y.n * F.in_x.n;
    GemmProb& b = p2[f];
    std::memset(&b, 0, sizeof(b));
    b.A = static_cast<char*>(dft_.p) + offBx[f] * esz; b.lda = F.in_x.n; b.sA = (long)F.out_x.n * F.in_x.n;
    b.B = static_cast<char*>(tbuf_.p) + offT[tof[f]] * esz; b.ldb = F.in_x.n; b.sB = nt;
    b.C = F.out; b.ldc = F.out_y.n; b.sC = (long)F.out_x.n * F.out_y.n;
    b.M = F.out_x.n; b.N = F.out_y.n; b.K = F.in_x.n;
    b.transC = 0;
    b.out_kind = F.out_kind == FK_PLANES ? OUT_PLANES : (F.out_kind == FK_C64 ? OUT_C64 : OUT_C128);
    b.terms = T;
    // s = dx dy / (i lf) = -i dx dy / lf
    b.alpha_re = 0.0; b.alpha_im = -F.in_scale * (F.in_x.dx * F.in_y.dx) / F.lambda_f;
    b.addend = F.add;
    b.mask_kind = F.mask_kind;
    if (F.mask_kind != MASK_NONE) {
      b.mask = F.mask; b.ldm = F.mrows;
      b.mm0 = F.col0; b.mn0 = F.row0; b.mM = F.mcols; b.mN = F.mrows;
    }
    flops2_ += 8.0 * (double)F.out_x.n * F.in_x.n * F.out_y.n;
    double eout = F.out_kind == FK_PLANES ? 4.0 * T : (F.out_kind == FK_C64 ? 8.0 : 16.0);
    bytes2_ += eout * F.out_x.n * F.out_y.n;
  }
  for (int u = 0; u < nt_; u++) {
    const MftField& F = fields[first[u]];
    bytes2_ += (double)esz * F.out_y.n * F.in_x.n;      // each T read once
  }
  flops_ = flops1_ + flops2_;
  // K range tables, one copy of each distinct table
  {
    std::map<std::vector<int>, size_t> tindex;
    std::map<size_t, unsigned> tag;
    std::vector<int> all;
    std::vector<long> off1(nt_, -1), off2(nf_, -1);
    auto place = [&](std::vector<int>&& t) -> long {
      if (t.empty()) return -1;
      auto it = tindex.find(t);
      if (it != tindex.end()) return (long)it->second;
      size_t o = all.size();
      unsigned h = 2166136261u;                    // FNV-1a over the entries
      for (int v : t) h = (h ^ (unsigned)v) * 16777619u;
      tag[o] = h;
      all.insert(all.end(), t.begin(), t.end());
      tindex.emplace(std::move(t), o);
      return (long)o;
    };
    // products each output needs (its block's K range): the flop count of a stage
    auto needed = [](const std::vector<int>& t, int M, int N) {
      const int nb = (N + kKrBlock - 1) / kKrBlock;
      double w = 0;
      for (size_t b = 0; 2 * b < t.size(); b++) {
        const int i = (int)(b / nb), j = (int)(b % nb);
        const int rows = std::min(kKrBlock, M - i * kKrBlock), cols = std::min(kKrBlock, N - j * kKrBlock);
        w += (double)rows * cols * std::max(0, t[2 * b + 1] - std::max(0, t[2 * b]));
      }
      return 8.0 * w;
    };
    for (int u = 0; u < nt_; u++) {
      std::vector<int> t = stage1_ranges(fields[first[u]]);
      if (!t.empty()) flops1_ += needed(t, p1[u].M, p1[u].N) - 8.0 * p1[u].M * p1[u].N * p1[u].K;
      off1[u] = place(std::move(t));
    }
    for (int f = 0; f < nf_; f++) {
      std::vector<int> t = stage2_ranges(fields[f]);
      if (!t.empty()) flops2_ += needed(t, p2[f].M, p2[f].N) - 8.0 * p2[f].M * p2[f].N * p2[f].K;
      off2[f] = place(std::move(t));
    }
    flops_ = flops1_ + flops2_;
    if (!all.empty()) {
      CSX_TRY(kr_.reserve(all.size() * sizeof(int)));
      CSX_HIP(hipMemcpyAsync(kr_.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice, s));
      CSX_HIP(hipStreamSynchronize(s));   // `all` is a host temporary
      auto bind = [&](GemmProb& g, long off) {
        if (off < 0) return;
        g.kr = kr_.as<int>() + off;
        g.kr_ld = (g.N + kKrBlock - 1) / kKrBlock;
        g.kr_tag = tag[(size_t)off];
      };
      for (int u = 0; u < nt_; u++) bind(p1[u], off1[u]);
      for (int f = 0; f < nf_; f++) bind(p2[f], off2[f]);
    }
  }
  // DFT kernels are read once per launch as well (each distinct one)
  {
    std::set<size_t> ay(offAy.begin(), offAy.end()), bx(offBx.begin(), offBx.end());
    for (int f = 0; f < nf_; f++) {
      if (ay.erase(offAy[f])) bytes1_ += (double)esz * fields[f].out_y.n * fields[f].in_y.n;
      if (bx.erase(offBx[f])) bytes2_ += (double)esz * fields[f].out_x.n * fields[f].in_x.n;
    }
  }
  // problems that share a DFT kernel are launched next to each other (an XCD works a
  // contiguous tile range, so its L2 keeps the kernel across those fields)
  auto by_kernel = [](std::vector<GemmProb>& v) {
    std::stable_sort(v.begin(), v.end(), [](const GemmProb& a, const GemmProb& b) { return a.A < b.A; });
  };
  by_kernel(p1);
  by_kernel(p2);
  spread_shapes(p1);
  spread_shapes(p2);
  // stage 2 reads stage 1's output, so stage 1 is tuned (and left written) first
  CSX_TRY(choose_variant(p1, p, probs1_, ws1_, s, &var1_, &ks1_, &tiles1_, arena == nullptr));
  CSX_TRY(choose_variant(p2, p, probs2_, ws2_, s, &var2_, &ks2_, &tiles2_, arena == nullptr));
  elems1_ = total_elems(p1);
  elems2_ = total_elems(p2);
  return CSX_OK;
}

// Choices looked up or timed in this process: {"file": path, "from_file": n, "timed": n,
// "choices": {"signature": [variant, ksplit, source], ...}} (source 0 file, 1 timed).
std::string tune_report() {
  TuneState& t = tune_state();
  std::lock_guard<std::mutex> lock(t.mu);
  load_tune_cache(t);
  int nf = 0, nt = 0;
  std::string c;
  for (const auto& kv : t.source) {
    const auto& ch = t.choices[kv.first];
    (kv.second ? nt : nf)++;
    char b[64];
    std::snprintf(b, sizeof b, "[%d, %d, %d]", ch.first, ch.second, kv.second);
    c += (c.empty() ? "\"" : ", \"") + kv.first + "\": " + b;
  }
  std::string path;
  for (char ch : t.path) { if (ch == '"' || ch == '\\') path += '\\'; path += ch; }
  return "{\"file\": \"" + path + "\", \"kernels\": \"" + t.digest + "\", \"entries\": " +
         std::to_string(t.choices.size()) + ", \"from_file\": " + std::to_string(nf) +
         ", \"timed\": " + std::to_string(nt) + ", \"choices\": {" + c + "}}";
}

static const char* prec_name(csx_precision p) {
  switch (p) {
    case CSX_FP64: return "fp64";
    case CSX_FP32: return "fp32";
    case CSX_BF16X3: return "bf16x3";
    case CSX_BF16X6: return "bf16x6";
  }
  return "?";
}

int MftPlan::run(hipStream_t s, KernelTimer* t) const {
  if (nf_ == 0) return CSX_OK;
  const bool timed = t && t->enabled();
  std::string base = std::string("gemm_") + prec_name(prec_);
  // one event pair around the two contractions (they run back to back; a split-K
  // finishing launch belongs to its contraction): the average launch duration is
  // the pair's time over 2, gaps included
  if (timed) CSX_TRY(t->begin(base, flops1_ + flops2_, bytes1_ + bytes2_, s, 2));
  CSX_TRY(launch_gemm(probs1_.as<GemmProb>(), nt_, tiles1_, prec_, s, var1_));
  if (ks1_ > 1) CSX_TRY(gemm_splitk_finish(probs1_.as<GemmProb>(), nt_, elems1_, s));
  CSX_TRY(launch_gemm(probs2_.as<GemmProb>(), nf_, tiles2_, prec_, s, var2_));
  if (ks2_ > 1) CSX_TRY(gemm_splitk_finish(probs2_.as<GemmProb>(), nf_, elems2_, s));
  if (timed) CSX_TRY(t->end(s));
  return CSX_OK;
}

// ------------------------------------------------------------ KernelTimer --
KernelTimer::~KernelTimer() {
  for (hipEvent_t e : all_) (void)hipEventDestroy(e);
}
void KernelTimer::enable(bool on) {
  on_ = on;
  stats_.clear();
  for (const Rec& r : open_) { free_.push_back(r.b); free_.push_back(r.e); }
  open_.clear();
}
int KernelTimer::get_event(hipEvent_t* e) {
  if (free_.empty()) {
    hipEvent_t n;
    // read only after the stream is synchronised: no system-scope fence, whose cache
    // writeback would stall the launches between the timed ones
    CSX_HIP(hipEventCreateWithFlags(&n, hipEventDisableSystemFence));
    all_.push_back(n);
    free_.push_back(n);
  }
  *e = free_.back();
  free_.pop_back();
  return CSX_OK;
}
int KernelTimer::begin(const std::string& name, double flops, double bytes, hipStream_t s, int launches) {
  Rec r{name, flops, bytes, nullptr, nullptr, launches};
  CSX_TRY(get_event(&r.b));
  CSX_TRY(get_event(&r.e));
  open_.push_back(r);
  CSX_HIP(hipEventRecord(r.b, s));
  return CSX_OK;
}
int KernelTimer::end(hipStream_t s) {
  if (open_.empty()) return fail(CSX_ERR_STATE, "KernelTimer::end without begin");
  CSX_HIP(hipEventRecord(open_.back().e, s));
  return CSX_OK;
}
int KernelTimer::collect() {
  for (const Rec& r : open_) {
    float ms = 0;
    CSX_HIP(hipEventElapsedTime(&ms, r.b, r.e));
    Stat& st = stats_[r.name];
    st.count += r.launches;
    st.ms += ms;
    st.flops += r.flops;
    st.bytes += r.bytes;
    free_.push_back(r.b);
    free_.push_back(r.e);
  }
  open_.clear();
  return CSX_OK;
}
std::string KernelTimer::json() const {
  std::string o = "{";
  bool first = true;
  for (const auto& kv : stats_) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s\"%s\": {\"launches\": %ld, \"ms_total\": %.9g, \"flops_total\": %.9g, \"bytes_total\": %.9g}",
                  first ? "" : ", ", kv.first.c_str(), kv.second.count, kv.second.ms, kv.second.flops, kv.second.bytes);
    o += buf;
    first = false;
  }
  return o + "}";
}

}  // namespace csx
