// Element chain engine: csim input scripts compiled into GPU stages.
#include "engine.hpp"

#include <complex>
#include <limits>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "gemm.hpp"
#include "ops.hpp"

namespace csx {

// ================================================================ parsing ==
// initCommandSet(file) (src/lib/csim_parser.cpp:60-82) and add_command (:93-118):
// leading blanks stripped, '#' lines and empty lines skipped, "key:arg" split at
// the first ':' (a line without ':' is a bare command).
Block parse_script_text(const std::string& text) {
  Block b;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    size_t p = line.find_first_not_of(" \t");
    if (p == std::string::npos) continue;
    line.erase(0, p);
    if (line[0] == '#') continue;
    size_t c = line.find(':');
    if (c == std::string::npos) b.cmds.push_back({line, ""});
    else b.cmds.push_back({line.substr(0, c), line.substr(c + 1)});
  }
  return b;
}

// find_command_blocks (src/lib/csim_parser.cpp:131-190): top-level start/end
// blocks (nested ones kept inside); commands outside any block form one extra
// block placed first.
std::vector<Block> find_blocks(const Block& b) {
  std::vector<Block> out;
  Block loose;
  bool has_loose = false;
  int depth = 0;
  for (const Cmd& c : b.cmds) {
    if (c.key == "start") {
      if (depth == 0) out.push_back(Block());
      else out.back().cmds.push_back(c);
      depth++;
    } else if (c.key == "end") {
      depth--;
      if (depth > 0) out.back().cmds.push_back(c);
    } else if (depth == 0) {
      loose.cmds.push_back(c);
      has_loose = true;
    } else {
      out.back().cmds.push_back(c);
    }
  }
  if (has_loose) out.insert(out.begin(), loose);
  return out;
}

static std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}
static std::vector<double> parse_list(const std::string& s) {
  std::vector<double> v;
  std::string t = s;
  for (char& ch : t) if (ch == ',') ch = ' ';
  std::istringstream in(t);
  double x;
  while (in >> x) v.push_back(x);
  return v;
}
// sscanf("%f, %f") as regionContrast::set does: values pass through float.
static std::vector<double> parse_float_pair(const std::string& s) {
  float a = 0, b = 0;
  std::sscanf(s.c_str(), "%f, %f", &a, &b);
  return {(double)a, (double)b};
}

// ---------------------------------------------------------------- FITS ---
int read_fits(const std::string& path, std::vector<double>& data, std::vector<long>& dims) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(CSX_ERR_ARG, "cannot open FITS file " + path);
  int bitpix = 0, naxis = 0;
  long ax[3] = {1, 1, 1};
  double bscale = 1, bzero = 0;
  bool end = false;
  char card[81];
  card[80] = 0;
  while (!end) {
    for (int i = 0; i < 36; i++) {
      if (!f.read(card, 80)) return fail(CSX_ERR_PARSE, "truncated FITS header in " + path);
      std::string k = trim(std::string(card, 8));
      std::string v = std::string(card + 10, 70);
      if (k == "END") end = true;
      else if (k == "BITPIX") bitpix = std::atoi(v.c_str());
      else if (k == "NAXIS") naxis = std::atoi(v.c_str());
      else if (k == "NAXIS1") ax[0] = std::atol(v.c_str());
      else if (k == "NAXIS2") ax[1] = std::atol(v.c_str());
      else if (k == "NAXIS3") ax[2] = std::atol(v.c_str());
      else if (k == "BSCALE") bscale = std::atof(v.c_str());
      else if (k == "BZERO") bzero = std::atof(v.c_str());
    }
  }
  if (naxis < 1 || naxis > 3) return fail(CSX_ERR_UNSUPPORTED, "FITS NAXIS must be 1..3");
  dims.assign(ax, ax + naxis);
  size_t n = 1;
  for (int i = 0; i < naxis; i++) n *= (size_t)ax[i];
  int bytes = std::abs(bitpix) / 8;
  std::vector<unsigned char> raw(n * bytes);
  if (!f.read(reinterpret_cast<char*>(raw.data()), raw.size())) return fail(CSX_ERR_PARSE, "truncated FITS data");
  data.resize(n);
  for (size_t i = 0; i < n; i++) {
    unsigned char* p = &raw[i * bytes];
    uint64_t u = 0;
    for (int b = 0; b < bytes; b++) u = (u << 8) | p[b];   // big endian
    double v;
    switch (bitpix) {
      case -64: { double d; std::memcpy(&d, &u, 8); v = d; break; }
      case -32: { uint32_t w = (uint32_t)u; float x; std::memcpy(&x, &w, 4); v = x; break; }
      case 32: v = (double)(int32_t)(uint32_t)u; break;
      case 16: v = (double)(int16_t)(uint16_t)u; break;
      case 8: v = (double)(uint8_t)u; break;
      default: return fail(CSX_ERR_UNSUPPORTED, "FITS BITPIX not supported");
    }
    data[i] = bzero + bscale * v;
  }
  return CSX_OK;
}

// save_mat (src/lib/csim_fits.cpp:480-493): BITPIX -64, NAXIS1 = columns, NAXIS2 =
// rows, the data row by row (the transposed column-major matrix), big-endian, header
// and data each padded to 2880-byte blocks. As write_fits_array (:101-109) does, an
// existing file is replaced and spaces in the name become underscores.
int write_fits_matrix(const std::string& name, const double* colmajor, int rows, int cols) {
  if (rows <= 0 || cols <= 0 || !colmajor) return fail(CSX_ERR_ARG, "write_fits: empty matrix");
  std::string path = name;
  if (!path.empty() && path[0] == '!') path.erase(0, 1);
  std::replace(path.begin(), path.end(), ' ', '_');
  std::string h;
  auto card = [&](const std::string& key, const std::string& val) {
    char c[81];
    std::snprintf(c, sizeof c, "%-8s= %20s", key.c_str(), val.c_str());
    std::string line(c);
    line.resize(80, ' ');
    h += line;
  };
  card("SIMPLE", "T");
  card("BITPIX", "-64");
  card("NAXIS", "2");
  card("NAXIS1", std::to_string(cols));
  card("NAXIS2", std::to_string(rows));
  card("EXTEND", "T");
  std::string end = "END";
  end.resize(80, ' ');
  h += end;
  h.resize((h.size() + 2879) / 2880 * 2880, ' ');
  std::vector<unsigned char> d((size_t)rows * cols * 8);
  for (int r = 0; r < rows; r++)
    for (int c = 0; c < cols; c++) {
      double v = colmajor[(size_t)c * rows + r];
      unsigned char b[8];
      std::memcpy(b, &v, 8);
      for (int k = 0; k < 8; k++) d[((size_t)r * cols + c) * 8 + k] = b[7 - k];
    }
  d.resize((d.size() + 2879) / 2880 * 2880, 0);
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(CSX_ERR_ARG, "cannot create FITS file " + path);
  f.write(h.data(), (std::streamsize)h.size());
  f.write(reinterpret_cast<const char*>(d.data()), (std::streamsize)d.size());
  if (!f) return fail(CSX_ERR_ARG, "short write to FITS file " + path);
  return CSX_OK;
}

ParseAudit*& parse_audit() {
  thread_local ParseAudit* a = nullptr;
  return a;
}

// FITS matrix -> column-major rows x cols (rows = NAXIS2, cols = NAXIS1), csim_fits.cpp:195-233.
int fits_matrix(const std::string& path, std::vector<double>& m, int& rows, int& cols, int& slices) {
  if (ParseAudit* a = parse_audit()) {
    a->files.push_back(path);
    m.clear();
    rows = cols = slices = 0;
    return CSX_OK;
  }
  std::vector<double> d;
  std::vector<long> dims;
  CSX_TRY(read_fits(path, d, dims));
  cols = (int)dims[0];
  rows = dims.size() > 1 ? (int)dims[1] : 1;
  slices = dims.size() > 2 ? (int)dims[2] : 1;
  if (dims.size() == 1) { rows = cols; cols = 1; }   // vectors: n x 1
  m.resize(d.size());
  for (int s = 0; s < slices; s++)
    for (int r = 0; r < rows; r++)
      for (int c = 0; c < cols; c++)
        m[(size_t)s * rows * cols + (size_t)c * rows + r] =
            dims.size() == 1 ? d[r] : d[(size_t)s * rows * cols + (size_t)r * cols + c];
  return CSX_OK;
}

// ================================================================ geometry ==
Geom geom_symmetric(int nx, int ny, double dx, double dy, double ox, double oy) {
  Geom g;
  g.nx = nx; g.ny = ny; g.dx = dx; g.dy = dy;
  g.x0 = -(dx * nx - dx) / 2.0 - ox;
  g.y0 = -(dy * ny - dy) / 2.0 - oy;
  return g;
}
Geom geom_offset(int nx, int ny, double dx, double dy) {
  Geom g;
  g.nx = nx; g.ny = ny; g.dx = dx; g.dy = dy;
  g.x0 = (0.0 - ((double)nx) / 2.) * dx;
  g.y0 = (0.0 - ((double)ny) / 2.) * dy;
  return g;
}
static csx_grid gx(const Geom& g) { return csx_grid{g.nx, g.x0, g.dx}; }
static csx_grid gy(const Geom& g) { return csx_grid{g.ny, g.y0, g.dy}; }

// A stage bracketed by HIP events of the chain's kernel timer while profiling is on
// (flops and bytes: the algorithmic work of one run of the stage).
static std::function<int(hipStream_t)> timed_stage(Chain& ch, const std::string& name, double flops, double bytes,
                                                   std::function<int(hipStream_t)> f) {
  Chain* chp = &ch;
  return [=](hipStream_t s) {
    KernelTimer* t = chp->timer();
    if (t) CSX_TRY(t->begin(name, flops, bytes, s));
    CSX_TRY(f(s));
    if (t) CSX_TRY(t->end(s));
    return CSX_OK;
  };
}

// ======================================================= fused pointwise ==
// Consecutive pointwise elements (and the point-source initialisation) are
// applied in one pass over the fields; any other element flushes them first.
// to_conv: a following MFT reads the result at the working precision, which the
// pass then writes straight into the conversion buffer (no separate split pass).
static int flush_pointwise(Chain& ch, Compile& c, bool to_conv = false) {
  if (c.pending.empty()) return CSX_OK;
  if (c.in_conv) return fail(CSX_ERR_STATE, "pointwise factors pending on converted fields");
  PwChain pc{};
  const int T = terms_of(ch.precision());
  const bool c64 = ch.precision() == CSX_FP32;
  if (to_conv && (T || c64)) {
    pc.out_kind = T ? PW_OUT_PLANES : PW_OUT_C64;
    pc.out = ch.c64_buf();
    pc.terms = T;
    c.in_conv = true;
  } else {
    pc.out_kind = PW_OUT_C128;
  }
  pc.nops = (int)c.pending.size();
  for (int i = 0; i < pc.nops; i++) pc.op[i] = c.pending[i];
  pc.rows = c.geom.ny;
  pc.cols = c.geom.nx;
  pc.pixel_scale = ch.efield().pixelScale;
  pc.src_params = ch.source_params(c.calibrating);
  pc.lambdas = c.calibrating ? ch.scratch("lam_cal").as<double>() : ch.lambdas_dev();
  double2* E = ch.field_buf(c.buf);
  int nf = ch.nfield_for(c);
  // fields are (candidate, source, lambda) with lambda fastest: a generating pass
  // over masks and mirrors only runs pixel-major (one phasor per wavelength and source)
  bool pm = pc.op[0].kind == PW_SOURCE && ch.nlam() <= kMaxPwLam && c.geom.npix() % 2 == 0;
  for (int i = 1; i < pc.nops; i++) pm = pm && (pc.op[i].kind == PW_MASK_1PI || pc.op[i].kind == PW_MIRROR);
  pc.nl = pm ? ch.nlam() : 0;
  std::string name = "pointwise x" + std::to_string(pc.nops) + (pc.out_kind != PW_OUT_C128 ? " -> working precision" : "");
  // bytes: every field written (and read unless generated), each real map read once
  const double npix = (double)c.geom.npix();
  const double osz = pc.out_kind == PW_OUT_C128 ? 16.0 : (pc.out_kind == PW_OUT_C64 ? 8.0 : 4.0 * T);
  int maps = 0;
  for (int i = 0; i < pc.nops; i++) maps += pc.op[i].kind == PW_MASK_1PI || pc.op[i].kind == PW_MIRROR;
  const double bytes = nf * npix * (osz + (pc.op[0].kind == PW_SOURCE ? 0.0 : 16.0)) + maps * npix * 8.0;
  c.stages.push_back(timed_stage(ch, "pointwise", 0.0, bytes, [=](hipStream_t s) { return pointwise_chain(E, nf, pc, s); }));
  c.stage_names.push_back(name);
  c.pending.clear();
  return CSX_OK;
}

// Copies the shared fields (one candidate's) to every candidate's slots of the
// current complex128 buffer; afterwards the buffers hold every candidate.
static int expand_candidates(Chain& ch, Compile& c) {
  const int all = ch.ncand_for(c);
  if (c.live <= 0 || c.live >= all) { c.live = 0; return CSX_OK; }
  CSX_TRY(flush_pointwise(ch, c));
  if (c.in_conv) return fail(CSX_ERR_STATE, "candidate expansion of converted fields");
  const size_t bytes = sizeof(double2) * (size_t)ch.nfield_for(c) * c.geom.npix();
  char* E = reinterpret_cast<char*>(ch.field_buf(c.buf));
  c.stages.push_back([=](hipStream_t s) {
    for (int k = 1; k < all; k++) CSX_HIP(hipMemcpyAsync(E + k * bytes, E, bytes, hipMemcpyDeviceToDevice, s));
    return CSX_OK;
  });
  c.stage_names.push_back("candidate copies");
  c.live = 0;
  return CSX_OK;
}

static int push_pointwise(Chain& ch, Compile& c, PwOp op) {
  if ((int)c.pending.size() >= kMaxPwOps) CSX_TRY(flush_pointwise(ch, c));
  c.pending.push_back(op);
  return CSX_OK;
}

// ================================================================ elements ==
bool Element::set_common(const std::string& key, const std::string& arg) {
  std::string a = trim(arg);
  if (key == "position") position = std::atof(a.c_str());
  else if (key == "name") name = a;
  else if (key == "disableForCalibration") disableForCalibration = (a == "on");
  else if (key == "draw" || key == "save" || key == "break") { /* host-side debugging hooks: no GPU effect */ }
  else return false;
  return true;
}
int Element::set_opt(Chain&, const std::string& n, const std::vector<double>& v) {
  if (n == "position" && !v.empty()) { position = v[0]; return CSX_OK; }   // celem.cpp:132-145
  return fail(CSX_ERR_ARG, "element '" + name + "' has no optimisation data '" + n + "'");
}
int Element::get_opt(Chain&, const std::string& n, std::vector<double>& v) {
  if (n == "position") { v.assign(1, position); return CSX_OK; }
  return fail(CSX_ERR_ARG, "element '" + name + "' has no optimisation data '" + n + "'");
}
int Element::set_array(Chain&, const std::string& key, const double*, int, int) {
  return fail(CSX_ERR_ARG, "element '" + name + "' (" + type + ") takes no array '" + key + "'");
}
int Element::set_complex_array(Chain&, const std::string& key, const double*, int, int, int) {
  return fail(CSX_ERR_ARG, "element '" + name + "' (" + type + ") takes no complex array '" + key + "'");
}

static int unknown_key(const std::string& type, const std::string& key) {
  if (ParseAudit* a = parse_audit()) {
    a->unknown_keys.push_back(type + ":" + key);
    return CSX_OK;
  }
  std::fprintf(stderr, "!!! %s bad set field name: %s\n", type.c_str(), key.c_str());
  return CSX_OK;   // csim prints and continues
}

// Real mask / mirror data shared helper: resampled to the current grid with
// interp2 from the initial efield grid when sizes differ (mask.cpp:28-35,
// mirror.cpp:26-31, csim_geom.cpp:331-339).
struct RealMap {
  std::vector<double> host;
  int rows = 0, cols = 0;
  std::map<long, DevBuf> dev;   // keyed by rows*65536+cols of the grid it is resampled to
  std::map<long, std::shared_ptr<const Support>> sup;   // its zero pattern on that grid
  std::map<long, DevBuf> dev32;   // a float copy when every value is exact in float (a GEMM epilogue reads half)
  std::shared_ptr<const Support> support_for(const Geom& cur) const {
    auto it = sup.find((long)cur.ny * 65536 + cur.nx);
    return it == sup.end() ? nullptr : it->second;
  }
  std::map<long, DevBuf> dev8;    // a byte copy when every value is 0 or 1
  const float* f32_for(const Geom& cur) const {
    auto it = dev32.find((long)cur.ny * 65536 + cur.nx);
    return it == dev32.end() ? nullptr : it->second.as<float>();
  }
  const unsigned char* u8_for(const Geom& cur) const {
    auto it = dev8.find((long)cur.ny * 65536 + cur.nx);
    return it == dev8.end() ? nullptr : it->second.as<unsigned char>();
  }
  int device_for(Chain& ch, const Geom& cur, const double** out) {
    long key = (long)cur.ny * 65536 + cur.nx;
    auto it = dev.find(key);
    if (it == dev.end()) {
      std::vector<double> m;
      const Efield& ef = ch.efield();
      if (rows == cur.ny && cols == cur.nx) {
        m = host;
      } else {
        if (ef.rows == cur.ny && ef.cols == cur.nx) {
          m = host;   // reference returns without resampling when the grids match in size
          if (rows != cur.ny || cols != cur.nx) return fail(CSX_ERR_ARG, "mask size does not match the field");
        } else {
          // separable linear interpolation from the initial grid (pixelX/pixelY)
          auto inX = grid_symmetric(ef.cols, ef.pixelScale), inY = grid_symmetric(ef.rows, ef.pixelScale);
          std::vector<double> outX(cur.nx), outY(cur.ny);
          for (int i = 0; i < cur.nx; i++) outX[i] = cur.x0 + i * cur.dx;
          for (int i = 0; i < cur.ny; i++) outY[i] = cur.y0 + i * cur.dy;
          if (rows != ef.rows || cols != ef.cols) return fail(CSX_ERR_ARG, "mask size does not match the efield");
          auto tx = interp_table(inX, outX), ty = interp_table(inY, outY);
          std::vector<double> s1((size_t)rows * cur.nx);
          for (int r = 0; r < rows; r++)
            for (int c = 0; c < cur.nx; c++) {
              const InterpTable& t = tx[c];
              s1[(size_t)c * rows + r] = t.lo < 0 ? 0.0
                  : (1.0 - t.w) * host[(size_t)t.lo * rows + r] + t.w * host[(size_t)t.hi * rows + r];
            }
          m.assign((size_t)cur.ny * cur.nx, 0.0);
          for (int c = 0; c < cur.nx; c++)
            for (int r = 0; r < cur.ny; r++) {
              const InterpTable& t = ty[r];
              m[(size_t)c * cur.ny + r] = t.lo < 0 ? 0.0
                  : (1.0 - t.w) * s1[(size_t)c * rows + t.lo] + t.w * s1[(size_t)c * rows + t.hi];
            }
        }
      }
      DevBuf b;
      CSX_TRY(b.reserve(m.size() * sizeof(double)));
      CSX_HIP(hipMemcpy(b.p, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice));
      sup[key] = Support::of_mask(m.data(), cur.ny, cur.nx);
      std::vector<float> m32(m.size());
      bool exact = true;
      for (size_t i = 0; i < m.size() && exact; i++) {
        m32[i] = (float)m[i];
        exact = (double)m32[i] == m[i];
      }
      bool binary = true;
      for (size_t i = 0; i < m.size() && binary; i++) binary = m[i] == 0.0 || m[i] == 1.0;
      if (binary) {
        std::vector<unsigned char> m8(m.size());
        for (size_t i = 0; i < m.size(); i++) m8[i] = m[i] == 1.0 ? 1 : 0;
        DevBuf b8;
        CSX_TRY(b8.reserve(m8.size()));
        CSX_HIP(hipMemcpy(b8.p, m8.data(), m8.size(), hipMemcpyHostToDevice));
        dev8[key] = std::move(b8);
      }
      if (exact) {
        DevBuf b32;
        CSX_TRY(b32.reserve(m32.size() * sizeof(float)));
        CSX_HIP(hipMemcpy(b32.p, m32.data(), m32.size() * sizeof(float), hipMemcpyHostToDevice));
        dev32[key] = std::move(b32);
      }
      it = dev.emplace(key, std::move(b)).first;
    }
    *out = it->second.as<double>();
    return CSX_OK;
  }
  int set(const double* d, int r, int c) {
    if (!d || r <= 0 || c <= 0) return fail(CSX_ERR_ARG, "empty array");
    host.assign(d, d + (size_t)r * c);
    rows = r; cols = c;
    dev.clear();
    sup.clear();
    dev32.clear();
    dev8.clear();
    return CSX_OK;
  }
  int load(const std::string& path) {
    int r, c, s;
    std::vector<double> m;
    CSX_TRY(fits_matrix(trim(path), m, r, c, s));
    if (parse_audit()) return CSX_OK;   // a device-free check: the name is recorded, nothing read
    return set(m.data(), r, c);
  }
};

// ---- mask (src/elements/mask.cpp): E *= (1+i) m
struct MaskElem : Element {
  RealMap map;
  int set(Chain&, const std::string& k, const std::string& a) override {
    if (set_common(k, a) || k == "mask") return CSX_OK;
    if (k == "filename") return map.load(a);
    return unknown_key(type, k);
  }
  int set_array(Chain& ch, const std::string& k, const double* d, int r, int c) override {
    if (k != "filename" && k != "mask") return Element::set_array(ch, k, d, r, c);
    ch.invalidate();
    return map.set(d, r, c);
  }
  bool keeps_support() const override { return true; }
  int pointwise_op(Chain& ch, const Geom& g, PwOp* op) override {
    if (map.host.empty()) return fail(CSX_ERR_STATE, "mask '" + name + "' has no data");
    const double* m = nullptr;
    CSX_TRY(map.device_for(ch, g, &m));
    *op = PwOp{PW_MASK_1PI, m, 0.0, 1};
    return 1;
  }
  int compile(Chain& ch, Compile& c, int) override {
    PwOp op;
    if (int rc = pointwise_op(ch, c.geom, &op); rc < 0) return rc;
    CSX_TRY(push_pointwise(ch, c, op));
    c.support = Support::intersect(c.support, map.support_for(c.geom));   // E (1+i) m is zero where m is
    return CSX_OK;
  }
};

// ---- mirror (src/elements/mirror.cpp): E *= exp(-sign 4 pi i h / lambda)
struct MirrorElem : Element {
  RealMap map;
  double sign = 1;
  int set(Chain&, const std::string& k, const std::string& a) override {
    if (set_common(k, a) || k == "mirror") return CSX_OK;
    if (k == "filename") return map.load(a);
    if (k == "sign") {
      double v = std::atof(a.c_str());
      if (v == -1 || v == 1) sign = v;
      else std::fprintf(stderr, "!!!! mirror: illegal sign, ignoring\n");
      return CSX_OK;
    }
    return unknown_key(type, k);
  }
  int set_array(Chain& ch, const std::string& k, const double* d, int r, int c) override {
    if (k != "filename" && k != "sag") return Element::set_array(ch, k, d, r, c);
    ch.invalidate();
    return map.set(d, r, c);
  }
  bool keeps_support() const override { return true; }   // a phase factor
  int pointwise_op(Chain& ch, const Geom& g, PwOp* op) override {
    if (map.host.empty()) return fail(CSX_ERR_STATE, "mirror '" + name + "' has no data");
    const double* h = nullptr;
    CSX_TRY(map.device_for(ch, g, &h));
    *op = PwOp{PW_MIRROR, h, sign, 1};
    return 1;
  }
  int compile(Chain& ch, Compile& c, int) override {
    PwOp op;
    if (int rc = pointwise_op(ch, c.geom, &op); rc < 0) return rc;
    return push_pointwise(ch, c, op);
  }
};

// ---- complexMask (src/elements/complexMask.cpp): centred window *= amp e^{-i ph}
struct ComplexMaskElem : Element {
  std::vector<double2> cube;   // rows x cols x slices (amp e^{-i ph})
  int rows = 0, cols = 0, slices = 0, maskIndex = 0;
  DevBuf dev;
  bool keeps_support() const override { return true; }   // a pointwise factor
  int set(Chain&, const std::string& k, const std::string& a) override {
    if (set_common(k, a) || k == "complexMask") return CSX_OK;
    if (k == "maskIndex") { maskIndex = std::atoi(a.c_str()); return CSX_OK; }
    if (k == "filename") {
      std::string s = a;
      size_t cpos = s.find(',');
      if (cpos == std::string::npos) return fail(CSX_ERR_PARSE, "complexMask filename needs amp,ph");
      std::vector<double> am, ph;
      int r, c, sl, r2, c2, sl2;
      CSX_TRY(fits_matrix(trim(s.substr(0, cpos)), am, r, c, sl));
      CSX_TRY(fits_matrix(trim(s.substr(cpos + 1)), ph, r2, c2, sl2));
      if (r != r2 || c != c2 || sl != sl2) return fail(CSX_ERR_ARG, "complexMask amp/ph sizes differ");
      rows = r; cols = c; slices = sl;
      cube.resize(am.size());
      for (size_t i = 0; i < am.size(); i++) cube[i] = make_double2(am[i] * std::cos(-ph[i]), am[i] * std::sin(-ph[i]));
      return CSX_OK;
    }
    return unknown_key(type, k);
  }
  int set_complex_array(Chain& ch, const std::string& k, const double* d, int r, int c, int slice) override {
    if (k != "mask") return Element::set_complex_array(ch, k, d, r, c, slice);
    if (slice < 0) return fail(CSX_ERR_ARG, "negative slice");
    if (rows != r || cols != c) { rows = r; cols = c; slices = 0; cube.clear(); }
    if (slice >= slices) { slices = slice + 1; cube.resize((size_t)rows * cols * slices); }
    std::memcpy(&cube[(size_t)slice * rows * cols], d, sizeof(double2) * rows * cols);
    ch.invalidate();
    return CSX_OK;
  }
  int compile(Chain& ch, Compile& c, int) override {
    if (skipped(c)) return CSX_OK;
    if (cube.empty() || maskIndex >= slices) return fail(CSX_ERR_STATE, "complexMask '" + name + "' has no data");
    CSX_TRY(dev.reserve(sizeof(double2) * rows * cols));
    CSX_HIP(hipMemcpy(dev.p, &cube[(size_t)maskIndex * rows * cols], sizeof(double2) * rows * cols,
                      hipMemcpyHostToDevice));
    // expand to a full-size factor once (centred window, 1 elsewhere)
    Geom g = c.geom;
    std::vector<double2> full((size_t)g.npix(), make_double2(1, 0));
    int r0 = g.ny / 2 - rows / 2, c0 = g.nx / 2 - cols / 2;
    for (int cc = 0; cc < cols; cc++)
      for (int rr = 0; rr < rows; rr++) {
        int R = r0 + rr, C = c0 + cc;
        if (R >= 0 && R < g.ny && C >= 0 && C < g.nx)
          full[(size_t)C * g.ny + R] = cube[(size_t)maskIndex * rows * cols + (size_t)cc * rows + rr];
      }
    CSX_TRY(dev.reserve(sizeof(double2) * full.size()));
    CSX_HIP(hipMemcpy(dev.p, full.data(), sizeof(double2) * full.size(), hipMemcpyHostToDevice));
    int nf = ch.nfield_for(c);
    return push_pointwise(ch, c, PwOp{PW_CFACTOR, dev.p, 0.0, nf});
  }
};

// ---- fresnelPropagator (src/elements/fresnelPropagator.cpp + csim_fft.cpp:371-392)
struct FresnelElem : Element {
  double sign = 1;
  int padFactor = 1;
  int set(Chain&, const std::string& k, const std::string& a) override {
    if (set_common(k, a) || k == "fresnelPropagator") return CSX_OK;
    if (k == "propagationSign") {
      double v = std::atof(a.c_str());
      if (v == -1 || v == 1) sign = v;
      return CSX_OK;
    }
    if (k == "padFactor") { padFactor = std::atoi(a.c_str()); return CSX_OK; }
    return unknown_key(type, k);
  }
  int compile(Chain& ch, Compile& c, int index) override {
    if (c.in_conv) return fail(CSX_ERR_STATE, "fresnelPropagator: fields at a working precision");
    Element* prev = ch.elem_at(index - 1);
    Element* next = ch.elem_at(index + 1);
    if (!prev || !next) return fail(CSX_ERR_STATE, "fresnelPropagator needs a previous and a next element");
    double z = sign * std::fabs(next->position - prev->position);
    if (c.geom.nx != c.geom.ny) return fail(CSX_ERR_UNSUPPORTED, "fresnelPropagator: square arrays only");
    int n = c.geom.nx;
    double apRad = ch.efield().arrayRadius;
    double2* E = ch.field_buf(c.buf);
    int nf = ch.nfield_for(c);
    Chain* chp = &ch;
    bool cal = c.calibrating;
    // algorithmic work: two 2-D FFTs of the (padded) grid, each field read and written once
    const int nn = padFactor == 1 ? n : 2 * n;
    const double fl = nf * 2.0 * 5.0 * (double)nn * nn * std::log2((double)nn * nn);
    const double by = nf * 32.0 * (double)n * n;
    if (padFactor == 1) {
      // the pointwise elements before the step (pending: point-source generation, masks,
      // mirrors) run in its first pass as it reads the fields, the masks and mirrors
      // right after it in its last pass as it writes them: no separate pointwise pass
      PwChain pre{}, post{};
      const bool has_pre = !c.pending.empty();
      if (has_pre) {
        pre.nops = (int)c.pending.size();
        for (int i = 0; i < pre.nops; i++) pre.op[i] = c.pending[i];
        c.pending.clear();
      }
      int claim = 0;
      for (int j = index + 1; j < ch.n_elems() && claim < kMaxPwOps; j++) {
        Element* e = ch.elem_at(j);
        if (e->skipped(c)) break;
        PwOp op;
        const int rc = e->pointwise_op(ch, c.geom, &op);
        if (rc < 0) return rc;
        if (rc == 0) break;
        post.op[claim++] = op;
      }
      post.nops = claim;
      for (PwChain* pc : {&pre, &post}) {
        pc->out_kind = PW_OUT_C128;
        pc->rows = n; pc->cols = n;
        pc->pixel_scale = ch.efield().pixelScale;
        pc->src_params = ch.source_params(cal);
        pc->lambdas = cal ? ch.scratch("lam_cal").as<double>() : ch.lambdas_dev();
      }
      const bool has_post = claim > 0;
      c.claimed = claim;
      const double by_f = by + (has_pre && pre.op[0].kind == PW_SOURCE ? -16.0 * nf * (double)n * n : 0.0);
      c.stages.push_back(timed_stage(ch, "fresnel_as", fl, by_f, [=](hipStream_t s) {
        const double* l = cal ? chp->scratch("lam_cal").as<double>() : chp->lambdas_dev();
        return fresnel_batched(E, n, nf, l, apRad, z, s, has_pre ? &pre : nullptr, has_post ? &post : nullptr);
      }));
      c.stage_names.push_back("fresnelPropagator " + name + (has_pre ? " (+ pointwise before)" : "") +
                              (has_post ? " (+ pointwise after)" : ""));
      return CSX_OK;
    } else {
      CSX_TRY(flush_pointwise(ch, c));
      // embed in a 2n zero pad with apRad * padFactor (fresnelPropagator.cpp:45-54)
      double2* P = ch.field_buf(1 - c.buf);
      double ap = padFactor * apRad;
      c.stages.push_back(timed_stage(ch, "fresnel_as", fl, by, [=](hipStream_t s) {
        const double* l = cal ? chp->scratch("lam_cal").as<double>() : chp->lambdas_dev();
        CSX_TRY(embed_center(E, n, n, P, 2 * n, 2 * n, nf, 0, s));
        CSX_TRY(fresnel_batched(P, 2 * n, nf, l, ap, z, s));
        return embed_center(P, 2 * n, 2 * n, E, n, n, nf, 1, s);
      }));
    }
    c.stage_names.push_back("fresnelPropagator " + name);
    return CSX_OK;
  }
};

// ---- downsample (src/elements/downsample.cpp)
struct DownsampleElem : Element {
  int factor = 1;
  DevBuf tables;
  int set(Chain&, const std::string& k, const std::string& a) override {
    if (set_common(k, a) || k == "downsample") return CSX_OK;
    if (k == "sampleFactor") { factor = std::atoi(a.c_str()); return CSX_OK; }
    return unknown_key(type, k);
  }
  int geometry(Chain&, const Geom& in, Geom& out) override {
    if (factor < 1) return fail(CSX_ERR_ARG, "downsample: sampleFactor must be >= 1");
    // newGeom.set_geometry(n/factor, ...): symmetric grid of pixel factor*d (downsample.cpp:35-40)
    out = geom_symmetric(in.nx / factor, in.ny / factor, factor * in.dx, factor * in.dy);
    return CSX_OK;
  }
  int compile(Chain& ch, Compile& c, int) override {
    CSX_TRY(flush_pointwise(ch, c));
    Geom out;
    CSX_TRY(geometry(ch, c.geom, out));
    if (c.geom.nx != c.geom.ny) return fail(CSX_ERR_UNSUPPORTED, "downsample: square arrays only");
    std::vector<double> inX(c.geom.nx), inY(c.geom.ny), outX(out.nx), outY(out.ny);
    for (int i = 0; i < c.geom.nx; i++) inX[i] = c.geom.x0 + i * c.geom.dx;
    for (int i = 0; i < c.geom.ny; i++) inY[i] = c.geom.y0 + i * c.geom.dy;
    for (int i = 0; i < out.nx; i++) outX[i] = out.x0 + i * out.dx;
    for (int i = 0; i < out.ny; i++) outY[i] = out.y0 + i * out.dy;
    auto tx = interp_table(inX, outX), ty = interp_table(inY, outY);
    CSX_TRY(tables.reserve(sizeof(InterpTable) * (tx.size() + ty.size())));
    CSX_HIP(hipMemcpy(tables.p, tx.data(), sizeof(InterpTable) * tx.size(), hipMemcpyHostToDevice));
    CSX_HIP(hipMemcpy(tables.as<InterpTable>() + tx.size(), ty.data(), sizeof(InterpTable) * ty.size(),
                      hipMemcpyHostToDevice));
    const InterpTable* dtx = tables.as<InterpTable>();
    const InterpTable* dty = dtx + tx.size();
    double2* in = ch.field_buf(c.buf);
    double2* o = ch.field_buf(1 - c.buf);
    int nf = ch.nfield_for(c);
    Geom g = c.geom;
    c.stages.push_back(timed_stage(ch, "downsample", 0.0, nf * 16.0 * ((double)g.npix() + (double)out.npix()),
                                   [=](hipStream_t s) {
                                     return resample_batched(in, g.ny, g.nx, nf, o, out.ny, out.nx, dty, dtx, s);
                                   }));
    c.stage_names.push_back("downsample " + name);
    c.buf = 1 - c.buf;
    c.geom = out;
    return CSX_OK;
  }
};

// ---- MFT helpers shared by the Fraunhofer-type elements
// One grouped MFT over all fields, from the current fields (geometry c.geom)
// to geometry gout[lambda] with lambda*f lf[lambda].  Pending pointwise factors
// are applied first.  Optional epilogue: per-field masks (FPM plane) or one real
// csim mask (1+i) m over the whole output grid.  out_conv writes the result at
// the working precision into the conversion buffer for a following MFT.
// gin_per_lam: the input grid of each wavelength's fields when it varies with lambda
// (csim's doubleFraunhoferFPM FPM plane, whose pixel is proportional to lambda); the
// current geometry c.geom otherwise.
static int compile_mft(Chain& ch, Compile& c, const std::vector<Geom>& gout_per_lam,
                       const std::vector<double>& lf_per_lam, const std::string& label,
                       const std::vector<MftField>* masks = nullptr, bool out_conv = false,
                       const double* epi_mask_1pi = nullptr, bool expand = false,
                       std::shared_ptr<const Support> epi_sup = nullptr, const float* epi_mask_f32 = nullptr,
                       const unsigned char* epi_mask_u8 = nullptr, const std::vector<Geom>* gin_per_lam = nullptr) {
  CSX_TRY(flush_pointwise(ch, c, /*to_conv=*/true));
  // expand: the fields are one candidate's copy; the MFT writes every candidate's
  // output (its own mask each) from that copy, whose first contraction the plan
  // then computes once
  const int nf_in = ch.nfield_for(c), nl = ch.nlam(), ns = ch.nsrc();
  const int nf = ch.nfield_out(c, expand);
  const int ns_in = c.live_src ? 1 : ns;
  const Efield& ef = ch.efield();
  const Geom gin = c.geom;
  const int T = terms_of(ch.precision());
  const bool c64 = ch.precision() == CSX_FP32;
  const bool mixed = c64 || T;
  if (out_conv && !mixed) out_conv = false;
  double2* in = ch.field_buf(c.buf);
  double2* out = ch.field_buf(1 - c.buf);
  char* conv = static_cast<char*>(ch.c64_buf());
  const size_t ein = T ? (size_t)4 * T : sizeof(float2);
  const int wkind = T ? FK_PLANES : (c64 ? FK_C64 : FK_C128);
  const bool from_conv = c.in_conv;
  // CSX_ZERO_SKIP=0 turns the zero-pattern tables off (the dense contractions)
  const char* zs = std::getenv("CSX_ZERO_SKIP");
  const bool zero_skip = !(zs && zs[0] == '0');
  std::vector<MftField> F(nf);
  const Geom& g0 = gout_per_lam[0];
  for (int f = 0; f < nf; f++) {
    int l = f % nl;
    const Geom& go = gout_per_lam[l];
    if (go.nx != g0.nx || go.ny != g0.ny) return fail(CSX_ERR_UNSUPPORTED, label + ": output size varies with lambda");
    MftField& m = F[f];
    // input field of output f = (candidate, source, lambda): the one shared copy of
    // the candidates and of the sources where the buffers carry one
    const int s = (f / nl) % ns, cand = f / (nl * ns);
    const size_t fi = (size_t)(((c.live > 0 ? 0 : cand) * ns_in + (c.live_src ? 0 : s)) * nl + l);
    if (fi >= (size_t)nf_in) return fail(CSX_ERR_STATE, label + ": input field out of range");
    if (c.live_src) {   // efield::init_point_sources (src/data/efield.cpp:179-206) folded in
      const double Lx = ef.pixelScale * gin.nx, Ly = ef.pixelScale * gin.ny;
      m.in_scale = ef.sources[s].flux;
      m.tilt_x = 2 * ef.sources[s].tipX / Lx;
      m.tilt_y = 2 * ef.sources[s].tiltY / Ly;
    }
    m.in = mixed ? (const void*)(conv + fi * gin.npix() * ein) : (const void*)(in + fi * gin.npix());
    m.in_kind = wkind;
    if (out_conv) {
      m.out = conv + (size_t)f * go.npix() * ein;
      m.out_kind = wkind;
    } else {
      m.out = out + (size_t)f * go.npix();
      m.out_kind = FK_C128;
    }
    const Geom& gi = gin_per_lam ? (*gin_per_lam)[l] : gin;
    if (gi.nx != gin.nx || gi.ny != gin.ny) return fail(CSX_ERR_STATE, label + ": input size varies with lambda");
    m.in_x = gx(gi); m.in_y = gy(gi); m.out_x = gx(go); m.out_y = gy(go);
    m.lambda_f = lf_per_lam[l];
    if (zero_skip) m.in_sup = c.support;
    if (masks) {
      const MftField& mk = (*masks)[f];
      m.mask_kind = mk.mask_kind; m.mask = mk.mask;
      m.mrows = mk.mrows; m.mcols = mk.mcols; m.row0 = mk.row0; m.col0 = mk.col0;
    } else if (epi_mask_1pi) {
      // the narrowest exact copy of the mask (bytes for a binary stop, then float)
      m.mask_kind = epi_mask_u8 ? MASK_REAL_1PI_U8 : (epi_mask_f32 ? MASK_REAL_1PI_F32 : MASK_REAL_1PI);
      m.mask = epi_mask_u8 ? (const void*)epi_mask_u8
                           : (epi_mask_f32 ? (const void*)epi_mask_f32 : (const void*)epi_mask_1pi);
      m.mrows = go.ny; m.mcols = go.nx; m.row0 = 0; m.col0 = 0;
      if (zero_skip) m.out_sup = epi_sup;
      m.out_read_sup = zero_skip && out_conv;   // the next MFT reads it through epi_sup
    }
  }
  auto plan = std::make_unique<MftPlan>();
  CSX_TRY(plan->build(F, ch.precision(), ch.stream()));
  const MftPlan* pp = plan.get();
  c.plans.push_back(std::move(plan));
  size_t nconv = (size_t)nf_in * gin.npix();
  long npix_in = gin.npix();
  if (expand) c.live = 0;
  c.live_src = 0;
  Chain* chp = &ch;
  if (!from_conv && (c64 || T)) {   // operands to the working precision (no fused producer)
    c.stages.push_back([=](hipStream_t s) {
      KernelTimer* t = chp->timer();
      if (c64) {
        if (t) CSX_TRY(t->begin("convert_c128_to_c64", 0.0, 24.0 * nconv, s));
        CSX_TRY(convert_c128_to_c64(in, conv, nconv, s));
      } else {
        if (t) CSX_TRY(t->begin("split_c128_to_planes", 0.0, (16.0 + 4.0 * T) * nconv, s));
        CSX_TRY(split_c128_to_planes(in, conv, npix_in, nf_in, T, s));
      }
      if (t) CSX_TRY(t->end(s));
      return CSX_OK;
    });
    c.stage_names.push_back(label + " operands");
  }
  // the contraction itself; its field list lets a chain group merge this launch
  c.mft_at[c.stages.size()] = F;
  c.stages.push_back([=](hipStream_t s) { return pp->run(s, chp->timer()); });
  c.stage_names.push_back(label);
  if (out_conv) {
    c.in_conv = true;
  } else {
    c.in_conv = false;
    c.buf = 1 - c.buf;
  }
  c.geom = g0;
  c.support = epi_mask_1pi ? epi_sup : nullptr;
  return CSX_OK;
}

// ---- fraunhoferFocal (src/elements/fraunhoferFocal.cpp)
struct FraunhoferFocalElem : Element {
  double referenceLambda = 0, samplesPerflD = 0, FOVflD = 0;
  int set(Chain&, const std::string& k, const std::string& a) override {
    if (set_common(k, a) || k == "fraunhoferFocal") return CSX_OK;
    if (k == "referenceLambda") referenceLambda = std::atof(a.c_str());
    else if (k == "samplesPerflD") samplesPerflD = std::atof(a.c_str());
    else if (k == "FOVflD") FOVflD = std::atof(a.c_str());
    else return unknown_key(type, k);
    return CSX_OK;
  }
  Geom out_geom(Chain& ch) const {   // set_geometry (:63-70)
    double scale = ch.telescope().loD(referenceLambda);
    int N = (int)std::ceil(FOVflD * samplesPerflD);
    return geom_symmetric(N, N, scale / samplesPerflD, scale / samplesPerflD);
  }
  int geometry(Chain& ch, const Geom&, Geom& out) override { out = out_geom(ch); return CSX_OK; }
  bool keeps_support() const override { return true; }
  int compile(Chain& ch, Compile& c, int) override {
    if (samplesPerflD <= 0 || FOVflD <= 0) return fail(CSX_ERR_ARG, "fraunhoferFocal: samplesPerflD/FOVflD unset");
    const Efield& ef = ch.efield();
    std::vector<Geom> go(ch.nlam(), out_geom(ch));
    std::vector<double> lf(ch.nlam());
    for (int l = 0; l < ch.nlam(); l++) lf[l] = ef.lambdas[l] * 2 * ef.beamRadius * ch.telescope().F;   // :38
    return compile_mft(ch, c, go, lf, "fraunhoferFocal " + name);
  }
};

// ---- zoomFftPropagator (src/elements/zoomFftPropagator.cpp:23-91)
struct ZoomFftPropElem : Element {
  int dir = 1;
  double zoomFactor = 1.0, padFactor = 1.0;
  int set(Chain&, const std::string& k, const std::string& a) override {
    if (set_common(k, a) || k == "zoomFftPropagator") return CSX_OK;
    if (k == "dir") dir = std::atoi(a.c_str());
    else if (k == "zoomFactor") zoomFactor = std::atof(a.c_str());
    else if (k == "padFactor") padFactor = std::atof(a.c_str());
    else return unknown_key(type, k);
    return CSX_OK;
  }
  int out_geom(Chain& ch, const Geom& in, Geom& out, double& lf) const {
    const Efield& ef = ch.efield();
    if (dir == 1) {
      double dxp = 2 * (ef.beamRadius / (ef.pixelScale * in.ny * zoomFactor)) * ef.lambdas[0] * 80;
      out = geom_offset((int)(padFactor * in.nx), (int)(padFactor * in.ny), dxp, dxp);
      lf = ef.lambdas[0] * 3.52;
    } else if (dir == -1) {
      out = geom_offset((int)(in.nx / padFactor), (int)(in.ny / padFactor), ef.pixelScale, ef.pixelScale);
      lf = -ef.lambdas[0] * 3.52;
    } else {
      return fail(CSX_ERR_ARG, "zoomFftPropagator: dir must be 1 or -1");
    }
    return CSX_OK;
  }
  int geometry(Chain& ch, const Geom& in, Geom& out) override {
    double lf;
    return out_geom(ch, in, out, lf);
  }
  bool keeps_support() const override { return true; }
  int compile(Chain& ch, Compile& c, int) override {
    if (skipped(c)) return CSX_OK;
    Geom out;
    double lf;
    CSX_TRY(out_geom(ch, c.geom, out, lf));
    if (out.nx != out.ny) return fail(CSX_ERR_UNSUPPORTED, "zoomFft output must be square");
    std::vector<Geom> go(ch.nlam(), out);
    std::vector<double> lfs(ch.nlam(), lf);   // the reference uses lambda[0] for every slice
    return compile_mft(ch, c, go, lfs, "zoomFftPropagator " + name);
  }
};

static bool is_mft_element(Element* e, const Compile& c);

// The S x S window of the FPM grid outside which every candidate's hex mask is 1.
struct BabinetWindow {
  bool ok = false;
  int S = 0, row0 = 0, col0 = 0, ncm = 1;
  const double2* masks = nullptr;   // per (candidate, lambda), S x S column-major
};

// ---- complexHexMaskFPM sub-element (src/elements/complexHexMaskFPM.cpp)
struct HexFpm {
  int maskNRings = -1, fpmArraySize = -1, nSubPix = -1, fpmMaskSize = -1, useOnlyThisHex = -1;
  int layout = 0;   // additive key hexLayout: csim (0) | standard (1)
  double hexStep = -1, fpmScaleFactor = -1, hexGap = -1, lambdaRef = -1, phaseSign = 1.0, fpmFRatio = -1;
  double fpmRadius = -1, fpmMaskPhysicalSize = -1, pixelScale = -1, zoomFactor = 1.0;
  std::vector<double> sags;
  HexRaster raster;
  bool rasterized = false;
  DevBuf d_nsub, d_hexnum, d_sags, d_masks;
  int set(const std::string& k, const std::string& a) {
    if (k == "complexHexMaskFPM" || k == "name" || k == "position" || k == "draw" || k == "save") return CSX_OK;
    if (k == "maskNRings") maskNRings = std::atoi(a.c_str());
    else if (k == "fpmArraySize") fpmArraySize = std::atoi(a.c_str());
    else if (k == "nSubPix") nSubPix = std::atoi(a.c_str());
    else if (k == "hexStep") hexStep = std::atof(a.c_str());
    else if (k == "fpmScaleFactor") fpmScaleFactor = std::atof(a.c_str());
    else if (k == "hexGap") hexGap = std::atof(a.c_str());
    else if (k == "lambdaRef") lambdaRef = std::atof(a.c_str());
    else if (k == "phaseSign") phaseSign = std::atof(a.c_str());
    else if (k == "fpmFRatio") fpmFRatio = std::atof(a.c_str());
    else if (k == "fpmRadius") fpmRadius = std::atof(a.c_str());
    else if (k == "fpmMaskSize") fpmMaskSize = std::atoi(a.c_str());
    else if (k == "fpmMaskPhysicalSize") fpmMaskPhysicalSize = std::atof(a.c_str());
    else if (k == "pixelScale") pixelScale = std::atof(a.c_str());
    else if (k == "hexLayout") layout = trim(a) == "standard" ? 1 : 0;
    else if (k == "sagFilename") {
      int r, c, s;
      std::vector<double> m;
      CSX_TRY(fits_matrix(trim(a), m, r, c, s));
      sags = m;
    } else if (k == "maskFilenameRoot") { /* the rasterisation is recomputed, not cached on disk */ }
    else if (k == "setSag") {
      int i = 0; float v = 0;
      std::sscanf(a.c_str(), "%d, %f", &i, &v);
      if (i < 0) return fail(CSX_ERR_ARG, "setSag: negative index");
      if ((size_t)i >= sags.size()) sags.resize(i + 1, 0.0);
      sags[i] = v;
    } else if (k == "useOnlyThisHex") useOnlyThisHex = std::atoi(a.c_str());
    else return unknown_key("complexHexMaskFPM", k);
    return CSX_OK;
  }
  // init_mask (:193-234): pixel scale, then the rasterisation.
  int init(const Efield& ef) {
    if (fpmMaskSize <= 0) return fail(CSX_ERR_ARG, "complexHexMaskFPM: fpmMaskSize unset");
    if (pixelScale == -1) {
      if (fpmMaskPhysicalSize > -1) pixelScale = fpmMaskPhysicalSize / fpmMaskSize;
      else if (zoomFactor > -1 && lambdaRef > -1 && fpmFRatio > -1)
        pixelScale = 2 * (ef.beamRadius / (ef.pixelScale * fpmMaskSize * zoomFactor)) * lambdaRef * fpmFRatio;
    }
    if (pixelScale <= 0) return fail(CSX_ERR_ARG, "complexHexMaskFPM: pixel scale undetermined");
    CSX_TRY(hex_rasterize(layout, maskNRings, hexStep > 0 ? hexStep : 1.0, fpmArraySize, fpmScaleFactor, hexGap,
                          fpmMaskSize, nSubPix, pixelScale, fpmRadius, raster));
    if (sags.size() < (size_t)raster.nhex) sags.resize(raster.nhex, 0.0);
    CSX_TRY(d_sags.reserve(sizeof(double) * sags.size()));
    CSX_HIP(hipMemcpy(d_sags.p, sags.data(), sizeof(double) * sags.size(), hipMemcpyHostToDevice));
    CSX_TRY(d_nsub.reserve(raster.nsub.size() * 4));
    CSX_TRY(d_hexnum.reserve(raster.hexnum.size() * 4));
    CSX_HIP(hipMemcpy(d_nsub.p, raster.nsub.data(), raster.nsub.size() * 4, hipMemcpyHostToDevice));
    CSX_HIP(hipMemcpy(d_hexnum.p, raster.hexnum.data(), raster.hexnum.size() * 4, hipMemcpyHostToDevice));
    rasterized = true;
    return CSX_OK;
  }
};

// ---- doubleFraunhoferFPM / doubleFraunhoferInterpFPM
// (src/elements/doubleFraunhoferFPM.cpp:23-110, doubleFraunhoferInterpFPM.cpp:24-125):
// MFT to the FPM plane, centred mask multiply (fused into the MFT epilogue),
// MFT back to the pupil grid.  Additive keys fpmSamplesPerflD / fpmFOVflD /
// referenceLambda place the FPM on a symmetric grid in lambda_ref/D units with
// lambda*f = lambda 2R F# (as fraunhoferFocal), and occulterRadiusFld builds an
// opaque disk mask analytically.
struct DoubleFraunhoferElem : Element {
  bool interp = false;   // doubleFraunhoferInterpFPM (hex mask)
  double zoomFactor = 1.0, padFactor = 1.0;
  int maskIndex = -1;
  // explicit geometry (additive)
  double fpmSamplesPerflD = 0, fpmFOVflD = 0, referenceLambda = 0, occulterRadiusFld = -1;
  // complex mask cube (maskFilename amp,ph or set_complex_array "mask")
  std::vector<double2> cube;
  int mrows = 0, mcols = 0, mslices = 0;
  DevBuf d_mask;
  HexFpm hex;
  bool has_hex = false;

  int set(Chain&, const std::string& k, const std::string& a) override {
    if (set_common(k, a) || k == "doubleFraunhoferFPM" || k == "doubleFraunhoferInterpFPM") return CSX_OK;
    if (k == "zoomFactor") zoomFactor = std::atof(a.c_str());
    else if (k == "padFactor") padFactor = std::atof(a.c_str());
    else if (k == "maskIndex") maskIndex = std::atoi(a.c_str());
    else if (k == "fpmSamplesPerflD") fpmSamplesPerflD = std::atof(a.c_str());
    else if (k == "fpmFOVflD") fpmFOVflD = std::atof(a.c_str());
    else if (k == "referenceLambda") referenceLambda = std::atof(a.c_str());
    else if (k == "occulterRadiusFld") occulterRadiusFld = std::atof(a.c_str());
    else if (k == "maskFilename") {
      size_t cpos = a.find(',');
      if (cpos == std::string::npos) return fail(CSX_ERR_PARSE, "maskFilename needs amp,ph");
      std::vector<double> am, ph;
      int r, c, s, r2, c2, s2;
      CSX_TRY(fits_matrix(trim(a.substr(0, cpos)), am, r, c, s));
      CSX_TRY(fits_matrix(trim(a.substr(cpos + 1)), ph, r2, c2, s2));
      if (r != r2 || c != c2 || s != s2) return fail(CSX_ERR_ARG, "maskFilename amp/ph sizes differ");
      mrows = r; mcols = c; mslices = s;
      cube.resize(am.size());
      for (size_t i = 0; i < am.size(); i++)   // initMask: amp % exp(-i ph)
        cube[i] = make_double2(am[i] * std::cos(-ph[i]), am[i] * std::sin(-ph[i]));
    } else return unknown_key(type, k);
    return CSX_OK;
  }
  int set_complex_array(Chain& ch, const std::string& k, const double* d, int r, int c, int slice) override {
    if (k != "mask") return Element::set_complex_array(ch, k, d, r, c, slice);
    if (slice < 0) return fail(CSX_ERR_ARG, "negative slice");
    if (mrows != r || mcols != c) { mrows = r; mcols = c; mslices = 0; cube.clear(); }
    if (slice >= mslices) { mslices = slice + 1; cube.resize((size_t)mrows * mcols * mslices); }
    std::memcpy(&cube[(size_t)slice * mrows * mcols], d, sizeof(double2) * mrows * mcols);
    ch.invalidate();
    return CSX_OK;
  }
  int set_opt(Chain& ch, const std::string& n, const std::vector<double>& v) override {
    if (has_hex && n == "sags") {
      // new sags are uploaded in place: the compiled stages read them from the device
      hex.sags = v;
      if (!disableForCalibration) ch.invalidate_calibration();
      if (hex.rasterized && (int)v.size() == hex.raster.nhex && hex.d_sags.bytes >= sizeof(double) * v.size()) {
        CSX_HIP(hipMemcpy(hex.d_sags.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
        return CSX_OK;
      }
      ch.invalidate();
      return CSX_OK;
    }
    if (has_hex && n == "useOnlyThisHex" && !v.empty()) { hex.useOnlyThisHex = (int)v[0]; ch.invalidate(); return CSX_OK; }
    return Element::set_opt(ch, n, v);
  }
  int get_opt(Chain& ch, const std::string& n, std::vector<double>& v) override {
    if (has_hex && n == "sags") { v = hex.sags; return CSX_OK; }
    if (has_hex && n == "useOnlyThisHex") { v.assign(1, hex.useOnlyThisHex); return CSX_OK; }
    return Element::get_opt(ch, n, v);
  }
  bool explicit_geom() const { return fpmSamplesPerflD > 0 && fpmFOVflD > 0; }
  // FPM-plane geometry and lambda*f per wavelength
  int fpm_geom(Chain& ch, const Geom& E, int l, Geom& g, double& lf) const {
    const Efield& ef = ch.efield();
    double lam = ef.lambdas[l];
    if (explicit_geom()) {
      double refl = referenceLambda > 0 ? referenceLambda : ef.lambdas[0];
      double pix = ch.telescope().loD(refl) / fpmSamplesPerflD;
      int N = (int)std::ceil(fpmFOVflD * fpmSamplesPerflD);
      g = geom_symmetric(N, N, pix, pix);
      lf = lam * 2 * ef.beamRadius * ch.telescope().F;
      return CSX_OK;
    }
    int nr = (int)(padFactor * E.ny), nc = (int)(padFactor * E.nx);
    double dxp;
    if (interp) {
      dxp = hex.pixelScale;   // init_geom: zoomGeom pixel = hex fpmScale
    } else {
      dxp = 2 * (ef.beamRadius / (ef.pixelScale * E.ny * zoomFactor)) * lam * 80;   // :51
    }
    g = geom_offset(nc, nr, dxp, dxp);
    lf = lam * 3.52;   // :45
    return CSX_OK;
  }
  int init(Chain& ch) override {
    if (has_hex) { hex.zoomFactor = zoomFactor; CSX_TRY(hex.init(ch.efield())); }
    return CSX_OK;
  }
  bool keeps_support() const override { return true; }
  int compile(Chain& ch, Compile& c, int index) override {
    if (skipped(c)) return CSX_OK;
    // the element holding the candidate sags expands the shared fields to every candidate
    const double* cs_dev = nullptr;
    int cs_nhex = 0, cs_ncand = 0;
    const bool expand = c.live > 0 && ch.candidate_sags(name, &cs_dev, &cs_nhex, &cs_ncand) == CSX_OK;
    const int nl = ch.nlam(), nf = ch.nfield_out(c, expand), ncand = ch.ncand_for(c);
    const Geom E = c.geom;
    std::vector<Geom> gf(nl);
    std::vector<double> lf(nl), lfo(nl);
    for (int l = 0; l < nl; l++) { CSX_TRY(fpm_geom(ch, E, l, gf[l], lf[l])); lfo[l] = -lf[l]; }
    // FPM masks: per field pointer into d_mask
    std::vector<MftField> mk(nf);
    Chain* chp = &ch;
    std::function<int(hipStream_t)> mask_stage;
    BabinetWindow bw;
    if (has_hex) {
      const int S = hex.raster.S;
      size_t plane = (size_t)S * S;
      CSX_TRY(hex.d_masks.reserve(sizeof(double2) * plane * ncand * nl));
      const double* sags_dev = nullptr;
      int nhex = hex.raster.nhex, ncd = 1;
      if (!c.calibrating && chp->candidate_sags(name, &sags_dev, &nhex, &ncd) == CSX_OK && sags_dev) {
        // candidate batch: sags already on the device
      } else {
        if (hex.d_sags.bytes < sizeof(double) * hex.sags.size()) {
          CSX_TRY(hex.d_sags.reserve(sizeof(double) * hex.sags.size()));
        }
        CSX_HIP(hipMemcpy(hex.d_sags.p, hex.sags.data(), sizeof(double) * hex.sags.size(), hipMemcpyHostToDevice));
        sags_dev = hex.d_sags.as<double>();
        nhex = (int)hex.sags.size();
        ncd = 1;
      }
      // lambdas of the wavelength axis (device)
      DevBuf& lamb = ch.scratch("hex_lam_" + name);
      CSX_TRY(lamb.reserve(sizeof(double) * nl));
      CSX_HIP(hipMemcpy(lamb.p, ch.efield().lambdas.data(), sizeof(double) * nl, hipMemcpyHostToDevice));
      const int32_t* dn = hex.d_nsub.as<int32_t>();
      const int32_t* dh = hex.d_hexnum.as<int32_t>();
      double2* dm = hex.d_masks.as<double2>();
      const double* dl = lamb.as<double>();
      int nsp = hex.raster.nsubpix, use = hex.useOnlyThisHex;
      double ps = hex.phaseSign;
      int ncm = std::min(ncd, ncand);
      mask_stage = [=](hipStream_t s) {
        return hex_masks(dn, dh, S, nsp, sags_dev, nhex, ncm, dl, nl, ps, use, dm, s);
      };
      for (int f = 0; f < nf; f++) {
        int l = f % nl, cand = f / (nf / ncand);
        if (cand >= ncm) cand = ncm - 1;
        MftField& m = mk[f];
        // response batch: a single-hex candidate passes nothing outside the mask
        // rectangle (that light belongs to the last, outside-the-hexes column)
        m.mask_kind = (use == kUseOnlyPerCandidate && f / (nf / ncand) < hex.raster.nhex) ? MASK_COMPLEX_CLIP
                                                                                          : MASK_COMPLEX;
        m.mask = dm + ((size_t)cand * nl + l) * plane;
        // applied as computed: doubleFraunhoferInterpFPM does not mirror the hex mask
        m.mrows = S; m.mcols = S;
        m.row0 = gf[l].ny / 2 - S / 2; m.col0 = gf[l].nx / 2 - S / 2;
      }
      // every candidate's mask is 1 outside the same window of every FPM grid
      bw.ok = use != kUseOnlyPerCandidate && S % 8 == 0;
      for (int l = 0; l < nl && bw.ok; l++)
        bw.ok = gf[l].ny == gf[0].ny && gf[l].nx == gf[0].nx && gf[0].ny >= S && gf[0].nx >= S;
      bw.S = S; bw.row0 = gf[0].ny / 2 - S / 2; bw.col0 = gf[0].nx / 2 - S / 2; bw.ncm = ncm; bw.masks = dm;
    } else if (!cube.empty() || occulterRadiusFld >= 0) {
      // one mask for every slice (maskIndex latches to the first slice, :86-87)
      std::vector<double2> m;
      int r = mrows, cc = mcols;
      if (occulterRadiusFld >= 0) {
        // analytic occulting disk on the FPM grid of lambda index 0
        const Geom& g = gf[0];
        double loD = ch.telescope().loD(referenceLambda > 0 ? referenceLambda : ch.efield().lambdas[0]);
        r = g.ny; cc = g.nx;
        m.assign((size_t)r * cc, make_double2(1, 0));
        for (int j = 0; j < cc; j++)
          for (int i = 0; i < r; i++) {
            double x = g.x0 + j * g.dx, y = g.y0 + i * g.dy;
            if (std::sqrt(x * x + y * y) / loD < occulterRadiusFld) m[(size_t)j * r + i] = make_double2(0, 0);
          }
      } else {
        int mi = maskIndex < 0 ? 0 : maskIndex;
        if (mi >= mslices) return fail(CSX_ERR_ARG, "doubleFraunhoferFPM: maskIndex out of range");
        m.assign(cube.begin() + (size_t)mi * r * cc, cube.begin() + (size_t)(mi + 1) * r * cc);
      }
      CSX_TRY(d_mask.reserve(sizeof(double2) * m.size()));
      CSX_HIP(hipMemcpy(d_mask.p, m.data(), sizeof(double2) * m.size(), hipMemcpyHostToDevice));
      for (int f = 0; f < nf; f++) {
        int l = f % nl;
        MftField& mm = mk[f];
        mm.mask_kind = MASK_COMPLEX;
        mm.mask = d_mask.p;
        mm.mrows = r; mm.mcols = cc;
        mm.row0 = gf[l].ny / 2 - r / 2; mm.col0 = gf[l].nx / 2 - cc / 2;
      }
    }
    if (mask_stage) { c.stages.push_back(mask_stage); c.stage_names.push_back("hex masks " + name); }
    // A candidate batch whose hex masks are transparent outside one window splits by
    // Babinet: the unmasked FPM field and its return MFT once per source and
    // wavelength, per candidate only the window's (M - 1) E (CSX_BABINET=0: off)
    const char* bz = std::getenv("CSX_BABINET");
    const bool babinet = expand && has_hex && bw.ok && ncand > 1 && !(bz && bz[0] == '0');
    // pupil -> FPM (mask in the epilogue) -> pupil grid of the input
    CSX_TRY(compile_mft(ch, c, gf, lf, type + " in " + name, babinet ? nullptr : &mk, /*out_conv=*/true, nullptr,
                        babinet ? false : expand));
    // a mask element right after folds into the epilogue of the return MFT
    const double* lyot = nullptr;
    std::shared_ptr<const Support> lyot_sup;
    const float* lyot32 = nullptr;
    const unsigned char* lyot8 = nullptr;
    int claim = 0;
    if (Element* nx = ch.elem_at(index + 1)) {
      if (auto* me = dynamic_cast<MaskElem*>(nx)) {
        if (!me->skipped(c) && !me->map.host.empty()) {
          CSX_TRY(me->map.device_for(ch, E, &lyot));
          lyot_sup = me->map.support_for(E);
          lyot32 = me->map.f32_for(E);
          lyot8 = me->map.u8_for(E);
          claim = 1;
        }
      }
    }
    bool next_mft = is_mft_element(ch.elem_at(index + 1 + claim), c);
    std::vector<Geom> back(nl, E);
    if (babinet) {
      CSX_TRY(compile_babinet(ch, c, E, gf, lfo, bw, next_mft, lyot, lyot32, lyot8, lyot_sup));
    } else {
      // the fields are on the FPM grid of their own wavelength (gf[l])
      CSX_TRY(compile_mft(ch, c, back, lfo, type + " out " + name, nullptr, next_mft, lyot, false, lyot_sup, lyot32,
                          lyot8, &gf));
    }
    c.claimed = claim;
    return CSX_OK;
  }

  // FPM -> Lyot of a candidate batch by Babinet.  The fields are the one shared copy's
  // FPM-plane fields (unmasked).  E_lyot[cand] = MFT(E) + MFT_window((M_cand - 1) E):
  // the first term once per (source, wavelength) into a scratch buffer stored like the
  // output, which the per-candidate window MFT adds in its epilogue before the Lyot mask.
  int compile_babinet(Chain& ch, Compile& c, const Geom& E, const std::vector<Geom>& gf,
                      const std::vector<double>& lfo, const BabinetWindow& bw, bool next_mft, const double* lyot,
                      const float* lyot32, const unsigned char* lyot8,
                      const std::shared_ptr<const Support>& lyot_sup) {
    const int T = terms_of(ch.precision());
    const bool c64 = ch.precision() == CSX_FP32, mixed = T || c64;
    const bool out_conv = next_mft && mixed;
    const int nl = ch.nlam(), ns = ch.nsrc();
    const int nf_in = ch.nfield_for(c), nf = ch.nfield_out(c, true);
    if (nf_in != ns * nl || c.live_src) return fail(CSX_ERR_STATE, "babinet: unexpected field layout");
    const size_t ew = T ? (size_t)4 * T : (c64 ? sizeof(float2) : sizeof(double2));   // working precision
    const int wkind = T ? FK_PLANES : (c64 ? FK_C64 : FK_C128);
    const size_t eo = out_conv ? ew : sizeof(double2);
    const int okind = out_conv ? wkind : FK_C128;
    const Geom& gF = c.geom;
    const char* Ein = mixed ? static_cast<const char*>(ch.c64_buf()) : reinterpret_cast<const char*>(ch.field_buf(c.buf));
    if (mixed != c.in_conv) return fail(CSX_ERR_STATE, "babinet: FPM fields not at the working precision");
    char* out = out_conv ? static_cast<char*>(ch.c64_buf()) : reinterpret_cast<char*>(ch.field_buf(1 - c.buf));
    DevBuf& base = ch.scratch("babinet base " + name + (c.calibrating ? " cal" : ""));
    DevBuf& win = ch.scratch("babinet window " + name + (c.calibrating ? " cal" : ""));
    const long wn = (long)bw.S * bw.S;
    CSX_TRY(base.reserve(eo * (size_t)nf_in * E.npix()));
    CSX_TRY(win.reserve(ew * (size_t)nf * wn));
    Chain* chp = &ch;
    // the unmasked return MFT, once per (source, lambda)
    std::vector<MftField> FB(nf_in);
    for (int f = 0; f < nf_in; f++) {
      const int l = f % nl;
      MftField& m = FB[f];
      m.in = Ein + (size_t)f * gF.npix() * ew; m.in_kind = wkind;
      m.out = static_cast<char*>(base.p) + (size_t)f * E.npix() * eo; m.out_kind = okind;
      m.in_x = gx(gf[l]); m.in_y = gy(gf[l]); m.out_x = gx(E); m.out_y = gy(E);
      m.lambda_f = lfo[l];
    }
    auto pb = std::make_unique<MftPlan>();
    CSX_TRY(pb->build(FB, ch.precision(), ch.stream()));
    const MftPlan* ppb = pb.get();
    c.plans.push_back(std::move(pb));
    c.mft_at[c.stages.size()] = FB;
    c.stages.push_back([=](hipStream_t s) { return ppb->run(s, chp->timer()); });
    c.stage_names.push_back(type + " out (shared) " + name);
    // per candidate: (M - 1) E on the window
    const long npF = gF.npix();
    const int rowsF = gF.ny;
    void* W = win.p;
    const BabinetWindow b = bw;
    c.stages.push_back([=](hipStream_t s) {
      return babinet_window(Ein, wkind, T, npF, rowsF, b.masks, b.S, b.row0, b.col0, nf, nl, ns, b.ncm, W, s);
    });
    c.stage_names.push_back("babinet window " + name);
    // window -> Lyot, plus the shared term, Lyot mask
    std::vector<MftField> FD(nf);
    for (int f = 0; f < nf; f++) {
      const int l = f % nl, fi = f % nf_in;
      MftField& m = FD[f];
      m.in = static_cast<const char*>(win.p) + (size_t)f * wn * ew; m.in_kind = wkind;
      m.out = out + (size_t)f * E.npix() * eo; m.out_kind = okind;
      m.in_x = csx_grid{bw.S, gf[l].x0 + bw.col0 * gf[l].dx, gf[l].dx};
      m.in_y = csx_grid{bw.S, gf[l].y0 + bw.row0 * gf[l].dy, gf[l].dy};
      m.out_x = gx(E); m.out_y = gy(E);
      m.lambda_f = lfo[l];
      m.add = static_cast<const char*>(base.p) + (size_t)fi * E.npix() * eo;
      if (lyot) {
        m.mask_kind = lyot8 ? MASK_REAL_1PI_U8 : (lyot32 ? MASK_REAL_1PI_F32 : MASK_REAL_1PI);
        m.mask = lyot8 ? (const void*)lyot8 : (lyot32 ? (const void*)lyot32 : (const void*)lyot);
        m.mrows = E.ny; m.mcols = E.nx; m.row0 = 0; m.col0 = 0;
        const char* zs = std::getenv("CSX_ZERO_SKIP");
        if (!(zs && zs[0] == '0')) {
          m.out_sup = lyot_sup;
          m.out_read_sup = out_conv;   // the next MFT reads it through lyot_sup
        }
      }
    }
    auto pd = std::make_unique<MftPlan>();
    CSX_TRY(pd->build(FD, ch.precision(), ch.stream()));
    const MftPlan* ppd = pd.get();
    c.plans.push_back(std::move(pd));
    c.mft_at[c.stages.size()] = FD;
    c.stages.push_back([=](hipStream_t s) { return ppd->run(s, chp->timer()); });
    c.stage_names.push_back(type + " out (window) " + name);
    c.live = 0;
    c.in_conv = out_conv;
    if (!out_conv) c.buf = 1 - c.buf;
    c.geom = E;
    c.support = lyot ? lyot_sup : nullptr;
    return CSX_OK;
  }
};

// ---- fpmPupToLyot (src/elements/fpmPupToLyot.cpp:24-167): the FPM applied as a
// convolution in the pupil on a 2n x 2n zero-padded grid,
//   E' = s * ifft2( fft2(fft_shift(z * zoomFft(K))) .* fft2(pad E) ) (+ pad E)  -> centre,
// K = mask - 1 with the Babinet masks (fpmCMCForPupToLyot, fpmIntHexCMCForPupToLyot,
// the hex mask mirrored left-right) or the opaque mask itself (fpmBinaryForPupToLyot),
// s = dx^2 / (-i fRatioSign lambda f).  Mask spectra are recomputed on every
// propagation (they follow the hex sags and the candidate batch).
struct FpmPupToLyotElem : Element {
  enum MaskType { MT_NONE = 0, MT_CMC = 1, MT_HEX = 2, MT_BINARY = 3 };
  bool calibration_aware() const override { return true; }   // the binary mask has a calibration mask
  int mtype = MT_NONE;
  double focalRatio = -1, fRatioSign = 1, zoomFftSign = 1;
  int babinet_set = -1;   // setBabinet (fpmPupToLyot.cpp:452-453); -1: the mask's own default
  // fpmCMCForPupToLyot
  std::vector<double2> cube;
  int mrows = 0, mcols = 0, mslices = 0, maskIndex = -1;
  double fpmMatScale = 0;
  // fpmIntHexCMCForPupToLyot
  HexFpm hex;
  double zoomFactor = -1;
  // fpmBinaryForPupToLyot
  std::vector<double> fpmMat, calibMat;
  int brows = 0, bcols = 0, crows = 0, ccols = 0;
  double referenceLambda = 0, innerRadiusFld = 0, outerRadiusFld = 0;
  // device buffers, [0] full run, [1] calibration run
  DevBuf d_src[2], d_K[2], d_H[2], d_P[2], d_scale[2], d_lam, d_Pu;

  int set(Chain&, const std::string& k, const std::string& a) override {
    if (set_common(k, a) || k == "fpmPupToLyot") return CSX_OK;
    if (k == "focalRatio") focalRatio = std::atof(a.c_str());
    else if (k == "fRatioSign") fRatioSign = std::atof(a.c_str());
    else if (k == "zoomFftSign") zoomFftSign = std::atof(a.c_str());
    else return unknown_key(type, k);
    return CSX_OK;
  }
  static int load_amp_ph(const std::string& a, bool reim, std::vector<double2>& out, int& r, int& c, int& s) {
    size_t cpos = a.find(',');
    if (cpos == std::string::npos) return fail(CSX_ERR_PARSE, "mask file names need the form first,second");
    std::vector<double> x, y;
    int r2, c2, s2;
    CSX_TRY(fits_matrix(trim(a.substr(0, cpos)), x, r, c, s));
    CSX_TRY(fits_matrix(trim(a.substr(cpos + 1)), y, r2, c2, s2));
    if (r != r2 || c != c2 || s != s2) return fail(CSX_ERR_ARG, "mask file sizes differ");
    out.resize(x.size());
    for (size_t i = 0; i < x.size(); i++)
      out[i] = reim ? make_double2(x[i], y[i])                                        // initMaskReIm
                    : make_double2(x[i] * std::cos(-y[i]), x[i] * std::sin(-y[i]));   // initMask: amp exp(-i ph)
    return CSX_OK;
  }
  // keys of the mask sub-block (fpmPupToLyot.cpp:250-330, 377-392, 463-496)
  int set_mask(const std::string& k, const std::string& a) {
    if (k == "fpmCMCForPupToLyot" || k == "fpmIntHexCMCForPupToLyot" || k == "fpmBinaryForPupToLyot") return CSX_OK;
    if (mtype == MT_CMC) {
      if (k == "maskIndex") maskIndex = std::atoi(a.c_str());
      else if (k == "fpmMatScale") fpmMatScale = std::atof(a.c_str());
      else if (k == "maskFilename") return load_amp_ph(a, false, cube, mrows, mcols, mslices);
      else if (k == "maskFilenameReIm") return load_amp_ph(a, true, cube, mrows, mcols, mslices);
      else return unknown_key("fpmCMCForPupToLyot", k);
    } else if (mtype == MT_HEX) {
      if (k == "zoomFactor") zoomFactor = std::atof(a.c_str());
      else return unknown_key("fpmIntHexCMCForPupToLyot", k);
    } else if (mtype == MT_BINARY) {
      int s = 0;
      if (k == "referenceLambda") referenceLambda = std::atof(a.c_str());
      else if (k == "innerRadiusFld") innerRadiusFld = std::atof(a.c_str());
      else if (k == "outerRadiusFld") outerRadiusFld = std::atof(a.c_str());
      else if (k == "maskFilename") return fits_matrix(trim(a), fpmMat, brows, bcols, s);
      else if (k == "calibFilename") return fits_matrix(trim(a), calibMat, crows, ccols, s);
      else return unknown_key("fpmBinaryForPupToLyot", k);
    }
    return CSX_OK;
  }
  int set_array(Chain& ch, const std::string& k, const double* d, int r, int c) override {
    if (mtype == MT_BINARY && (k == "maskFilename" || k == "calibFilename")) {
      if (!d || r <= 0 || c <= 0) return fail(CSX_ERR_ARG, "empty array");
      auto& dst = k == "maskFilename" ? fpmMat : calibMat;
      dst.assign(d, d + (size_t)r * c);
      if (k == "maskFilename") { brows = r; bcols = c; } else { crows = r; ccols = c; }
      ch.invalidate();
      return CSX_OK;
    }
    return Element::set_array(ch, k, d, r, c);
  }
  int set_complex_array(Chain& ch, const std::string& k, const double* d, int r, int c, int slice) override {
    if (mtype != MT_CMC || k != "mask") return Element::set_complex_array(ch, k, d, r, c, slice);
    if (slice < 0) return fail(CSX_ERR_ARG, "negative slice");
    if (mrows != r || mcols != c) { mrows = r; mcols = c; mslices = 0; cube.clear(); }
    if (slice >= mslices) { mslices = slice + 1; cube.resize((size_t)mrows * mcols * mslices); }
    std::memcpy(&cube[(size_t)slice * mrows * mcols], d, sizeof(double2) * mrows * mcols);
    ch.invalidate();
    return CSX_OK;
  }
  int set_opt(Chain& ch, const std::string& n, const std::vector<double>& v) override {
    if (n == "reinit") return CSX_OK;   // spectra are rebuilt on every propagation
    if (mtype == MT_HEX && n == "setBabinet" && !v.empty()) { babinet_set = v[0] != 0; ch.invalidate(); return CSX_OK; }
    if (mtype == MT_HEX && n == "sags") {
      hex.sags = v;
      ch.invalidate_calibration();
      if (hex.rasterized && (int)v.size() == hex.raster.nhex && hex.d_sags.bytes >= sizeof(double) * v.size()) {
        CSX_HIP(hipMemcpy(hex.d_sags.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
        return CSX_OK;
      }
      ch.invalidate();
      return CSX_OK;
    }
    if (mtype == MT_HEX && n == "useOnlyThisHex" && !v.empty()) { hex.useOnlyThisHex = (int)v[0]; ch.invalidate(); return CSX_OK; }
    return Element::set_opt(ch, n, v);
  }
  int get_opt(Chain& ch, const std::string& n, std::vector<double>& v) override {
    if (mtype == MT_HEX && n == "sags") { v = hex.sags; return CSX_OK; }
    if (mtype == MT_HEX && n == "useOnlyThisHex") { v.assign(1, hex.useOnlyThisHex); return CSX_OK; }
    return Element::get_opt(ch, n, v);
  }
  int init(Chain& ch) override {
    if (mtype == MT_HEX) {
      if (zoomFactor <= 0) return fail(CSX_ERR_ARG, "fpmIntHexCMCForPupToLyot: zoomFactor must be set before the hex mask");
      hex.zoomFactor = zoomFactor;
      CSX_TRY(hex.init(ch.efield()));
    }
    return CSX_OK;
  }
  // mask grid (maskGeom.set_xy) and lambda * f of slice l (set_geometry)
  int mask_geom(Chain& ch, Geom& g, std::vector<double>& lf) const {
    const Efield& ef = ch.efield();
    const int nl = ch.nlam();
    lf.resize(nl);
    if (mtype == MT_CMC || mtype == MT_HEX) {
      if (focalRatio <= 0) return fail(CSX_ERR_ARG, "fpmPupToLyot: focalRatio unset");
      if (mtype == MT_CMC) {
        if (cube.empty() || fpmMatScale <= 0) return fail(CSX_ERR_ARG, "fpmCMCForPupToLyot: mask or fpmMatScale unset");
        g = geom_symmetric(mcols, mrows, fpmMatScale, fpmMatScale);
      } else {
        g = geom_symmetric(hex.raster.S, hex.raster.S, hex.pixelScale, hex.pixelScale);
      }
      double cFRatio = focalRatio * 2 * ef.beamRadius;
      for (int l = 0; l < nl; l++) lf[l] = fRatioSign * ef.lambdas[l] * cFRatio;
    } else if (mtype == MT_BINARY) {
      if (fpmMat.empty() || outerRadiusFld <= 0) return fail(CSX_ERR_ARG, "fpmBinaryForPupToLyot: mask or outerRadiusFld unset");
      double flD = ch.telescope().loD(referenceLambda);
      double pix = 2 * outerRadiusFld * flD / brows;
      g = geom_symmetric(bcols, brows, pix, pix);
      for (int l = 0; l < nl; l++) lf[l] = fRatioSign * ef.lambdas[l] * ch.telescope().fLength;
    } else {
      return fail(CSX_ERR_ARG, "fpmPupToLyot: no mask sub-block");
    }
    return CSX_OK;
  }
  int compile(Chain& ch, Compile& c, int) override {
    const bool binary = mtype == MT_BINARY;
    // doBabinet: on for the complex-cube and hex masks unless set off (set_babinet(true)
    // in their constructors); it both subtracts 1 from the mask and adds E back
    const bool babinet = babinet_set >= 0 ? babinet_set != 0 : !binary;
    if (skipped(c) && !binary && babinet) return CSX_OK;   // calibration K = exp(0) - 1 = 0: E unchanged
    const bool calib_ones = skipped(c) && !binary;         // Babinet off: calibration K = exp(0) = 1
    const bool use_calib = binary && skipped(c);
    if (use_calib && calibMat.empty()) return fail(CSX_ERR_ARG, "fpmBinaryForPupToLyot: calibFilename unset");
    if (use_calib && (crows != brows || ccols != bcols)) return fail(CSX_ERR_ARG, "fpmBinaryForPupToLyot: calib size");
    const Geom E = c.geom;
    if (E.nx != E.ny || (E.nx & 1)) return fail(CSX_ERR_UNSUPPORTED, "fpmPupToLyot: fields must be square and even");
    const int n = E.nx, n2 = 2 * n;
    if (n2 > 4096 || (n2 & (n2 - 1))) return fail(CSX_ERR_UNSUPPORTED, "fpmPupToLyot: 2n must be a power of two <= 4096");
    CSX_TRY(flush_pointwise(ch, c));
    if (c.in_conv) return fail(CSX_ERR_STATE, "fpmPupToLyot: fields at working precision");
    // holding the candidate sags: the shared fields (one candidate's) are transformed
    // once and every candidate's convolution starts from that spectrum; otherwise a
    // shared prefix ends here
    const double* cs_dev = nullptr;
    int cs_nhex = 0, cs_ncand = 0;
    const bool expand = c.live > 0 && ch.candidate_sags(name, &cs_dev, &cs_nhex, &cs_ncand) == CSX_OK;
    if (!expand) CSX_TRY(expand_candidates(ch, c));
    const int nf_in = ch.nfield_for(c);
    Geom mg;
    std::vector<double> lf;
    CSX_TRY(mask_geom(ch, mg, lf));
    const int nl = ch.nlam(), nf = ch.nfield_all(c), ncand = ch.ncand_for(c), nsrc = ch.nsrc();
    const int which = c.calibrating ? 1 : 0;
    const long S2 = (long)mg.nx * mg.ny, P2 = (long)n2 * n2;
    // source masks and the number of mask candidates
    int ncm = 1;
    const double2* src = nullptr;
    long src_stride = 0;
    int fl = 0, minus1 = 0;
    std::function<int(hipStream_t)> hex_stage;
    if (calib_ones) {
      std::vector<double2> m((size_t)S2, make_double2(1.0, 0.0));
      CSX_TRY(d_src[which].reserve(sizeof(double2) * S2));
      CSX_HIP(hipMemcpy(d_src[which].p, m.data(), sizeof(double2) * S2, hipMemcpyHostToDevice));
      src = d_src[which].as<double2>(); src_stride = 0;
    } else if (mtype == MT_HEX) {
      const int S = hex.raster.S;
      const double* sags_dev = nullptr;
      int nhex = hex.raster.nhex, ncd = 1;
      if (!(!c.calibrating && ch.candidate_sags(name, &sags_dev, &nhex, &ncd) == CSX_OK && sags_dev)) {
        if (hex.d_sags.bytes < sizeof(double) * hex.sags.size()) CSX_TRY(hex.d_sags.reserve(sizeof(double) * hex.sags.size()));
        CSX_HIP(hipMemcpy(hex.d_sags.p, hex.sags.data(), sizeof(double) * hex.sags.size(), hipMemcpyHostToDevice));
        sags_dev = hex.d_sags.as<double>();
        nhex = (int)hex.sags.size();
        ncd = 1;
      }
      ncm = std::min(ncd, ncand);
      CSX_TRY(d_src[which].reserve(sizeof(double2) * S2 * ncm * nl));
      CSX_TRY(d_lam.reserve(sizeof(double) * nl));
      CSX_HIP(hipMemcpy(d_lam.p, ch.efield().lambdas.data(), sizeof(double) * nl, hipMemcpyHostToDevice));
      const int32_t* dn = hex.d_nsub.as<int32_t>();
      const int32_t* dh = hex.d_hexnum.as<int32_t>();
      double2* dm = d_src[which].as<double2>();
      const double* dl = d_lam.as<double>();
      int nsp = hex.raster.nsubpix, use = hex.useOnlyThisHex, ncmc = ncm;
      double ps = hex.phaseSign;
      hex_stage = [=](hipStream_t s) { return hex_masks(dn, dh, S, nsp, sags_dev, nhex, ncmc, dl, nl, ps, use, dm, s); };
      src = dm; src_stride = S2; fl = 1; minus1 = 1;
    } else {
      std::vector<double2> m((size_t)S2);
      if (mtype == MT_CMC) {
        int mi = maskIndex < 0 ? 0 : maskIndex;   // maskIndex latches to the first slice executed
        if (mi >= mslices) return fail(CSX_ERR_ARG, "fpmCMCForPupToLyot: maskIndex out of range");
        std::copy(cube.begin() + (size_t)mi * S2, cube.begin() + (size_t)(mi + 1) * S2, m.begin());
        minus1 = 1;
      } else {
        const std::vector<double>& b = use_calib ? calibMat : fpmMat;
        for (long i = 0; i < S2; i++) m[i] = make_double2(b[i], 0.0);
      }
      CSX_TRY(d_src[which].reserve(sizeof(double2) * S2));
      CSX_HIP(hipMemcpy(d_src[which].p, m.data(), sizeof(double2) * S2, hipMemcpyHostToDevice));
      src = d_src[which].as<double2>(); src_stride = 0;   // the same mask for every slice
    }
    if (mtype != MT_BINARY) minus1 = babinet ? 1 : 0;
    // kernel inputs K (one per (mask candidate, lambda)), their transforms H
    const int nk = ncm * nl;
    CSX_TRY(d_K[which].reserve(sizeof(double2) * S2 * nk));
    CSX_TRY(d_H[which].reserve(sizeof(double2) * P2 * nk));
    CSX_TRY(d_P[which].reserve(sizeof(double2) * P2 * nf));
    // s = dx^2 / (-i fRatioSign lambda f) per slice (fftMHatTimesfftPaddedEScale)
    std::vector<double2> sc(nl);
    const std::complex<double> i1(0, 1);
    for (int l = 0; l < nl; l++) {
      std::complex<double> v = std::pow(E.dx, 2) / (-i1 * fRatioSign * lf[l]);
      sc[l] = make_double2(v.real(), v.imag());
    }
    CSX_TRY(d_scale[which].reserve(sizeof(double2) * nl));
    CSX_HIP(hipMemcpy(d_scale[which].p, sc.data(), sizeof(double2) * nl, hipMemcpyHostToDevice));
    // zoomFft of K from the mask grid to the padded pupil grid, FP64
    const Geom pg = geom_symmetric(n2, n2, E.dx, E.dy);
    std::vector<MftField> F(nk);
    double2* dK = d_K[which].as<double2>();
    double2* dH = d_H[which].as<double2>();
    for (int k = 0; k < nk; k++) {
      MftField& m = F[k];
      m.in = dK + (size_t)k * S2; m.out = dH + (size_t)k * P2;
      m.in_kind = FK_C128; m.out_kind = FK_C128;
      m.in_x = gx(mg); m.in_y = gy(mg); m.out_x = gx(pg); m.out_y = gy(pg);
      m.lambda_f = lf[k % nl];
    }
    auto plan = std::make_unique<MftPlan>();
    CSX_TRY(plan->build(F, CSX_FP64, ch.stream()));
    const MftPlan* pp = plan.get();
    c.plans.push_back(std::move(plan));
    // Babinet rows: every mask, or in the response batch (candidate c uses only hex c)
    // only the last candidate, whose mask is the area outside the hexes
    // (makeHexCFpmResponse.cpp:150-162: setBabinet for the last column only)
    const bool resp = mtype == MT_HEX && hex.useOnlyThisHex == kUseOnlyPerCandidate;
    const int big = std::numeric_limits<int>::max();
    const int rows = mg.ny, cols = mg.nx, fli = fl;
    const int mi1 = !minus1 ? big : (resp ? hex.raster.nhex * nl : 0);
    const double zs = zoomFftSign;
    Chain* chp = &ch;
    c.stages.push_back([=](hipStream_t s) {
      if (hex_stage) CSX_TRY(hex_stage(s));
      // src_stride 0: the same mask for every slice
      CSX_TRY(fpm_prep(src, src_stride, rows, cols, nk, fli, mi1, zs, dK, s));
      CSX_TRY(pp->run(s, chp->timer()));
      return fft2_batched(dH, n2, nk, -1, s);
    });
    c.stage_names.push_back("fpmPupToLyot mask spectra " + name);
    double2* Ef = ch.field_buf(c.buf);
    double2* P = d_P[which].as<double2>();
    const double2* dsc = d_scale[which].as<double2>();
    const int bab = !babinet ? big : (resp ? hex.raster.nhex * nsrc * nl : 0);
    const int per_cand = nsrc * nl, ncmc = ncm;
    if (expand) {
      // spectra of the shared fields in their own buffer; every candidate's output goes
      // to the other field buffer (its Babinet term reads the shared input)
      CSX_TRY(d_Pu.reserve(sizeof(double2) * P2 * nf_in));
      double2* Pu = d_Pu.as<double2>();
      double2* Eo = ch.field_buf(1 - c.buf);
      c.stages.push_back([=](hipStream_t s) {
        CSX_TRY(embed_center(Ef, n, n, Pu, n2, n2, nf_in, 0, s));
        CSX_TRY(fft2_batched(Pu, n2, nf_in, -1, s));
        CSX_TRY(spectrum_mul(P, Pu, nf_in, n2, nf, dH, per_cand, nl, ncmc, s));
        CSX_TRY(fft2_batched(P, n2, nf, +1, s));
        return crop_add(Eo, Ef, nf_in, n, P, n2, nf, nl, dsc, bab, s);
      });
      c.buf = 1 - c.buf;
      c.live = 0;
    } else {
      c.stages.push_back([=](hipStream_t s) {
        CSX_TRY(embed_center(Ef, n, n, P, n2, n2, nf, 0, s));
        CSX_TRY(fft2_batched(P, n2, nf, -1, s));
        CSX_TRY(spectrum_mul(P, P, nf, n2, nf, dH, per_cand, nl, ncmc, s));
        CSX_TRY(fft2_batched(P, n2, nf, +1, s));
        return crop_add(Ef, Ef, nf, n, P, n2, nf, nl, dsc, bab, s);
      });
    }
    c.stage_names.push_back("fpmPupToLyot convolution " + name);
    return CSX_OK;
  }
};

// An element whose first GPU work is an MFT over the current fields.
static bool is_mft_element(Element* e, const Compile& c) {
  if (!e || e->skipped(c)) return false;
  return dynamic_cast<FraunhoferFocalElem*>(e) || dynamic_cast<ZoomFftPropElem*>(e) ||
         dynamic_cast<DoubleFraunhoferElem*>(e);
}

// ---- element factory (src/coronagraph/coronagraph.cpp:67-155)
int Chain::make_element(const Block& b, std::unique_ptr<Element>& out) {
  const std::string t = b.cmds.empty() ? "" : b.cmds[0].key;
  std::unique_ptr<Element> e;
  if (t == "mask") e.reset(new MaskElem());
  else if (t == "mirror") e.reset(new MirrorElem());
  else if (t == "complexMask") e.reset(new ComplexMaskElem());
  else if (t == "fresnelPropagator") e.reset(new FresnelElem());
  else if (t == "downsample") e.reset(new DownsampleElem());
  else if (t == "fraunhoferFocal") e.reset(new FraunhoferFocalElem());
  else if (t == "zoomFftPropagator") e.reset(new ZoomFftPropElem());
  else if (t == "doubleFraunhoferFPM" || t == "doubleFraunhoferInterpFPM") {
    auto* d = new DoubleFraunhoferElem();
    d->interp = t == "doubleFraunhoferInterpFPM";
    e.reset(d);
  } else if (t == "fpmPupToLyot") {
    e.reset(new FpmPupToLyotElem());
  } else {
    if (ParseAudit* a = parse_audit()) {
      a->unknown_elements.push_back(t);
      return CSX_OK;                       // out stays empty: the element is skipped
    }
    std::fprintf(stderr, "!!! coronagraph init unknown object name: %s\n", t.c_str());
    return fail(CSX_ERR_UNSUPPORTED, "element type not supported: " + t);
  }
  e->type = t;
  e->name = t;
  if (auto* d = dynamic_cast<DoubleFraunhoferElem*>(e.get())) {
    // own keys and a complexHexMaskFPM sub-block (doubleFraunhoferInterpFPM::init)
    std::vector<Block> sub = find_blocks(b);
    for (const Block& sb : sub) {
      if (sb.cmds.empty()) continue;
      if (sb.cmds[0].key == "complexHexMaskFPM") {
        d->has_hex = true;
        for (const Cmd& c : sb.cmds) CSX_TRY(d->hex.set(c.key, c.arg));
      } else {
        for (const Cmd& c : sb.cmds) CSX_TRY(e->set(*this, c.key, c.arg));
      }
    }
  } else if (auto* p = dynamic_cast<FpmPupToLyotElem*>(e.get())) {
    // own keys, then one mask sub-block (fpmPupToLyot::init, :172-195), which may
    // hold a complexHexMaskFPM block after its own keys
    for (const Block& sb : find_blocks(b)) {
      if (sb.cmds.empty()) continue;
      const std::string& h = sb.cmds[0].key;
      if (h == "fpmCMCForPupToLyot" || h == "fpmIntHexCMCForPupToLyot" || h == "fpmBinaryForPupToLyot") {
        p->mtype = h == "fpmCMCForPupToLyot" ? FpmPupToLyotElem::MT_CMC
                 : (h == "fpmIntHexCMCForPupToLyot" ? FpmPupToLyotElem::MT_HEX : FpmPupToLyotElem::MT_BINARY);
        for (const Block& mb : find_blocks(sb)) {
          if (mb.cmds.empty()) continue;
          if (mb.cmds[0].key == "complexHexMaskFPM") {
            for (const Cmd& c : mb.cmds) CSX_TRY(p->hex.set(c.key, c.arg));
          } else {
            for (const Cmd& c : mb.cmds) CSX_TRY(p->set_mask(c.key, c.arg));
          }
        }
      } else {
        for (const Cmd& c : sb.cmds) CSX_TRY(e->set(*this, c.key, c.arg));
      }
    }
  } else {
    for (const Cmd& c : b.cmds) CSX_TRY(e->set(*this, c.key, c.arg));
  }
  out = std::move(e);
  return CSX_OK;
}

// ================================================================== chain ==
Element* Chain::elem_by_name(const std::string& n) {
  for (auto& e : elems_)
    if (e->name == n) return e.get();
  return nullptr;
}

int Chain::parse_top(const std::vector<Block>& blocks) {
  for (const Block& b : blocks) {
    if (b.cmds.empty()) continue;
    const std::string& t = b.cmds[0].key;
    if (t == "telescope") {   // telescope.cpp:21-46
      for (const Cmd& c : b.cmds) {
        std::string a = trim(c.arg);
        if (c.key == "magnification") tel_.mag = std::atof(a.c_str());
        else if (c.key == "primaryDiameter") tel_.D = std::atof(a.c_str());
        else if (c.key == "primaryFRatio" || c.key == "primaryfRatio") tel_.F = std::atof(a.c_str());
      }
      if (tel_.D == 0 || tel_.F == 0) return fail(CSX_ERR_PARSE, "telescope: primaryDiameter and primaryFRatio required");
      tel_.fLength = tel_.F * tel_.D;
    } else if (t == "efield") {   // efield.cpp:246-349
      int nl_size = 0;
      for (const Cmd& c : b.cmds) {
        const std::string& k = c.key;
        const std::string a = trim(c.arg);
        if (k == "efield" || k == "name" || k == "outputDirectory") continue;
        if (k == "setSize") {
          int r = 0, cc = 0, n = 0;
          std::sscanf(a.c_str(), "%d, %d, %d", &r, &cc, &n);
          ef_.rows = r; ef_.cols = cc; nl_size = n;
          ef_.arraySizePixels = r;            // efield::set_size (:442-452) sets it to the rows
        } else if (k == "constantValue") { ef_.constant = true; ef_.constantValue = std::atof(a.c_str()); }
        else if (k == "pointSource") {
          double fl = 0, tx = 0, ty = 0;
          std::sscanf(a.c_str(), "%lf, %lf, %lf", &fl, &tx, &ty);
          ef_.sources.push_back({std::sqrt(fl), tx * tel_.mag, ty * tel_.mag});
        } else if (k == "referenceWavelength") ef_.referenceWavelength = std::atof(a.c_str());
        else if (k == "wavelength") ef_.lambdas.assign(1, std::atof(a.c_str()));
        else if (k == "wavelengthRange") {
          double s0 = 0, s1 = 0, d = 0;
          std::sscanf(a.c_str(), "%lf, %lf, %lf", &s0, &s1, &d);
          int n = (int)std::round((s1 - s0) / d) + 1;
          ef_.lambdas.clear();
          for (int i = 0; i < n; i++) ef_.lambdas.push_back(s0 + i * d);
          ef_.arraySizePixels = ef_.rows;     // set_wavelength_range ends in set_size (:365)
        } else if (k == "wavelengthBandwidth") {
          double cw = 0, fr = 0; int n = 0;
          std::sscanf(a.c_str(), "%lf, %lf, %d", &cw, &fr, &n);
          double s0 = cw - cw * fr / 2, s1 = cw + cw * fr / 2, d = (s1 - s0) / ((double)n - 1);
          ef_.lambdas.clear();
          for (int i = 0; i < n; i++) ef_.lambdas.push_back(s0 + i * d);
          ef_.arraySizePixels = ef_.rows;     // set_wavelength_bandwidth: set_size (:388)
        } else if (k == "wavelengthList") {
          ef_.lambdas = parse_list(a);
          ef_.arraySizePixels = ef_.rows;     // set_wavelength_list: set_size (:416)
        }
        else if (k == "beamRadiusPhysical") ef_.beamRadius = std::atof(a.c_str());
        else if (k == "pixelScale") ef_.pixelScale = std::atof(a.c_str());
        else if (k == "arraySizePixels") ef_.arraySizePixels = std::atoi(a.c_str());
        else if (k == "filename") {
          size_t cpos = a.find(',');
          if (cpos == std::string::npos) return fail(CSX_ERR_PARSE, "efield filename needs re,im");
          std::vector<double> re, im;
          int r, cc, s, r2, c2, s2;
          CSX_TRY(fits_matrix(trim(a.substr(0, cpos)), re, r, cc, s));
          CSX_TRY(fits_matrix(trim(a.substr(cpos + 1)), im, r2, c2, s2));
          if (parse_audit()) continue;        // names recorded, nothing read
          ef_.rows = r; ef_.cols = cc;
          ef_.arraySizePixels = r;            // efield::init(amp, ph) -> set_size (:229)
          for (int sl = 0; sl < s; sl++) {
            std::vector<double2> f((size_t)r * cc);
            for (size_t i = 0; i < f.size(); i++)
              f[i] = make_double2(re[(size_t)sl * r * cc + i], im[(size_t)sl * r * cc + i]);
            ef_.fields[sl] = f;
          }
        } else if (ParseAudit* au = parse_audit()) {
          au->unknown_keys.push_back("efield:" + k);
        } else {
          std::fprintf(stderr, "!!!!! efield set: unknown fieldName: %s\n", k.c_str());
        }
      }
      if (ef_.lambdas.empty()) return fail(CSX_ERR_PARSE, "efield: no wavelength");
      if (nl_size && (int)ef_.lambdas.size() != nl_size && ef_.lambdas.size() == 1)
        ef_.lambdas.resize(nl_size, ef_.lambdas[0]);
      const bool audit = parse_audit() != nullptr;   // files not read: sizes unknown
      if (ef_.beamRadius == 0 || ef_.pixelScale == 0 || (ef_.arraySizePixels == 0 && !audit))
        return fail(CSX_ERR_PARSE, "efield: beamRadiusPhysical, pixelScale and arraySizePixels required");
      ef_.beamSizePixels = std::ceil(2 * ef_.beamRadius / ef_.pixelScale);          // :424
      ef_.arrayRadius = ef_.beamRadius * ef_.arraySizePixels / ef_.beamSizePixels;  // :427
      if (!audit && (ef_.rows <= 0 || ef_.cols <= 0)) return fail(CSX_ERR_PARSE, "efield: setSize required");
    } else if (t == "coronagraph") {
      std::vector<Block> sub = find_blocks(b);
      for (const Block& sb : sub) {
        if (sb.cmds.empty() || sb.cmds[0].key == "coronagraph") continue;
        std::unique_ptr<Element> e;
        CSX_TRY(make_element(sb, e));
        if (e) elems_.push_back(std::move(e));
      }
    } else if (t == "regionContrast") {   // regionContrast.cpp:36-87
      region_.present = true;
      for (const Cmd& c : b.cmds) {
        if (c.key == "referenceLambda") region_.refLambda = std::atof(c.arg.c_str());
        else if (c.key == "xlimit") { auto v = parse_float_pair(c.arg); region_.x1 = v[0]; region_.x2 = v[1]; }
        else if (c.key == "anglesInDegrees") { auto v = parse_float_pair(c.arg); region_.a1 = v[0]; region_.a2 = v[1]; }
        else if (c.key == "radii") { auto v = parse_float_pair(c.arg); region_.r1 = v[0]; region_.r2 = v[1]; }
      }
    } else {
      // other top-level blocks (optimisers, response builders, contrast curves) are
      // driven from the host API, not from the propagation chain
    }
  }
  return CSX_OK;
}

int Chain::load_script(const std::string& text) {
  CSX_TRY(parse_only(text));
  for (auto& e : elems_) CSX_TRY(e->init(*this));
  return CSX_OK;
}

int Chain::parse_only(const std::string& text) {
  Block all = parse_script_text(text);
  std::vector<Block> blocks = find_blocks(all);
  // the telescope must be known before efield keys use the magnification
  std::vector<Block> ordered;
  for (const Block& b : blocks) if (!b.cmds.empty() && b.cmds[0].key == "telescope") ordered.push_back(b);
  for (const Block& b : blocks) if (b.cmds.empty() || b.cmds[0].key != "telescope") ordered.push_back(b);
  CSX_TRY(parse_top(ordered));
  if (tel_.D == 0) return fail(CSX_ERR_PARSE, "script has no telescope block");
  if (ef_.rows == 0 && !(parse_audit() && ef_.pixelScale > 0)) return fail(CSX_ERR_PARSE, "script has no efield block");
  return CSX_OK;
}

int Chain::set_array(const std::string& elem, const std::string& key, const double* d, int rows, int cols) {
  Element* e = elem_by_name(elem);
  if (!e) return fail(CSX_ERR_ARG, "no element named " + elem);
  invalidate();
  return e->set_array(*this, key, d, rows, cols);
}

int Chain::set_complex_array(const std::string& elem, const std::string& key, const double* d, int rows, int cols,
                             int slice) {
  if (elem == "efield" && key == "field") {
    if (rows != ef_.rows || cols != ef_.cols) return fail(CSX_ERR_ARG, "field size does not match setSize");
    if (slice < 0) return fail(CSX_ERR_ARG, "negative slice");
    std::vector<double2> f((size_t)rows * cols);
    std::memcpy(f.data(), d, sizeof(double2) * f.size());
    ef_.fields[slice] = std::move(f);
    invalidate();
    return CSX_OK;
  }
  Element* e = elem_by_name(elem);
  if (!e) return fail(CSX_ERR_ARG, "no element named " + elem);
  invalidate();
  return e->set_complex_array(*this, key, d, rows, cols, slice);
}

int Chain::set_opt(const std::string& elem, const std::string& n, const std::vector<double>& v) {
  Element* e = elem_by_name(elem);
  if (!e) return fail(CSX_ERR_ARG, "no element named " + elem);
  return e->set_opt(*this, n, v);
}
int Chain::get_opt(const std::string& elem, const std::string& n, std::vector<double>& v) {
  Element* e = elem_by_name(elem);
  if (!e) return fail(CSX_ERR_ARG, "no element named " + elem);
  return e->get_opt(*this, n, v);
}

int Chain::set_candidates(const std::string& elem, const double* sags, int nhex, int ncand) {
  Element* e = elem_by_name(elem);
  const HexFpm* hx = nullptr;
  if (auto* d = dynamic_cast<DoubleFraunhoferElem*>(e)) hx = d->has_hex ? &d->hex : nullptr;
  else if (auto* p = dynamic_cast<FpmPupToLyotElem*>(e)) hx = p->mtype == FpmPupToLyotElem::MT_HEX ? &p->hex : nullptr;
  if (!hx) return fail(CSX_ERR_ARG, "no hex-FPM element named " + elem);
  if (ncand < 0 || (ncand > 0 && (!sags || nhex <= 0))) return fail(CSX_ERR_ARG, "bad candidate batch");
  if (ncand > 0 && nhex != hx->raster.nhex)
    return fail(CSX_ERR_ARG, "candidate vectors must have one sag per hex (" + std::to_string(hx->raster.nhex) + ")");
  bool shape_change = ncand != ncand_;
  if (ncand > 0) {
    CSX_TRY(cand_sags_.reserve(sizeof(double) * (size_t)nhex * ncand));
    CSX_HIP(hipMemcpyAsync(cand_sags_.p, sags, sizeof(double) * (size_t)nhex * ncand, hipMemcpyHostToDevice, stream()));
    CSX_HIP(hipStreamSynchronize(stream()));
  }
  ncand_ = ncand;
  cand_elem_ = elem;
  cand_nhex_ = nhex;
  if (shape_change) invalidate();
  return CSX_OK;
}

int Chain::candidate_sags(const std::string& elem, const double** dev, int* nhex, int* ncand) {
  if (ncand_ <= 0 || elem != cand_elem_) { *dev = nullptr; return CSX_ERR_STATE; }
  *dev = cand_sags_.as<double>();
  *nhex = cand_nhex_;
  *ncand = ncand_;
  return CSX_OK;
}

int Chain::compile_one(bool calibrating, Compile& c) {
  c.calibrating = calibrating;
  c.stages.clear();
  c.stage_names.clear();
  c.plans.clear();
  c.mft_at.clear();
  c.geom = geom_symmetric(ef_.cols, ef_.rows, ef_.pixelScale, ef_.pixelScale);
  c.buf = 0;
  // a candidate batch carries one copy of the fields up to the element holding the
  // candidate sags (everything before it is the same for every candidate)
  c.live = 0;
  if (!calibrating && ncand() > 1)
    for (auto& e : elems_)
      if (e->name == cand_elem_ && !e->skipped(c)) c.live = 1;
  // several point sources and nothing but masks and mirrors before the first MFT:
  // the buffers carry one unit source up to it and it applies every source
  // (CSX_SOURCE_FOLD=0 generates every source's field instead)
  c.live_src = 0;
  const char* sf = std::getenv("CSX_SOURCE_FOLD");
  if (!calibrating && ef_.fields.empty() && nsrc() > 1 && !(sf && sf[0] == '0'))
    for (auto& e : elems_) {
      if (dynamic_cast<MaskElem*>(e.get()) || dynamic_cast<MirrorElem*>(e.get())) continue;
      if (is_mft_element(e.get(), c)) c.live_src = 1;
      break;
    }
  // initial fields (efield::init_point_sources / init / filename slices)
  const int nf = nfield_for(c), nl = nlam(), ns = nsrc();
  const long npix = c.geom.npix();
  double2* E0 = field_buf(0);
  c.pending.clear();
  DevBuf& prm = calibrating ? src_prm_cal_ : src_prm_;
  std::vector<double> p((size_t)3 * nf);
  for (int f = 0; f < nf; f++) {
    int s = (f / nl) % ns;
    if (c.live_src) { p[3 * f] = 1.0; p[3 * f + 1] = 0; p[3 * f + 2] = 0; }
    else if (!ef_.sources.empty()) { p[3 * f] = ef_.sources[s].flux; p[3 * f + 1] = ef_.sources[s].tipX; p[3 * f + 2] = ef_.sources[s].tiltY; }
    else { p[3 * f] = ef_.constant ? ef_.constantValue : 1.0; p[3 * f + 1] = 0; p[3 * f + 2] = 0; }
  }
  CSX_TRY(prm.reserve(sizeof(double) * p.size()));
  CSX_HIP(hipMemcpy(prm.p, p.data(), sizeof(double) * p.size(), hipMemcpyHostToDevice));
  if (!ef_.fields.empty()) {
    // fields bound per lambda slice (one source), replicated over candidates
    DevBuf& fb = scratch(calibrating ? "init_fields_cal" : "init_fields");
    CSX_TRY(fb.reserve(sizeof(double2) * npix * nl));
    for (int l = 0; l < nl; l++) {
      auto it = ef_.fields.find(l);
      if (it == ef_.fields.end()) it = ef_.fields.begin();
      CSX_HIP(hipMemcpy(fb.as<double2>() + (size_t)l * npix, it->second.data(), sizeof(double2) * npix,
                        hipMemcpyHostToDevice));
    }
    const double2* src = fb.as<double2>();
    c.stages.push_back([=](hipStream_t s) {
      for (int f0 = 0; f0 < nf; f0 += nl)
        CSX_HIP(hipMemcpyAsync(E0 + (size_t)f0 * npix, src, sizeof(double2) * npix * nl, hipMemcpyDeviceToDevice, s));
      return CSX_OK;
    });
    c.stage_names.push_back("efield init");
  } else {
    c.pending.push_back(PwOp{PW_SOURCE, nullptr, 0.0, 1});   // fused with the following factors
  }
  c.in_conv = false;
  c.claimed = 0;
  c.support.reset();
  for (int i = 0; i < (int)elems_.size(); i++) {
    Element* e = elems_[i].get();
    if (c.claimed > 0) { c.claimed--; continue; }
    if (e->skipped(c) && !e->calibration_aware()) continue;
    if (!e->keeps_support()) c.support.reset();
    CSX_TRY(e->compile(*this, c, i));
  }
  CSX_TRY(flush_pointwise(*this, c));
  if (c.in_conv) return fail(CSX_ERR_STATE, "chain ends on converted fields");
  if (c.live_src) return fail(CSX_ERR_STATE, "point sources were not applied");
  CSX_TRY(expand_candidates(*this, c));
  return CSX_OK;
}

int Chain::compile_all() {
  if (compiled_) return CSX_OK;
  CSX_HIP(hipSetDevice(ctx_->device));
  // pass 1: largest field over both chains
  long maxe = 0;
  for (int cal = 0; cal < 2; cal++) {
    Compile dummy;
    dummy.calibrating = cal;
    Geom g = geom_symmetric(ef_.cols, ef_.rows, ef_.pixelScale, ef_.pixelScale);
    long nf = nfield_for(dummy);
    maxe = std::max(maxe, g.npix() * nf);
    for (auto& e : elems_) {
      if (e->skipped(dummy)) continue;
      Geom o;
      CSX_TRY(e->geometry(*this, g, o));
      if (auto* d = dynamic_cast<DoubleFraunhoferElem*>(e.get())) {
        for (int l = 0; l < nlam(); l++) {
          Geom gf; double lf;
          CSX_TRY(d->fpm_geom(*this, g, l, gf, lf));
          maxe = std::max(maxe, gf.npix() * nf);
        }
      }
      if (auto* fr = dynamic_cast<FresnelElem*>(e.get()))
        if (fr->padFactor != 1) maxe = std::max(maxe, 4 * g.npix() * nf);
      g = o;
      maxe = std::max(maxe, g.npix() * nf);
    }
  }
  max_field_elems_ = maxe;
  CSX_TRY(bufs_[0].reserve(sizeof(double2) * maxe));
  CSX_TRY(bufs_[1].reserve(sizeof(double2) * maxe));
  if (prec_ == CSX_FP32) CSX_TRY(c64_.reserve(sizeof(float2) * maxe));
  else if (prec_ != CSX_FP64) CSX_TRY(c64_.reserve((size_t)4 * terms_of(prec_) * maxe));
  // per-field wavelengths (candidate-major, then source, then lambda)
  for (int cal = 0; cal < 2; cal++) {
    Compile dummy;
    dummy.calibrating = cal;
    int nf = nfield_for(dummy);
    std::vector<double> lam(nf);
    for (int f = 0; f < nf; f++) lam[f] = ef_.lambdas[f % nlam()];
    DevBuf& b = cal ? scratch("lam_cal") : lam_dev_;
    CSX_TRY(b.reserve(sizeof(double) * nf));
    CSX_HIP(hipMemcpy(b.p, lam.data(), sizeof(double) * nf, hipMemcpyHostToDevice));
  }
  CSX_TRY(compile_one(false, plan_full_));
  out_geom_ = plan_full_.geom; out_buf_ = plan_full_.buf;
  CSX_TRY(compile_one(true, plan_cal_));
  out_geom_cal_ = plan_cal_.geom; out_buf_cal_ = plan_cal_.buf;
  compiled_ = true;
  generation_++;
  region_key_.clear();
  have_full_ = have_cal_ = false;
  calib_valid_ = false;
  return CSX_OK;
}

Chain::~Chain() {
  drop_graphs();
  if (contrast_host_) (void)hipHostFree(contrast_host_);
  if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
}

void Chain::drop_graphs() {
  for (GraphState& g : graph_) {
    if (g.e) (void)hipGraphExecDestroy(g.e);
    if (g.g) (void)hipGraphDestroy(g.g);
    g = GraphState();
  }
}

int Chain::run_stages(Compile& c, hipStream_t s) {
  for (size_t i = 0; i < c.stages.size(); i++) {
    int rc = c.stages[i](s);
    if (rc != CSX_OK) return fail(rc, "stage '" + c.stage_names[i] + "': " + csx_last_error());
  }
  return CSX_OK;
}

int Chain::propagate(bool calibrating, bool sync) {
  CSX_TRY(compile_all());
  CSX_HIP(hipSetDevice(ctx_->device));
  Compile& c = calibrating ? plan_cal_ : plan_full_;
  GraphState& gs = graph_[calibrating ? 1 : 0];
  // the kernel timer's per-launch events need eager launches (event records
  // captured into a graph do not keep their place between the kernels)
  const bool graphable = use_graph_ && !timer_.enabled();
  if (graphable && !gs.e && gs.ran_eager) {
    // capture after one eager run (lazily created tables exist by then) on a
    // private stream: the legacy default stream cannot be captured
    if (!cap_stream_) CSX_HIP(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking));
    CSX_HIP(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeRelaxed));
    int rc = run_stages(c, cap_stream_);
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(cap_stream_, &g);
    if (rc != CSX_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (ec != hipSuccess) return fail(CSX_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
    gs.g = g;
    CSX_HIP(hipGraphInstantiate(&gs.e, gs.g, nullptr, nullptr, 0));
  }
  if (graphable && gs.e) {
    CSX_HIP(hipGraphLaunch(gs.e, stream()));
  } else {
    CSX_TRY(run_stages(c, stream()));
    gs.ran_eager = true;
  }
  if (sync) {
    CSX_HIP(hipStreamSynchronize(stream()));
    CSX_TRY(finish_timing());
  }
  if (calibrating) { have_cal_ = true; have_full_ = false; } else { have_full_ = true; have_cal_ = false; }
  return CSX_OK;
}

// The fields linearOptimizer::compute_jacobian / compute_objective_vector read
// (src/optimization/linearOptimizer.cpp:420-505): every candidate's field at the
// region pixels, slice-major like the response matrix.
int Chain::region_fields(const csx_region* r, double2* d_out, int* npix, int* ncand_out, int* nslice) {
  CSX_TRY(compile_all());
  std::vector<int32_t> idx;
  CSX_TRY(region_pixels(r, idx));
  if (npix) *npix = (int)idx.size();
  if (ncand_out) *ncand_out = ncand();
  if (nslice) *nslice = nsrc() * nlam();
  if (!d_out) return CSX_OK;
  if (!have_full_) return fail(CSX_ERR_STATE, "region fields: no full propagation has run");
  DevBuf& di = scratch("region_field_idx");
  CSX_TRY(di.reserve(sizeof(int32_t) * idx.size()));
  CSX_HIP(hipMemcpyAsync(di.p, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice, stream()));
  CSX_TRY(gather_response(field_buf(out_buf_), out_geom_.npix(), di.as<int32_t>(), (int)idx.size(), ncand(),
                          nsrc() * nlam(), d_out, stream()));
  CSX_HIP(hipStreamSynchronize(stream()));
  return CSX_OK;
}

int Chain::get_field(int cand, int src, int lam, double* out) {
  bool cal = have_cal_;
  if (!have_full_ && !have_cal_) return fail(CSX_ERR_STATE, "no propagation has run");
  int ncd = cal ? 1 : ncand();
  if (cand < 0 || cand >= ncd || src < 0 || src >= nsrc() || lam < 0 || lam >= nlam())
    return fail(CSX_ERR_ARG, "field index out of range");
  const Geom& g = cal ? out_geom_cal_ : out_geom_;
  int b = cal ? out_buf_cal_ : out_buf_;
  size_t f = ((size_t)cand * nsrc() + src) * nlam() + lam;
  CSX_HIP(hipMemcpy(out, field_buf(b) + f * g.npix(), sizeof(double2) * g.npix(), hipMemcpyDeviceToHost));
  return CSX_OK;
}

int Chain::get_intensity(int cand, double* out) {
  bool cal = have_cal_;
  if (!have_full_ && !have_cal_) return fail(CSX_ERR_STATE, "no propagation has run");
  int ncd = cal ? 1 : ncand();
  if (cand < 0 || cand >= ncd) return fail(CSX_ERR_ARG, "candidate out of range");
  const Geom& g = cal ? out_geom_cal_ : out_geom_;
  int b = cal ? out_buf_cal_ : out_buf_;
  DevBuf& I = scratch("inten_tmp");
  CSX_TRY(I.reserve(sizeof(double) * g.npix()));
  int per = nsrc() * nlam();
  CSX_TRY(intensity_groups(field_buf(b) + (size_t)cand * per * g.npix(), g.npix(), 1, per, I.as<double>(), stream()));
  CSX_HIP(hipMemcpyAsync(out, I.p, sizeof(double) * g.npix(), hipMemcpyDeviceToHost, stream()));
  CSX_HIP(hipStreamSynchronize(stream()));
  return CSX_OK;
}

// Calibration (unocculted) run and its peak, cached until the chain changes.
int Chain::ensure_calibration() {
  CSX_TRY(compile_all());
  if (calib_valid_) return CSX_OK;
  const int per = nsrc() * nlam();
  CSX_TRY(propagate(true));
  const Geom gc = out_geom_cal_;
  CSX_TRY(inten_cal_.reserve(sizeof(double) * gc.npix()));
  CSX_TRY(peak_.reserve(sizeof(double)));
  CSX_TRY(intensity_groups(field_buf(out_buf_cal_), gc.npix(), 1, per, inten_cal_.as<double>(), stream()));
  CSX_TRY(max_groups(inten_cal_.as<double>(), gc.npix(), 1, peak_.as<double>(), stream()));
  CSX_HIP(hipMemcpyAsync(&calib_peak_, peak_.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
  CSX_HIP(hipStreamSynchronize(stream()));
  calib_valid_ = true;
  return CSX_OK;
}

int Chain::calibration_intensity(std::vector<double>& out, int* rows, int* cols) {
  CSX_TRY(ensure_calibration());
  const Geom gc = out_geom_cal_;
  out.resize((size_t)gc.npix());
  CSX_HIP(hipMemcpy(out.data(), inten_cal_.p, sizeof(double) * out.size(), hipMemcpyDeviceToHost));
  *rows = gc.ny;
  *cols = gc.nx;
  return CSX_OK;
}

// Output pixels of the contrast region (regionContrast::get_region_pixels, :89-101),
// column-major (arma::find) order; r overrides the script's regionContrast block.
int Chain::region_pixels(const csx_region* r, std::vector<int32_t>& idx) {
  Region R = region_;
  if (r) {
    R.r1 = r->radius1; R.r2 = r->radius2; R.a1 = r->angle1; R.a2 = r->angle2;
    R.x1 = r->xlim1; R.x2 = r->xlim2; R.refLambda = r->reference_lambda > 0 ? r->reference_lambda : -1.0;
  }
  double refl = R.refLambda;
  if (refl == -1.0) {
    double sum = 0;
    for (double l : ef_.lambdas) sum += l;
    refl = sum / nlam();
  }
  const double loD = tel_.loD(refl);
  const Geom g = out_geom_;
  idx.clear();
  for (int c = 0; c < g.nx; c++)
    for (int rr = 0; rr < g.ny; rr++) {
      double XX = g.x0 + c * g.dx, YY = g.y0 + rr * g.dy;
      double RR = std::sqrt(std::pow(XX, 2) + std::pow(YY, 2)), TT = std::atan2(YY, XX);
      if ((XX / loD >= R.x1) && (XX / loD <= R.x2) && (RR / loD >= R.r1) && (RR / loD <= R.r2) &&
          (TT > R.a1 * M_PI / 180) && (TT < R.a2 * M_PI / 180))
        idx.push_back(c * g.ny + rr);
    }
  if (idx.empty()) return fail(CSX_ERR_ARG, "contrast region contains no pixel");
  return CSX_OK;
}

// Region mean of sum_{sources, lambdas} |E|^2 / peak per candidate from the fields
// of the last full propagation (regionContrast::compute_contrast, :145-205).
int Chain::reduce_contrast(const csx_region* r, double* out) {
  Region R = region_;
  if (r) {
    R.r1 = r->radius1; R.r2 = r->radius2; R.a1 = r->angle1; R.a2 = r->angle2;
    R.x1 = r->xlim1; R.x2 = r->xlim2; R.refLambda = r->reference_lambda > 0 ? r->reference_lambda : -1.0;
  }
  const int per = nsrc() * nlam();
  const Geom g = out_geom_;
  int nc = ncand();
  CSX_TRY(inten_.reserve(sizeof(double) * g.npix() * nc));
  CSX_TRY(intensity_groups(field_buf(out_buf_), g.npix(), nc, per, inten_.as<double>(), stream()));
  // region pixels (regionContrast::get_region_pixels, :89-101), column-major order;
  // recomputed only when the region or the output grid changes
  double refl = R.refLambda;
  if (refl == -1.0) {
    double sum = 0;
    for (double l : ef_.lambdas) sum += l;
    refl = sum / nlam();
  }
  const double loD = tel_.loD(refl);
  const std::vector<double> key = {R.r1, R.r2, R.a1, R.a2, R.x1, R.x2, loD, (double)g.nx, (double)g.ny, g.x0, g.y0,
                                   g.dx, g.dy};
  if (key != region_key_) {
    std::vector<int32_t> idx;
    CSX_TRY(region_pixels(r, idx));
    CSX_TRY(region_idx_.reserve(sizeof(int32_t) * idx.size()));
    CSX_HIP(hipMemcpy(region_idx_.p, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice));
    n_region_ = (int)idx.size();
    region_key_ = key;
  }
  CSX_TRY(contrast_.reserve(sizeof(double) * nc));
  CSX_TRY(region_mean_groups(inten_.as<double>(), g.npix(), nc, region_idx_.as<int32_t>(), n_region_,
                             peak_.as<double>(), 0.0, contrast_.as<double>(), stream()));
  // through a pinned buffer: a direct DMA of a few bytes instead of a staged copy
  if (contrast_host_n_ < nc) {
    if (contrast_host_) (void)hipHostFree(contrast_host_);
    contrast_host_ = nullptr;
    contrast_host_n_ = 0;
    CSX_HIP(hipHostMalloc((void**)&contrast_host_, sizeof(double) * nc, hipHostMallocDefault));
    contrast_host_n_ = nc;
  }
  CSX_HIP(hipMemcpyAsync(contrast_host_, contrast_.p, sizeof(double) * nc, hipMemcpyDeviceToHost, stream()));
  CSX_HIP(hipStreamSynchronize(stream()));
  std::memcpy(out, contrast_host_, sizeof(double) * nc);
  return CSX_OK;
}

int Chain::contrast(const csx_region* r, double* out, csx_report* rep) {
  CSX_TRY(compile_all());
  auto t0 = std::chrono::steady_clock::now();
  CSX_TRY(ensure_calibration());
  auto t1 = std::chrono::steady_clock::now();
  CSX_TRY(propagate(false, false));       // the reduction below is queued behind it
  CSX_TRY(reduce_contrast(r, out));        // ends with a stream synchronisation
  auto t2 = std::chrono::steady_clock::now();
  CSX_TRY(finish_timing());
  if (rep) {
    rep->calib_peak = calib_peak_;
    rep->n_region_pixels = n_region_;
    rep->seconds_calibration = std::chrono::duration<double>(t1 - t0).count();
    rep->seconds_propagation = std::chrono::duration<double>(t2 - t1).count();
  }
  return CSX_OK;
}

// contrastCurve::make_contrast_curve (src/tools/contrastCurve.cpp:68-158); its
// pixelSampling is 1 (contrastCurve.hpp:15) and it reads the middle wavelength
// lambdaData[(int) round(n_slices/2)] (integer division first).
int Chain::contrast_curve(int cand, double* pos, double* avg, int capacity, int* rows) {
  CSX_TRY(compile_all());
  const Geom g = out_geom_;
  const int pixel_sampling = 1;
  const int ns = (g.nx - 2) / pixel_sampling;
  const int nrow = std::max(ns - 3, 0);
  *rows = nrow;
  if (!pos) return CSX_OK;
  if (nrow == 0) return fail(CSX_ERR_ARG, "contrast curve: output grid too small");
  if (capacity < nrow) return fail(CSX_ERR_ARG, "contrast curve: capacity below the row count");
  if (cand < 0 || cand >= ncand()) return fail(CSX_ERR_ARG, "contrast curve: candidate out of range");
  CSX_TRY(ensure_calibration());
  CSX_TRY(propagate(false, false));
  const int per = nsrc() * nlam(), nc = ncand();
  CSX_TRY(inten_.reserve(sizeof(double) * g.npix() * nc));
  CSX_TRY(intensity_groups(field_buf(out_buf_), g.npix(), nc, per, inten_.as<double>(), stream()));
  const double loD = tel_.loD(ef_.lambdas[nlam() / 2]);
  // arma::linspace(0, max(pixelX / loD), ns)
  const double pmax = std::max(g.x0, g.x0 + (g.nx - 1) * g.dx) / loD;
  std::vector<double> smp(ns);
  if (ns == 1) smp[0] = pmax;
  for (int i = 0; ns > 1 && i < ns; i++) smp[i] = i == ns - 1 ? pmax : i * (pmax / (ns - 1));
  DevBuf& ds = scratch("curve_samples");
  DevBuf& dout = scratch("curve_out");
  CSX_TRY(ds.reserve(sizeof(double) * ns));
  CSX_TRY(dout.reserve(sizeof(double) * nrow));
  CSX_HIP(hipMemcpyAsync(ds.p, smp.data(), sizeof(double) * ns, hipMemcpyHostToDevice, stream()));
  CSX_TRY(annulus_means(inten_.as<double>() + (size_t)cand * g.npix(), g.nx, g.ny, g.x0, g.dx, g.y0, g.dy, loD,
                        ds.as<double>() + 1, nrow, peak_.as<double>(), dout.as<double>(), stream()));
  CSX_HIP(hipMemcpyAsync(avg, dout.p, sizeof(double) * nrow, hipMemcpyDeviceToHost, stream()));
  CSX_HIP(hipStreamSynchronize(stream()));
  CSX_TRY(finish_timing());
  for (int k = 0; k < nrow; k++) pos[k] = smp[k + 2];
  return CSX_OK;
}

int Chain::get_normalized_image(int cand, double* out) {
  if (calib_peak_ == 0) return fail(CSX_ERR_STATE, "run csx_chain_contrast first");
  if (cand < 0 || cand >= ncand()) return fail(CSX_ERR_ARG, "candidate out of range");
  const Geom g = out_geom_;
  CSX_HIP(hipMemcpy(out, inten_.as<double>() + (size_t)cand * g.npix(), sizeof(double) * g.npix(),
                    hipMemcpyDeviceToHost));
  for (long i = 0; i < g.npix(); i++) out[i] /= calib_peak_;
  return CSX_OK;
}

int Chain::field_profiles(const int32_t* bins, int nbins, double* out) {
  if (!have_full_ && !have_cal_) return fail(CSX_ERR_STATE, "no propagation has run");
  bool cal = have_cal_;
  const Geom& g = cal ? out_geom_cal_ : out_geom_;
  int b = cal ? out_buf_cal_ : out_buf_;
  int nf = (cal ? 1 : ncand()) * nsrc() * nlam();
  DevBuf& db = scratch("profile_bins");
  DevBuf& dout = scratch("profile_out");
  CSX_TRY(db.reserve(sizeof(int32_t) * g.npix()));
  CSX_TRY(dout.reserve(sizeof(double) * (size_t)nf * nbins));
  CSX_HIP(hipMemcpyAsync(db.p, bins, sizeof(int32_t) * g.npix(), hipMemcpyHostToDevice, stream()));
  CSX_TRY(bin_sums(field_buf(b), g.npix(), nf, db.as<int32_t>(), nbins, dout.as<double>(), stream()));
  CSX_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * (size_t)nf * nbins, hipMemcpyDeviceToHost, stream()));
  CSX_HIP(hipStreamSynchronize(stream()));
  return CSX_OK;
}

int Chain::info(csx_chain_info* o) {
  CSX_TRY(compile_all());
  o->n_sources = nsrc();
  o->n_lambdas = nlam();
  o->n_candidates = ncand();
  o->in_rows = ef_.rows;
  o->in_cols = ef_.cols;
  o->out_rows = out_geom_.ny;
  o->out_cols = out_geom_.nx;
  o->out_pixel_size = out_geom_.dx;
  double refl = region_.refLambda;
  if (refl == -1.0) { double s = 0; for (double l : ef_.lambdas) s += l; refl = s / nlam(); }
  o->lambda_over_d = tel_.loD(refl);
  return CSX_OK;
}


// ======================================================= response matrix ==
// makeHexCFpmResponse::compute_response (src/tools/makeHexCFpmResponse.cpp:91-190) as
// one batch: with every sag zero, candidate c < n_hex transmits only hex c and the
// last candidate only the area outside the hexes (with the Babinet term for the
// convolution FPM); the region pixels of each field form column c of M.  The
// objective of any sag vector is then |M [exp(i phaseSign 4 pi s / lambda); 1]|^2
// (optHexRespObjective, src/optimization/optObjective.cpp:24-69).
int Chain::hex_response(const std::string& elem, const csx_region* r, double2* d_M, int* npix, int* ncol,
                        int* nslice, double* calib) {
  Element* e = elem_by_name(elem);
  const HexFpm* hx = nullptr;
  if (auto* d = dynamic_cast<DoubleFraunhoferElem*>(e)) hx = d->has_hex ? &d->hex : nullptr;
  else if (auto* p = dynamic_cast<FpmPupToLyotElem*>(e)) hx = p->mtype == FpmPupToLyotElem::MT_HEX ? &p->hex : nullptr;
  if (!hx) return fail(CSX_ERR_ARG, "no hex-FPM element named " + elem);
  const int nhex = hx->raster.nhex;
  CSX_TRY(compile_all());
  std::vector<int32_t> idx;
  CSX_TRY(region_pixels(r, idx));
  if (npix) *npix = (int)idx.size();
  if (ncol) *ncol = nhex + 1;
  if (nslice) *nslice = nsrc() * nlam();
  if (!d_M) return CSX_OK;
  std::vector<double> keep_use;
  CSX_TRY(e->get_opt(*this, "useOnlyThisHex", keep_use));
  const int keep_ncand = ncand_, keep_nhex = cand_nhex_;
  std::vector<double> keep_sags;
  if (keep_ncand > 0) {
    keep_sags.resize((size_t)keep_nhex * keep_ncand);
    CSX_HIP(hipMemcpy(keep_sags.data(), cand_sags_.p, sizeof(double) * keep_sags.size(), hipMemcpyDeviceToHost));
  }
  const std::string keep_elem = cand_elem_;
  std::vector<double> zeros((size_t)nhex * (nhex + 1), 0.0);
  int rc = e->set_opt(*this, "useOnlyThisHex", std::vector<double>{(double)kUseOnlyPerCandidate});
  if (rc == CSX_OK) rc = set_candidates(elem, zeros.data(), nhex, nhex + 1);
  if (rc == CSX_OK) rc = ensure_calibration();
  if (rc == CSX_OK) rc = propagate(false);
  if (rc == CSX_OK) {
    DevBuf& di = scratch("response_idx");
    rc = di.reserve(sizeof(int32_t) * idx.size());
    if (rc == CSX_OK && hipMemcpy(di.p, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(CSX_ERR_HIP, "response: index upload");
    if (rc == CSX_OK)
      rc = gather_response(field_buf(out_buf_), out_geom_.npix(), di.as<int32_t>(), (int)idx.size(), nhex + 1,
                           nsrc() * nlam(), d_M, stream());
    if (rc == CSX_OK && hipStreamSynchronize(stream()) != hipSuccess) rc = fail(CSX_ERR_HIP, "response: sync");
  }
  if (calib) *calib = calib_peak_;
  // restore the element's hex selection and the candidate batch
  int rc2 = e->set_opt(*this, "useOnlyThisHex", keep_use);
  int rc3 = keep_ncand > 0 ? set_candidates(keep_elem, keep_sags.data(), keep_nhex, keep_ncand)
                           : set_candidates(elem, nullptr, 0, 0);
  if (rc != CSX_OK) return rc;
  if (rc2 != CSX_OK) return rc2;
  return rc3;
}

// ============================================================ chain group ==
int ChainGroup::n_candidates() const {
  int n = 0;
  for (Chain* c : ch_) n += c->ncand();
  return n;
}

int ChainGroup::build() {
  if (ch_.empty()) return fail(CSX_ERR_ARG, "chain group: no chains");
  for (Chain* c : ch_) CSX_TRY(c->compile());
  bool same = gen_.size() == ch_.size();
  for (size_t k = 0; same && k < ch_.size(); k++) same = gen_[k] == ch_[k]->generation();
  if (same && !stages_.empty()) return CSX_OK;
  Chain* c0 = ch_[0];
  const Compile& p0 = c0->plan(false);
  for (Chain* c : ch_) {
    const Compile& pk = c->plan(false);
    if (c->ctx() != c0->ctx()) return fail(CSX_ERR_ARG, "chain group: chains on different contexts");
    if (c->precision() != c0->precision()) return fail(CSX_ERR_ARG, "chain group: precisions differ");
    if (pk.stages.size() != p0.stages.size()) return fail(CSX_ERR_ARG, "chain group: element lists differ");
    for (size_t i = 0; i < p0.stages.size(); i++)
      if ((pk.mft_at.count(i) != 0) != (p0.mft_at.count(i) != 0))
        return fail(CSX_ERR_ARG, "chain group: element lists differ at stage " + std::to_string(i));
  }
  stages_.clear();
  plans_.clear();
  merged_ = 0;
  for (size_t i = 0; i < p0.stages.size(); i++) {
    if (p0.mft_at.count(i)) {
      std::vector<MftField> F;
      for (Chain* c : ch_) {
        const auto& f = c->plan(false).mft_at.at(i);
        F.insert(F.end(), f.begin(), f.end());
      }
      auto plan = std::make_unique<MftPlan>();
      CSX_TRY(plan->build(F, c0->precision(), c0->stream()));
      const MftPlan* pp = plan.get();
      plans_.push_back(std::move(plan));
      Chain* tc = c0;
      stages_.push_back([pp, tc](hipStream_t s) { return pp->run(s, tc->timer()); });
      merged_++;
    } else {
      std::vector<std::function<int(hipStream_t)>> fns;
      for (Chain* c : ch_) fns.push_back(c->plan(false).stages[i]);
      stages_.push_back([fns](hipStream_t s) {
        for (const auto& f : fns) CSX_TRY(f(s));
        return CSX_OK;
      });
    }
  }
  gen_.clear();
  for (Chain* c : ch_) gen_.push_back(c->generation());
  return CSX_OK;
}

int ChainGroup::propagate(bool sync) {
  CSX_TRY(build());
  hipStream_t s = ch_[0]->stream();
  for (size_t i = 0; i < stages_.size(); i++) {
    int rc = stages_[i](s);
    if (rc != CSX_OK) return fail(rc, "group stage " + std::to_string(i) + ": " + csx_last_error());
  }
  if (sync) CSX_HIP(hipStreamSynchronize(s));
  for (Chain* c : ch_) {
    c->mark_propagated(false);
    if (sync) CSX_TRY(c->finish_timing());
  }
  return CSX_OK;
}

int ChainGroup::contrast(const csx_region* r, double* out) {
  for (Chain* c : ch_) CSX_TRY(c->ensure_calibration());   // each chain's own (cached) calibration
  CSX_TRY(propagate(false));               // the reductions are queued behind it on the same stream
  size_t off = 0;
  for (Chain* c : ch_) {
    CSX_TRY(c->reduce_contrast(r, out + off));
    off += c->ncand();
  }
  for (Chain* c : ch_) CSX_TRY(c->finish_timing());
  return CSX_OK;
}

}  // namespace csx
