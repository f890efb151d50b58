// Element chain: csim's input-script element API (src/elements/celem.hpp,
// src/coronagraph/coronagraph.cpp) compiled into a list of GPU stages that run
// over every field of a batch (candidate x source x wavelength).
#pragma once
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "hostgeom.hpp"
#include "mft.hpp"
#include "ops.hpp"

namespace csx {

// ------------------------------------------------------------- script text --
// initCommandSet / find_command_blocks (src/lib/csim_parser.cpp:60-190).
struct Cmd { std::string key, arg; };
struct Block { std::vector<Cmd> cmds; };
Block parse_script_text(const std::string& text);
std::vector<Block> find_blocks(const Block& b);

// Minimal FITS reader (BITPIX 8/16/32/-32/-64, NAXIS <= 3): values in FITS order
// (NAXIS1 fastest).  csim reads matrices as rows = NAXIS2, cols = NAXIS1.
int read_fits(const std::string& path, std::vector<double>& data, std::vector<long>& dims);

// Device-free script check (csx_script_parse): while an audit is installed on this
// thread, FITS matrix loads record the file name and return an empty matrix, unknown
// keys and element types are recorded (and parsing goes on, as csim prints and goes
// on), and nothing touches a device.
struct ParseAudit {
  std::vector<std::string> files, unknown_keys, unknown_elements;
};
ParseAudit*& parse_audit();

// --------------------------------------------------------------- geometry --
// arrayGeom with uniform grids: x along columns, y along rows.
struct Geom {
  int nx = 0, ny = 0;
  double x0 = 0, y0 = 0, dx = 0, dy = 0;
  int rows() const { return ny; }
  int cols() const { return nx; }
  long npix() const { return (long)nx * ny; }
};
Geom geom_symmetric(int nx, int ny, double dx, double dy, double ox = 0, double oy = 0);
Geom geom_offset(int nx, int ny, double dx, double dy);

struct Telescope {
  double D = 0, F = 0, mag = 1, fLength = 0;
  double loD(double lam) const { return mag * fLength * lam / D; }   // telescope.cpp:58-60
};

struct Source { double flux, tipX, tiltY; };

struct Efield {
  int rows = 0, cols = 0;
  std::vector<double> lambdas;
  double beamRadius = 0, pixelScale = 0, beamSizePixels = 0, arrayRadius = 0, referenceWavelength = 0;
  int arraySizePixels = 0;
  std::vector<Source> sources;
  bool constant = false;
  double constantValue = 0;
  // complex fields bound from memory / FITS, one per (source, lambda) slice
  std::map<int, std::vector<double2>> fields;
  int nsources() const { return std::max<int>(1, std::max<int>((int)sources.size(), fields.empty() ? 0 : 1)); }
};

struct Region {
  double r1 = 1.0, r2 = 2.0, a1 = 0.0, a2 = 30.0, x1 = -1e16, x2 = 1e16, refLambda = -1.0;
  bool present = false;
};

class Chain;

// Compile-time state threaded through the elements.
struct Compile {
  bool calibrating = false;
  Geom geom;            // geometry of the current fields
  int buf = 0;          // ping-pong buffer holding the fields (complex128)
  bool in_conv = false; // fields live in the conversion buffer at the working precision instead
  int claimed = 0;      // following elements already folded into the last one compiled
  std::vector<PwOp> pending;   // pointwise factors not yet applied (one fused pass)
  std::vector<std::function<int(hipStream_t)>> stages;
  std::vector<std::string> stage_names;
  std::vector<std::unique_ptr<MftPlan>> plans;
  std::map<size_t, std::vector<MftField>> mft_at;   // stage index -> fields of a plain MFT launch
  // candidate copies the buffers hold (0: all).  Before the element that takes the
  // candidate sags every candidate's fields are identical, so they are carried once
  // (live = 1) and that element expands them to every candidate.
  int live = 0;
  // 1: the buffers hold one unit point source per (candidate copy, wavelength); the
  // next MFT applies every source (flux and tilt folded into its DFT kernels)
  int live_src = 0;
  // zero pattern of the current fields (null: dense): masks narrow it, an MFT reads
  // it for its input and leaves its epilogue mask's; every other element clears it
  std::shared_ptr<const Support> support;
};

// ---------------------------------------------------------------- elements --
struct Element {
  std::string type, name;
  double position = 0;
  bool disableForCalibration = false;
  virtual ~Element() = default;
  // celem::set (src/elements/celem.cpp:97-161) then the element's own keys.
  bool set_common(const std::string& key, const std::string& arg);
  virtual int set(Chain& ch, const std::string& key, const std::string& arg) = 0;
  virtual int init(Chain& ch) { (void)ch; return CSX_OK; }
  // Output geometry for an input geometry (no GPU work).
  virtual int geometry(Chain& ch, const Geom& in, Geom& out) { (void)ch; out = in; return CSX_OK; }
  // Appends the GPU stages of this element.
  virtual int compile(Chain& ch, Compile& c, int index) = 0;
  virtual int set_opt(Chain& ch, const std::string& name, const std::vector<double>& v);
  virtual int get_opt(Chain& ch, const std::string& name, std::vector<double>& v);
  virtual int set_array(Chain& ch, const std::string& key, const double* d, int rows, int cols);
  virtual int set_complex_array(Chain& ch, const std::string& key, const double* d, int rows, int cols, int slice);
  bool skipped(const Compile& c) const { return disableForCalibration && c.calibrating; }
  // true: compile() is called for the calibration run even when skipped(c) (the
  // element then applies its own calibration behaviour)
  virtual bool calibration_aware() const { return false; }
  // true: compile() keeps Compile::support valid for the fields it leaves (a
  // pointwise factor, or an MFT that reads and resets it); false: it is cleared first
  virtual bool keeps_support() const { return false; }
  // A pointwise element (mask, mirror) on fields of geometry g as one factor of a fused
  // pass: 1 (op filled), 0 (not a pointwise element), or a negative error code.
  virtual int pointwise_op(Chain& ch, const Geom& g, PwOp* op) { (void)ch; (void)g; (void)op; return 0; }
};

class Chain {
 public:
  explicit Chain(csx_ctx* ctx) : ctx_(ctx) {}
  ~Chain();
  Chain(const Chain&) = delete;
  Chain& operator=(const Chain&) = delete;
  int load_script(const std::string& text);
  // the script's blocks parsed into the telescope, efield, elements and region, with
  // no element init (no device work, no cached data); load_script = this + init
  int parse_only(const std::string& text);
  // sum over sources and wavelengths of the calibration fields' |E|^2 (the
  // calibration run cached like csx_chain_contrast's), out_rows x out_cols column-major
  int calibration_intensity(std::vector<double>& out, int* rows, int* cols);
  int set_array(const std::string& elem, const std::string& key, const double* d, int rows, int cols);
  int set_complex_array(const std::string& elem, const std::string& key, const double* d, int rows, int cols,
                        int slice);
  int set_opt(const std::string& elem, const std::string& name, const std::vector<double>& v);
  int get_opt(const std::string& elem, const std::string& name, std::vector<double>& v);
  int set_candidates(const std::string& elem, const double* sags, int nhex, int ncand);
  int set_precision(csx_precision p) { if (p != prec_) { prec_ = p; invalidate(); } return CSX_OK; }
  // sync = false leaves the work queued on the context's stream (the caller
  // synchronises, then calls finish_timing)
  int propagate(bool calibrating, bool sync = true);
  int finish_timing() { return timer_.enabled() ? timer_.collect() : CSX_OK; }
  int get_field(int cand, int src, int lam, double* out);
  int get_intensity(int cand, double* out);
  int contrast(const csx_region* r, double* out, csx_report* rep);
  int ensure_calibration();
  int reduce_contrast(const csx_region* r, double* out);
  int region_pixels(const csx_region* r, std::vector<int32_t>& idx);
  // region-pixel fields of every candidate after the last full propagation
  int region_fields(const csx_region* r, double2* d_out, int* npix, int* ncand_out, int* nslice);
  int hex_response(const std::string& elem, const csx_region* r, double2* d_M, int* npix, int* ncol, int* nslice,
                   double* calib);
  // ---- chain groups (one launch per contraction over several chains)
  Compile& plan(bool calibrating) { return calibrating ? plan_cal_ : plan_full_; }
  int compile() { return compile_all(); }
  long generation() const { return generation_; }     // bumps on every recompilation
  void mark_propagated(bool calibrating) {
    if (calibrating) { have_cal_ = true; have_full_ = false; } else { have_full_ = true; have_cal_ = false; }
  }
  double calib_peak() const { return calib_peak_; }
  int n_region_pixels() const { return n_region_; }
  int get_normalized_image(int cand, double* out);
  // pos/avg null: only *rows is set (no propagation)
  int contrast_curve(int cand, double* pos, double* avg, int capacity, int* rows);
  int field_profiles(const int32_t* bins, int nbins, double* out);
  int info(csx_chain_info* out);
  int set_profiling(bool on) { timer_.enable(on); return CSX_OK; }
  int set_graph(bool on) { use_graph_ = on; if (!on) drop_graphs(); return CSX_OK; }
  std::string profile_json() const { return timer_.json(); }
  KernelTimer* timer() { return timer_.enabled() ? &timer_ : nullptr; }

  // ---- shared with elements
  csx_ctx* ctx() const { return ctx_; }
  hipStream_t stream() const { return ctx_->stream; }
  const Telescope& telescope() const { return tel_; }
  const Efield& efield() const { return ef_; }
  csx_precision precision() const { return prec_; }
  int ncand() const { return std::max(1, ncand_); }
  int ncand_for(const Compile& c) const { return c.calibrating ? 1 : ncand(); }
  // fields in the buffers at this point of the compile / with every candidate
  int nfield_for(const Compile& c) const {
    return (c.live > 0 ? c.live : ncand_for(c)) * (c.live_src ? 1 : ef_.nsources()) * (int)ef_.lambdas.size();
  }
  // fields an MFT writes from the current ones (every source; every candidate if it expands them)
  int nfield_out(const Compile& c, bool expand) const {
    return (expand || c.live <= 0 ? ncand_for(c) : c.live) * ef_.nsources() * (int)ef_.lambdas.size();
  }
  int nfield_all(const Compile& c) const { return ncand_for(c) * ef_.nsources() * (int)ef_.lambdas.size(); }
  int nlam() const { return (int)ef_.lambdas.size(); }
  int nsrc() const { return ef_.nsources(); }
  double2* field_buf(int i) const { return bufs_[i].as<double2>(); }
  void* c64_buf() const { return c64_.p; }
  const double* lambdas_dev() const { return lam_dev_.as<double>(); }   // per field (candidate-major)
  const double* source_params(bool cal) const { return (cal ? src_prm_cal_ : src_prm_).as<double>(); }
  Element* elem_by_name(const std::string& n);
  Element* elem_at(int i) { return (i >= 0 && i < (int)elems_.size()) ? elems_[i].get() : nullptr; }
  int n_elems() const { return (int)elems_.size(); }
  void invalidate() { compiled_ = false; calib_valid_ = false; drop_graphs(); }
  void invalidate_calibration() { calib_valid_ = false; }
  DevBuf& scratch(const std::string& key) { return scratch_[key]; }
  int candidate_sags(const std::string& elem, const double** dev, int* nhex, int* ncand);

 private:
  int compile_all();
  int compile_one(bool calibrating, Compile& c);
  int run_stages(Compile& c, hipStream_t s);
  void drop_graphs();
  int parse_top(const std::vector<Block>& blocks);
  int make_element(const Block& b, std::unique_ptr<Element>& out);

  csx_ctx* ctx_;
  Telescope tel_;
  Efield ef_;
  Region region_;
  std::vector<std::unique_ptr<Element>> elems_;
  csx_precision prec_ = CSX_FP64;
  int ncand_ = 0;
  std::string cand_elem_;
  DevBuf cand_sags_;
  int cand_nhex_ = 0;

  bool compiled_ = false;
  Compile plan_full_, plan_cal_;
  DevBuf bufs_[2], c64_, lam_dev_, lam_cal_dev_, src_prm_, src_prm_cal_;
  std::map<std::string, DevBuf> scratch_;
  long max_field_elems_ = 0;
  Geom out_geom_, out_geom_cal_;
  int out_buf_ = 0, out_buf_cal_ = 0;
  bool have_full_ = false, have_cal_ = false;
  DevBuf inten_, inten_cal_, peak_, region_idx_, contrast_;
  double* contrast_host_ = nullptr;   // pinned landing buffer of the per-candidate contrasts
  int contrast_host_n_ = 0;
  double calib_peak_ = 0;
  bool calib_valid_ = false;
  int n_region_ = 0;
  std::vector<double> region_key_;
  long generation_ = 0;
  KernelTimer timer_;
  // graph replay of each plan ([0] full, [1] calibration)
  struct GraphState { hipGraph_t g = nullptr; hipGraphExec_t e = nullptr; bool ran_eager = false; };
  bool use_graph_ = false;
  GraphState graph_[2];
  hipStream_t cap_stream_ = nullptr;
};


// Several compiled chains with identical element lists (the same script at
// different pupil sizes, e.g. 512^2 and 1024^2 DE candidates) propagated together:
// every contraction stage of all chains is one grouped GEMM launch over uneven
// problem sizes; the other stages run chain by chain.
// FITS matrices in csim's convention (column-major rows x cols [x slices]).
int fits_matrix(const std::string& path, std::vector<double>& m, int& rows, int& cols, int& slices);
int write_fits_matrix(const std::string& path, const double* colmajor, int rows, int cols);

class ChainGroup {
 public:
  explicit ChainGroup(std::vector<Chain*> chains) : ch_(std::move(chains)) {}
  int propagate(bool sync = true);
  int contrast(const csx_region* r, double* out);   // out: candidates of every chain, chain by chain
  int n_candidates() const;
  int merged_launches() const { return merged_; }

 private:
  int build();
  std::vector<Chain*> ch_;
  std::vector<long> gen_;
  std::vector<std::function<int(hipStream_t)>> stages_;
  std::vector<std::unique_ptr<MftPlan>> plans_;
  int merged_ = 0;
};

}  // namespace csx
