// Matrix Fourier transform plans: csim zoomFft (src/lib/csim_fft.cpp:670-826) as
// two grouped complex GEMMs over a whole batch of fields.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"

namespace csx {

// Per-launch timing with HIP events recorded on the launch stream; enabled for
// measurement runs only.  Event pairs are pooled and read after a synchronise.
class KernelTimer {
 public:
  ~KernelTimer();
  void enable(bool on);
  bool enabled() const { return on_; }
  // `launches` kernels of `name` run back to back between begin and end
  int begin(const std::string& name, double flops, double bytes, hipStream_t s, int launches = 1);
  int end(hipStream_t s);
  int collect();                        // after the stream is synchronised
  std::string json() const;
  struct Stat { long count = 0; double ms = 0, flops = 0, bytes = 0; };
  const std::map<std::string, Stat>& stats() const { return stats_; }
 private:
  struct Rec { std::string name; double flops, bytes; hipEvent_t b, e; int launches; };   // launches between two events
  int get_event(hipEvent_t* e);
  bool on_ = false;
  std::vector<hipEvent_t> free_, all_;
  std::vector<Rec> open_;
  std::map<std::string, Stat> stats_;
};

// One field of a grouped MFT.  Fields are column-major (rows = y).
//   E_out = dx_in dy_in / (i lf) * A_y * E_in * B_x^T
// computed y-first:  T^T = A_y E  (stage 1, row-major out_y x in_x)  then
// X_out = B_x (T^T)^T = E_out^T  (stage 2), with X = E^T the column-major field
// read as a row-major matrix.  Both stages are NT GEMMs (K-contiguous operands).
// Element storage of a field operand.
enum FieldKind : int { FK_C128 = 0, FK_C64 = 1, FK_PLANES = 2 };
int terms_of(csx_precision p);   // BF16 terms per value (0 for FP64 / FP32)

// Zero pattern of a column-major ny x nx array: per column x, the rows [lo, hi]
// outside which it holds exact zeros (lo > hi: the column is zero).  Masks give it
// to the fields they multiply; the contractions skip the K slabs it leaves zero.
struct Support {
  int ny = 0, nx = 0;
  std::vector<int> lo, hi;
  static std::shared_ptr<const Support> of_mask(const double* colmajor, int ny, int nx);
  // rows of b within those of a (the fields a masked by b); null is dense
  static std::shared_ptr<const Support> intersect(const std::shared_ptr<const Support>& a,
                                                  const std::shared_ptr<const Support>& b);
  bool dense() const;
};

struct MftField {
  const void* in = nullptr;    // FieldKind in_kind: complex128, complex64 or BF16 planes
  void* out = nullptr;         // FieldKind out_kind
  int in_kind = FK_C128, out_kind = FK_C128;
  csx_grid in_x{}, in_y{}, out_x{}, out_y{};
  double lambda_f = 0;
  // optional mask applied to the output (E_out(r, c) *= mask(r - row0, c - col0))
  int mask_kind = MASK_NONE;
  const void* mask = nullptr;  // complex128 / double, column-major mrows x mcols
  int mrows = 0, mcols = 0, row0 = 0, col0 = 0;
  // optional zero patterns (sizes in_y x in_x and out_y x out_x; null: dense): the
  // input field's, and the output's when the epilogue mask zeroes it there
  std::shared_ptr<const Support> in_sup, out_sup;
  // the output feeds an MFT that reads it only within out_sup (its first contraction's
  // K ranges): zero blocks that range never reaches are not written at all
  bool out_read_sup = false;
  // point source applied by this MFT to its input: E_in * in_scale *
  // exp(2 pi i (tilt_x x + tilt_y y)) (cycles per unit length).  With the DFT
  // kernel's exp(-2 pi i xo x / lf) the phase ramp is the kernel of an output grid
  // shifted by tilt lf, so the sources of one input share its first contraction.
  double in_scale = 1.0, tilt_x = 0.0, tilt_y = 0.0;
  // optional addend stored like the output (out_kind, out_y x out_x per plane):
  // E_out = mask (s A_y E_in B_x^T + add)
  const void* add = nullptr;
};

class MftPlan {
 public:
  // Builds DFT kernel matrices (FP64 phases, stored at the working precision),
  // the stage-1 workspace and both problem arrays.  Field pointers are baked in.
  // With an arena every device buffer is a slice of it (no allocation) and no tile
  // variant is timed: each launch takes the tuner's persisted choice, else variant 0,
  // never split K.
  int build(const std::vector<MftField>& fields, csx_precision p, hipStream_t s, Arena* arena = nullptr);
  // Bytes an arena must hold for build() over `batch` fields of these grids (an upper
  // bound: every field with its own DFT kernels and first contraction).
  static size_t arena_bytes(int in_x, int in_y, int out_x, int out_y, int batch, csx_precision p);
  // Enqueues both GEMM stages (no allocation, no host sync: graph-capturable).
  int run(hipStream_t s, KernelTimer* timer = nullptr) const;
  size_t workspace_bytes() const { return dft_.bytes + tbuf_.bytes; }
  int variant(int stage) const { return stage == 0 ? var1_ : var2_; }
  int n_fields() const { return nf_; }
  double flops() const { return flops_; }   // algorithmic, 8 per complex MAC
  double flops_stage(int i) const { return i == 0 ? flops1_ : flops2_; }
  // algorithmic HBM bytes of each stage: operands read once, output written once
  double bytes_stage(int i) const { return i == 0 ? bytes1_ : bytes2_; }

 private:
  csx_precision prec_ = CSX_FP64;
  DevBuf dft_, tbuf_, probs1_, probs2_, specs_, ws1_, ws2_, kr_;
  int nf_ = 0, nt_ = 0, tiles1_ = 0, tiles2_ = 0;   // fields, distinct first contractions
  int var1_ = 0, var2_ = 0;   // tile variant of each stage (index into the precision's list)
  int ks1_ = 1, ks2_ = 1;     // K splits of each stage (split-K workspace ws1_/ws2_)
  long elems1_ = 0, elems2_ = 0;
  double flops_ = 0, flops1_ = 0, flops2_ = 0, bytes1_ = 0, bytes2_ = 0;
};

int launch_gemm(const GemmProb* probs_dev, int n, int tiles, csx_precision p, hipStream_t s, int variant = 0);
// JSON summary of the tuner's choices in this process (file, counts, per signature).
std::string tune_report();

}  // namespace csx
