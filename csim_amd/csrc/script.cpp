// csim's top-level driver and the tools it runs, on the batched engine:
//   the script loop              src/csim.cpp:54-108 (blocks run in order)
//   responseData files           src/data/responseData.cpp:175-273
//   makeHexCFpmResponse          src/tools/makeHexCFpmResponse.cpp:91-190
//   regionContrast / contrast    src/tools/regionContrast.cpp:145-205,
//   curve files                  src/tools/contrastCurve.cpp:68-158
//   differential evolution       src/optimization/differentialEvolution/
//                                differentialEvolutionOptimizer.cpp:66-455, deStrategy.cpp:23-158
//   optHexRespObjective          src/optimization/optObjective.cpp:24-107
// Every candidate population is evaluated as one batch: the response objective as one
// batched complex GEMM (csx_hex_resp_objective), a chain objective as one candidate
// batch through the element chain.  Random numbers come from Philox4x32-10: the
// reference's arma::randu / shuffle / rand() streams are not reproduced (trajectories
// differ, the update rules do not).
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <fstream>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <sys/types.h>

#include "engine.hpp"
#include "linopt.hpp"

namespace csx {

static std::string strip(const std::string& s) {
  size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}
// sscanf("%f, %f"): csim reads its pairs through float
static void float_pair(const std::string& s, double& a, double& b) {
  float x = 0, y = 0;
  std::sscanf(s.c_str(), "%f, %f", &x, &y);
  a = x;
  b = y;
}
static std::string join_path(const std::string& dir, const std::string& name) {
  if (dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}
// mkdir -p
static int make_dirs(const std::string& dir) {
  if (dir.empty()) return CSX_OK;
  std::string cur;
  std::stringstream ss(dir);
  std::string part;
  if (dir[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return fail(CSX_ERR_ARG, "cannot create directory " + cur);
  }
  return CSX_OK;
}

// --------------------------------------------------------------- FITS ---
// write_fits_array (src/lib/csim_fits.cpp:101-109) for BITPIX -64: data in FITS order
// (NAXIS1 fastest), big-endian, header and data padded to 2880-byte blocks.
static int write_fits_f64(const std::string& name, const double* data, const std::vector<long>& dims) {
  std::string path = name;
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
  card("NAXIS", std::to_string(dims.size()));
  size_t n = 1;
  for (size_t i = 0; i < dims.size(); i++) {
    card("NAXIS" + std::to_string(i + 1), std::to_string(dims[i]));
    n *= (size_t)dims[i];
  }
  card("EXTEND", "T");
  std::string end = "END";
  end.resize(80, ' ');
  h += end;
  h.resize((h.size() + 2879) / 2880 * 2880, ' ');
  std::vector<unsigned char> d(n * 8);
  for (size_t i = 0; i < n; i++) {
    unsigned char b[8];
    std::memcpy(b, &data[i], 8);
    for (int k = 0; k < 8; k++) d[i * 8 + k] = b[7 - k];
  }
  d.resize((d.size() + 2879) / 2880 * 2880, 0);
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(CSX_ERR_ARG, "cannot create FITS file " + path);
  f.write(h.data(), (std::streamsize)h.size());
  f.write(reinterpret_cast<const char*>(d.data()), (std::streamsize)d.size());
  if (!f) return fail(CSX_ERR_ARG, "short write to FITS file " + path);
  return CSX_OK;
}
// save_vec (csim_fits.cpp:567-579): NAXIS = 1
static int save_vector(const std::string& name, const std::vector<double>& v) {
  return write_fits_f64(name, v.data(), {(long)v.size()});
}
// save_mat of a column-major rows x cols matrix: NAXIS1 = cols, rows in order
static int save_matrix(const std::string& name, const double* colmajor, int rows, int cols) {
  std::vector<double> t((size_t)rows * cols);
  for (int r = 0; r < rows; r++)
    for (int c = 0; c < cols; c++) t[(size_t)r * cols + c] = colmajor[(size_t)c * rows + r];
  return write_fits_f64(name, t.data(), {(long)cols, (long)rows});
}

// -------------------------------------------------------- responseData ---
// M[s][p]: a rows x cols x slices cube (pixels x response columns x wavelengths),
// column-major per slice, stored for source s, polarisation p at index s * npol + p.
struct ResponseData {
  int rows = 0, cols = 0, slices = 0, npol = 1, nsrc = 1;
  double calib = 0;
  std::vector<double> lambdas;
  std::vector<double2> M;
  size_t cube() const { return (size_t)rows * cols * slices; }
};

// responseData::save (:175-189), save_cube (csim_fits.cpp:235-286: each slice
// transposed, NAXIS1 = columns), save_problem_params (:220-233).
static int response_save(const std::string& dir, const ResponseData& r) {
  CSX_TRY(make_dirs(dir));
  {
    std::FILE* f = std::fopen(join_path(dir, "responseParams.txt").c_str(), "w");
    if (!f) return fail(CSX_ERR_ARG, "cannot write responseParams.txt in " + dir);
    std::fprintf(f, "nRows = %d, nCols = %d, nSlices = %d, nPolarizations = %d, nSources = %d\n", r.rows, r.cols,
                 r.slices, r.npol, r.nsrc);
    std::fclose(f);
  }
  std::vector<double> re(r.cube()), im(r.cube());
  for (int s = 0; s < r.nsrc; s++)
    for (int p = 0; p < r.npol; p++) {
      const double2* M = r.M.data() + (size_t)(s * r.npol + p) * r.cube();
      for (int sl = 0; sl < r.slices; sl++)
        for (int row = 0; row < r.rows; row++)
          for (int c = 0; c < r.cols; c++) {
            const double2 v = M[(size_t)sl * r.rows * r.cols + (size_t)c * r.rows + row];
            const size_t o = (size_t)sl * r.rows * r.cols + (size_t)row * r.cols + c;
            re[o] = v.x;
            im[o] = v.y;
          }
      const std::string root = join_path(dir, "responseData_s" + std::to_string(s) + "_p" + std::to_string(p));
      const std::vector<long> dims = {(long)r.cols, (long)r.rows, (long)r.slices};
      CSX_TRY(write_fits_f64(root + "_re.fits", re.data(), dims));
      CSX_TRY(write_fits_f64(root + "_im.fits", im.data(), dims));
    }
  std::FILE* f = std::fopen(join_path(dir, "problemParams.txt").c_str(), "w");
  if (!f) return fail(CSX_ERR_ARG, "cannot write problemParams.txt in " + dir);
  std::fprintf(f, "calibMaxIntensity: %lf\n", r.calib);
  std::fprintf(f, "wavelengths: ");
  for (double l : r.lambdas) std::fprintf(f, "%g ", l);
  std::fprintf(f, "\n");
  std::fclose(f);
  return CSX_OK;
}

// responseData::load (:191-218) and read_problem_params (:235-273): the sizes line, the
// _re/_im cubes of every (source, polarisation), the calibration peak and every
// wavelength listed after "wavelengths:".
static int response_load(const std::string& dir, ResponseData& r) {
  std::FILE* f = std::fopen(join_path(dir, "responseParams.txt").c_str(), "r");
  if (!f) return fail(CSX_ERR_ARG, "cannot open responseParams.txt in " + dir);
  int n = std::fscanf(f, "nRows = %d, nCols = %d, nSlices = %d, nPolarizations = %d, nSources = %d", &r.rows,
                      &r.cols, &r.slices, &r.npol, &r.nsrc);
  std::fclose(f);
  if (n != 5 || r.rows <= 0 || r.cols <= 0 || r.slices <= 0 || r.npol <= 0 || r.nsrc <= 0)
    return fail(CSX_ERR_PARSE, "responseParams.txt: bad sizes line in " + dir);
  r.M.assign((size_t)r.nsrc * r.npol * r.cube(), make_double2(0, 0));
  for (int s = 0; s < r.nsrc; s++)
    for (int p = 0; p < r.npol; p++) {
      const std::string root = join_path(dir, "responseData_s" + std::to_string(s) + "_p" + std::to_string(p));
      std::vector<double> re, im;
      std::vector<long> dr, di;
      CSX_TRY(read_fits(root + "_re.fits", re, dr));
      CSX_TRY(read_fits(root + "_im.fits", im, di));
      const long sl = dr.size() > 2 ? dr[2] : 1;
      if (dr.size() < 2 || dr != di || dr[0] != r.cols || dr[1] != r.rows || sl != r.slices)
        return fail(CSX_ERR_PARSE, "response cube sizes differ from responseParams.txt: " + root);
      double2* M = r.M.data() + (size_t)(s * r.npol + p) * r.cube();
      for (int k = 0; k < r.slices; k++)
        for (int row = 0; row < r.rows; row++)
          for (int c = 0; c < r.cols; c++) {
            const size_t o = (size_t)k * r.rows * r.cols + (size_t)row * r.cols + c;
            M[(size_t)k * r.rows * r.cols + (size_t)c * r.rows + row] = make_double2(re[o], im[o]);
          }
    }
  std::ifstream pf(join_path(dir, "problemParams.txt"));
  if (!pf) return fail(CSX_ERR_ARG, "cannot open problemParams.txt in " + dir);
  std::string line;
  r.lambdas.clear();
  bool have_calib = false;
  while (std::getline(pf, line)) {
    if (line.rfind("calibMaxIntensity:", 0) == 0) {
      r.calib = std::atof(line.c_str() + 18);
      have_calib = true;
    } else if (line.rfind("wavelengths:", 0) == 0) {
      std::istringstream in(line.substr(12));
      double l;
      while (in >> l) r.lambdas.push_back(l);
    }
  }
  if (!have_calib || (int)r.lambdas.size() != r.slices)
    return fail(CSX_ERR_PARSE, "problemParams.txt: calibration peak or one wavelength per slice missing in " + dir);
  return CSX_OK;
}

// --------------------------------------------------------------- random ---
// Philox4x32-10 stream: uniform doubles in [0, 1) (53 bits), standard normals by
// Box-Muller, Fisher-Yates shuffles.
class DeRandom {
 public:
  explicit DeRandom(uint64_t seed) : seed_(seed) {}
  double uniform() {
    if (have_ == 0) {
      const uint32_t key[2] = {(uint32_t)seed_, (uint32_t)(seed_ >> 32)};
      const uint32_t ctr[4] = {(uint32_t)blk_, (uint32_t)(blk_ >> 32), 0x44450001u, 0u};
      uint32_t r[4];
      csx_philox4x32_10(ctr, key, r);
      blk_++;
      buf_[0] = (double)((((uint64_t)r[0] << 32) | r[1]) >> 11) * 0x1.0p-53;
      buf_[1] = (double)((((uint64_t)r[2] << 32) | r[3]) >> 11) * 0x1.0p-53;
      have_ = 2;
    }
    return buf_[2 - have_--];
  }
  double normal() {
    const double u0 = uniform(), u1 = uniform();
    return std::sqrt(-2.0 * std::log(1.0 - u0)) * std::cos(2.0 * M_PI * u1);
  }
  std::vector<int> permutation(int n) {
    std::vector<int> v(n);
    for (int i = 0; i < n; i++) v[i] = i;
    for (int i = n - 1; i > 0; i--) std::swap(v[i], v[(int)(uniform() * (i + 1))]);
    return v;
  }

 private:
  uint64_t seed_, blk_ = 0;
  double buf_[2] = {0, 0};
  int have_ = 0;
};

// ---------------------------------------------------------------- DE ---
// A population is D x P column-major (one candidate per column, as csim's arma::mat).
using Objective = std::function<int(const std::vector<double>& pop, int P, std::vector<double>& val)>;

enum DeStrategyKind { DE_NONE = 0, DE_RAND1, DE_LOCAL_TO_BEST1, DE_EITHER_OR };

struct DeResult {
  double best = 0;
  std::vector<double> best_vector;
  int iterations = 0, retries = 0;
  long evaluations = 0;
};

class DeOptimizer {
 public:
  int dim = 0;
  // defaults of differentialEvolutionOptimizer.hpp
  DeStrategyKind start = DE_NONE, base = DE_NONE, end = DE_NONE, reinit = DE_NONE;
  int population_size = 0;
  std::vector<double> pop_mean;
  bool have_mean = false;
  double init_spread = 0.1, reinit_compression = 1.0, reinit_spread = 0.0;
  double glo = 0, ghi = 0, ilo = 0, ihi = 0;
  bool randn_init = false;
  int max_iterations = 1000000;
  double stop_objective = 0.0, stop_spread = 0.0, reinit_changefrac = 0.0;
  int max_small_change = 5, max_retries = 5;
  double f_weight = 0.8, crossover = 1.0;
  double base_switch = 0.0, end_switch = 0.0;
  int output_interval = 100;
  std::string outdir;
  bool save_state = true;
  uint64_t seed = 20261206;
  std::vector<std::string> unknown;

  // differentialEvolutionOptimizer::set (:66-197)
  int set(const std::string& k, const std::string& arg) {
    const std::string a = strip(arg);
    auto strat = [&](DeStrategyKind& s) {
      if (a == "deRand1") s = DE_RAND1;
      else if (a == "deLocalToBest1") s = DE_LOCAL_TO_BEST1;
      else if (a == "eitherOr") s = DE_EITHER_OR;
      else return fail(CSX_ERR_PARSE, "differentialEvolutionOptimizer: unknown strategy " + a);
      return CSX_OK;
    };
    if (k == "differentialEvolutionOptimizer") return CSX_OK;
    if (k == "startStrategy") return strat(start);
    if (k == "baseStrategy") return strat(base);
    if (k == "endStrategy") return strat(end);
    if (k == "reinitStrategy") return strat(reinit);
    if (k == "populationSize") population_size = std::atoi(a.c_str());
    else if (k == "initPopulationMean") {
      if (a == "zero") {
        pop_mean.assign(dim, 0.0);
      } else {
        int r = 0, c = 0, s = 0;
        CSX_TRY(fits_matrix(a, pop_mean, r, c, s));
      }
      have_mean = true;
    } else if (k == "initSpread") init_spread = std::atof(a.c_str());
    else if (k == "reinitCompression") reinit_compression = std::atof(a.c_str());
    else if (k == "reinitSpread") reinit_spread = std::atof(a.c_str());
    else if (k == "globalBounds") float_pair(a, glo, ghi);
    else if (k == "initBounds") float_pair(a, ilo, ihi);
    else if (k == "initType") {
      if (a == "uniformRandom") randn_init = false;
      else if (a == "normalRandom") randn_init = true;
      else unknown.push_back("initType:" + a);
    } else if (k == "maxIterations") max_iterations = std::atoi(a.c_str());
    else if (k == "stopObjectiveValue") stop_objective = std::atof(a.c_str());
    else if (k == "stopPopulationSpread") stop_spread = std::atof(a.c_str());
    else if (k == "reinitChangefracThreshold") reinit_changefrac = std::atof(a.c_str());
    else if (k == "maxSmallChange") max_small_change = std::atoi(a.c_str());
    else if (k == "maxRetries") max_retries = std::atoi(a.c_str());
    else if (k == "fWeight") f_weight = std::atof(a.c_str());
    else if (k == "crossoverProb") crossover = std::atof(a.c_str());
    else if (k == "baseStrategySpreadSwitch") base_switch = std::atof(a.c_str());
    else if (k == "endStrategySpreadSwitch") end_switch = std::atof(a.c_str());
    else if (k == "outputInterval") output_interval = std::max(1, std::atoi(a.c_str()));
    else if (k == "outputDirectory") outdir = a;
    else if (k == "saveState") save_state = std::atoi(a.c_str()) != 0;
    else if (k == "seed") seed = std::strtoull(a.c_str(), nullptr, 10);   // additive: the Philox stream
    else unknown.push_back(k);
    return CSX_OK;
  }

  int optimize(const Objective& f, DeResult& res) {
    const int D = dim, P = population_size;
    if (D <= 0 || P < 4) return fail(CSX_ERR_ARG, "differential evolution: need a dimension and populationSize >= 4");
    DeStrategyKind strategy = start ? start : (base ? base : end);
    if (!strategy) return fail(CSX_ERR_PARSE, "differential evolution: no strategy given");
    if (!reinit) reinit = base ? base : strategy;
    if (!base) base = strategy;
    if (!end) end = strategy;
    if (have_mean && (int)pop_mean.size() != D)
      return fail(CSX_ERR_ARG, "initPopulationMean has " + std::to_string(pop_mean.size()) + " values, the problem " +
                                   std::to_string(D));
    if (!have_mean) pop_mean.assign(D, 0.0);
    if (!outdir.empty()) CSX_TRY(make_dirs(outdir));
    std::FILE* best_fid = std::fopen(join_path(outdir, "optBestVal.txt").c_str(), "w");
    std::FILE* hist_fid = std::fopen(join_path(outdir, "optValHistory.txt").c_str(), "w");
    if (!best_fid || !hist_fid) {
      if (best_fid) std::fclose(best_fid);
      if (hist_fid) std::fclose(hist_fid);
      return fail(CSX_ERR_ARG, "differential evolution: cannot write in " + outdir);
    }
    int rc = run(f, res, strategy, best_fid, hist_fid);
    std::fclose(best_fid);
    std::fclose(hist_fid);
    return rc;
  }

 private:
  DeRandom rng_{20261206};
  std::vector<double> pop_, val_, cand_, cval_, mutated_, original_, best_vec_, gbest_vec_;
  double best_ = 1e15, gbest_ = 1e15, spread_mean_ = 0, spread_max_ = 0;
  int iteration_ = 0;
  long evals_ = 0;

  double& at(std::vector<double>& m, int r, int c) { return m[(size_t)c * dim + r]; }

  int evaluate(const Objective& f, const std::vector<double>& pop, std::vector<double>& val) {
    val.assign(population_size, 0.0);
    CSX_TRY(f(pop, population_size, val));
    evals_ += population_size;
    for (double v : val)
      if (!std::isfinite(v)) return fail(CSX_ERR_NUMERIC, "differential evolution: objective returned a non-finite value");
    return CSX_OK;
  }
  // init_population (:332-358): origin + spread initSpread (U - 1/2 | N) (initUpper - initLower)
  void init_population(const std::vector<double>& origin, bool randn, double spread) {
    pop_.resize((size_t)dim * population_size);
    for (int c = 0; c < population_size; c++)
      for (int r = 0; r < dim; r++) {
        const double z = randn ? rng_.normal() : rng_.uniform() - 0.5;
        at(pop_, r, c) = origin[r] + spread * init_spread * z * (ihi - ilo);
      }
  }
  // makePopulationPermutations (deStrategy.cpp:23-57): slice i takes the members of one
  // shuffle, and each next slice's order is the previous rotated by si2[i] + 1.
  std::vector<std::vector<int>> permutations(int n) {
    const int P = population_size;
    std::vector<int> si2 = rng_.permutation(n), si1 = rng_.permutation(P);
    std::vector<std::vector<int>> out;
    for (int i = 0; i < n; i++) {
      out.push_back(si1);
      if (i < n - 1) {
        const int k = si2[i] + 1;
        std::vector<int> nx(P);
        for (int j = 0; j < P; j++) nx[j] = si1[((j - k) % P + P) % P];
        si1.swap(nx);
      }
    }
    return out;
  }
  // deRand1 / deLocalToBest1 / eitherOr::execute (deStrategy.cpp:63-158)
  void mutate(DeStrategyKind s) {
    const int P = population_size;
    mutated_.resize(pop_.size());
    original_.resize(pop_.size());
    const double F = f_weight;
    if (s == DE_LOCAL_TO_BEST1) {
      auto p = permutations(2);
      for (int c = 0; c < P; c++)
        for (int r = 0; r < dim; r++) {
          const double x = at(pop_, r, c);
          at(mutated_, r, c) = x + F * (best_vec_[r] - x) + F * (at(pop_, r, p[0][c]) - at(pop_, r, p[1][c]));
          at(original_, r, c) = x;
        }
      return;
    }
    auto p = permutations(3);
    const bool second = s == DE_EITHER_OR && rng_.uniform() >= 0.5;
    for (int c = 0; c < P; c++)
      for (int r = 0; r < dim; r++) {
        const double a = at(pop_, r, p[0][c]), b = at(pop_, r, p[1][c]), o = at(pop_, r, p[2][c]);
        at(mutated_, r, c) = second ? o + 0.5 * (F + 1.0) * (a + b - 2 * o) : o + F * (a - b);
        at(original_, r, c) = o;
      }
  }
  // do_crossover (:360-369)
  void crossover_step() {
    cand_ = mutated_;
    if (crossover < 1.0)
      for (size_t i = 0; i < cand_.size(); i++)
        if (!(rng_.uniform() < crossover)) cand_[i] = pop_[i];
  }
  // enforce_boundaries (:395-434): bounce back inside [mean + initBounds] clipped to the
  // global bounds, towards the member the mutation started from
  void enforce_boundaries() {
    for (int c = 0; c < population_size; c++)
      for (int r = 0; r < dim; r++) {
        const double up = std::min(pop_mean[r] + ihi, ghi), lo = std::max(pop_mean[r] + ilo, glo);
        double& x = at(cand_, r, c);
        if (x < lo) x = lo + rng_.uniform() * (at(original_, r, c) - lo);
        else if (x > up) x = up + rng_.uniform() * (at(original_, r, c) - up);
      }
  }
  // compute_population_statistics (:458-471)
  void statistics() {
    const int P = population_size;
    pop_mean.assign(dim, 0.0);
    for (int c = 0; c < P; c++)
      for (int r = 0; r < dim; r++) pop_mean[r] += at(pop_, r, c);
    for (int r = 0; r < dim; r++) pop_mean[r] /= P;
    double sum = 0, mx = 0;
    for (int c = 0; c < P; c++) {
      double d2 = 0;
      for (int r = 0; r < dim; r++) {
        const double d = at(pop_, r, c) - pop_mean[r];
        d2 += d * d;
      }
      sum += std::sqrt(d2);
      mx = std::max(mx, std::sqrt(d2));
    }
    spread_mean_ = sum / P;
    spread_max_ = mx;
  }
  // find_best_member (:371-393): the generation's best, and the global best written out
  int find_best(std::FILE* best_fid) {
    int bi = 0;
    for (int c = 1; c < population_size; c++)
      if (val_[c] < val_[bi]) bi = c;
    if (val_[bi] < best_) {
      best_ = val_[bi];
      best_vec_.assign(pop_.begin() + (size_t)bi * dim, pop_.begin() + (size_t)(bi + 1) * dim);
      if (best_ < gbest_) {
        gbest_ = best_;
        gbest_vec_ = best_vec_;
        CSX_TRY(save_vector(join_path(outdir, "bestOptVector.fits"), gbest_vec_));
        std::fprintf(best_fid, "%d %g\n", iteration_, gbest_);
        std::fflush(best_fid);
      }
    }
    return CSX_OK;
  }

  // optimize (:216-330)
  int run(const Objective& f, DeResult& res, DeStrategyKind strategy, std::FILE* best_fid, std::FILE* hist_fid) {
    rng_ = DeRandom(seed);
    best_ = gbest_ = 1e15;
    iteration_ = 0;
    evals_ = 0;
    int retries = 0, small = 0;
    double prev_best = 0;
    bool sel_base = false, sel_end = false, do_reinit = false;
    init_population(pop_mean, randn_init, 1.0);
    CSX_TRY(evaluate(f, pop_, val_));
    CSX_TRY(find_best(best_fid));
    statistics();
    if (save_state) CSX_TRY(save_matrix(join_path(outdir, "initial_population.fits"), pop_.data(), dim, population_size));
    auto stop = [&]() { return iteration_ > max_iterations || best_ < stop_objective || spread_max_ < stop_spread; };
    while ((!stop() && retries <= max_retries) || iteration_ == 0) {
      if (small >= max_small_change || do_reinit || spread_mean_ < reinit_spread) {
        best_ = 1e15;
        strategy = reinit;
        init_population(gbest_vec_, false, reinit_compression);
        CSX_TRY(evaluate(f, pop_, val_));
        CSX_TRY(find_best(best_fid));
        small = 0;
        do_reinit = false;
        retries++;
      }
      mutate(strategy);
      crossover_step();
      enforce_boundaries();
      CSX_TRY(evaluate(f, cand_, cval_));      // update_population (:437-455)
      for (int c = 0; c < population_size; c++)
        if (cval_[c] < val_[c]) {
          std::copy(cand_.begin() + (size_t)c * dim, cand_.begin() + (size_t)(c + 1) * dim, pop_.begin() + (size_t)c * dim);
          val_[c] = cval_[c];
        }
      statistics();
      CSX_TRY(find_best(best_fid));
      if (iteration_ % output_interval == 0) {
        if (spread_mean_ < base_switch && !sel_base) { strategy = base; sel_base = true; do_reinit = true; }
        if (spread_mean_ < end_switch && !sel_end) { strategy = end; sel_end = true; do_reinit = true; }
        if (std::fabs((best_ - prev_best) / best_) < reinit_changefrac) small++;
        else small = 0;
        prev_best = best_;
        std::fprintf(hist_fid, "%d %g\n", iteration_, best_);
        std::fflush(hist_fid);
        if (save_state)
          CSX_TRY(save_matrix(join_path(outdir, "population_it_" + std::to_string(iteration_) + ".fits"), pop_.data(),
                              dim, population_size));
      }
      iteration_++;
    }
    res.best = gbest_;
    res.best_vector = gbest_vec_;
    res.iterations = iteration_;
    res.retries = retries;
    res.evaluations = evals_;
    return CSX_OK;
  }
};

// --------------------------------------------------------- objectives ---
// optHexRespObjective (optObjective.cpp:24-107) over a loaded responseData: every
// (source, polarisation, wavelength) slice as one problem of one batched GEMM.
struct RespObjective {
  csx_ctx* ctx = nullptr;
  ResponseData r;
  int use_max = 0;
  double phase_sign = 1.0;
  csx_precision prec = CSX_FP64;
  DevBuf dM, dX, dobj;
  std::vector<double> slice_lambdas;

  int set(const std::string& k, const std::string& arg) {
    const std::string a = strip(arg);
    if (k == "optHexRespObjective") return CSX_OK;
    if (k == "responseDirectory") return response_load(a, r);
    if (k == "objectiveType") {
      if (a == "mean") use_max = 0;
      else if (a == "max") use_max = 1;
      else return fail(CSX_ERR_PARSE, "optHexRespObjective objectiveType must be mean or max");
      return CSX_OK;
    }
    if (k == "phaseSign") { phase_sign = std::atof(a.c_str()); return CSX_OK; }
    if (k == "precision") {     // additive: the GEMM precision
      if (a == "fp64") prec = CSX_FP64;
      else if (a == "fp32") prec = CSX_FP32;
      else if (a == "bf16x3") prec = CSX_BF16X3;
      else if (a == "bf16x6") prec = CSX_BF16X6;
      else return fail(CSX_ERR_PARSE, "optHexRespObjective precision: fp64, fp32, bf16x3 or bf16x6");
      return CSX_OK;
    }
    std::fprintf(stderr, "!!! optHexRespObjective bad set field name: %s\n", k.c_str());
    return CSX_OK;
  }
  int dim() const { return r.cols - 1; }   // the last response column is not optimised
  int upload() {
    if (r.M.empty()) return fail(CSX_ERR_ARG, "optHexRespObjective: no responseDirectory");
    CSX_HIP(hipSetDevice(ctx->device));
    CSX_TRY(dM.reserve(sizeof(double2) * r.M.size()));
    CSX_HIP(hipMemcpy(dM.p, r.M.data(), sizeof(double2) * r.M.size(), hipMemcpyHostToDevice));
    slice_lambdas.clear();
    for (int sp = 0; sp < r.nsrc * r.npol; sp++)
      for (int k = 0; k < r.slices; k++) slice_lambdas.push_back(r.lambdas[k]);
    return CSX_OK;
  }
  int operator()(const std::vector<double>& pop, int P, std::vector<double>& val) {
    CSX_HIP(hipSetDevice(ctx->device));
    CSX_TRY(dX.reserve(sizeof(double) * pop.size()));
    CSX_TRY(dobj.reserve(sizeof(double) * P));
    CSX_HIP(hipMemcpy(dX.p, pop.data(), sizeof(double) * pop.size(), hipMemcpyHostToDevice));
    CSX_TRY(csx_hex_resp_objective(ctx, dM.p, r.rows, r.cols, r.nsrc * r.npol * r.slices, slice_lambdas.data(),
                                   phase_sign, r.calib, dX.as<double>(), dim(), P, use_max, dobj.as<double>(), prec));
    val.resize(P);
    CSX_HIP(hipMemcpy(val.data(), dobj.p, sizeof(double) * P, hipMemcpyDeviceToHost));
    return CSX_OK;
  }
};

// The full-propagation objective (the batched engine's own): every candidate's sag
// vector through the chain as one candidate batch, the objective its region contrast.
struct ChainObjective {
  Chain* chain = nullptr;
  std::string component;
  csx_region region{};
  bool have_region = false;
  int operator()(const std::vector<double>& pop, int P, std::vector<double>& val) {
    const int D = (int)(pop.size() / P);
    CSX_TRY(chain->set_candidates(component, pop.data(), D, P));
    val.resize(P);
    return chain->contrast(have_region ? &region : nullptr, val.data(), nullptr);
  }
};

static void region_keys(const Block& b, csx_region& r) {
  r = csx_region{1.0, 2.0, 0.0, 30.0, -1e16, 1e16, -1.0};   // regionContrast.hpp defaults
  for (const Cmd& c : b.cmds) {
    if (c.key == "referenceLambda") r.reference_lambda = std::atof(c.arg.c_str());
    else if (c.key == "xlimit") float_pair(c.arg, r.xlim1, r.xlim2);
    else if (c.key == "anglesInDegrees") float_pair(c.arg, r.angle1, r.angle2);
    else if (c.key == "radii") float_pair(c.arg, r.radius1, r.radius2);
  }
}
static std::string block_key(const Block& b, const std::string& k, const std::string& dflt = "") {
  for (const Cmd& c : b.cmds)
    if (c.key == k) return strip(c.arg);
  return dflt;
}

// A DE block: the objective sub-block first (it fixes the dimension), then the
// optimiser's own keys (differentialEvolutionOptimizer::init, :27-55).
static int run_de_block(csx_ctx* ctx, Chain* chain, const Block& b, std::string& log, DeResult& res) {
  std::vector<Block> sub = find_blocks(b);
  DeOptimizer de;
  std::unique_ptr<RespObjective> ro;
  std::unique_ptr<ChainObjective> co;
  const Block* own = nullptr;
  for (const Block& sb : sub) {
    if (sb.cmds.empty()) continue;
    const std::string& h = sb.cmds[0].key;
    if (h == "optHexRespObjective") {
      if (!ctx) return fail(CSX_ERR_ARG, "optHexRespObjective needs a device context");
      ro.reset(new RespObjective());
      ro->ctx = ctx;
      for (const Cmd& c : sb.cmds) CSX_TRY(ro->set(c.key, c.arg));
      CSX_TRY(ro->upload());
      de.dim = ro->dim();
    } else if (h == "optChainObjective") {    // additive: DE over full propagations
      if (!chain) return fail(CSX_ERR_ARG, "optChainObjective needs the script's telescope, efield and coronagraph");
      co.reset(new ChainObjective());
      co->chain = chain;
      co->component = block_key(sb, "component");
      for (const Block& rb : find_blocks(sb))
        if (!rb.cmds.empty() && rb.cmds[0].key == "regionContrast") { region_keys(rb, co->region); co->have_region = true; }
      std::vector<double> sags;
      CSX_TRY(chain->get_opt(co->component, "sags", sags));
      de.dim = (int)sags.size();
    } else {
      own = &sb;
    }
  }
  if (!ro && !co) return fail(CSX_ERR_PARSE, "differentialEvolutionOptimizer needs an optHexRespObjective or optChainObjective block");
  if (own)
    for (const Cmd& c : own->cmds) CSX_TRY(de.set(c.key, c.arg));
  for (const std::string& u : de.unknown) log += "!!! differentialEvolutionOptimizer bad set field name: " + u + "\n";
  Objective f = ro ? Objective([&](const std::vector<double>& p, int P, std::vector<double>& v) { return (*ro)(p, P, v); })
                   : Objective([&](const std::vector<double>& p, int P, std::vector<double>& v) { return (*co)(p, P, v); });
  auto t0 = std::chrono::steady_clock::now();
  CSX_TRY(de.optimize(f, res));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  char line[256];
  std::snprintf(line, sizeof line, "iteration = %d, bestVal = %g, retryCount = %d, evaluations = %ld, %.3f s\n",
                res.iterations, res.best, res.retries, res.evaluations, secs);
  log += line;
  return CSX_OK;
}

// makeHexCFpmResponse::compute_response (:91-190): the response columns as one
// candidate batch (Chain::hex_response), written as a responseData directory.
static int run_response_block(Chain& ch, const Block& b, std::string& log) {
  std::string comp, data = "sags", outdir;
  csx_region reg{};
  bool have_region = false;
  for (const Block& sb : find_blocks(b)) {
    if (sb.cmds.empty()) continue;
    if (sb.cmds[0].key == "regionContrast") {
      region_keys(sb, reg);
      have_region = true;
    } else {
      for (const Cmd& c : sb.cmds) {
        if (c.key == "dataToOptimize") {
          const size_t p = c.arg.find(',');
          comp = strip(c.arg.substr(0, p));
          if (p != std::string::npos) data = strip(c.arg.substr(p + 1));
        } else if (c.key == "outputDirectory") outdir = strip(c.arg);
      }
    }
  }
  if (comp.empty() || outdir.empty())
    return fail(CSX_ERR_PARSE, "makeHexCFpmResponse needs dataToOptimize and outputDirectory");
  if (data != "sags") return fail(CSX_ERR_UNSUPPORTED, "makeHexCFpmResponse: only sags responses (" + data + ")");
  int npix = 0, ncol = 0, nsl = 0;
  double calib = 0;
  const csx_region* rp = have_region ? &reg : nullptr;
  CSX_TRY(ch.hex_response(comp, rp, nullptr, &npix, &ncol, &nsl, &calib));
  DevBuf dM;
  CSX_TRY(dM.reserve(sizeof(double2) * (size_t)npix * ncol * nsl));
  CSX_TRY(ch.hex_response(comp, rp, dM.as<double2>(), &npix, &ncol, &nsl, &calib));
  ResponseData r;
  r.rows = npix; r.cols = ncol; r.slices = ch.nlam(); r.npol = 1; r.nsrc = ch.nsrc();
  r.calib = calib;
  r.lambdas = ch.efield().lambdas;
  r.M.resize((size_t)npix * ncol * nsl);
  CSX_HIP(hipMemcpy(r.M.data(), dM.p, sizeof(double2) * r.M.size(), hipMemcpyDeviceToHost));
  CSX_TRY(response_save(outdir, r));
  log += "response: " + std::to_string(npix) + " pixels x " + std::to_string(ncol) + " columns x " +
         std::to_string(nsl) + " slices, calibMaxIntensity " + std::to_string(calib) + " -> " + outdir + "\n";
  return CSX_OK;
}

// regionContrast::compute_contrast (:145-205): calibration_PSF.fits, normalized_PSF.fits
// and the region file (angles, radii, x limits, contrast).
static int run_region_block(Chain& ch, const Block& b, const std::string& workdir, std::string& log) {
  csx_region reg{};
  region_keys(b, reg);
  double c = 0;
  csx_report rep{};
  CSX_TRY(ch.contrast(&reg, &c, &rep));
  std::vector<double> cal, img;
  int rows = 0, cols = 0;
  CSX_TRY(ch.calibration_intensity(cal, &rows, &cols));
  CSX_TRY(save_matrix(join_path(workdir, "calibration_PSF.fits"), cal.data(), rows, cols));
  csx_chain_info inf{};
  CSX_TRY(ch.info(&inf));
  img.resize((size_t)inf.out_rows * inf.out_cols);
  CSX_TRY(ch.get_normalized_image(0, img.data()));
  CSX_TRY(save_matrix(join_path(workdir, "normalized_PSF.fits"), img.data(), inf.out_rows, inf.out_cols));
  char line[128];
  std::snprintf(line, sizeof line, "region contrast: %.10e\n", c);
  log += line;
  const std::string fn = block_key(b, "filename");
  if (!fn.empty()) {
    std::FILE* f = std::fopen(join_path(workdir, fn).c_str(), "w");
    if (!f) return fail(CSX_ERR_ARG, "regionContrast: cannot write " + fn);
    std::fprintf(f, "%f, %f\n", reg.angle1, reg.angle2);
    std::fprintf(f, "%f, %f\n", reg.radius1, reg.radius2);
    std::fprintf(f, "%f, %f\n", reg.xlim1, reg.xlim2);
    std::fprintf(f, "%e\n", c);
    std::fclose(f);
  }
  return CSX_OK;
}

// contrastCurve::make_contrast_curve (:68-158): the PSF files and the curve rows
// written "%f %e" from the second row on.
static int run_curve_block(Chain& ch, const Block& b, const std::string& workdir, std::string& log) {
  const std::string fn = block_key(b, "filename");
  if (fn.empty()) return fail(CSX_ERR_PARSE, "makeContrastCurve needs a filename");
  int nrow = 0;
  CSX_TRY(ch.contrast_curve(0, nullptr, nullptr, 0, &nrow));
  std::vector<double> pos(nrow), avg(nrow);
  CSX_TRY(ch.contrast_curve(0, pos.data(), avg.data(), nrow, &nrow));
  std::vector<double> cal;
  int rows = 0, cols = 0;
  CSX_TRY(ch.calibration_intensity(cal, &rows, &cols));
  CSX_TRY(save_matrix(join_path(workdir, "calibration_PSF.fits"), cal.data(), rows, cols));
  csx_chain_info inf{};
  CSX_TRY(ch.info(&inf));
  std::vector<double> img((size_t)inf.out_rows * inf.out_cols);
  CSX_TRY(ch.get_normalized_image(0, img.data()));
  CSX_TRY(save_matrix(join_path(workdir, "normalized_PSF.fits"), img.data(), inf.out_rows, inf.out_cols));
  std::FILE* f = std::fopen(join_path(workdir, fn).c_str(), "w");
  if (!f) return fail(CSX_ERR_ARG, "makeContrastCurve: cannot write " + fn);
  for (int i = 0; i < nrow; i++) std::fprintf(f, "%f %e\n", pos[i], avg[i]);
  std::fclose(f);
  log += "contrast curve: " + std::to_string(nrow) + " rows -> " + fn + "\n";
  return CSX_OK;
}

// ------------------------------------------------------- linear optimiser ---
// linearOptimizer (src/optimization/linearOptimizer.cpp:72-505) on the batched chain:
// the Jacobian's n + 1 propagations are one candidate batch (x, then x + dx e_i), the
// normal equations (J^T J + mu^2 I) d = J^T e are formed on the device
// (linear_normal_equations) and solved on the host by Cholesky for each mu of
// compute_step's scan (1e7 to 3e8 by 1e7), whose contrasts are one more candidate batch
// read in csim's order: the last mu that lowered the contrast is kept.
struct LinearOptimizer {
  std::string comp, data = "sags", outdir;
  csx_region region{};
  bool have_region = false;
  double jacobian_dx = 1e-8, glo = 0, ghi = 0, stop_contrast = 1e-9, rel_tol = 1e-5, delta_tol = 1e-9;
  int max_iterations = 20;
  bool random_init = false, have_init = false;
  double init_value = 0;
  double mu0 = 1e7, mu1 = 3e8, mu_step = 1e7;   // compute_step's scan (:375)
  std::vector<std::string> unknown;

  // linearOptimizer::set (:72-158)
  void set(const std::string& k, const std::string& arg) {
    const std::string a = strip(arg);
    if (k == "linearOptimizer") return;
    if (k == "dataToOptimize") {
      const size_t p = a.find(',');
      comp = strip(a.substr(0, p));
      if (p != std::string::npos) data = strip(a.substr(p + 1));
    } else if (k == "initOptVec") {
      have_init = true;
      random_init = a == "random";
      float v = 0;
      if (!random_init) std::sscanf(a.c_str(), "%f", &v);
      init_value = v;
    } else if (k == "contrastRelTolerance") rel_tol = std::atof(a.c_str());
    else if (k == "jacobianDx") jacobian_dx = std::atof(a.c_str());
    else if (k == "globalBounds") float_pair(a, glo, ghi);
    else if (k == "maxIterations") max_iterations = std::atoi(a.c_str());
    else if (k == "stopContrast") stop_contrast = std::atof(a.c_str());
    else if (k == "deltaOptVecTolerance") delta_tol = std::atof(a.c_str());
    else if (k == "outputDirectory") outdir = a;
    else if (k == "muScan") {                     // additive: start, stop, step of the mu scan
      std::sscanf(a.c_str(), "%lf, %lf, %lf", &mu0, &mu1, &mu_step);
    } else if (k == "regularizationMu" || k == "stepSizeScale" || k == "outputInterval" || k == "saveState") {
      // read by csim but without effect on the step it takes (the mu scan sets mu)
    } else unknown.push_back(k);
  }

  // contrast of every row of X (ncand x n) as one candidate batch
  int contrasts(Chain& ch, const std::vector<double>& X, int ncand, int n, std::vector<double>& c) {
    CSX_TRY(ch.set_candidates(comp, X.data(), n, ncand));
    c.assign(ncand, 0.0);
    return ch.contrast(have_region ? &region : nullptr, c.data(), nullptr);
  }

  // (A + mu^2 I) d = b by Cholesky (A symmetric positive semi-definite, mu > 0)
  static bool solve(const std::vector<double>& A, const std::vector<double>& b, double mu, int n,
                    std::vector<double>& d) {
    std::vector<double> L(A);
    for (int i = 0; i < n; i++) L[(size_t)i * n + i] += mu * mu;
    for (int j = 0; j < n; j++) {
      double s = L[(size_t)j * n + j];
      for (int k = 0; k < j; k++) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
      if (!(s > 0)) return false;
      const double ljj = std::sqrt(s);
      L[(size_t)j * n + j] = ljj;
      for (int i = j + 1; i < n; i++) {
        double t = L[(size_t)i * n + j];
        for (int k = 0; k < j; k++) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
        L[(size_t)i * n + j] = t / ljj;
      }
    }
    d.assign(n, 0.0);
    for (int i = 0; i < n; i++) {            // L y = b
      double t = b[i];
      for (int k = 0; k < i; k++) t -= L[(size_t)i * n + k] * d[k];
      d[i] = t / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {       // L^T d = y
      double t = d[i];
      for (int k = i + 1; k < n; k++) t -= L[(size_t)k * n + i] * d[k];
      d[i] = t / L[(size_t)i * n + i];
    }
    return true;
  }

  // optimize (:192-300) with the stopping rule of stopping_criteria (:302-313)
  int run(Chain& ch, std::string& log, double* best_out, std::vector<double>& best_x) {
    if (comp.empty()) return fail(CSX_ERR_PARSE, "linearOptimizer needs dataToOptimize");
    if (data != "sags") return fail(CSX_ERR_UNSUPPORTED, "linearOptimizer: only hex-FPM sags (" + data + ")");
    std::vector<double> x;
    CSX_TRY(ch.get_opt(comp, data, x));
    const int n = (int)x.size();
    if (n == 0) return fail(CSX_ERR_ARG, "linearOptimizer: empty optimisation vector");
    if (have_init) {
      DeRandom rng(20261208);
      for (double& v : x) v = random_init ? ghi * (rng.uniform() - 0.5) : init_value;
    }
    if (!outdir.empty()) CSX_TRY(make_dirs(outdir));
    std::FILE* best_fid = std::fopen(join_path(outdir, "optBestVal.txt").c_str(), "w");
    std::FILE* hist_fid = std::fopen(join_path(outdir, "optValHistory.txt").c_str(), "w");
    if (!best_fid || !hist_fid) {
      if (best_fid) std::fclose(best_fid);
      if (hist_fid) std::fclose(hist_fid);
      return fail(CSX_ERR_ARG, "linearOptimizer: cannot write in " + outdir);
    }
    int rc = iterate(ch, x, n, log, best_fid, hist_fid, best_out, best_x);
    std::fclose(best_fid);
    std::fclose(hist_fid);
    return rc;
  }

 private:
  int iterate(Chain& ch, std::vector<double>& x, int n, std::string& log, std::FILE* best_fid, std::FILE* hist_fid,
              double* best_out, std::vector<double>& best_x) {
    auto t0 = std::chrono::steady_clock::now();
    auto secs = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    double best = 1e15, last = 0.0, delta_norm = 1e300;
    DevBuf dF, dA, db;
    CSX_TRY(dA.reserve(sizeof(double) * (size_t)n * n));
    CSX_TRY(db.reserve(sizeof(double) * n));
    std::vector<double> X((size_t)(n + 1) * n), A((size_t)n * n), b(n), c;
    for (int it = 0;; it++) {
      // the Jacobian batch: x, then x + dx e_i
      for (int r = 0; r <= n; r++) {
        std::copy(x.begin(), x.end(), X.begin() + (size_t)r * n);
        if (r > 0) X[(size_t)r * n + r - 1] += jacobian_dx;
      }
      CSX_TRY(contrasts(ch, X, n + 1, n, c));
      const double contrast = c[0];
      int npix = 0, nc = 0, nsl = 0;
      CSX_TRY(ch.region_fields(have_region ? &region : nullptr, nullptr, &npix, &nc, &nsl));
      CSX_TRY(dF.reserve(sizeof(double2) * (size_t)nsl * nc * npix));
      CSX_TRY(ch.region_fields(have_region ? &region : nullptr, dF.as<double2>(), &npix, &nc, &nsl));
      std::fprintf(hist_fid, "%g %g\n", secs(), contrast);
      std::fflush(hist_fid);
      if (contrast < best) {
        best = contrast;
        best_x = x;
        std::fprintf(best_fid, "%g %g\n", secs(), contrast);
        std::fflush(best_fid);
        CSX_TRY(save_vector(join_path(outdir, "bestOptVector.fits"), best_x));
      }
      const bool stop = std::fabs(last - contrast) < rel_tol || contrast < stop_contrast || delta_norm < delta_tol ||
                        it > max_iterations;
      last = contrast;
      char line[160];
      std::snprintf(line, sizeof line, "evaluation %d, %.3f s, contrast = %.6e\n", it, secs(), contrast);
      log += line;
      if (stop) break;
      // compute_step: normal equations, then the mu scan as one batch
      CSX_TRY(linear_normal_equations(dF.p, nsl, n, npix, jacobian_dx, dA.as<double>(), db.as<double>(), ch.stream()));
      CSX_HIP(hipMemcpyAsync(A.data(), dA.p, sizeof(double) * A.size(), hipMemcpyDeviceToHost, ch.stream()));
      CSX_HIP(hipMemcpyAsync(b.data(), db.p, sizeof(double) * n, hipMemcpyDeviceToHost, ch.stream()));
      CSX_HIP(hipStreamSynchronize(ch.stream()));
      std::vector<double> mus, D;
      if (!(mu_step > 0) || !(mu0 > 0)) return fail(CSX_ERR_ARG, "linearOptimizer: muScan needs positive values");
      for (int k = 0; mu0 + k * mu_step <= mu1 + 0.5 * mu_step; k++) mus.push_back(mu0 + k * mu_step);
      std::vector<double> d;
      for (double mu : mus) {
        if (!solve(A, b, mu, n, d)) return fail(CSX_ERR_NUMERIC, "linearOptimizer: normal equations not positive");
        D.insert(D.end(), d.begin(), d.end());
      }
      const int nm = (int)mus.size();
      std::vector<double> Xm((size_t)nm * n);
      for (int k = 0; k < nm; k++)
        for (int i = 0; i < n; i++) Xm[(size_t)k * n + i] = x[i] + D[(size_t)k * n + i];
      CSX_TRY(contrasts(ch, Xm, nm, n, c));
      double lastc = 1e6;
      int kb = 0;
      for (int k = 0; k < nm; k++) {
        if (c[k] < lastc) { lastc = c[k]; kb = k; }
        else break;
      }
      double dn = 0;
      for (int i = 0; i < n; i++) {
        const double di = D[(size_t)kb * n + i];
        dn += di * di;
        x[i] = std::min(ghi, std::max(glo, x[i] + di));
      }
      delta_norm = std::sqrt(dn);
      CSX_TRY(save_vector(join_path(outdir, "deltaOptVec_" + std::to_string(it) + ".fits"),
                          std::vector<double>(D.begin() + (size_t)kb * n, D.begin() + (size_t)(kb + 1) * n)));
      std::snprintf(line, sizeof line, "  selected mu = %g, step norm %.3e\n", mus[kb], delta_norm);
      log += line;
    }
    *best_out = best;
    // the chain returns to the single sag set, holding the best vector
    CSX_TRY(ch.set_candidates(comp, nullptr, 0, 0));
    return ch.set_opt(comp, data, best_x);
  }
};

static int run_linear_block(Chain& ch, const Block& b, std::string& log) {
  LinearOptimizer lo;
  for (const Block& sb : find_blocks(b)) {
    if (sb.cmds.empty()) continue;
    if (sb.cmds[0].key == "regionContrast") {
      region_keys(sb, lo.region);
      lo.have_region = true;
    } else {
      for (const Cmd& c : sb.cmds) lo.set(c.key, c.arg);
    }
  }
  for (const std::string& u : lo.unknown) log += "!!! linearOptimizer bad set field name: " + u + "\n";
  double best = 0;
  std::vector<double> x;
  CSX_TRY(lo.run(ch, log, &best, x));
  char line[96];
  std::snprintf(line, sizeof line, "linear optimiser: best contrast %.6e\n", best);
  log += line;
  return CSX_OK;
}

// The script loop of src/csim.cpp:54-108.  The telescope, efield, coronagraph and
// regionContrast blocks define the chain (parsed once, as csim's globals); the tool
// blocks run in script order.  Relative output paths are taken as given (csim's
// working directory); workdir is where the PSF files and region/curve files go.
static int run_script(csx_ctx* ctx, const std::string& text, const std::string& workdir, std::string& log) {
  std::vector<Block> blocks = find_blocks(parse_script_text(text));
  bool has_chain = false;
  for (const Block& b : blocks)
    if (!b.cmds.empty() && b.cmds[0].key == "coronagraph") has_chain = true;
  std::unique_ptr<Chain> ch;
  if (has_chain) {
    ch.reset(new Chain(ctx));
    CSX_TRY(ch->load_script(text));
  }
  auto need_chain = [&](const std::string& what) {
    return ch ? CSX_OK : fail(CSX_ERR_STATE, what + " needs a telescope, efield and coronagraph");
  };
  for (const Block& b : blocks) {
    if (b.cmds.empty()) continue;
    const std::string& t = b.cmds[0].key;
    if (t == "telescope" || t == "efield" || t == "coronagraph") continue;
    if (t == "break") break;
    if (t == "execute") {
      CSX_TRY(need_chain(t));
      auto t0 = std::chrono::steady_clock::now();
      CSX_TRY(ch->propagate(false));
      char line[96];
      std::snprintf(line, sizeof line, "csim execution time: %.6f seconds\n",
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
      log += line;
    } else if (t == "regionContrast") {
      CSX_TRY(need_chain(t));
      CSX_TRY(run_region_block(*ch, b, workdir, log));
    } else if (t == "makeContrastCurve") {
      CSX_TRY(need_chain(t));
      CSX_TRY(run_curve_block(*ch, b, workdir, log));
    } else if (t == "makeHexCFpmResponse") {
      CSX_TRY(need_chain(t));
      CSX_TRY(run_response_block(*ch, b, log));
    } else if (t == "differentialEvolutionOptimizer") {
      DeResult res;
      CSX_TRY(run_de_block(ctx, ch.get(), b, log, res));
    } else if (t == "nloptOptimizer") {
      return fail(CSX_ERR_UNSUPPORTED, "nloptOptimizer: the nlopt library is not part of this build");
    } else if (t == "linearOptimizer") {
      CSX_TRY(need_chain(t));
      CSX_TRY(run_linear_block(*ch, b, log));
    } else {
      log += "!!! unknown top-level block: " + t + "\n";
    }
  }
  return CSX_OK;
}

static void json_str(std::string& o, const std::string& s) {
  o += '"';
  for (char c : s) {
    if (c == '"' || c == '\\') { o += '\\'; o += c; }
    else if ((unsigned char)c < 0x20) { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", c); o += b; }
    else o += c;
  }
  o += '"';
}
static void json_list(std::string& o, const std::vector<std::string>& v) {
  o += '[';
  for (size_t i = 0; i < v.size(); i++) {
    if (i) o += ", ";
    json_str(o, v[i]);
  }
  o += ']';
}
static void json_block(std::string& o, const Block& b) {
  o += "{\"type\": ";
  json_str(o, b.cmds.empty() ? "" : b.cmds[0].key);
  std::vector<std::string> keys;
  int depth = 0;
  for (size_t i = 1; i < b.cmds.size(); i++) {
    const std::string& k = b.cmds[i].key;
    if (k == "start") depth++;
    else if (k == "end") depth--;
    else if (depth == 0) keys.push_back(k);
  }
  o += ", \"keys\": ";
  json_list(o, keys);
  o += ", \"blocks\": [";
  // find_blocks puts the block's own depth-0 lines (its type line among them) first
  const std::vector<Block> sub = find_blocks(b);
  for (size_t i = 1; i < sub.size(); i++) {
    if (i > 1) o += ", ";
    json_block(o, sub[i]);
  }
  o += "]}";
}

// Device-free check of a script: the block structure, the chain's elements as the
// engine builds them (no init, no file read, no device), unknown keys and element
// types, and the files the script names.
static std::string parse_report(const std::string& text) {
  ParseAudit audit;
  parse_audit() = &audit;
  Chain ch(nullptr);
  int rc = CSX_OK;
  std::string err;
  std::vector<Block> blocks = find_blocks(parse_script_text(text));
  bool has_chain = false;
  for (const Block& b : blocks)
    if (!b.cmds.empty() && b.cmds[0].key == "coronagraph") has_chain = true;
  if (has_chain) {
    rc = ch.parse_only(text);
    if (rc != CSX_OK) err = csx_last_error();
  }
  parse_audit() = nullptr;
  std::string o = "{\"ok\": ";
  o += rc == CSX_OK ? "true" : "false";
  o += ", \"error\": ";
  json_str(o, err);
  o += ", \"blocks\": [";
  for (size_t i = 0; i < blocks.size(); i++) {
    if (i) o += ", ";
    json_block(o, blocks[i]);
  }
  o += "], \"elements\": [";
  if (has_chain)
    for (int i = 0; i < ch.n_elems(); i++) {
      if (i) o += ", ";
      o += "{\"type\": ";
      json_str(o, ch.elem_at(i)->type);
      o += ", \"name\": ";
      json_str(o, ch.elem_at(i)->name);
      o += "}";
    }
  o += "], \"unknown_keys\": ";
  json_list(o, audit.unknown_keys);
  o += ", \"unknown_elements\": ";
  json_list(o, audit.unknown_elements);
  o += ", \"files\": ";
  json_list(o, audit.files);
  std::vector<std::string> tools;
  for (const Block& b : blocks)
    if (!b.cmds.empty()) tools.push_back(b.cmds[0].key);
  o += ", \"top_level\": ";
  json_list(o, tools);
  o += "}";
  return o;
}

static int copy_out(const std::string& s, char* buf, size_t cap) {
  if (buf && cap) {
    const size_t n = std::min(cap - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return (int)std::min<size_t>(s.size(), 0x7fffffff);
}

}  // namespace csx

using namespace csx;

#define CSX_GUARD_BEGIN try {
#define CSX_GUARD_END                                                               \
  }                                                                                 \
  catch (const std::bad_alloc&) { return fail(CSX_ERR_NOMEM, "host allocation failed"); } \
  catch (const std::exception& e) { return fail(CSX_ERR_STATE, e.what()); }          \
  catch (...) { return fail(CSX_ERR_STATE, "unknown exception"); }

extern "C" {

int csx_script_parse(const char* text, char* json, size_t cap) {
  CSX_GUARD_BEGIN
  if (!text) return fail(CSX_ERR_ARG, "csx_script_parse: null text");
  return copy_out(parse_report(text), json, cap);
  CSX_GUARD_END
}

int csx_run_script(csx_ctx* ctx, const char* text, const char* workdir, char* log, size_t cap) {
  CSX_GUARD_BEGIN
  if (!ctx || !text) return fail(CSX_ERR_ARG, "csx_run_script: null argument");
  CSX_HIP(hipSetDevice(ctx->device));
  std::string out;
  int rc = run_script(ctx, text, workdir ? workdir : "", out);
  copy_out(out, log, cap);
  return rc == CSX_OK ? (int)std::min<size_t>(out.size(), 0x7fffffff) : rc;
  CSX_GUARD_END
}

int csx_response_save(const char* dir, const double* M, int rows, int cols, int slices, int npol, int nsrc,
                      double calib, const double* lambdas) {
  CSX_GUARD_BEGIN
  if (!dir || !M || !lambdas || rows <= 0 || cols <= 0 || slices <= 0 || npol <= 0 || nsrc <= 0)
    return fail(CSX_ERR_ARG, "csx_response_save: bad argument");
  ResponseData r;
  r.rows = rows; r.cols = cols; r.slices = slices; r.npol = npol; r.nsrc = nsrc;
  r.calib = calib;
  r.lambdas.assign(lambdas, lambdas + slices);
  r.M.resize((size_t)nsrc * npol * r.cube());
  std::memcpy(r.M.data(), M, sizeof(double2) * r.M.size());
  return response_save(dir, r);
  CSX_GUARD_END
}

int csx_response_load(const char* dir, double* M, size_t capacity, int* sizes, double* calib, double* lambdas,
                      int lambda_capacity) {
  CSX_GUARD_BEGIN
  if (!dir || !sizes) return fail(CSX_ERR_ARG, "csx_response_load: null argument");
  ResponseData r;
  CSX_TRY(response_load(dir, r));
  sizes[0] = r.rows; sizes[1] = r.cols; sizes[2] = r.slices; sizes[3] = r.npol; sizes[4] = r.nsrc;
  if (calib) *calib = r.calib;
  if (lambdas) {
    if (lambda_capacity < r.slices) return fail(CSX_ERR_ARG, "csx_response_load: lambda capacity below nSlices");
    std::copy(r.lambdas.begin(), r.lambdas.end(), lambdas);
  }
  if (M) {
    if (capacity < r.M.size()) return fail(CSX_ERR_ARG, "csx_response_load: capacity below the response size");
    std::memcpy(M, r.M.data(), sizeof(double2) * r.M.size());
  }
  return CSX_OK;
  CSX_GUARD_END
}

int csx_de_optimize(const char* de_block_text, int dim, csx_objective_fn f, void* user, double* best_vector,
                    double* best_value, int* iterations) {
  CSX_GUARD_BEGIN
  if (!de_block_text || !f || dim <= 0) return fail(CSX_ERR_ARG, "csx_de_optimize: bad argument");
  DeOptimizer de;
  de.dim = dim;
  Block all = parse_script_text(de_block_text);
  std::vector<Block> blocks = find_blocks(all);
  for (const Block& b : blocks)
    for (const Block& sb : find_blocks(b))
      if (!sb.cmds.empty() && sb.cmds[0].key == "differentialEvolutionOptimizer")
        for (const Cmd& c : sb.cmds) CSX_TRY(de.set(c.key, c.arg));
  Objective obj = [&](const std::vector<double>& pop, int P, std::vector<double>& val) {
    val.assign(P, 0.0);
    const int rc = f(user, pop.data(), dim, P, val.data());
    return rc < 0 ? fail(CSX_ERR_ARG, "csx_de_optimize: the objective returned " + std::to_string(rc)) : CSX_OK;
  };
  DeResult res;
  CSX_TRY(de.optimize(obj, res));
  if (best_vector) std::copy(res.best_vector.begin(), res.best_vector.end(), best_vector);
  if (best_value) *best_value = res.best;
  if (iterations) *iterations = res.iterations;
  return CSX_OK;
  CSX_GUARD_END
}

}  // extern "C"
