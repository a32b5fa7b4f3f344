// ftlm_host.h -- host side of the drop-in ftlm_n_vs_mu / ftlm_ed_n_vs_mu
// executables: command line, checkpoint files, thermodynamic accumulation
// and CSV output, with the reference's formats and messages.
#pragma once
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace ftlmh {

[[noreturn]] inline void die(const std::string& m) { throw std::runtime_error(m); }

// Params (main.cpp:32-61) plus the engine's additive flags.
struct Options {
  int lx = 2, ly = 2;
  double tx = 1.0, ty = 1.0, tp = 0.0, phix = 1.0, phiy = 1.0, u = 8.0, v = 0.0;
  double beta = 2.0;
  std::vector<double> betas;
  double mu_min = -4.0, mu_max = 10.0;
  int mu_count = 61, samples = 5, steps = 80, threads = 0;
  std::size_t threshold = 256;
  std::uint64_t seed = 12345;
  std::size_t max_sector_dim = 2000000;
  std::string output = "n_vs_mu.csv", checkpoint, trace_csv, debug_csv;
  int dbg_nup = -1, dbg_ndn = -1, dbg_mx = -1, dbg_my = -1;
  // engine flags (not in the reference)
  int device = 0;
  int reduction = 1;  // 0 serial, 1 tree
  bool per_vector = false;
  std::string timing_json;
  // packed small blocks: FTLM blocks of dimension <= pack_max_dim run
  // batched in persistent launches (0 disables packing)
  std::int64_t pack_max_dim = 131072;
  int pack_width = 0, pack_wg = 0;  // 0: tuned on the first pack
  double pack_budget_gb = 0.0;      // device bytes per pack (0: engine default)
  // twist scan: phix, phiy in {0, 1/N, ..., (N-1)/N}, one CSV per twist
  int twist_grid = 0;
  // start vectors drawn on host threads even with tree reductions (which
  // otherwise draw the same stream on the GPU)
  bool host_starts = false;
  // buffers timed for the Lanczos vectors when they are (re)allocated (the
  // engine's placement tuning; 0 off)
  int vector_placement = 8;
  // twist scan over listed (phix, phiy) pairs, one CSV per twist
  std::vector<std::pair<double, double>> twist_list;
  // resumable runs: the grand-canonical sums of finished sectors kept in a
  // small file (--partial-sums), and a wall-time budget after which the run
  // stops between sectors (--stop-after, with --partial-sums or
  // --checkpoint)
  std::string partial_sums;
  double stop_after = 0.0;
};

// parse_args (main.cpp:257-368). Returns false when --help was printed.
bool parse_options(int argc, char** argv, Options& o, bool ed_tool);
void print_help(const char* argv0);
int requested_threads(const Options& o);

// checkpoint (main.cpp:370-412, 1269-1398)
struct BlockKey {
  int nup, ndn, mx, my;
  bool operator==(const BlockKey& b) const { return nup == b.nup && ndn == b.ndn && mx == b.mx && my == b.my; }
};
struct BlockKeyHash {
  std::size_t operator()(const BlockKey& k) const {
    std::size_t h = (std::size_t)(k.nup * 131 + k.ndn);
    h = h * 1000003u ^ (std::size_t)(k.mx * 257 + k.my);
    return h;
  }
};
struct Spectrum {
  std::vector<double> eig, w;  // w empty for exact blocks
  double prefactor = 1.0;
  std::uint64_t dim = 0;
  int particles = 0;
  bool exact = false;
};
std::string metadata_text(const Options& o);
void check_or_write_metadata(const Options& o);
void ensure_header(const std::string& path);
// Reads every complete record; `valid_bytes` receives the offset just past
// the last complete record (0 when the file is absent or empty).
std::unordered_map<BlockKey, Spectrum, BlockKeyHash> read_checkpoint(const std::string& path,
                                                                    std::uint64_t* valid_bytes);
// Cut a torn tail (a partial record left by an interrupted run) so that new
// records follow the last complete one.
void truncate_to(const std::string& path, std::uint64_t bytes);
void append_record(const std::string& path, const BlockKey& k, const Spectrum& s);

// streaming log-sum-exp accumulation (main.cpp:1172-1219)
struct Thermo {
  std::vector<double> emin, z, n1, n2;
  explicit Thermo(std::size_t nmu = 0)
      : emin(nmu, std::numeric_limits<double>::infinity()), z(nmu, 0.0), n1(nmu, 0.0), n2(nmu, 0.0) {}
};
void accumulate(const double* eig, const double* w, std::size_t cnt, double prefactor, int particles, double beta,
                const std::vector<double>& mu, Thermo& t, int threads);

struct BlockThermo {
  double z = 0.0, n_total = 0.0, n2_total = 0.0, min_shifted = std::numeric_limits<double>::infinity();
};
// raw_block_thermo_from_ftlm / _from_exact (main.cpp:1115-1170)
BlockThermo raw_thermo(const double* eig, const double* w, std::size_t cnt, double prefactor, int particles,
                       double beta, double mu);

std::vector<double> linspace(double a, double b, int n);

// Grand-canonical sums of the sectors finished so far, per (twist, beta)
// stream, with the first sector still to run (--partial-sums). Written
// atomically after every sector; doubles in hexadecimal, so a resumed run
// continues from exactly the values an uninterrupted one would hold.
struct PartialState {
  std::string meta;          // identity of the run (refuses a different one)
  int next_nup = 0, next_ndn = 0;
  int runs = 0;              // processes that contributed
  double elapsed_s = 0.0;    // their summed wall time
  std::int64_t blocks = 0;   // blocks summed
  int ntwist = 0, nbeta = 0, nmu = 0;
  std::vector<Thermo> th;    // [twist * nbeta + beta]
};
std::string partial_meta(const Options& o, const std::vector<std::pair<double, double>>& twists);
// false when the file does not exist; dies on a malformed file
bool read_partial(const std::string& path, PartialState& s);
void write_partial(const std::string& path, const PartialState& s);
// fold the sums of a later set of blocks (same beta) into a running state:
// both are rescaled to the smaller of their minima (main.cpp:1195-1216
// applies the same rescale block by block)
void merge_thermo(Thermo& run, const Thermo& part, double beta);

struct Results {
  std::vector<double> n, cc, kappa, z, logz;
};
Results finalize(const Thermo& t, double beta, int sites);
void write_results(const std::string& path, const std::vector<double>& mu, const Results& r);
void write_multi_beta(const std::string& path, const std::vector<double>& betas, const std::vector<double>& mu,
                      const std::vector<Results>& r);

}  // namespace ftlmh
