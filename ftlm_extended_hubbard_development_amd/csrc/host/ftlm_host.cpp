// ftlm_host.cpp -- host side of the drop-in executables (see ftlm_host.h).
//
// Formats, messages and the floating-point order of the thermodynamic
// accumulation follow the reference (src/main.cpp, lines cited per function)
// so that a user switching executables sees the same command line, the same
// CSV columns and a checkpoint file either program can resume.
#include "ftlm_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <thread>

#include <sys/stat.h>
#include <unistd.h>

namespace ftlmh {

namespace {

template <typename T>
T parse_as(const std::string& text) {
  std::istringstream in(text);
  T value{};
  in >> value;
  if (!in || !in.eof()) die("Failed to parse value: " + text);
  return value;
}

std::vector<double> parse_list(const std::string& text) {
  std::vector<double> out;
  std::istringstream in(text);
  std::string item;
  while (std::getline(in, item, ',')) {
    if (item.empty()) die("Empty entry in comma-separated value list: " + text);
    out.push_back(parse_as<double>(item));
  }
  if (out.empty()) die("Expected at least one value in comma-separated list: " + text);
  return out;
}

const char kFileMagic[8] = {'F', 'T', 'L', 'M', 'C', 'P', '1', '\n'};
const char kRecMagic[8] = {'B', 'L', 'K', 'R', 'E', 'C', '1', '\n'};

template <typename T>
bool get(std::istream& in, T& v) {
  return static_cast<bool>(in.read(reinterpret_cast<char*>(&v), sizeof(T)));
}
template <typename T>
void put(std::ostream& out, const T& v) {
  out.write(reinterpret_cast<const char*>(&v), sizeof(T));
}

}  // namespace

void print_help(const char* argv0) {
  std::cout << "Usage: " << argv0 << " [options]\n"
            << "  --lx N                lattice size in x\n"
            << "  --ly N                lattice size in y\n"
            << "  --tx X                hopping along x bonds\n"
            << "  --ty Y                hopping along y bonds\n"
            << "  --tp X                second-neighbor diagonal hopping\n"
            << "  --phix X              homogeneous x twist in units of 2*pi\n"
            << "  --phiy Y              homogeneous y twist in units of 2*pi\n"
            << "  --u U                 onsite interaction\n"
            << "  --v V                 nearest-neighbor interaction\n"
            << "  --beta B              inverse temperature\n"
            << "  --beta-list LIST      comma-separated inverse temperatures\n"
            << "  --mu-min X            minimum chemical potential\n"
            << "  --mu-max X            maximum chemical potential\n"
            << "  --mu-count N          number of mu points\n"
            << "  --samples N           FTLM random vectors per k block\n"
            << "  --lanczos-steps N     Lanczos steps per random vector\n"
            << "  --threads N           host threads (0 uses the hardware default)\n"
            << "  --exact-block-threshold N\n"
            << "                        use exact diagonalization for k blocks with dimension <= N\n"
            << "  --seed N              random seed\n"
            << "  --max-sector-dim N    abort if a full sector exceeds this size\n"
            << "  --checkpoint PATH     append/resume compact per-block spectra\n"
            << "  --trace-partition-csv PATH\n"
            << "                        write per-block FTLM vs ED thermodynamic traces\n"
            << "  --debug-block-nup N   selected block debug Nup\n"
            << "  --debug-block-ndown N selected block debug Ndown\n"
            << "  --debug-block-mx N    selected block debug mx\n"
            << "  --debug-block-my N    selected block debug my\n"
            << "  --debug-block-csv PATH\n"
            << "                        write selected-block FTLM vs ED comparison\n"
            << "  --output PATH         CSV output path\n"
            << "  --device N            HIP device index (GPU engine)\n"
            << "  --reduction MODE      Lanczos reductions: tree (default) or serial\n"
            << "                        (serial sums in row order, bit-identical to the CPU code)\n"
            << "  --per-vector          run each random vector alone (plain per-vector GPU FTLM)\n"
            << "  --host-starts         draw start vectors on host threads with tree reductions too\n"
            << "                        (tree reductions draw the same seeded stream on the GPU)\n"
            << "  --timing-json PATH    write per-block device timings\n"
            << "  --pack-max-dim N      FTLM blocks up to N rows run packed in persistent launches\n"
            << "                        (default 131072; 0 runs every block on its own)\n"
            << "  --pack-width W        columns per packed work item: 4, 8, 16, 32, 64 (default: tuned)\n"
            << "  --pack-wg N           packed workgroups per CU (default: tuned)\n"
            << "  --pack-budget-gb G    device memory per pack in GiB (default 64)\n"
            << "  --twist-grid N        run the twists phix, phiy in {0, 1/N, ..., (N-1)/N} together;\n"
            << "                        one CSV per twist, OUTPUT with .phixA_phiyB before the extension\n"
            << "  --twist-list L        like --twist-grid for the listed twists PX:PY,PX:PY,...\n"
            << "  --partial-sums PATH   keep the grand-canonical sums of finished sectors in PATH and\n"
            << "                        resume from it (a few KB; the CSV is written when all are done)\n"
            << "  --stop-after S        stop between sectors before S seconds of wall time are exceeded\n"
            << "                        (predicted from this run's rate); rerun the command to continue\n"
            << "  --vector-placement N  buffers timed for the Lanczos vectors when they are allocated\n"
            << "                        (default 8; 0 keeps the first allocation)\n"
            << "  --help                show this message\n";
}

bool parse_options(int argc, char** argv, Options& p, bool ed_tool) {
  (void)ed_tool;
  for (int i = 1; i < argc; ++i) {
    const std::string arg = argv[i];
    auto next = [&](const std::string& name) {
      if (i + 1 >= argc) die("Missing value for " + name);
      return std::string(argv[++i]);
    };
    if (arg == "--help") {
      print_help(argv[0]);
      return false;
    } else if (arg == "--lx") p.lx = parse_as<int>(next(arg));
    else if (arg == "--ly") p.ly = parse_as<int>(next(arg));
    else if (arg == "--tx") p.tx = parse_as<double>(next(arg));
    else if (arg == "--ty") p.ty = parse_as<double>(next(arg));
    else if (arg == "--tp") p.tp = parse_as<double>(next(arg));
    else if (arg == "--phix") p.phix = parse_as<double>(next(arg));
    else if (arg == "--phiy") p.phiy = parse_as<double>(next(arg));
    else if (arg == "--u") p.u = parse_as<double>(next(arg));
    else if (arg == "--v") p.v = parse_as<double>(next(arg));
    else if (arg == "--beta") p.beta = parse_as<double>(next(arg));
    else if (arg == "--beta-list") p.betas = parse_list(next(arg));
    else if (arg == "--mu-min") p.mu_min = parse_as<double>(next(arg));
    else if (arg == "--mu-max") p.mu_max = parse_as<double>(next(arg));
    else if (arg == "--mu-count") p.mu_count = parse_as<int>(next(arg));
    else if (arg == "--samples") p.samples = parse_as<int>(next(arg));
    else if (arg == "--lanczos-steps") p.steps = parse_as<int>(next(arg));
    else if (arg == "--threads") p.threads = parse_as<int>(next(arg));
    else if (arg == "--exact-block-threshold") p.threshold = parse_as<std::size_t>(next(arg));
    else if (arg == "--seed") p.seed = parse_as<std::uint64_t>(next(arg));
    else if (arg == "--max-sector-dim") p.max_sector_dim = parse_as<std::size_t>(next(arg));
    else if (arg == "--checkpoint") p.checkpoint = next(arg);
    else if (arg == "--trace-partition-csv") p.trace_csv = next(arg);
    else if (arg == "--debug-block-nup") p.dbg_nup = parse_as<int>(next(arg));
    else if (arg == "--debug-block-ndown") p.dbg_ndn = parse_as<int>(next(arg));
    else if (arg == "--debug-block-mx") p.dbg_mx = parse_as<int>(next(arg));
    else if (arg == "--debug-block-my") p.dbg_my = parse_as<int>(next(arg));
    else if (arg == "--debug-block-csv") p.debug_csv = next(arg);
    else if (arg == "--output") p.output = next(arg);
    else if (arg == "--device") p.device = parse_as<int>(next(arg));
    else if (arg == "--reduction") {
      const std::string m = next(arg);
      if (m == "tree") p.reduction = 1;
      else if (m == "serial") p.reduction = 0;
      else die("--reduction must be tree or serial");
    } else if (arg == "--per-vector") p.per_vector = true;
    else if (arg == "--host-starts") p.host_starts = true;
    else if (arg == "--timing-json") p.timing_json = next(arg);
    else if (arg == "--pack-max-dim") p.pack_max_dim = parse_as<std::int64_t>(next(arg));
    else if (arg == "--pack-width") p.pack_width = parse_as<int>(next(arg));
    else if (arg == "--pack-wg") p.pack_wg = parse_as<int>(next(arg));
    else if (arg == "--pack-budget-gb") p.pack_budget_gb = parse_as<double>(next(arg));
    else if (arg == "--twist-grid") p.twist_grid = parse_as<int>(next(arg));
    else if (arg == "--partial-sums") p.partial_sums = next(arg);
    else if (arg == "--stop-after") p.stop_after = parse_as<double>(next(arg));
    else if (arg == "--vector-placement") p.vector_placement = parse_as<int>(next(arg));
    else if (arg == "--twist-list") {
      std::stringstream ss(next(arg));
      std::string item;
      p.twist_list.clear();
      while (std::getline(ss, item, ',')) {
        const std::size_t colon = item.find(':');
        if (colon == std::string::npos) die("--twist-list entries must be PX:PY, got: " + item);
        p.twist_list.push_back({parse_as<double>(item.substr(0, colon)), parse_as<double>(item.substr(colon + 1))});
      }
      if (p.twist_list.empty()) die("--twist-list must name at least one twist.");
    }
    else die("Unknown argument: " + arg);
  }
  if (p.lx <= 0 || p.ly <= 0) die("Lattice dimensions must be positive.");
  if (p.lx * p.ly > 63) die("This implementation supports at most 63 sites.");
  if (p.betas.empty()) p.betas.push_back(p.beta);
  else p.beta = p.betas.front();
  for (double b : p.betas)
    if (b <= 0.0) die("All beta values must be positive.");
  if (p.samples <= 0 || p.steps <= 0) die("--samples and --lanczos-steps must be positive.");
  if (p.threads < 0) die("--threads must be non-negative.");
  if (p.mu_count < 2) die("--mu-count must be at least 2.");
  if (p.pack_max_dim < 0) die("--pack-max-dim must be non-negative.");
  if (p.pack_width != 0 && p.pack_width != 4 && p.pack_width != 8 && p.pack_width != 16 && p.pack_width != 32 &&
      p.pack_width != 64)
    die("--pack-width must be 4, 8, 16, 32 or 64.");
  if (p.pack_wg < 0 || p.pack_budget_gb < 0.0) die("--pack-wg and --pack-budget-gb must be non-negative.");
  if (p.twist_grid < 0) die("--twist-grid must be non-negative.");
  if (p.vector_placement != 0 && (p.vector_placement < 3 || p.vector_placement > 16))
    die("--vector-placement must be 0 or 3 to 16.");
  if (p.twist_grid > 0 && !p.twist_list.empty()) die("--twist-grid and --twist-list are exclusive.");
  if (p.stop_after < 0.0) die("--stop-after must be non-negative.");
  if (p.stop_after > 0.0 && p.partial_sums.empty() && p.checkpoint.empty())
    die("--stop-after needs --partial-sums or --checkpoint to resume from.");
  if (!p.partial_sums.empty() && !p.checkpoint.empty()) die("--partial-sums and --checkpoint are exclusive.");
  const bool partial = p.dbg_nup >= 0 || p.dbg_ndn >= 0 || p.dbg_mx >= 0 || p.dbg_my >= 0 || !p.debug_csv.empty();
  const bool full = p.dbg_nup >= 0 && p.dbg_ndn >= 0 && p.dbg_mx >= 0 && p.dbg_my >= 0;
  if (partial && !full)
    die("Selected-block debug requires --debug-block-nup/--debug-block-ndown/--debug-block-mx/--debug-block-my "
        "together.");
  return true;
}

int requested_threads(const Options& o) {
  if (o.threads > 0) return o.threads;
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int n = std::atoi(e);
    if (n > 0) return n;
  }
  const unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

// checkpoint_metadata_text (main.cpp:370-387)
std::string metadata_text(const Options& o) {
  std::ostringstream out;
  out << std::setprecision(17) << "lx=" << o.lx << "\n"
      << "ly=" << o.ly << "\n"
      << "tx=" << o.tx << "\n"
      << "ty=" << o.ty << "\n"
      << "tp=" << o.tp << "\n"
      << "phix=" << o.phix << "\n"
      << "phiy=" << o.phiy << "\n"
      << "u=" << o.u << "\n"
      << "v=" << o.v << "\n"
      << "samples=" << o.samples << "\n"
      << "lanczos_steps=" << o.steps << "\n"
      << "exact_block_threshold=" << o.threshold << "\n"
      << "seed=" << o.seed << "\n";
  return out.str();
}

// validate_or_write_checkpoint_metadata (main.cpp:389-412)
void check_or_write_metadata(const Options& o) {
  if (o.checkpoint.empty()) return;
  const std::string path = o.checkpoint + ".meta";
  const std::string want = metadata_text(o);
  {
    std::ifstream in(path);
    if (in) {
      const std::string have((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
      if (have != want) die("Checkpoint metadata does not match current run: " + path);
      return;
    }
  }
  std::ofstream out(path);
  if (!out) die("Failed to write checkpoint metadata: " + path);
  out << want;
}

// ensure_checkpoint_header (main.cpp:1279-1295)
void ensure_header(const std::string& path) {
  if (path.empty()) return;
  {
    std::ifstream in(path, std::ios::binary);
    if (in && in.peek() != std::ifstream::traits_type::eof()) return;
  }
  std::ofstream out(path, std::ios::binary);
  if (!out) die("Failed to create checkpoint file: " + path);
  out.write(kFileMagic, 8);
}

// read_checkpoint (main.cpp:1297-1361): stops at the first incomplete or
// unrecognised record; later records of the same block replace earlier ones
std::unordered_map<BlockKey, Spectrum, BlockKeyHash> read_checkpoint(const std::string& path,
                                                                    std::uint64_t* valid_bytes) {
  std::unordered_map<BlockKey, Spectrum, BlockKeyHash> out;
  if (valid_bytes) *valid_bytes = 0;
  if (path.empty()) return out;
  std::ifstream in(path, std::ios::binary);
  if (!in) return out;
  char magic[8];
  if (!in.read(magic, 8)) return out;
  if (std::memcmp(magic, kFileMagic, 8) != 0) die("Checkpoint has an unrecognized format: " + path);
  std::uint64_t good = 8;
  while (true) {
    char rm[8];
    if (!in.read(rm, 8)) break;
    if (std::memcmp(rm, kRecMagic, 8) != 0) break;
    BlockKey k{};
    Spectrum s;
    std::uint8_t ex = 0;
    std::uint64_t dim = 0, ne = 0, nw = 0;
    if (!get(in, k.nup) || !get(in, k.ndn) || !get(in, k.mx) || !get(in, k.my) || !get(in, s.particles) ||
        !get(in, dim) || !get(in, s.prefactor) || !get(in, ex) || !get(in, ne) || !get(in, nw))
      break;
    if (ne > 100000000ULL || nw > 100000000ULL) die("Checkpoint record is implausibly large: " + path);
    s.dim = dim;
    s.exact = ex != 0;
    s.eig.resize(ne);
    s.w.resize(nw);
    if (!in.read(reinterpret_cast<char*>(s.eig.data()), (std::streamsize)(8 * ne)) ||
        !in.read(reinterpret_cast<char*>(s.w.data()), (std::streamsize)(8 * nw)))
      break;
    out[k] = std::move(s);
    good += 8 + 5 * 4 + 8 + 8 + 1 + 8 + 8 + 8 * (ne + nw);
  }
  if (valid_bytes) *valid_bytes = good;
  return out;
}

void truncate_to(const std::string& path, std::uint64_t bytes) {
  if (path.empty() || bytes < 8) return;
  struct stat st;
  if (stat(path.c_str(), &st) != 0) return;
  if ((std::uint64_t)st.st_size > bytes) {
    if (truncate(path.c_str(), (off_t)bytes) != 0) die("Failed to truncate torn checkpoint tail: " + path);
  }
}

// append_checkpoint_record (main.cpp:1363-1398)
void append_record(const std::string& path, const BlockKey& k, const Spectrum& s) {
  if (path.empty()) return;
  ensure_header(path);
  std::ofstream out(path, std::ios::binary | std::ios::app);
  if (!out) die("Failed to append checkpoint file: " + path);
  out.write(kRecMagic, 8);
  put(out, k.nup);
  put(out, k.ndn);
  put(out, k.mx);
  put(out, k.my);
  put(out, s.particles);
  put(out, s.dim);
  put(out, s.prefactor);
  const std::uint8_t ex = s.exact ? 1 : 0;
  put(out, ex);
  const std::uint64_t ne = s.eig.size(), nw = s.w.size();
  put(out, ne);
  put(out, nw);
  out.write(reinterpret_cast<const char*>(s.eig.data()), (std::streamsize)(8 * ne));
  out.write(reinterpret_cast<const char*>(s.w.data()), (std::streamsize)(8 * nw));
  out.flush();
}

// accumulate_block_thermo (main.cpp:1172-1219). The mu points are
// independent chains, so they may run on several threads without changing a
// single rounding.
void accumulate(const double* eig, const double* w, std::size_t cnt, double pref, int particles, double beta,
                const std::vector<double>& mu, Thermo& t, int threads) {
  const double N = static_cast<double>(particles);
  auto one = [&](std::size_t k) {
    double bmin = std::numeric_limits<double>::infinity();
    for (std::size_t i = 0; i < cnt; ++i) bmin = std::min(bmin, eig[i] - mu[k] * N);
    if (!std::isfinite(bmin)) return;
    if (!std::isfinite(t.emin[k])) {
      t.emin[k] = bmin;
    } else if (bmin < t.emin[k]) {
      const double sc = std::exp(-beta * (t.emin[k] - bmin));
      t.z[k] *= sc;
      t.n1[k] *= sc;
      t.n2[k] *= sc;
      t.emin[k] = bmin;
    }
    for (std::size_t i = 0; i < cnt; ++i) {
      const double sh = eig[i] - mu[k] * N - t.emin[k];
      const double wt = pref * (w ? w[i] : 1.0) * std::exp(-beta * sh);
      t.z[k] += wt;
      t.n1[k] += N * wt;
      t.n2[k] += N * N * wt;
    }
  };
  const std::size_t nmu = mu.size();
  const int nt = (int)std::min<std::size_t>((std::size_t)std::max(1, threads), nmu);
  if (nt <= 1 || cnt * nmu < 200000) {
    for (std::size_t k = 0; k < nmu; ++k) one(k);
    return;
  }
  std::vector<std::thread> th;
  for (int j = 0; j < nt; ++j)
    th.emplace_back([&, j]() {
      for (std::size_t k = (std::size_t)j; k < nmu; k += (std::size_t)nt) one(k);
    });
  for (auto& x : th) x.join();
}

// raw_block_thermo_from_ftlm / raw_block_thermo_from_exact (main.cpp:1115-1170)
BlockThermo raw_thermo(const double* eig, const double* w, std::size_t cnt, double pref, int particles,
                       double beta, double mu) {
  BlockThermo t;
  const double N = static_cast<double>(particles);
  for (std::size_t i = 0; i < cnt; ++i) t.min_shifted = std::min(t.min_shifted, eig[i] - mu * N);
  if (!std::isfinite(t.min_shifted)) return t;
  for (std::size_t i = 0; i < cnt; ++i) {
    const double sh = eig[i] - mu * N - t.min_shifted;
    const double wt = w ? pref * w[i] * std::exp(-beta * sh) : std::exp(-beta * sh);
    t.z += wt;
    t.n_total += N * wt;
    t.n2_total += N * N * wt;
  }
  return t;
}

// linspace (main.cpp:1400-1407)
std::vector<double> linspace(double a, double b, int n) {
  std::vector<double> v((std::size_t)n);
  const double st = (b - a) / static_cast<double>(n - 1);
  for (int i = 0; i < n; ++i) v[(std::size_t)i] = a + st * static_cast<double>(i);
  return v;
}

// main.cpp:1770-1788 / 1805-1823
Results finalize(const Thermo& t, double beta, int sites) {
  Results r;
  const std::size_t n = t.z.size();
  r.n.resize(n);
  r.cc.resize(n);
  r.kappa.resize(n);
  r.z.resize(n);
  r.logz.resize(n);
  for (std::size_t k = 0; k < n; ++k) {
    r.z[k] = t.z[k];
    r.logz[k] = std::log(t.z[k]) - beta * t.emin[k];
    const double s = static_cast<double>(sites);
    const double mean = t.n1[k] / t.z[k];
    const double mean2 = t.n2[k] / t.z[k];
    const double cc = (mean2 - mean * mean) / s;
    r.n[k] = mean / s;
    r.cc[k] = std::max(0.0, cc);
    r.kappa[k] = beta * r.cc[k];
  }
  return r;
}

// write_results (main.cpp:1409-1430)
void write_results(const std::string& path, const std::vector<double>& mu, const Results& r) {
  std::ofstream out(path);
  if (!out) die("Failed to open output file: " + path);
  out << "mu,n,charge_correlation,compressibility,partition_like,log_partition\n";
  out << std::setprecision(15);
  for (std::size_t i = 0; i < mu.size(); ++i)
    out << mu[i] << "," << r.n[i] << "," << r.cc[i] << "," << r.kappa[i] << "," << r.z[i] << "," << r.logz[i]
        << "\n";
}

// write_multi_beta_results (main.cpp:1432-1455)
void write_multi_beta(const std::string& path, const std::vector<double>& betas, const std::vector<double>& mu,
                      const std::vector<Results>& r) {
  std::ofstream out(path);
  if (!out) die("Failed to open output file: " + path);
  out << "beta,mu,n,charge_correlation,compressibility,partition_like,log_partition\n";
  out << std::setprecision(15);
  for (std::size_t b = 0; b < betas.size(); ++b)
    for (std::size_t i = 0; i < mu.size(); ++i)
      out << betas[b] << "," << mu[i] << "," << r[b].n[i] << "," << r[b].cc[i] << "," << r[b].kappa[i] << ","
          << r[b].z[i] << "," << r[b].logz[i] << "\n";
}

std::string partial_meta(const Options& o, const std::vector<std::pair<double, double>>& twists) {
  std::ostringstream out;
  out << metadata_text(o) << std::setprecision(17) << "betas=";
  for (std::size_t i = 0; i < o.betas.size(); ++i) out << (i ? "," : "") << o.betas[i];
  out << "\nmu=" << o.mu_min << "," << o.mu_max << "," << o.mu_count << "\nreduction="
      << (o.reduction == 0 ? "serial" : "tree") << "\ntwists=";
  for (std::size_t i = 0; i < twists.size(); ++i) out << (i ? ";" : "") << twists[i].first << ":" << twists[i].second;
  out << "\n";
  return out.str();
}

namespace {
const char* kPartialMagic = "FTLM_PARTIAL_SUMS 1";

std::string hexd(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%a", v);
  return b;
}

double parse_hexd(const std::string& t, const std::string& path) {
  char* end = nullptr;
  const double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str() || *end != '\0') die("Malformed partial-sum file: " + path);
  return v;
}
}  // namespace

bool read_partial(const std::string& path, PartialState& s) {
  std::ifstream in(path);
  if (!in) return false;
  auto bad = [&]() { die("Malformed partial-sum file: " + path); };
  std::string line, key;
  if (!std::getline(in, line) || line != kPartialMagic) bad();
  int nmeta = 0;
  if (!(in >> key >> nmeta) || key != "meta_lines" || nmeta < 0) bad();
  std::getline(in, line);
  s.meta.clear();
  for (int i = 0; i < nmeta; ++i) {
    if (!std::getline(in, line)) bad();
    s.meta += line + "\n";
  }
  std::string el;
  if (!(in >> key >> s.next_nup >> s.next_ndn) || key != "next") bad();
  if (!(in >> key >> s.runs) || key != "runs") bad();
  if (!(in >> key >> el) || key != "elapsed") bad();
  s.elapsed_s = parse_hexd(el, path);
  if (!(in >> key >> s.blocks) || key != "blocks") bad();
  if (!(in >> key >> s.ntwist >> s.nbeta >> s.nmu) || key != "shape") bad();
  if (s.ntwist <= 0 || s.nbeta <= 0 || s.nmu <= 0) bad();
  s.th.assign((std::size_t)s.ntwist * s.nbeta, Thermo((std::size_t)s.nmu));
  for (Thermo& t : s.th)
    for (int k = 0; k < s.nmu; ++k) {
      std::string a, b, c, d;
      if (!(in >> a >> b >> c >> d)) bad();
      t.emin[k] = parse_hexd(a, path);
      t.z[k] = parse_hexd(b, path);
      t.n1[k] = parse_hexd(c, path);
      t.n2[k] = parse_hexd(d, path);
    }
  if (!(in >> key) || key != "end") bad();
  return true;
}

void write_partial(const std::string& path, const PartialState& s) {
  const std::string tmp = path + ".tmp";
  {
    std::ofstream out(tmp);
    if (!out) die("Failed to write partial-sum file: " + tmp);
    int nmeta = 0;
    for (char ch : s.meta) nmeta += ch == '\n';
    out << kPartialMagic << "\nmeta_lines " << nmeta << "\n" << s.meta << "next " << s.next_nup << " "
        << s.next_ndn << "\nruns " << s.runs << "\nelapsed " << hexd(s.elapsed_s) << "\nblocks " << s.blocks
        << "\nshape " << s.ntwist << " " << s.nbeta << " " << s.nmu << "\n";
    for (const Thermo& t : s.th)
      for (int k = 0; k < s.nmu; ++k)
        out << hexd(t.emin[k]) << " " << hexd(t.z[k]) << " " << hexd(t.n1[k]) << " " << hexd(t.n2[k]) << "\n";
    out << "end\n";
    out.flush();
    if (!out) die("Failed to write partial-sum file: " + tmp);
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) die("Failed to replace partial-sum file: " + path);
}

void merge_thermo(Thermo& run, const Thermo& part, double beta) {
  for (std::size_t k = 0; k < run.z.size(); ++k) {
    if (!(part.z[k] > 0.0)) continue;
    if (!(run.z[k] > 0.0)) {
      run.emin[k] = part.emin[k];
      run.z[k] = part.z[k];
      run.n1[k] = part.n1[k];
      run.n2[k] = part.n2[k];
      continue;
    }
    const double e = std::min(run.emin[k], part.emin[k]);
    const double fr = std::exp(-beta * (run.emin[k] - e));
    const double fp = std::exp(-beta * (part.emin[k] - e));
    run.z[k] = run.z[k] * fr + part.z[k] * fp;
    run.n1[k] = run.n1[k] * fr + part.n1[k] * fp;
    run.n2[k] = run.n2[k] * fr + part.n2[k] * fp;
    run.emin[k] = e;
  }
}

}  // namespace ftlmh

// ---- C hooks for the CPU tests of the host layer (no GPU involved) --------
extern "C" {

static thread_local std::string g_host_err;

const char* ftlmh_last_error(void) { return g_host_err.c_str(); }

// 0 = parsed, 1 = error (message in ftlmh_last_error), 2 = --help
int ftlmh_parse(int argc, char** argv) {
  try {
    ftlmh::Options o;
    return ftlmh::parse_options(argc, argv, o, false) ? 0 : 2;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return 1;
  }
}

int ftlmh_metadata(int argc, char** argv, char* buf, int cap) {
  try {
    ftlmh::Options o;
    ftlmh::parse_options(argc, argv, o, false);
    const std::string s = ftlmh::metadata_text(o);
    if ((int)s.size() >= cap) return -1;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return (int)s.size();
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
}

// number of records loaded (-1 on error); valid receives the byte offset
// just past the last complete record
long long ftlmh_checkpoint_count(const char* path, unsigned long long* valid) {
  try {
    std::uint64_t v = 0;
    const auto m = ftlmh::read_checkpoint(path, &v);
    if (valid) *valid = v;
    return (long long)m.size();
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
}

// fetch one record: returns eigen count (weights count in *nw) or -1
long long ftlmh_checkpoint_get(const char* path, int nup, int ndn, int mx, int my, double* eig, double* w,
                               long long* nw, double* pref, unsigned long long* dim, int* particles, int* exact) {
  try {
    const auto m = ftlmh::read_checkpoint(path, nullptr);
    const auto it = m.find(ftlmh::BlockKey{nup, ndn, mx, my});
    if (it == m.end()) return -1;
    const auto& s = it->second;
    if (eig) std::memcpy(eig, s.eig.data(), 8 * s.eig.size());
    if (w) std::memcpy(w, s.w.data(), 8 * s.w.size());
    if (nw) *nw = (long long)s.w.size();
    if (pref) *pref = s.prefactor;
    if (dim) *dim = s.dim;
    if (particles) *particles = s.particles;
    if (exact) *exact = s.exact;
    return (long long)s.eig.size();
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
}

// open for resume: read, then cut a torn tail; returns records loaded
long long ftlmh_checkpoint_resume(const char* path) {
  try {
    std::uint64_t v = 0;
    ftlmh::ensure_header(path);
    const auto m = ftlmh::read_checkpoint(path, &v);
    ftlmh::truncate_to(path, v);
    return (long long)m.size();
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
}

int ftlmh_checkpoint_append(const char* path, int nup, int ndn, int mx, int my, int particles,
                            unsigned long long dim, double pref, int exact, long long ne, const double* eig,
                            long long nw, const double* w) {
  try {
    ftlmh::Spectrum s;
    s.particles = particles;
    s.dim = dim;
    s.prefactor = pref;
    s.exact = exact != 0;
    s.eig.assign(eig, eig + ne);
    s.w.assign(w, w + nw);
    ftlmh::append_record(path, ftlmh::BlockKey{nup, ndn, mx, my}, s);
    return 0;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return 1;
  }
}

void ftlmh_accumulate(const double* eig, const double* w, long long cnt, double pref, int particles, double beta,
                      const double* mu, int nmu, double* emin, double* z, double* n1, double* n2, int threads) {
  ftlmh::Thermo t((std::size_t)nmu);
  std::vector<double> m(mu, mu + nmu);
  std::memcpy(t.emin.data(), emin, 8 * (size_t)nmu);
  std::memcpy(t.z.data(), z, 8 * (size_t)nmu);
  std::memcpy(t.n1.data(), n1, 8 * (size_t)nmu);
  std::memcpy(t.n2.data(), n2, 8 * (size_t)nmu);
  ftlmh::accumulate(eig, w, (std::size_t)cnt, pref, particles, beta, m, t, threads);
  std::memcpy(emin, t.emin.data(), 8 * (size_t)nmu);
  std::memcpy(z, t.z.data(), 8 * (size_t)nmu);
  std::memcpy(n1, t.n1.data(), 8 * (size_t)nmu);
  std::memcpy(n2, t.n2.data(), 8 * (size_t)nmu);
}

int ftlmh_write_results(const char* path, int nbeta, const double* betas, int nmu, const double* mu, const double* n,
                        const double* cc, const double* kappa, const double* z, const double* logz) {
  try {
    std::vector<double> M(mu, mu + nmu);
    auto take = [&](int b) {
      ftlmh::Results r;
      r.n.assign(n + (size_t)b * nmu, n + (size_t)(b + 1) * nmu);
      r.cc.assign(cc + (size_t)b * nmu, cc + (size_t)(b + 1) * nmu);
      r.kappa.assign(kappa + (size_t)b * nmu, kappa + (size_t)(b + 1) * nmu);
      r.z.assign(z + (size_t)b * nmu, z + (size_t)(b + 1) * nmu);
      r.logz.assign(logz + (size_t)b * nmu, logz + (size_t)(b + 1) * nmu);
      return r;
    };
    if (nbeta <= 1) {
      ftlmh::write_results(path, M, take(0));
    } else {
      std::vector<ftlmh::Results> rs;
      for (int b = 0; b < nbeta; ++b) rs.push_back(take(b));
      ftlmh::write_multi_beta(path, std::vector<double>(betas, betas + nbeta), M, rs);
    }
    return 0;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return 1;
  }
}

// Partial-sum files: sums[(t * nb + b) * 4 * nmu + q * nmu + k], q = emin, z,
// n1, n2.
int ftlmh_partial_write(const char* path, const char* meta, int next_nup, int next_ndn, int runs, double elapsed,
                        long long blocks, int nt, int nb, int nmu, const double* sums) {
  try {
    ftlmh::PartialState s;
    s.meta = meta;
    s.next_nup = next_nup;
    s.next_ndn = next_ndn;
    s.runs = runs;
    s.elapsed_s = elapsed;
    s.blocks = blocks;
    s.ntwist = nt;
    s.nbeta = nb;
    s.nmu = nmu;
    s.th.assign((std::size_t)nt * nb, ftlmh::Thermo((std::size_t)nmu));
    for (std::size_t i = 0; i < s.th.size(); ++i)
      for (int k = 0; k < nmu; ++k) {
        const double* q = sums + i * 4 * nmu;
        s.th[i].emin[k] = q[k];
        s.th[i].z[k] = q[nmu + k];
        s.th[i].n1[k] = q[2 * nmu + k];
        s.th[i].n2[k] = q[3 * nmu + k];
      }
    ftlmh::write_partial(path, s);
    return 0;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
}

// 1 read, 0 absent, -1 error; meta is NUL-terminated within cap bytes
int ftlmh_partial_read(const char* path, char* meta, int cap, int* head, double* elapsed, long long* blocks,
                       double* sums, long long cap_sums) {
  try {
    ftlmh::PartialState s;
    if (!ftlmh::read_partial(path, s)) return 0;
    if ((int)s.meta.size() >= cap) throw std::runtime_error("meta buffer too small");
    std::memcpy(meta, s.meta.c_str(), s.meta.size() + 1);
    head[0] = s.next_nup;
    head[1] = s.next_ndn;
    head[2] = s.runs;
    head[3] = s.ntwist;
    head[4] = s.nbeta;
    head[5] = s.nmu;
    *elapsed = s.elapsed_s;
    *blocks = s.blocks;
    if ((long long)s.th.size() * 4 * s.nmu > cap_sums) throw std::runtime_error("sums buffer too small");
    for (std::size_t i = 0; i < s.th.size(); ++i)
      for (int k = 0; k < s.nmu; ++k) {
        double* q = sums + i * 4 * s.nmu;
        q[k] = s.th[i].emin[k];
        q[s.nmu + k] = s.th[i].z[k];
        q[2 * s.nmu + k] = s.th[i].n1[k];
        q[3 * s.nmu + k] = s.th[i].n2[k];
      }
    return 1;
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
}

// run, part: [4 * nmu] planes emin, z, n1, n2; run is updated in place
void ftlmh_merge(int nmu, double beta, double* run, const double* part) {
  ftlmh::Thermo a((std::size_t)nmu), b((std::size_t)nmu);
  for (int k = 0; k < nmu; ++k) {
    a.emin[k] = run[k], a.z[k] = run[nmu + k], a.n1[k] = run[2 * nmu + k], a.n2[k] = run[3 * nmu + k];
    b.emin[k] = part[k], b.z[k] = part[nmu + k], b.n1[k] = part[2 * nmu + k], b.n2[k] = part[3 * nmu + k];
  }
  ftlmh::merge_thermo(a, b, beta);
  for (int k = 0; k < nmu; ++k)
    run[k] = a.emin[k], run[nmu + k] = a.z[k], run[2 * nmu + k] = a.n1[k], run[3 * nmu + k] = a.n2[k];
}

}  // extern "C"
