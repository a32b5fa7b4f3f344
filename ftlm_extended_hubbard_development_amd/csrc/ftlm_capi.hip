// ftlm_capi.hip -- C ABI of the FTLM engine (declared in include/ftlm.h).
//
// Host orchestration of the device kernels: sector and block tables,
// Hamiltonian application, the batched Lanczos driver, the device Jacobi
// tridiagonal solve and the batched exact eigensolver. Everything numeric
// runs on the GPU except two inputs that must match the reference bit for
// bit and are therefore produced by the same host library calls as there:
// the twist / momentum phases (glibc cos, sin; main.cpp:544-577, 736-743)
// and the seeded Gaussian start vectors (libstdc++ mt19937_64 +
// normal_distribution; main.cpp:1015-1025).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ftlm.h"
#include "ftlm_launch.h"

using namespace ftlm;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                               \
  do {                                                                                               \
    const hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                            \
      throw Error(e_ == hipErrorOutOfMemory ? FTLM_ERR_MEMORY : FTLM_ERR_DEVICE,                     \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                                \
  } while (0)

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return FTLM_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return FTLM_ERR_MEMORY;
  } catch (const std::exception& e) {
    g_err = e.what();
    return FTLM_ERR_ARG;
  }
}

// grow-only device buffer
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void* get(size_t n) {
    if (n == 0) n = 8;
    if (n > bytes) {
      if (p) HIPCHECK(hipFree(p));
      p = nullptr;
      bytes = 0;
      HIPCHECK(hipMalloc(&p, n));
      bytes = n;
    }
    return p;
  }
  template <typename T>
  T* as(size_t count) {
    return static_cast<T*>(get(count * sizeof(T)));
  }
  template <typename T>
  T* ptr() const {
    return static_cast<T*>(p);
  }
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
  }
};

constexpr double kPi = 3.141592653589793238462643383279502884;

uint64_t binom_u64(int n, int k) {
  if (k < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  unsigned __int128 r = 1;
  for (int i = 1; i <= k; ++i) r = r * (unsigned)(n - k + i) / (unsigned)i;
  if (r > (unsigned __int128)UINT64_MAX) return UINT64_MAX;
  return (uint64_t)r;
}

// complex product with separately rounded products, like the reference
inline void cmul_host(double ar, double ai, double br, double bi, double& cr, double& ci) {
  cr = ar * br - ai * bi;
  ci = ar * bi + ai * br;
}

// per-momentum tables of one block
struct BlockTables {
  int mx = -1, my = -1;
  int64_t dim = 0;
  DevBuf offs, nt_loc, nt_dnf, nt_list, nt_listoff, ptab, phase, act, counts, nt_count, tmp;
  BlockView view{};
};

// bump allocator over large device chunks (kept across packs)
struct Arena {
  std::vector<std::pair<void*, size_t>> chunks;
  size_t cur = 0, off = 0, used = 0;
  Arena() = default;
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  ~Arena() {
    for (auto& ch : chunks) (void)hipFree(ch.first);
  }
  void reset() {
    cur = 0;
    off = 0;
    used = 0;
  }
  void* alloc(size_t n) {
    n = (n + 255) & ~(size_t)255;
    if (n == 0) n = 256;
    while (cur < chunks.size()) {
      if (off + n <= chunks[cur].second) {
        void* p = static_cast<char*>(chunks[cur].first) + off;
        off += n;
        used += n;
        return p;
      }
      ++cur;
      off = 0;
    }
    const size_t sz = std::max(n, (size_t)1 << 30);
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, sz));
    chunks.push_back({p, sz});
    cur = chunks.size() - 1;
    off = n;
    used += n;
    return p;
  }
};

// grow-only pinned host buffer (hipHostMalloc)
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void* get(size_t n) {
    if (n > bytes) {
      if (p) HIPCHECK(hipHostFree(p));
      p = nullptr;
      bytes = 0;
      HIPCHECK(hipHostMalloc(&p, n, hipHostMallocDefault));
      bytes = n;
    }
    return p;
  }
};

// blocks batched into one persistent launch (ftlm_pack_*)
struct PackState {
  bool open = false;
  int R = 0, steps = 0;
  uint64_t budget = 0;
  Arena arena;
  std::vector<PackEntry> entries;
  std::vector<double> cost;       // per entry: nnz + 2 dim (rows gathered per column and step)
  std::vector<double> alg_bytes;  // per entry: kdim * dim * R * 16 * (hbar + 2)
  std::vector<std::pair<std::pair<uint64_t, int64_t>, const double2*>> starts;  // (seed, dim) -> start vectors
  int64_t max_dim = 0;
  // results of the last run
  bool has_result = false;
  int rcap = 0;
  std::vector<double> eig, w;
  std::vector<int32_t> m;
  // tuned launch shape
  int best_w = 0, best_wg = 0;
  // reductions of the next launches: 1 tree (QL runs), 0 row order (Jacobi runs)
  int tree_red = 1;
  DevBuf d_entries, d_items, d_nitems, queue, fail, alpha, beta, mcount, eig_d, w_d, scratch, iscratch, slots;
  PinnedBuf stage;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> exec_key;
  PackState() = default;
  PackState(const PackState&) = delete;
  PackState& operator=(const PackState&) = delete;
  ~PackState() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

// Exact blocks of any sector and twist queued for one batched eigensolve:
// each block's Hermitised dense matrix is built when it is added (its
// sector tables are current then) into a growing device arena.
struct ExactQueue {
  BlockTables tab;              // tables of the block being added
  DevBuf mats, work, eig, moff_d, dims_d, eoff_d, view_d, moff1_d;
  size_t mats_cap = 0;          // doubles held by mats
  uint64_t budget = 0;          // bytes of matrices per run
  std::vector<int64_t> moff{0}, dims, eoff{0};
  std::vector<double> eig_host;
  int64_t solved = 0;           // entries solved by earlier runs
  bool has_result = false;
};

}  // namespace

struct ftlm_ctx {
  int device = 0;
  int cus = 256;
  hipStream_t stream = nullptr;
  ftlm_model model{};
  LatticeConst g{};
  DevBuf g_dev, binom_dev;
  std::vector<uint64_t> binom_host;
  double amp_re[8], amp_im[8];
  // sector
  bool has_sector = false;
  int nup = -1, ndn = -1;
  int64_t n_upc = 0, n_dnc = 0;
  int n_canon = 0, n_nt = 0;
  int64_t parents = 0;
  DevBuf is_canon, min_rank, upc_tfirst, upc_tsign, upc_amask, cidx, cnt, upc_canon, can_bits, can_stab, nt_flag,
      nt_index, can_nt, nt_to_c;
  // blocks
  BlockTables cur;
  std::vector<std::unique_ptr<BlockTables>> pool;
  // Lanczos
  DevBuf vec[3], part, counter, st_alpha, st_beta, st_scale, st_sprev, st_bprev, st_dot, st_nrm, st_done, st_mcount,
      ritz_eig, ritz_w, ritz_scratch, ritz_iscratch, ritz_fail, pvec[3], pristine;
  int last_reduce = FTLM_REDUCE_TREE;
  // placement of the Lanczos vectors: when they are (re)allocated at
  // place_min_bytes or more, the next batched run times the hop-list apply
  // from place_candidates buffers and keeps the fastest two as the gather
  // sources (0 candidates: off)
  int place_candidates = 8;
  uint64_t place_min_bytes = (uint64_t)256 << 20;
  bool place_pending = false;
  std::vector<double> place_ms;  // last tuning: apply ms per candidate
  std::vector<int> place_pick;   // candidates chosen for vec[0], vec[1], vec[2]
  int prepared_R = 0;
  int prepared_layout = -1;  // 0: [dim][R], 1: [R][dim]
  // H-apply configuration: path 0 automatic (hop list), 1 hop list, 2 on the
  // fly; column tile width -1 automatic, 0 untiled, W > 0 tiles of W
  int apply_path = 0;
  int apply_w = -1;
  // R = 1 (per-vector) runs: the on-the-fly kernel is the faster plain path
  int apply_path_r1 = 2;
  // hop list of the current block
  DevBuf hl_rowptr, hl_col, hl_code, hl_spec, hl_diag, hl_cnt, hl_scnt, hl_srowptr, hl_scan_tmp;
  int64_t hl_key = -1;
  // column-tile tasks of the hop-list gather (TileTask), per block, R and W
  DevBuf tile_tasks;
  int64_t tile_key_gen = -1, n_tile_tasks = 0;
  int tile_key_R = 0, tile_key_W = 0;
  int64_t block_gen = 0;
  HopList hl{};
  DevBuf item_tiles;
  int plan_key_w = -2, plan_key_R = 0;
  int64_t plan_key_dim = -1;
  int plan_key_mx = -1, plan_key_my = -1, plan_key_nup = -1, plan_key_ndn = -1;
  ApplyPlan plan{};
  ftlm_stats stats{};
  // exact
  DevBuf ex_mats, ex_work, ex_eig, ex_moff, ex_dims, ex_eoff, ex_views;
  // packed blocks
  PackState pack;
  // thermodynamic sums
  DevBuf th_e, th_w, th_n, th_beta, th_mu, th_emin, th_part, th_out;
  // queued exact blocks
  ExactQueue exq;
  // where seeded start vectors are drawn: FTLM_START_HOST or FTLM_START_DEVICE
  int start_source = FTLM_START_HOST;
  DevBuf start_norms, start_one;

  ~ftlm_ctx() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// hopping directions with non-zero amplitude, and their amplitudes
// -t * twist (main.cpp:544-577, 683-687) for fermion sign +1
void set_amplitudes(ftlm_ctx* c) {
  const ftlm_model& m = c->model;
  LatticeConst& g = c->g;
  const double ax = 2.0 * kPi * m.phix / (double)m.lx;
  const double ay = 2.0 * kPi * m.phiy / (double)m.ly;
  const double ang[8] = {ax, -ax, ay, -ay, ax + ay, -ax + ay, ax - ay, -ax - ay};
  g.ndir = 0;
  for (int d = 0; d < 8; ++d) {
    const double t = d < 2 ? m.tx : (d < 4 ? m.ty : m.tp);
    const double a = -t * 1.0;
    const double cr = std::cos(ang[d]), ci = std::sin(ang[d]);
    c->amp_re[d] = a * cr;
    c->amp_im[d] = a * ci;
    if (t != 0.0) g.dirs[g.ndir++] = (int8_t)d;
  }
}

void build_lattice(ftlm_ctx* c) {
  const ftlm_model& m = c->model;
  LatticeConst& g = c->g;
  std::memset(&g, 0, sizeof(g));
  g.lx = m.lx;
  g.ly = m.ly;
  g.L = m.lx * m.ly;
  const int lx = m.lx, ly = m.ly;
  auto at = [lx, ly](int x, int y) { return ((y % ly + ly) % ly) * lx + ((x % lx + lx) % lx); };
  std::vector<std::pair<int, int>> bonds;
  for (int y = 0; y < ly; ++y)
    for (int x = 0; x < lx; ++x) {
      const int s = at(x, y);
      const int nbv[8] = {at(x + 1, y),     at(x - 1, y),     at(x, y + 1),     at(x, y - 1),
                          at(x + 1, y + 1), at(x - 1, y + 1), at(x + 1, y - 1), at(x - 1, y - 1)};
      for (int d = 0; d < 8; ++d) g.nb[s][d] = (uint8_t)nbv[d];
      bonds.push_back({s, at(x + 1, y)});
      bonds.push_back({s, at(x, y + 1)});
    }
  std::sort(bonds.begin(), bonds.end());
  bonds.erase(std::unique(bonds.begin(), bonds.end()), bonds.end());
  g.nbonds = (int)bonds.size();
  for (size_t i = 0; i < bonds.size(); ++i) {
    g.bonds[i][0] = (uint8_t)bonds[i].first;
    g.bonds[i][1] = (uint8_t)bonds[i].second;
  }
  g.full = g.L == 64 ? ~0ull : ((1ull << g.L) - 1);
  for (int k = 0; k <= lx; ++k) {
    uint64_t mk = 0;
    for (int y = 0; y < ly; ++y)
      for (int x = k; x < lx; ++x) mk |= 1ull << (y * lx + x);
    g.col_ge[k] = mk;
  }
  for (int k = 0; k <= ly; ++k) {
    uint64_t mk = 0;
    for (int y = 0; y < k; ++y)
      for (int x = 0; x < lx; ++x) mk |= 1ull << (y * lx + x);
    g.rows_lt[k] = mk;
  }
  g.U = m.u;
  g.V = m.v;
  set_amplitudes(c);
}

void upload_ctx_tables(ftlm_ctx* c) {
  HIPCHECK(hipMemcpy(c->g_dev.get(sizeof(LatticeConst)), &c->g, sizeof(LatticeConst), hipMemcpyHostToDevice));
  c->binom_host.assign(64 * 64, 0);
  for (int k = 0; k < 64; ++k)
    for (int p = 0; p < 64; ++p) c->binom_host[k * 64 + p] = binom_u64(p, k);
  HIPCHECK(hipMemcpy(c->binom_dev.get(64 * 64 * 8), c->binom_host.data(), 64 * 64 * 8, hipMemcpyHostToDevice));
}

const LatticeConst* gdev(ftlm_ctx* c) { return c->g_dev.ptr<LatticeConst>(); }

void sync(ftlm_ctx* c) { HIPCHECK(hipStreamSynchronize(c->stream)); }

template <typename T>
T read1(ftlm_ctx* c, const void* dptr) {
  T v;
  HIPCHECK(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  return v;
}

void need_sector(ftlm_ctx* c) {
  if (!c->has_sector) throw Error(FTLM_ERR_STATE, "no sector selected (call ftlm_sector first)");
}
void need_block(ftlm_ctx* c) {
  need_sector(c);
  if (c->cur.mx < 0) throw Error(FTLM_ERR_STATE, "no block selected (call ftlm_block first)");
}

void build_sector(ftlm_ctx* c, int nup, int ndn) {
  const int L = c->g.L;
  if (nup < 0 || ndn < 0 || nup > L || ndn > L) throw Error(FTLM_ERR_ARG, "particle numbers out of range");
  const uint64_t nu = binom_u64(L, nup), nd = binom_u64(L, ndn);
  if (nu > (uint64_t)INT32_MAX || nd > (uint64_t)INT32_MAX)
    throw Error(FTLM_ERR_LIMIT, "spin configuration count exceeds 2^31");
  c->has_sector = false;
  c->cur.mx = c->cur.my = -1;
  ++c->block_gen;
  c->nup = nup;
  c->ndn = ndn;
  c->n_upc = (int64_t)nu;
  c->n_dnc = (int64_t)nd;
  hipStream_t s = c->stream;
  const uint64_t* binom = c->binom_dev.ptr<uint64_t>();
  uint8_t* is_canon = c->is_canon.as<uint8_t>(nu);
  int64_t* min_rank = c->min_rank.as<int64_t>(nu);
  uint8_t* tfirst = c->upc_tfirst.as<uint8_t>(nu);
  int8_t* tsg = c->upc_tsign.as<int8_t>(nu);
  uint64_t* amask = c->upc_amask.as<uint64_t>(nu);
  int32_t* cidx = c->cidx.as<int32_t>(nu);
  int64_t* cnt = c->cnt.as<int64_t>(4);
  launch_up_table(gdev(c), binom, nup, (int64_t)nu, is_canon, min_rank, tfirst, tsg, amask, s);
  launch_segscan(is_canon, (int64_t)nu, 1, cidx, cnt, s);
  HIPCHECK(hipGetLastError());
  c->n_canon = (int)read1<int64_t>(c, cnt);
  const int nc = c->n_canon;
  int32_t* upc_canon = c->upc_canon.as<int32_t>(nu);
  uint64_t* can_bits = c->can_bits.as<uint64_t>(nc);
  uint64_t* can_stab = c->can_stab.as<uint64_t>(nc);
  uint8_t* nt_flag = c->nt_flag.as<uint8_t>(nc);
  launch_canon_fill((int64_t)nu, nup, L, binom, is_canon, cidx, min_rank, amask, upc_canon, can_bits, can_stab,
                    nt_flag, s);
  int32_t* nt_index = c->nt_index.as<int32_t>(nc);
  launch_segscan(nt_flag, nc, 1, nt_index, cnt, s);
  HIPCHECK(hipGetLastError());
  c->n_nt = (int)read1<int64_t>(c, cnt);
  int32_t* can_nt = c->can_nt.as<int32_t>(nc);
  int32_t* nt_to_c = c->nt_to_c.as<int32_t>(c->n_nt > 0 ? c->n_nt : 1);
  launch_nt_fill(nc, nt_index, can_nt, nt_to_c, s);
  HIPCHECK(hipGetLastError());
  c->has_sector = true;
}

BlockView base_view(ftlm_ctx* c) {
  BlockView v{};
  v.nup = c->nup;
  v.ndn = c->ndn;
  v.n_canon = c->n_canon;
  v.n_dnc = c->n_dnc;
  v.upc_canon = c->upc_canon.ptr<int32_t>();
  v.upc_tfirst = c->upc_tfirst.ptr<uint8_t>();
  v.upc_tsign = c->upc_tsign.ptr<int8_t>();
  v.upc_amask = c->upc_amask.ptr<uint64_t>();
  v.can_bits = c->can_bits.ptr<uint64_t>();
  v.can_stab = c->can_stab.ptr<uint64_t>();
  v.can_nt = c->can_nt.ptr<int32_t>();
  v.binom = c->binom_dev.ptr<uint64_t>();
  return v;
}

// Per-momentum tables (main.cpp:736-805). Returns the block dimension.
void build_block(ftlm_ctx* c, BlockTables& b, int mx, int my) {
  const LatticeConst& g = c->g;
  if (mx < 0 || my < 0 || mx >= g.lx || my >= g.ly) throw Error(FTLM_ERR_ARG, "momentum out of range");
  hipStream_t s = c->stream;
  const int L = g.L;
  // momentum phases e^{i 2 pi (mx dx/lx + my dy/ly)} for shift j = dy*lx+dx
  std::vector<double> ph(2 * L), pt(16 * L);
  for (int j = 0; j < L; ++j) {
    const int dx = j % g.lx, dy = j / g.lx;
    const double arg =
        2.0 * kPi * (static_cast<double>(mx * dx) / static_cast<double>(g.lx) +
                     static_cast<double>(my * dy) / static_cast<double>(g.ly));
    ph[2 * j] = std::cos(arg);
    ph[2 * j + 1] = std::sin(arg);
  }
  for (int d = 0; d < 8; ++d)
    for (int j = 0; j < L; ++j)
      cmul_host(c->amp_re[d], c->amp_im[d], ph[2 * j], ph[2 * j + 1], pt[2 * (d * L + j)], pt[2 * (d * L + j) + 1]);
  double* phd = b.phase.as<double>(2 * L);
  double* ptd = b.ptab.as<double>(16 * L);
  HIPCHECK(hipMemcpyAsync(phd, ph.data(), ph.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(ptd, pt.data(), pt.size() * 8, hipMemcpyHostToDevice, s));
  const int nnt = c->n_nt;
  const int64_t ndc = c->n_dnc;
  const int64_t ntot = (int64_t)nnt * ndc;
  int32_t* nt_loc = b.nt_loc.as<int32_t>(ntot > 0 ? ntot : 1);
  double* nt_dnf = b.nt_dnf.as<double>(ntot > 0 ? ntot : 1);
  int64_t* nt_count = b.nt_count.as<int64_t>(nnt + 1);
  int64_t* listoff = b.nt_listoff.as<int64_t>(nnt + 1);
  int64_t list_total = 0;
  if (nnt > 0) {
    uint8_t* act = b.act.as<uint8_t>(ntot);
    launch_nt_active(gdev(c), c->binom_dev.ptr<uint64_t>(), c->ndn, ndc, nnt, c->nt_to_c.ptr<int32_t>(),
                     c->can_bits.ptr<uint64_t>(), c->can_stab.ptr<uint64_t>(), phd, act, nt_dnf, s);
    launch_segscan(act, ndc, nnt, nt_loc, nt_count, s);
    launch_prefix64(nt_count, nnt, listoff, s);
    HIPCHECK(hipGetLastError());
    list_total = read1<int64_t>(c, listoff + nnt);
  }
  int32_t* list = b.nt_list.as<int32_t>(list_total > 0 ? list_total : 1);
  if (nnt > 0) launch_nt_list(nnt, ndc, nt_loc, listoff, list, s);
  int64_t* counts = b.counts.as<int64_t>(c->n_canon);
  int64_t* offs = b.offs.as<int64_t>(c->n_canon + 1);
  launch_row_counts(c->n_canon, ndc, c->can_nt.ptr<int32_t>(), nt_count, counts, s);
  launch_prefix64(counts, c->n_canon, offs, s);
  HIPCHECK(hipGetLastError());
  b.dim = read1<int64_t>(c, offs + c->n_canon);
  b.mx = mx;
  b.my = my;
  BlockView v = base_view(c);
  v.dim = b.dim;
  v.offs = offs;
  v.nt_loc = nt_loc;
  v.nt_dnf = nt_dnf;
  v.nt_list = list;
  v.nt_listoff = listoff;
  v.ptab = ptd;
  b.view = v;
}

// Number of translation orbits (the reference's parents.size()):
// trivial-stabiliser canonical ups own every down; an nt up owns the downs
// that are minimal under its stabiliser. Counted on the host from the
// device tables (only the few nt ups need work).
int64_t count_parents(ftlm_ctx* c) {
  if (c->n_nt == 0) return (int64_t)c->n_canon * c->n_dnc;
  std::vector<uint64_t> cs(c->n_canon);
  std::vector<int32_t> n2c(c->n_nt);
  HIPCHECK(hipMemcpy(cs.data(), c->can_stab.p, 8 * c->n_canon, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(n2c.data(), c->nt_to_c.p, 4 * c->n_nt, hipMemcpyDeviceToHost));
  const LatticeConst& g = c->g;
  const int L = g.L;
  auto tr = [&](uint64_t b, int i) {
    const int dx = i % g.lx, dy = i / g.lx;
    if (dx) b = ((b << dx) & g.col_ge[dx]) | ((b >> (g.lx - dx)) & ~g.col_ge[dx] & g.full);
    if (dy) {
      const int sft = g.lx * dy;
      b = ((b << sft) | (b >> (L - sft))) & g.full;
    }
    return b;
  };
  std::vector<uint64_t> downs;
  downs.reserve(c->n_dnc);
  if (c->ndn == 0) {
    downs.push_back(0);
  } else {
    uint64_t st = (1ull << c->ndn) - 1;
    const uint64_t lim = 1ull << L;
    while (st < lim) {
      downs.push_back(st);
      const uint64_t low = st & (~st + 1);
      const uint64_t rip = st + low;
      if (rip == 0) break;
      st = (((rip ^ st) >> 2) / low) | rip;
    }
  }
  int64_t reps = 0;
  for (int x = 0; x < c->n_nt; ++x) {
    const uint64_t stab = cs[n2c[x]] & ~1ull;
    for (uint64_t d : downs) {
      bool rep = true;
      uint64_t m = stab;
      while (m) {
        const int i = __builtin_ctzll(m);
        m &= m - 1;
        if (tr(d, i) < d) {
          rep = false;
          break;
        }
      }
      reps += rep;
    }
  }
  return (int64_t)(c->n_canon - c->n_nt) * c->n_dnc + reps;
}

// ---- seeded start vectors (host; bit-identical to main.cpp:1015-1025) ----
uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// one sample: Gaussian pairs with the real part taken from the SECOND draw
// (g++ evaluates Complex(normal(rng), normal(rng)) right to left), then
// division by the serially summed norm
void gen_start(uint64_t bseed, int sample, int64_t dim, double* out) {
  std::mt19937_64 rng(mix64(bseed ^ (uint64_t)sample));
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int64_t i = 0; i < dim; ++i) {
    const double first = nd(rng);
    const double second = nd(rng);
    out[2 * i] = second;
    out[2 * i + 1] = first;
  }
  double s = 0.0;
  for (int64_t i = 0; i < dim; ++i) s += out[2 * i] * out[2 * i] + out[2 * i + 1] * out[2 * i + 1];
  const double nrm = std::sqrt(s);
  if (nrm <= 0.0) throw Error(FTLM_ERR_ARG, "Attempted to normalize a zero vector.");
  for (int64_t i = 0; i < 2 * dim; ++i) out[i] /= nrm;
}

// Host threads for start-vector generation: FTLM_HOST_THREADS, else
// OMP_NUM_THREADS, else the hardware count, capped at 64.
int host_threads() {
  for (const char* var : {"FTLM_HOST_THREADS", "OMP_NUM_THREADS"}) {
    if (const char* e = std::getenv(var)) {
      const int n = std::atoi(e);
      if (n > 0) return std::min(n, 64);
    }
  }
  const unsigned n = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(n, 64u));
}

// Generate samples r0..r0+cnt-1 into host memory (sample-major), threaded
// over samples.
void gen_starts(uint64_t bseed, int r0, int cnt, int64_t dim, double* host) {
  const int nt = std::min(host_threads(), cnt);
  std::vector<std::thread> th;
  std::vector<std::string> errs(nt);
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      try {
        for (int j = t; j < cnt; j += nt) gen_start(bseed, r0 + j, dim, host + 2 * (int64_t)j * dim);
      } catch (const std::exception& e) {
        errs[t] = e.what();
      }
    });
  for (auto& x : th) x.join();
  for (auto& e : errs)
    if (!e.empty()) throw Error(FTLM_ERR_ARG, e);
}

// device vector buffers for R columns of the current block
void ensure_vectors(ftlm_ctx* c, int R) {
  const size_t bytes = (size_t)c->cur.dim * R * 16;
  bool grown = false;
  for (int i = 0; i < 3; ++i) {
    grown = grown || bytes > c->vec[i].bytes;
    c->vec[i].get(bytes);
  }
  if (grown && c->place_candidates >= 3 && bytes >= c->place_min_bytes) c->place_pending = true;
}

// upload seeded start vectors of samples sample0..sample0+R-1 into layout 0
// ([dim][R] in vec[0]) or layout 1 (sample-major [R][dim] in vec[0])
void prepare_starts(ftlm_ctx* c, int R, uint64_t bseed, int layout, int sample0 = 0) {
  const int64_t dim = c->cur.dim;
  ensure_vectors(c, R);
  if (c->start_source == FTLM_START_DEVICE) {
    launch_start_vectors(bseed, sample0, 0, R, R, dim, layout, c->vec[0].ptr<double>(), c->start_norms.as<double>(R),
                         c->stream);
    HIPCHECK(hipGetLastError());
    sync(c);
    c->prepared_R = R;
    c->prepared_layout = layout;
    return;
  }
  // pinned staging of at most ~2 GB (or one sample)
  const int64_t fit = std::max<int64_t>(1, (int64_t)(2ll << 30) / std::max<int64_t>(1, dim * 16));
  const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)R, (int64_t)host_threads(), fit}));
  double* host = nullptr;
  HIPCHECK(hipHostMalloc(&host, (size_t)chunk * dim * 16, hipHostMallocDefault));
  std::unique_ptr<double, void (*)(double*)> guard(host, [](double* p) { (void)hipHostFree(p); });
  double* x = c->vec[0].ptr<double>();
  double* stage = c->vec[2].ptr<double>();
  for (int r0 = 0; r0 < R; r0 += chunk) {
    const int cnt = std::min(chunk, R - r0);
    gen_starts(bseed, sample0 + r0, cnt, dim, host);
    if (layout == 1) {
      HIPCHECK(hipMemcpyAsync(x + 2 * (int64_t)r0 * dim, host, (size_t)cnt * dim * 16, hipMemcpyHostToDevice,
                              c->stream));
    } else {
      HIPCHECK(hipMemcpyAsync(stage, host, (size_t)cnt * dim * 16, hipMemcpyHostToDevice, c->stream));
      launch_transpose_in(dim, R, r0, cnt, stage, x, c->stream);
    }
    sync(c);
  }
  c->prepared_R = R;
  c->prepared_layout = layout;
}

LanczosState lanczos_state(ftlm_ctx* c, int R, int steps) {
  LanczosState st;
  st.alpha = c->st_alpha.as<double>((size_t)steps * R);
  st.beta = c->st_beta.as<double>((size_t)steps * R);
  st.scale = c->st_scale.as<double>(R);
  st.sprev = c->st_sprev.as<double>(R);
  st.bprev = c->st_bprev.as<double>(R);
  st.dot = c->st_dot.as<double>(R);
  st.nrm = c->st_nrm.as<double>(R);
  st.done = c->st_done.as<int32_t>(R);
  st.mcount = c->st_mcount.as<int32_t>(R);
  return st;
}

// Column tile width for R vectors on the current block. Automatic choice:
// the widest power of two (>= 8) whose largest group slab rows * W * 16 B
// stays within 2 MB of the XCD's 4 MB L2; untiled when that covers R.
int auto_tile(ftlm_ctx* c, int R, int64_t max_group) {
  if (R < 16) return 0;
  int w = 8;
  while (w * 2 < R && (double)max_group * (w * 2) * 16.0 <= 2.0 * 1024 * 1024) w *= 2;
  return w >= R ? 0 : w;
}

const ApplyPlan& apply_plan(ftlm_ctx* c, int R, int w_req) {
  BlockTables& b = c->cur;
  std::vector<int64_t> offs(c->n_canon + 1);
  int w = w_req;
  if (w < 0 || (w > 0 && w >= R) || c->plan_key_dim != b.dim || c->plan_key_mx != b.mx || c->plan_key_my != b.my ||
      c->plan_key_nup != c->nup || c->plan_key_ndn != c->ndn || c->plan_key_R != R || c->plan_key_w != w_req) {
    HIPCHECK(hipMemcpy(offs.data(), b.offs.p, 8 * (c->n_canon + 1), hipMemcpyDeviceToHost));
  } else {
    return c->plan;
  }
  int64_t maxg = 0;
  for (int k = 0; k < c->n_canon; ++k) maxg = std::max(maxg, offs[k + 1] - offs[k]);
  if (w < 0) w = auto_tile(c, R, maxg);
  if (w <= 0 || w >= R) {
    c->plan = plan_apply(c->g, c->nup, c->ndn, b.dim, R, c->cus);
    c->plan.W = 0;
  } else {
    std::vector<int64_t> it;
    c->plan = plan_apply_tiled(c->g, c->nup, c->ndn, offs.data(), c->n_canon, R, w, c->cus, it);
    int64_t* d = c->item_tiles.as<int64_t>(it.size());
    HIPCHECK(hipMemcpy(d, it.data(), it.size() * 8, hipMemcpyHostToDevice));
    c->plan.item_tiles = d;
  }
  c->plan_key_dim = b.dim;
  c->plan_key_mx = b.mx;
  c->plan_key_my = b.my;
  c->plan_key_nup = c->nup;
  c->plan_key_ndn = c->ndn;
  c->plan_key_R = R;
  c->plan_key_w = w_req;
  return c->plan;
}

// Hop list of the current block (built on the device once per block).
const HopList& ensure_hops(ftlm_ctx* c) {
  if (c->hl_key == c->block_gen) return c->hl;
  BlockTables& b = c->cur;
  const int64_t n = b.dim;
  hipStream_t s = c->stream;
  int64_t* cnt = c->hl_cnt.as<int64_t>(n + 1);
  int64_t* scnt = c->hl_scnt.as<int64_t>(n + 1);
  int64_t* rp = c->hl_rowptr.as<int64_t>(n + 1);
  int64_t* srp = c->hl_srowptr.as<int64_t>(n + 1);
  launch_hops_count(gdev(c), b.view, cnt, scnt, s);
  int64_t* tmp = c->hl_scan_tmp.as<int64_t>((size_t)prefix64_tmp_words(n));
  launch_prefix64_tiled(cnt, n, rp, tmp, s);
  launch_prefix64_tiled(scnt, n, srp, tmp, s);
  HIPCHECK(hipGetLastError());
  const int64_t nnz = read1<int64_t>(c, rp + n);
  const int64_t nsp = read1<int64_t>(c, srp + n);
  if (nnz > INT64_C(0x7fffffff) * 8) throw Error(FTLM_ERR_LIMIT, "hop list too large");
  int32_t* col = c->hl_col.as<int32_t>(nnz + 1);
  uint32_t* code = c->hl_code.as<uint32_t>(nnz + 1);
  double* spec = c->hl_spec.as<double>(2 * nsp + 2);
  double* diag = c->hl_diag.as<double>(n + 1);
  launch_hops_fill(gdev(c), b.view, rp, srp, col, code, spec, diag, s);
  HIPCHECK(hipGetLastError());
  sync(c);
  c->hl.row_ptr = rp;
  c->hl.col = col;
  c->hl.code = code;
  c->hl.spec = spec;
  c->hl.diag = diag;
  c->hl.ptab = b.view.ptab;
  c->hl_key = c->block_gen;
  return c->hl;
}

struct ApplyCfg {
  int path;  // 1 hop list, 2 on the fly
  int w;     // 0 untiled, else tile width
};

ApplyCfg resolve_cfg(ftlm_ctx* c, int R, int path, int w) {
  ApplyCfg a;
  if (R == 1 && path == c->apply_path) path = c->apply_path_r1;
  a.path = path == 2 ? 2 : 1;
  if (w < 0) {
    std::vector<int64_t> offs(c->n_canon + 1);
    HIPCHECK(hipMemcpy(offs.data(), c->cur.offs.p, 8 * (c->n_canon + 1), hipMemcpyDeviceToHost));
    int64_t maxg = 0;
    for (int k = 0; k < c->n_canon; ++k) maxg = std::max(maxg, offs[k + 1] - offs[k]);
    w = a.path == 1 ? 0 : auto_tile(c, R, maxg);
  }
  a.w = (w >= R) ? 0 : w;
  return a;
}

// y = H x for R columns with configuration cfg; partials/dot_out may be null.
// Returns true when <x, y> was reduced into dot_out by the apply itself.
// Tasks of the lane-dense tiled gather for W in {8, 16, 32, 64}: runs of
// tile_task_rows(W) rows of one canonical-up group and one column tile,
// ordered (group, tile, rows).
const TileTask* ensure_tile_tasks(ftlm_ctx* c, int R, int W, int64_t* ntask) {
  if (c->tile_key_gen != c->block_gen || c->tile_key_R != R || c->tile_key_W != W) {
    std::vector<int64_t> offs(c->n_canon + 1);
    HIPCHECK(hipMemcpy(offs.data(), c->cur.offs.p, 8 * (c->n_canon + 1), hipMemcpyDeviceToHost));
    const int J = (R + W - 1) / W;
    const int ch = tile_task_rows(W);
    std::vector<TileTask> t;
    for (int g = 0; g < c->n_canon; ++g)
      for (int j = 0; j < J; ++j)
        for (int64_t r0 = offs[g]; r0 < offs[g + 1]; r0 += ch)
          t.push_back({r0, (int32_t)std::min<int64_t>(ch, offs[g + 1] - r0), j});
    TileTask* d = c->tile_tasks.as<TileTask>(std::max<size_t>(t.size(), 1));
    if (!t.empty()) HIPCHECK(hipMemcpy(d, t.data(), t.size() * sizeof(TileTask), hipMemcpyHostToDevice));
    c->n_tile_tasks = (int64_t)t.size();
    c->tile_key_gen = c->block_gen;
    c->tile_key_R = R;
    c->tile_key_W = W;
  }
  *ntask = c->n_tile_tasks;
  return c->tile_tasks.ptr<TileTask>();
}

bool lane_dense_tile(int R, int w) { return w > 0 && w < R && (w == 8 || w == 16 || w == 32 || w == 64); }

bool do_apply(ftlm_ctx* c, const ApplyCfg& cfg, int R, const double* x, double* y, double* partials,
              unsigned* counter, double* dot_out) {
  const int64_t dim = c->cur.dim;
  if (cfg.path == 1 && R > 1 && lane_dense_tile(R, cfg.w)) {
    const HopList& h = ensure_hops(c);
    int64_t nt = 0;
    const TileTask* tasks = ensure_tile_tasks(c, R, cfg.w, &nt);
    launch_apply_hl_tiles(h, R, cfg.w, tasks, nt, c->cus, x, y, c->stream);
    return false;
  }
  if (cfg.path == 1) {
    const HopList& h = ensure_hops(c);
    const bool fused = cfg.w == 0 || R == 1;
    launch_apply_hl(h, dim, R, cfg.w, c->cur.offs.ptr<int64_t>(), c->n_canon, c->cus, x, y,
                    fused ? partials : nullptr, counter, dot_out, c->stream);
    return fused && partials;
  }
  if (R == 1) {
    launch_apply_r1(gdev(c), c->g.L, c->cur.view, x, y, c->cus, partials, counter, dot_out, c->stream);
    return partials != nullptr;
  }
  const ApplyPlan& p = apply_plan(c, R, cfg.w);
  launch_apply(p, gdev(c), c->cur.view, x, y, partials, counter, dot_out, c->stream);
  return partials != nullptr;
}

// distinct non-zero-amplitude neighbour sites of one site (translation
// invariant), for the algorithmic byte count
double hbar(ftlm_ctx* c) {
  const LatticeConst& g = c->g;
  std::vector<int> sites;
  for (int di = 0; di < g.ndir; ++di) {
    const int t = g.nb[0][g.dirs[di]];
    if (t != 0 && std::find(sites.begin(), sites.end(), t) == sites.end()) sites.push_back(t);
  }
  const double d = (double)sites.size();
  const int L = g.L;
  if (L < 2) return 0.0;
  return d * ((double)c->nup * (L - c->nup) + (double)c->ndn * (L - c->ndn)) / (double)(L - 1);
}

// The Lanczos recurrence over R columns x (layout [dim][R]); x is the
// normalised start in vec-buffer `xi`. Results stay on the device in the
// LanczosState; `timed` records per-apply events into c->stats.
void lanczos_run(ftlm_ctx* c, int R, int steps, int reduce, double* x0, double* x1, double* w, LanczosState st,
                 bool r1_kernel, bool timed) {
  const int64_t dim = c->cur.dim;
  const int kdim = (int)std::min<int64_t>(steps, dim);
  hipStream_t s = c->stream;
  if (r1_kernel && R != 1) throw Error(FTLM_ERR_ARG, "one-row kernel needs R = 1");
  const ApplyCfg cfg = resolve_cfg(c, R, c->apply_path, c->apply_w);
  if (cfg.path == 1) ensure_hops(c);
  double* partials = c->part.as<double>((size_t)std::max(c->cus * 8, 8192) * R + 64);
  unsigned* counter = c->counter.as<unsigned>(4);
  HIPCHECK(hipMemsetAsync(counter, 0, 16, s));
  HIPCHECK(hipMemsetAsync(x1, 0, (size_t)dim * R * 16, s));
  launch_lanczos_init(R, st, s);
  const int rgrid = (int)std::max<int64_t>(1, std::min<int64_t>((dim * R + 255) / 256, (int64_t)c->cus * 8));
  std::vector<hipEvent_t> ev;
  if (timed) {
    ev.resize(2 * kdim + 2);
    for (auto& e : ev) HIPCHECK(hipEventCreate(&e));
    HIPCHECK(hipEventRecord(ev[2 * kdim], s));
  }
  const bool tree = reduce == FTLM_REDUCE_TREE;
  for (int k = 0; k < kdim; ++k) {
    if (timed) HIPCHECK(hipEventRecord(ev[2 * k], s));
    const bool fused = do_apply(c, cfg, R, x0, w, tree ? partials : nullptr, counter, st.dot);
    if (timed) HIPCHECK(hipEventRecord(ev[2 * k + 1], s));
    if (tree) {
      if (!fused) launch_dot_tree(dim, R, x0, w, partials, counter, st.dot, rgrid, s);
      launch_recur_tree(dim, R, w, x0, x1, k, kdim, st, partials, counter, rgrid, s);
    } else {
      launch_dot_serial(dim, R, x0, w, st.dot, s);
      launch_recur_serial(dim, R, w, x0, x1, st.dot, st.bprev, st.done, s);
      launch_norm_serial(dim, R, x1, st.nrm, s);
      launch_finish_serial(R, k, kdim, st, s);
      launch_normalize(dim, R, x1, st.beta + (size_t)k * R, st.done, k, st.mcount, s);
    }
    std::swap(x0, x1);
  }
  HIPCHECK(hipGetLastError());
  if (timed) {
    HIPCHECK(hipEventRecord(ev[2 * kdim + 1], s));
    sync(c);
    double tot = 0.0;
    for (int k = 0; k < kdim; ++k) {
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
      tot += ms;
    }
    float all = 0.f;
    HIPCHECK(hipEventElapsedTime(&all, ev[2 * kdim], ev[2 * kdim + 1]));
    c->stats.ms_apply += tot;
    c->stats.ms_total += all;
    c->stats.n_apply += kdim;
    c->stats.bytes_apply += (double)kdim * (double)dim * R * 16.0 * (hbar(c) + 2.0);
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
}

// Placement of the Lanczos vectors. The batched hop-list gather reads ~35
// random rows of x per row of y, and how fast the HBM path serves them
// depends on where x physically lies: the same apply on the same data takes
// from ~117 to ~132 ms on the 4x4 (8,8) block depending on the buffer x is
// in, reproducibly for that buffer, and the output buffer moves it too.
// Called once after the vectors were (re)allocated: time one apply from each
// of the three vector buffers and up to place_candidates - 3 extra ones
// (what free memory allows, minus a reserve), keep the fastest two as vec[0]
// and vec[1] (the buffers the recurrence alternates as the gather source),
// then time the applies from both into each remaining candidate and keep the
// fastest as vec[2] (the output); free the rest. vec[0] holds the prepared start vectors; they are
// copied into the chosen buffer. Results do not depend on the choice.
void tune_placement(ftlm_ctx* c, int R) {
  c->place_pending = false;
  const ApplyCfg cfg = resolve_cfg(c, R, c->apply_path, c->apply_w);
  if (cfg.path != 1 || cfg.w != 0 || R < 2 || c->cur.dim == 0) return;
  const int64_t dim = c->cur.dim;
  const size_t need = (size_t)dim * R * 16;
  const size_t cap = c->vec[0].bytes;
  if (c->vec[1].bytes != cap || c->vec[2].bytes != cap || cap < need) return;
  const HopList& h = ensure_hops(c);
  size_t fr = 0, tot = 0;
  HIPCHECK(hipMemGetInfo(&fr, &tot));
  const size_t reserve = std::max<size_t>((size_t)4 << 30, tot / 20);
  std::vector<std::unique_ptr<DevBuf>> cand;
  for (int i = 0; i < 3; ++i) {
    cand.push_back(std::make_unique<DevBuf>());
    cand.back()->swap(c->vec[i]);
  }
  // an error before the choice gives the context its own buffers back
  struct Restore {
    ftlm_ctx* c;
    std::vector<std::unique_ptr<DevBuf>>& cand;
    bool armed = true;
    ~Restore() {
      if (armed)
        for (int i = 0; i < 3; ++i) c->vec[i].swap(*cand[i]);
    }
  } restore{c, cand};
  for (int i = 3; i < c->place_candidates; ++i) {
    if (fr < reserve + cap) break;
    void* p = nullptr;
    if (hipMalloc(&p, cap) != hipSuccess) {
      (void)hipGetLastError();
      break;
    }
    cand.push_back(std::make_unique<DevBuf>());
    cand.back()->p = p;
    cand.back()->bytes = cap;
    fr -= cap;
  }
  const int n = (int)cand.size();
  hipStream_t s = c->stream;
  // finite contents in every buffer but the start vectors (apply input only)
  for (int i = 1; i < n; ++i) launch_zero(cand[i]->ptr<double>(), (int64_t)(need / 8), s);
  const int64_t* offs = c->cur.offs.ptr<int64_t>();
  auto apply = [&](int xi, int yi) {
    launch_apply_hl(h, dim, R, 0, offs, c->n_canon, c->cus, cand[xi]->ptr<double>(), cand[yi]->ptr<double>(),
                    nullptr, nullptr, nullptr, s);
  };
  apply(0, 1);  // warm-up
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  auto timed = [&](int xi, int yi) {
    HIPCHECK(hipEventRecord(e0, s));
    apply(xi, yi);
    HIPCHECK(hipEventRecord(e1, s));
    HIPCHECK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHECK(hipEventElapsedTime(&t, e0, e1));
    return (double)t;
  };
  // the gather sources: each candidate as x (never writing cand[0], which
  // holds the start vectors)
  std::vector<double> ms(n);
  for (int i = 0; i < n; ++i) ms[i] = timed(i, i == 1 ? 2 : 1);
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ms[a] < ms[b]; });
  // the output: the remaining candidate with the least time of the applies
  // from both sources into it
  int out = -1;
  double best = 1e300;
  for (int k = 2; k < n; ++k) {
    const int j = order[k];
    if (j == 0) continue;  // keep the start vectors until they are copied
    const double t = timed(order[0], j) + timed(order[1], j);
    if (t < best) {
      best = t;
      out = j;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHECK(hipGetLastError());
  if (out < 0) out = order[2];
  if (order[0] != 0)
    HIPCHECK(hipMemcpyAsync(cand[order[0]]->p, cand[0]->p, need, hipMemcpyDeviceToDevice, s));
  sync(c);
  const int pick[3] = {order[0], order[1], out};
  restore.armed = false;
  for (int i = 0; i < 3; ++i) c->vec[i].swap(*cand[pick[i]]);
  c->place_ms = ms;
  c->place_pick.assign(pick, pick + 3);
}

// Lanczos of R columns stored in vec[0] (layout 0) or, per vector, of the
// sample-major columns of vec[0] (layout 1). Leaves alpha/beta/mcount of all
// R columns in the ctx state arrays ([steps][R]).
void run_all(ftlm_ctx* c, int R, int steps, int reduce, bool per_vector, bool timed) {
  const int64_t dim = c->cur.dim;
  LanczosState st = lanczos_state(c, R, steps);
  if (!per_vector) {
    if (c->place_pending) tune_placement(c, R);
    lanczos_run(c, R, steps, reduce, c->vec[0].ptr<double>(), c->vec[1].ptr<double>(), c->vec[2].ptr<double>(), st,
                false, timed);
    return;
  }
  // one column at a time through the R = 1 kernels
  double* p0 = c->pvec[0].as<double>((size_t)dim * 2);
  double* p1 = c->pvec[1].as<double>((size_t)dim * 2);
  double* p2 = c->pvec[2].as<double>((size_t)dim * 2);
  LanczosState one;
  DevBuf a1, b1, sc1, sp1, bp1, d1, n1, dn1, m1;
  one.alpha = a1.as<double>(steps);
  one.beta = b1.as<double>(steps);
  one.scale = sc1.as<double>(1);
  one.sprev = sp1.as<double>(1);
  one.bprev = bp1.as<double>(1);
  one.dot = d1.as<double>(1);
  one.nrm = n1.as<double>(1);
  one.done = dn1.as<int32_t>(1);
  one.mcount = m1.as<int32_t>(1);
  for (int r = 0; r < R; ++r) {
    HIPCHECK(hipMemcpyAsync(p0, c->vec[0].ptr<double>() + 2 * (int64_t)r * dim, (size_t)dim * 16,
                            hipMemcpyDeviceToDevice, c->stream));
    lanczos_run(c, 1, steps, reduce, p0, p1, p2, one, true, timed);
    // scatter column r of alpha/beta/mcount
    HIPCHECK(hipMemcpy2DAsync(st.alpha + r, R * 8, one.alpha, 8, 8, steps, hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipMemcpy2DAsync(st.beta + r, R * 8, one.beta, 8, 8, steps, hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(st.mcount + r, one.mcount, 4, hipMemcpyDeviceToDevice, c->stream));
  }
  sync(c);
}

// method 0: the reference's Jacobi (bit-identical); 1: implicit QL
void ritz_device(ftlm_ctx* c, int R, int steps, const double* alpha, const double* beta, const int32_t* mcount,
                 int wdiv, int method) {
  double* eig = c->ritz_eig.as<double>((size_t)R * steps);
  double* w = c->ritz_w.as<double>((size_t)R * steps);
  if (method == 0) {
    const size_t words = (size_t)steps * (steps + 1) / 2 + 3 * (size_t)steps;
    double* scratch = c->ritz_scratch.as<double>(words * R);
    launch_ritz(R, steps, alpha, beta, mcount, wdiv, eig, w, scratch, c->stream);
  } else {
    double* scratch = c->ritz_scratch.as<double>((size_t)3 * steps * R);
    int32_t* isc = c->ritz_iscratch.as<int32_t>((size_t)steps * R + 1);
    int32_t* fail = c->ritz_fail.as<int32_t>(1);
    HIPCHECK(hipMemsetAsync(fail, 0, 4, c->stream));
    launch_ritz_ql(R, steps, alpha, beta, mcount, wdiv, eig, w, scratch, isc, fail, c->stream);
    int32_t nf = 0;
    HIPCHECK(hipMemcpyAsync(&nf, fail, 4, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    if (nf) throw Error(FTLM_ERR_DEVICE, "tridiagonal QL did not converge");
  }
  HIPCHECK(hipGetLastError());
}

// copy Ritz values/weights of R columns back, concatenated in sample order
int64_t gather_ritz(ftlm_ctx* c, int R, int steps, double* eig, double* w) {
  std::vector<double> he((size_t)R * steps), hw((size_t)R * steps);
  std::vector<int32_t> m(R);
  HIPCHECK(hipMemcpyAsync(he.data(), c->ritz_eig.p, he.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(hw.data(), c->ritz_w.p, hw.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(m.data(), c->st_mcount.p, 4 * R, hipMemcpyDeviceToHost, c->stream));
  sync(c);
  int64_t n = 0;
  for (int r = 0; r < R; ++r) {
    for (int j = 0; j < m[r]; ++j) {
      if (eig) eig[n] = he[(size_t)r * steps + j];
      if (w) w[n] = hw[(size_t)r * steps + j];
      ++n;
    }
  }
  return n;
}

void check_R(int R) {
  if (R < 1 || R > 1024) throw Error(FTLM_ERR_LIMIT, "R must be in [1, 1024]");
}

// Columns one batched launch covers: the apply kernels hold at most 4 x 64
// columns per wavefront (hop list and on the fly) and their fused dot stages
// 256 per workgroup. ftlm_ftlm_block runs larger R in batches of this many
// samples (the samples are independent, main.cpp:1011-1014).
constexpr int kMaxBatchR = 256;

void check_batch_R(int R) {
  if (R < 1 || R > kMaxBatchR)
    throw Error(FTLM_ERR_LIMIT, "R must be in [1, 256] for one batched call (ftlm_ftlm_block splits larger R)");
}

// widest column tile each apply path handles (hop list: 4 x 64 columns per
// wavefront; on-the-fly tiles: 2 x 64)
constexpr int kMaxTileHopList = 256;
constexpr int kMaxTileOnTheFly = 128;

// ---- packed blocks ---------------------------------------------------------
constexpr uint64_t kPackDefaultBudget = (uint64_t)64 << 30;
constexpr uint64_t kExactDefaultBudget = (uint64_t)16 << 30;  // bytes of dense matrices per exact-queue run
constexpr size_t kPackSlotLimit = (size_t)32 << 30;  // bytes of per-workgroup vectors kept between runs

size_t al256(size_t n) { return (n + 255) & ~(size_t)255; }

int pack_width(const PackState& p, int width) {
  int W = width > 0 ? width : p.best_w;
  if (W <= 0) W = p.R >= 16 ? 16 : (p.R > 8 ? 16 : (p.R > 4 ? 8 : 4));
  if (W != 4 && W != 8 && W != 16 && W != 32 && W != 64) throw Error(FTLM_ERR_ARG, "pack width must be 4, 8, 16, 32 or 64");
  return W;
}

// items (entry, column group) in decreasing cost; ties keep entry order.
// stride > 1 keeps every stride-th entry in decreasing cost (the first, most
// costly one included): a sample with the pack's mix of block sizes.
std::vector<PackItem> pack_items(const PackState& p, int W, int stride = 1) {
  std::vector<char> keep(p.entries.size(), 1);
  if (stride > 1) {
    std::vector<int32_t> order(p.entries.size());
    for (size_t e = 0; e < order.size(); ++e) order[e] = (int32_t)e;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return p.cost[a] > p.cost[b]; });
    for (size_t i = 0; i < order.size(); ++i) keep[order[i]] = i % stride == 0;
  }
  std::vector<PackItem> items;
  for (size_t e = 0; e < p.entries.size(); ++e)
    if (keep[e])
      for (int c0 = 0; c0 < p.R; c0 += W) items.push_back({(int32_t)e, c0, std::min(W, p.R - c0), 0});
  std::stable_sort(items.begin(), items.end(), [&](const PackItem& a, const PackItem& b) {
    return p.cost[a.entry] * a.w > p.cost[b.entry] * b.w;
  });
  return items;
}

// Launch state for one run: uploads the tables and returns the kernel args.
// Device bytes the per-workgroup slots may take in one launch: half of what
// is free (counting what the slots hold already), between 32 and 128 GiB.
// Wide column groups of the largest packed blocks need this to fill the GPU;
// ftlm_pack_run gives anything above 32 GiB back after the run.
size_t pack_slot_limit(ftlm_ctx* c) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return kPackSlotLimit;
  const size_t avail = free_b + c->pack.slots.bytes;
  return std::min(std::max(kPackSlotLimit, avail / 2), (size_t)128 << 30);
}

PackArgs pack_args(ftlm_ctx* c, int W, int* grid, int kmax, int sample = 1) {
  PackState& p = c->pack;
  const std::vector<PackItem> items = pack_items(p, W, sample);
  const int ncols = (int)p.entries.size() * p.R;
  const int rcap = std::max(p.rcap, (ncols + 1023) / 1024 * 1024);
  p.rcap = rcap;
  const int64_t stride = std::max<int64_t>(1, p.max_dim) * W;
  const size_t per_wg = (size_t)3 * stride * 16;
  int g = std::max(1, std::min(*grid, (int)std::min<size_t>(pack_slot_limit(c) / per_wg, (size_t)1 << 20)));
  g = std::min<int64_t>(g, (int64_t)items.size());
  *grid = g;
  PackArgs a{};
  PackEntry* de = p.d_entries.as<PackEntry>(p.entries.size());
  PackItem* di = p.d_items.as<PackItem>(items.size());
  int32_t* dn = p.d_nitems.as<int32_t>(1);
  HIPCHECK(hipMemcpyAsync(de, p.entries.data(), p.entries.size() * sizeof(PackEntry), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHECK(hipMemcpyAsync(di, items.data(), items.size() * sizeof(PackItem), hipMemcpyHostToDevice, c->stream));
  const int32_t nitems = (int32_t)items.size();
  HIPCHECK(hipMemcpyAsync(dn, &nitems, 4, hipMemcpyHostToDevice, c->stream));
  a.entries = de;
  a.items = di;
  a.n_items = dn;
  a.queue = p.queue.as<unsigned>(4);
  a.fail = p.fail.as<int32_t>(4);
  a.slots = p.slots.as<double2>((size_t)g * 3 * stride);
  a.slot_stride = stride;
  a.rcap = rcap;
  a.kmax = kmax;
  a.tree = p.tree_red;
  a.alpha = p.alpha.as<double>((size_t)p.steps * rcap);
  a.beta = p.beta.as<double>((size_t)p.steps * rcap);
  a.mcount = p.mcount.as<int32_t>(rcap);
  sync(c);  // the host item vector goes out of scope
  return a;
}

void pack_ritz(ftlm_ctx* c, const PackArgs& a, int method) {
  PackState& p = c->pack;
  const int rcap = a.rcap, steps = p.steps;
  double* eig = p.eig_d.as<double>((size_t)rcap * steps);
  double* w = p.w_d.as<double>((size_t)rcap * steps);
  if (method == 0) {
    const size_t words = (size_t)steps * (steps + 1) / 2 + 3 * (size_t)steps;
    double* scratch = p.scratch.as<double>(words * 8 <= 60 * 1024 ? 1 : words * rcap);
    launch_ritz(rcap, steps, a.alpha, a.beta, a.mcount, p.R, eig, w, scratch, c->stream);
  } else {
    double* scratch = p.scratch.as<double>((size_t)3 * steps * rcap);
    int32_t* isc = p.iscratch.as<int32_t>((size_t)steps * rcap + 1);
    launch_ritz_ql(rcap, steps, a.alpha, a.beta, a.mcount, p.R, eig, w, scratch, isc, a.fail, c->stream);
  }
}

// One run of the whole pack through a HIP graph (reset, persistent Lanczos,
// Ritz), instantiated once per launch shape and buffer set and replayed for
// later packs with the same shape; returns the graph's wall time in ms.
double pack_run(ftlm_ctx* c, int width, int wg_per_cu, int method) {
  PackState& p = c->pack;
  if (!p.open) throw Error(FTLM_ERR_STATE, "no pack (call ftlm_pack_begin first)");
  if (method != 0 && method != 1) throw Error(FTLM_ERR_ARG, "method must be 0 (Jacobi) or 1 (QL)");
  p.has_result = false;
  p.tree_red = method == 1 ? 1 : 0;
  const int W = pack_width(p, width);
  if (p.entries.empty()) {
    p.has_result = true;
    return 0.0;
  }
  int wg = wg_per_cu > 0 ? wg_per_cu : (p.best_wg > 0 ? p.best_wg : 2);
  int grid = c->cus * wg;
  const PackArgs a = pack_args(c, W, &grid, 0);
  // the Ritz buffers must exist before capture (no allocation inside it)
  const int rcap = a.rcap, steps = p.steps;
  p.eig_d.as<double>((size_t)rcap * steps);
  p.w_d.as<double>((size_t)rcap * steps);
  const size_t words = (size_t)steps * (steps + 1) / 2 + 3 * (size_t)steps;
  if (method == 0) {
    p.scratch.as<double>(words * 8 <= 60 * 1024 ? 1 : words * rcap);
  } else {
    p.scratch.as<double>((size_t)3 * steps * rcap);
    p.iscratch.as<int32_t>((size_t)steps * rcap + 1);
  }
  std::vector<uint64_t> key = {(uint64_t)W, (uint64_t)grid, (uint64_t)rcap, (uint64_t)steps, (uint64_t)method,
                               (uint64_t)p.R, (uint64_t)a.slot_stride, (uint64_t)(uintptr_t)a.entries,
                               (uint64_t)(uintptr_t)a.items, (uint64_t)(uintptr_t)a.slots,
                               (uint64_t)(uintptr_t)a.alpha, (uint64_t)(uintptr_t)a.beta,
                               (uint64_t)(uintptr_t)a.mcount, (uint64_t)(uintptr_t)p.eig_d.p,
                               (uint64_t)(uintptr_t)p.w_d.p, (uint64_t)(uintptr_t)p.scratch.p,
                               (uint64_t)(uintptr_t)p.iscratch.p};
  if (!p.exec || key != p.exec_key) {
    if (p.exec) {
      HIPCHECK(hipGraphExecDestroy(p.exec));
      p.exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    launch_pack_reset(a, rcap, c->stream);
    launch_pack_lanczos(a, W, grid, c->stream);
    pack_ritz(c, a, method);
    const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
    if (ce != hipSuccess) throw Error(FTLM_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&p.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      p.exec = nullptr;
      throw Error(FTLM_ERR_DEVICE, std::string("graph instantiation failed: ") + hipGetErrorString(ie));
    }
    p.exec_key = key;
  }
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  HIPCHECK(hipEventRecord(e0, c->stream));
  HIPCHECK(hipGraphLaunch(p.exec, c->stream));
  HIPCHECK(hipEventRecord(e1, c->stream));
  sync(c);
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int32_t nf = 0;
  if (method == 1) {
    HIPCHECK(hipMemcpy(&nf, a.fail, 4, hipMemcpyDeviceToHost));
    if (nf) throw Error(FTLM_ERR_DEVICE, "tridiagonal QL did not converge");
  }
  const int ncols = (int)p.entries.size() * p.R;
  p.eig.resize((size_t)ncols * steps);
  p.w.resize((size_t)ncols * steps);
  p.m.resize(ncols);
  HIPCHECK(hipMemcpy(p.eig.data(), p.eig_d.p, p.eig.size() * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(p.w.data(), p.w_d.p, p.w.size() * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(p.m.data(), a.mcount, 4 * (size_t)ncols, hipMemcpyDeviceToHost));
  p.has_result = true;
  c->stats = ftlm_stats{};
  c->stats.ms_total = ms;
  c->stats.ms_apply = ms;
  c->stats.n_apply = 0;
  for (size_t e = 0; e < p.entries.size(); ++e) {
    c->stats.n_apply += p.entries[e].kdim;
    c->stats.bytes_apply += p.alg_bytes[e];
  }
  return ms;
}

// Timing of the persistent kernel alone over `kmax` steps on every
// sample-th entry (tuning).
double pack_time(ftlm_ctx* c, int W, int wg_per_cu, int kmax, int sample) {
  int grid = c->cus * wg_per_cu;
  const PackArgs a = pack_args(c, W, &grid, kmax, sample);
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  launch_pack_reset(a, a.rcap, c->stream);
  HIPCHECK(hipEventRecord(e0, c->stream));
  launch_pack_lanczos(a, W, grid, c->stream);
  HIPCHECK(hipEventRecord(e1, c->stream));
  HIPCHECK(hipGetLastError());
  sync(c);
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->pack.has_result = false;
  return ms;
}

// Add the current block: hop list into the pack's arena, start vectors
// generated on the host (shared by entries of equal seed and dimension, e.g.
// the twists of one block). Returns false when the budget is exceeded.
bool pack_add(ftlm_ctx* c, uint64_t bseed, int64_t* index) {
  PackState& p = c->pack;
  if (!p.open) throw Error(FTLM_ERR_STATE, "no pack (call ftlm_pack_begin first)");
  need_block(c);
  BlockTables& b = c->cur;
  const int64_t n = b.dim;
  if (n <= 0) throw Error(FTLM_ERR_ARG, "empty block");
  if ((int64_t)(p.entries.size() + 1) * p.R > INT32_MAX / 2) throw Error(FTLM_ERR_LIMIT, "too many pack columns");
  hipStream_t s = c->stream;
  int64_t* cnt = c->hl_cnt.as<int64_t>(n + 1);
  int64_t* scnt = c->hl_scnt.as<int64_t>(n + 1);
  int64_t* rp_tmp = c->hl_rowptr.as<int64_t>(n + 1);
  int64_t* srp = c->hl_srowptr.as<int64_t>(n + 1);
  c->hl_key = -1;  // the ctx hop-list buffers are reused here
  launch_hops_count(gdev(c), b.view, cnt, scnt, s);
  int64_t* tmp = c->hl_scan_tmp.as<int64_t>((size_t)prefix64_tmp_words(n));
  launch_prefix64_tiled(cnt, n, rp_tmp, tmp, s);
  launch_prefix64_tiled(scnt, n, srp, tmp, s);
  HIPCHECK(hipGetLastError());
  int64_t tot[2];
  HIPCHECK(hipMemcpyAsync(&tot[0], rp_tmp + n, 8, hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(&tot[1], srp + n, 8, hipMemcpyDeviceToHost, s));
  sync(c);
  const int64_t nnz = tot[0], nsp = tot[1];
  if (nnz > INT32_MAX) throw Error(FTLM_ERR_LIMIT, "hop list too large for a pack entry");
  const int L = c->g.L;
  const double2* start = nullptr;
  for (auto& st : p.starts)
    if (st.first.first == bseed && st.first.second == n) start = st.second;
  const size_t need = al256(8 * (size_t)(n + 1)) + 2 * al256(4 * (size_t)(nnz + 1)) +
                      al256(16 * (size_t)(nsp + 1)) + al256(8 * (size_t)n) + al256(128 * (size_t)L) +
                      (start ? 0 : al256((size_t)16 * n * p.R));
  if (!p.entries.empty() && p.arena.used + need > p.budget) return false;
  int64_t* rp = static_cast<int64_t*>(p.arena.alloc(8 * (size_t)(n + 1)));
  int32_t* col = static_cast<int32_t*>(p.arena.alloc(4 * (size_t)(nnz + 1)));
  uint32_t* code = static_cast<uint32_t*>(p.arena.alloc(4 * (size_t)(nnz + 1)));
  double* spec = static_cast<double*>(p.arena.alloc(16 * (size_t)(nsp + 1)));
  double* diag = static_cast<double*>(p.arena.alloc(8 * (size_t)n));
  double* ptab = static_cast<double*>(p.arena.alloc(128 * (size_t)L));
  HIPCHECK(hipMemcpyAsync(rp, rp_tmp, 8 * (size_t)(n + 1), hipMemcpyDeviceToDevice, s));
  HIPCHECK(hipMemcpyAsync(ptab, b.view.ptab, 128 * (size_t)L, hipMemcpyDeviceToDevice, s));
  launch_hops_fill(gdev(c), b.view, rp, srp, col, code, spec, diag, s);
  HIPCHECK(hipGetLastError());
  if (!start && c->start_source == FTLM_START_DEVICE) {
    double2* st = static_cast<double2*>(p.arena.alloc((size_t)16 * n * p.R));
    launch_start_vectors(bseed, 0, 0, p.R, p.R, n, 1, reinterpret_cast<double*>(st), c->start_norms.as<double>(p.R),
                         s);
    HIPCHECK(hipGetLastError());
    start = st;
    p.starts.push_back({{bseed, n}, st});
  }
  if (!start) {
    double2* st = static_cast<double2*>(p.arena.alloc((size_t)16 * n * p.R));
    const int64_t fit = std::max<int64_t>(1, (int64_t)(1ll << 30) / std::max<int64_t>(1, n * 16));
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)p.R, fit));
    double* host = static_cast<double*>(p.stage.get((size_t)chunk * n * 16));
    for (int r0 = 0; r0 < p.R; r0 += chunk) {
      const int cnt_r = std::min(chunk, p.R - r0);
      sync(c);  // staging reuse
      gen_starts(bseed, r0, cnt_r, n, host);
      HIPCHECK(hipMemcpyAsync(st + (int64_t)r0 * n, host, (size_t)cnt_r * n * 16, hipMemcpyHostToDevice, s));
    }
    start = st;
    p.starts.push_back({{bseed, n}, st});
  }
  sync(c);
  PackEntry e{};
  e.row_ptr = rp;
  e.col = col;
  e.code = code;
  e.spec = reinterpret_cast<const double2*>(spec);
  e.diag = diag;
  e.ptab = reinterpret_cast<const double2*>(ptab);
  e.start = start;
  e.dim = n;
  e.kdim = (int32_t)std::min<int64_t>(p.steps, n);
  e.col0 = (int32_t)(p.entries.size() * p.R);
  if (index) *index = (int64_t)p.entries.size();
  p.entries.push_back(e);
  p.cost.push_back((double)nnz + 2.0 * (double)n);
  p.alg_bytes.push_back((double)e.kdim * (double)n * p.R * 16.0 * (hbar(c) + 2.0));
  p.max_dim = std::max(p.max_dim, n);
  p.has_result = false;
  return true;
}

}  // namespace

void ftlm::set_last_error(const char* msg) { g_err = msg; }

// ============================================================================
extern "C" {

const char* ftlm_last_error(void) { return g_err.c_str(); }


uint64_t ftlm_block_seed(uint64_t seed, int nup, int ndn, int mx, int my) {
  uint64_t s = mix64(seed);
  s ^= mix64((uint64_t)(nup + 1));
  s ^= mix64((uint64_t)(ndn + 17));
  s ^= mix64((uint64_t)(mx + 257));
  s ^= mix64((uint64_t)(my + 65537));
  return mix64(s);
}

int ftlm_create(const ftlm_model* model, int device, ftlm_ctx** out) {
  return guarded([&] {
    if (!model || !out) throw Error(FTLM_ERR_ARG, "null argument");
    *out = nullptr;
    if (model->lx <= 0 || model->ly <= 0) throw Error(FTLM_ERR_ARG, "Lattice dimensions must be positive.");
    if (model->lx * model->ly > 63) throw Error(FTLM_ERR_LIMIT, "This implementation supports at most 63 sites.");
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
      throw Error(FTLM_ERR_DEVICE, "no HIP device available (the FTLM engine runs on the GPU only)");
    if (device < 0 || device >= n) throw Error(FTLM_ERR_DEVICE, "device index out of range");
    auto c = std::make_unique<ftlm_ctx>();
    c->device = device;
    HIPCHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, device));
    c->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    // device address of the fixed scratch array of ftlm_thermo_noalloc on
    // this device (resolved once per device, with the device current)
    HIPCHECK((hipError_t)thermo_scratch_setup(device));
    c->model = *model;
    build_lattice(c.get());
    upload_ctx_tables(c.get());
    *out = c.release();
  });
}

void ftlm_destroy(ftlm_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  delete ctx;
}

int ftlm_sector(ftlm_ctx* ctx, int nup, int ndn, uint64_t* full_dim, int64_t* parents) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    HIPCHECK(hipSetDevice(ctx->device));
    build_sector(ctx, nup, ndn);
    const uint64_t fd = binom_u64(ctx->g.L, nup) * binom_u64(ctx->g.L, ndn);
    if (full_dim) *full_dim = fd;
    if (parents) *parents = count_parents(ctx);
  });
}

int ftlm_block(ftlm_ctx* ctx, int mx, int my, int64_t* dim) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    need_sector(ctx);
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->prepared_R = 0;
    ++ctx->block_gen;
    build_block(ctx, ctx->cur, mx, my);
    if (dim) *dim = ctx->cur.dim;
  });
}

int ftlm_block_reps(ftlm_ctx* ctx, uint64_t* up, uint64_t* dn, int32_t* deg, double* dnf) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    need_block(ctx);
    const int64_t n = ctx->cur.dim;
    if (n == 0) return;
    DevBuf a, b, d, f;
    uint64_t* du = a.as<uint64_t>(n);
    uint64_t* dd = b.as<uint64_t>(n);
    int32_t* dg = d.as<int32_t>(n);
    double* df = f.as<double>(n);
    launch_block_reps(gdev(ctx), ctx->cur.view, du, dd, dg, df, ctx->stream);
    HIPCHECK(hipGetLastError());
    if (up) HIPCHECK(hipMemcpyAsync(up, du, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (dn) HIPCHECK(hipMemcpyAsync(dn, dd, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (deg) HIPCHECK(hipMemcpyAsync(deg, dg, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (dnf) HIPCHECK(hipMemcpyAsync(dnf, df, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
  });
}

int ftlm_block_csr(ftlm_ctx* ctx, int all_dirs, int64_t* nnz, int32_t* row_ptr, int32_t* col, double* val,
                   double* diag) {
  return guarded([&] {
    if (!ctx || !nnz) throw Error(FTLM_ERR_ARG, "null argument");
    need_block(ctx);
    const int64_t n = ctx->cur.dim;
    if (n == 0) {
      *nnz = 0;
      if (row_ptr) row_ptr[0] = 0;
      return;
    }
    DevBuf cnt, rp, cl, vl, dg;
    int32_t* dc = cnt.as<int32_t>(n);
    launch_csr_count(gdev(ctx), ctx->cur.view, all_dirs, dc, ctx->stream);
    std::vector<int32_t> hc(n);
    HIPCHECK(hipMemcpyAsync(hc.data(), dc, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    std::vector<int64_t> hp(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) hp[i + 1] = hp[i] + hc[i];
    *nnz = hp[n];
    if (!col) return;
    if (hp[n] > INT32_MAX) throw Error(FTLM_ERR_LIMIT, "nnz exceeds 2^31");
    int64_t* drp = rp.as<int64_t>(n + 1);
    HIPCHECK(hipMemcpyAsync(drp, hp.data(), (n + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    int32_t* dcol = cl.as<int32_t>(hp[n] + 1);
    double* dval = vl.as<double>(2 * hp[n] + 2);
    double* ddg = dg.as<double>(n);
    launch_csr_fill(gdev(ctx), ctx->cur.view, all_dirs, drp, dcol, dval, ddg, ctx->stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(col, dcol, hp[n] * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (val) HIPCHECK(hipMemcpyAsync(val, dval, hp[n] * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (diag) HIPCHECK(hipMemcpyAsync(diag, ddg, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    if (row_ptr)
      for (int64_t i = 0; i <= n; ++i) row_ptr[i] = (int32_t)hp[i];
  });
}

int ftlm_apply(ftlm_ctx* ctx, int R, int kernel, const double* x, double* y) {
  return guarded([&] {
    if (!ctx || !x || !y) throw Error(FTLM_ERR_ARG, "null argument");
    need_block(ctx);
    check_batch_R(R);
    if (kernel == 1 && R != 1) throw Error(FTLM_ERR_ARG, "kernel 1 handles R = 1 only");
    if (kernel != 0 && kernel != 1) throw Error(FTLM_ERR_ARG, "kernel must be 0 (configured) or 1 (one row per thread)");
    const int64_t n = ctx->cur.dim;
    if (n == 0) return;
    ensure_vectors(ctx, R);
    double* dx = ctx->vec[0].ptr<double>();
    double* dy = ctx->vec[1].ptr<double>();
    HIPCHECK(hipMemcpyAsync(dx, x, (size_t)n * R * 16, hipMemcpyHostToDevice, ctx->stream));
    if (kernel == 1) {
      launch_apply_r1(gdev(ctx), ctx->g.L, ctx->cur.view, dx, dy, ctx->cus, nullptr, nullptr, nullptr, ctx->stream);
    } else {
      do_apply(ctx, resolve_cfg(ctx, R, ctx->apply_path, ctx->apply_w), R, dx, dy, nullptr, nullptr, nullptr);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(y, dy, (size_t)n * R * 16, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    ctx->prepared_R = 0;
  });
}

int ftlm_start_vector(ftlm_ctx* ctx, uint64_t block_seed, int sample, double* out) {
  return guarded([&] {
    if (!ctx || !out) throw Error(FTLM_ERR_ARG, "null argument");
    need_block(ctx);
    if (sample < 0) throw Error(FTLM_ERR_ARG, "sample must be non-negative");
    if (ctx->start_source == FTLM_START_DEVICE) {
      HIPCHECK(hipSetDevice(ctx->device));
      const int64_t dim = ctx->cur.dim;
      double* d = ctx->start_one.as<double>(2 * (size_t)dim);
      launch_start_vectors(block_seed, sample, 0, 1, 1, dim, 1, d, ctx->start_norms.as<double>(1), ctx->stream);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(out, d, 16 * (size_t)dim, hipMemcpyDeviceToHost, ctx->stream));
      sync(ctx);
      return;
    }
    gen_start(block_seed, sample, ctx->cur.dim, out);
  });
}

int ftlm_set_start_source(ftlm_ctx* ctx, int source) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    if (source != FTLM_START_HOST && source != FTLM_START_DEVICE) throw Error(FTLM_ERR_ARG, "unknown start source");
    if (source != ctx->start_source) {
      ctx->start_source = source;
      ctx->prepared_R = 0;  // prepared start vectors came from the other source
      ctx->pack.starts.clear();
    }
  });
}

int ftlm_lanczos(ftlm_ctx* ctx, int R, int steps, const double* start, uint64_t block_seed, int reduce,
                 int per_vector, double* alpha, double* beta, int32_t* m) {
  return guarded([&] {
    if (!ctx || !alpha || !beta || !m) throw Error(FTLM_ERR_ARG, "null argument");
    need_block(ctx);
    if (per_vector) check_R(R);
    else check_batch_R(R);
    if (steps < 1) throw Error(FTLM_ERR_ARG, "--samples and --lanczos-steps must be positive.");
    const int64_t n = ctx->cur.dim;
    if (n == 0) throw Error(FTLM_ERR_ARG, "empty block");
    HIPCHECK(hipSetDevice(ctx->device));
    const int layout = per_vector ? 1 : 0;
    if (start) {
      ensure_vectors(ctx, R);
      if (layout == 0) {
        HIPCHECK(hipMemcpyAsync(ctx->vec[0].p, start, (size_t)n * R * 16, hipMemcpyHostToDevice, ctx->stream));
      } else {
        std::vector<double> tr((size_t)n * R * 2);
        for (int r = 0; r < R; ++r)
          for (int64_t i = 0; i < n; ++i) {
            tr[2 * ((size_t)r * n + i)] = start[2 * ((size_t)i * R + r)];
            tr[2 * ((size_t)r * n + i) + 1] = start[2 * ((size_t)i * R + r) + 1];
          }
        HIPCHECK(hipMemcpyAsync(ctx->vec[0].p, tr.data(), tr.size() * 8, hipMemcpyHostToDevice, ctx->stream));
      }
      sync(ctx);
    } else {
      prepare_starts(ctx, R, block_seed, layout);
    }
    run_all(ctx, R, steps, reduce, per_vector != 0, false);
    HIPCHECK(hipMemcpyAsync(alpha, ctx->st_alpha.p, (size_t)steps * R * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(beta, ctx->st_beta.p, (size_t)steps * R * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(m, ctx->st_mcount.p, (size_t)R * 4, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    ctx->prepared_R = 0;
  });
}

int ftlm_ritz(ftlm_ctx* ctx, int R, int steps, const double* alpha, const double* beta, const int32_t* m, int wdiv,
              int method, double* eig, double* w) {
  return guarded([&] {
    if (!ctx || !alpha || !beta || !m || !eig || !w) throw Error(FTLM_ERR_ARG, "null argument");
    check_R(R);
    DevBuf da, db, dm;
    double* a = da.as<double>((size_t)steps * R);
    double* b = db.as<double>((size_t)steps * R);
    int32_t* mm = dm.as<int32_t>(R);
    for (int r = 0; r < R; ++r)
      if (m[r] < 0 || m[r] > steps) throw Error(FTLM_ERR_ARG, "m out of range");
    HIPCHECK(hipMemcpyAsync(a, alpha, (size_t)steps * R * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(b, beta, (size_t)steps * R * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(mm, m, (size_t)R * 4, hipMemcpyHostToDevice, ctx->stream));
    ritz_device(ctx, R, steps, a, b, mm, wdiv, method);
    HIPCHECK(hipMemcpyAsync(eig, ctx->ritz_eig.p, (size_t)R * steps * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(w, ctx->ritz_w.p, (size_t)R * steps * 8, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
  });
}

int ftlm_ftlm_block(ftlm_ctx* ctx, int R, int steps, uint64_t block_seed, int reduce, int per_vector, double* eig,
                    double* w, int64_t* count) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    need_block(ctx);
    check_R(R);
    if (steps < 1) throw Error(FTLM_ERR_ARG, "--samples and --lanczos-steps must be positive.");
    if (ctx->cur.dim == 0) {
      if (count) *count = 0;
      return;
    }
    HIPCHECK(hipSetDevice(ctx->device));
    const int kdim = (int)std::min<int64_t>(steps, ctx->cur.dim);
    // samples r0..r0+cnt-1 per batch; weights are divided by the full R
    // (main.cpp:1090-1094) and concatenated in sample order. Batched runs
    // record their H-apply device time (ftlm_last_stats).
    ctx->stats = ftlm_stats{};
    int64_t n = 0;
    for (int r0 = 0; r0 < R; r0 += kMaxBatchR) {
      const int cnt = std::min(kMaxBatchR, R - r0);
      prepare_starts(ctx, cnt, block_seed, per_vector ? 1 : 0, r0);
      run_all(ctx, cnt, steps, reduce, per_vector != 0, per_vector == 0);
      // serial runs keep the reference's Jacobi (bit-identical Ritz pairs);
      // tree runs use QL. The state arrays are [steps][cnt]; rows >= kdim
      // unused.
      ritz_device(ctx, cnt, kdim, ctx->st_alpha.ptr<double>(), ctx->st_beta.ptr<double>(),
                  ctx->st_mcount.ptr<int32_t>(), R, reduce == FTLM_REDUCE_SERIAL ? 0 : 1);
      n += gather_ritz(ctx, cnt, kdim, eig ? eig + n : nullptr, w ? w + n : nullptr);
    }
    if (count) *count = n;
    ctx->prepared_R = 0;
  });
}

int ftlm_prepare(ftlm_ctx* ctx, int R, uint64_t block_seed) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    need_block(ctx);
    check_batch_R(R);
    HIPCHECK(hipSetDevice(ctx->device));
    prepare_starts(ctx, R, block_seed, 0);
    // keep the start vectors resident so that every timed run starts from
    // the same input
    const size_t bytes = (size_t)ctx->cur.dim * R * 16;
    HIPCHECK(hipMemcpyAsync(ctx->pristine.get(bytes), ctx->vec[0].p, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    sync(ctx);
  });
}

int ftlm_run(ftlm_ctx* ctx, int R, int steps, int reduce, int per_vector_column) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    need_block(ctx);
    if (ctx->prepared_R != R || ctx->prepared_layout != 0) throw Error(FTLM_ERR_STATE, "call ftlm_prepare first");
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->stats = ftlm_stats{};
    ctx->last_reduce = reduce;
    const size_t bytes = (size_t)ctx->cur.dim * R * 16;
    if (per_vector_column < 0) {
      HIPCHECK(hipMemcpyAsync(ctx->vec[0].p, ctx->pristine.p, bytes, hipMemcpyDeviceToDevice, ctx->stream));
      run_all(ctx, R, steps, reduce, false, true);
    } else {
      // plain per-vector mode on one column of the prepared block: extract
      // it and run the R = 1 kernels
      const int64_t dim = ctx->cur.dim;
      if (per_vector_column >= R) throw Error(FTLM_ERR_ARG, "column out of range");
      double* p0 = ctx->pvec[0].as<double>((size_t)dim * 2);
      double* p1 = ctx->pvec[1].as<double>((size_t)dim * 2);
      double* p2 = ctx->pvec[2].as<double>((size_t)dim * 2);
      HIPCHECK(hipMemcpy2DAsync(p0, 16, ctx->pristine.ptr<double>() + 2 * per_vector_column, (size_t)R * 16, 16, dim,
                                hipMemcpyDeviceToDevice, ctx->stream));
      LanczosState one = lanczos_state(ctx, 1, steps);
      lanczos_run(ctx, 1, steps, reduce, p0, p1, p2, one, true, true);
    }
  });
}

int ftlm_finish(ftlm_ctx* ctx, int R, int steps, double* eig, double* w, int64_t* count) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    need_block(ctx);
    const int kdim = (int)std::min<int64_t>(steps, ctx->cur.dim);
    ritz_device(ctx, R, kdim, ctx->st_alpha.ptr<double>(), ctx->st_beta.ptr<double>(), ctx->st_mcount.ptr<int32_t>(),
                R, ctx->last_reduce == FTLM_REDUCE_SERIAL ? 0 : 1);
    const int64_t n = gather_ritz(ctx, R, kdim, eig, w);
    if (count) *count = n;
  });
}

int ftlm_last_stats(ftlm_ctx* ctx, ftlm_stats* out) {
  return guarded([&] {
    if (!ctx || !out) throw Error(FTLM_ERR_ARG, "null argument");
    *out = ctx->stats;
  });
}

int ftlm_exact_blocks(ftlm_ctx* ctx, int nblk, const int32_t* mx, const int32_t* my, double* eig_out,
                      int64_t* offsets) {
  return guarded([&] {
    if (!ctx || (nblk > 0 && (!mx || !my || !eig_out || !offsets))) throw Error(FTLM_ERR_ARG, "null argument");
    need_sector(ctx);
    HIPCHECK(hipSetDevice(ctx->device));
    if (nblk <= 0) {
      if (offsets) offsets[0] = 0;
      return;
    }
    while ((int)ctx->pool.size() < nblk) ctx->pool.emplace_back(new BlockTables);
    std::vector<BlockView> views(nblk);
    std::vector<int64_t> moff(nblk + 1, 0), dims(nblk), eoff(nblk + 1, 0);
    int64_t maxd = 0;
    for (int i = 0; i < nblk; ++i) {
      build_block(ctx, *ctx->pool[i], mx[i], my[i]);
      views[i] = ctx->pool[i]->view;
      dims[i] = ctx->pool[i]->dim;
      if (dims[i] > 8192) throw Error(FTLM_ERR_LIMIT, "exact diagonalisation is limited to dim <= 8192");
      moff[i + 1] = moff[i] + dims[i] * dims[i];
      eoff[i + 1] = eoff[i] + dims[i];
      maxd = std::max(maxd, dims[i]);
    }
    hipStream_t s = ctx->stream;
    double* mats = ctx->ex_mats.as<double>(2 * (size_t)moff[nblk] + 2);
    double* work = ctx->ex_work.as<double>(6 * (size_t)eoff[nblk] + 6);
    double* eig = ctx->ex_eig.as<double>((size_t)eoff[nblk] + 1);
    int64_t* dmoff = ctx->ex_moff.as<int64_t>(nblk + 1);
    int64_t* ddims = ctx->ex_dims.as<int64_t>(nblk);
    int64_t* deoff = ctx->ex_eoff.as<int64_t>(nblk + 1);
    BlockView* dviews = ctx->ex_views.as<BlockView>(nblk);
    HIPCHECK(hipMemcpyAsync(dmoff, moff.data(), (nblk + 1) * 8, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(ddims, dims.data(), nblk * 8, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(deoff, eoff.data(), (nblk + 1) * 8, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(dviews, views.data(), nblk * sizeof(BlockView), hipMemcpyHostToDevice, s));
    launch_zero(mats, 2 * moff[nblk], s);
    launch_dense_build(gdev(ctx), dviews, nblk, dmoff, mats, maxd, s);
    launch_herm_eig(dmoff, ddims, nblk, mats, work, eig, deoff, s);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(eig_out, eig, eoff[nblk] * 8, hipMemcpyDeviceToHost, s));
    sync(ctx);
    for (int i = 0; i <= nblk; ++i) offsets[i] = eoff[i];
    for (int64_t i = 0; i < eoff[nblk]; ++i)
      if (std::isnan(eig_out[i])) throw Error(FTLM_ERR_DEVICE, "exact eigensolver did not converge");
  });
}

int ftlm_exact_queue_begin(ftlm_ctx* ctx, uint64_t budget_bytes) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    ExactQueue& q = ctx->exq;
    q.budget = budget_bytes ? budget_bytes : kExactDefaultBudget;
    q.moff.assign(1, 0);
    q.dims.clear();
    q.eoff.assign(1, 0);
    q.eig_host.clear();
    q.solved = 0;
    q.has_result = false;
  });
}

int ftlm_exact_queue_add(ftlm_ctx* ctx, int mx, int my, int64_t* index) {
  bool added = true;
  const int rc = guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    need_sector(ctx);
    HIPCHECK(hipSetDevice(ctx->device));
    ExactQueue& q = ctx->exq;
    if (q.budget == 0) q.budget = kExactDefaultBudget;
    build_block(ctx, q.tab, mx, my);
    const int64_t n = q.tab.dim;
    if (n <= 0) throw Error(FTLM_ERR_ARG, "empty block");
    if (n > 8192) throw Error(FTLM_ERR_LIMIT, "exact diagonalisation is limited to dim <= 8192");
    const int64_t o = q.moff.back();
    const size_t need = 2 * (size_t)(o + n * n);  // doubles
    if (!q.dims.empty() && need * 8 > q.budget) {
      added = false;
      return;
    }
    hipStream_t s = ctx->stream;
    if (need > q.mats_cap) {  // grow, keeping the matrices built so far
      const size_t cap = std::max(need, 2 * q.mats_cap);
      DevBuf grown;
      double* p = grown.as<double>(cap);
      if (o > 0) HIPCHECK(hipMemcpyAsync(p, q.mats.ptr<double>(), 16 * (size_t)o, hipMemcpyDeviceToDevice, s));
      sync(ctx);
      std::swap(q.mats.p, grown.p);
      std::swap(q.mats.bytes, grown.bytes);
      q.mats_cap = cap;
    }
    double* A = q.mats.ptr<double>() + 2 * o;
    launch_zero(A, 2 * n * n, s);
    const int64_t mo[2] = {0, n * n};
    int64_t* dmo = q.moff1_d.as<int64_t>(2);
    BlockView* dv = q.view_d.as<BlockView>(1);
    HIPCHECK(hipMemcpyAsync(dmo, mo, sizeof(mo), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(dv, &q.tab.view, sizeof(BlockView), hipMemcpyHostToDevice, s));
    launch_dense_build(gdev(ctx), dv, 1, dmo, A, n, s);
    HIPCHECK(hipGetLastError());
    sync(ctx);  // the staging above and the block tables are reused by the next add
    q.dims.push_back(n);
    q.moff.push_back(o + n * n);
    q.eoff.push_back(q.eoff.back() + n);
    q.has_result = false;
    if (index) *index = (int64_t)q.dims.size() - 1;
  });
  if (rc != FTLM_OK) return rc;
  return added ? FTLM_OK : FTLM_PACK_FULL;
}

int ftlm_exact_queue_run(ftlm_ctx* ctx) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    HIPCHECK(hipSetDevice(ctx->device));
    ExactQueue& q = ctx->exq;
    // entries added since the last run (the solver overwrites its matrices,
    // so solved entries keep their eigenvalues and are not solved again)
    const int64_t nb = (int64_t)q.dims.size();
    const int64_t first = q.solved;
    q.eig_host.resize((size_t)q.eoff.back(), 0.0);
    if (nb > first) {
      if (nb - first > INT32_MAX) throw Error(FTLM_ERR_LIMIT, "too many queued blocks");
      hipStream_t s = ctx->stream;
      int64_t* dmoff = q.moff_d.as<int64_t>(nb + 1);
      int64_t* ddims = q.dims_d.as<int64_t>(nb);
      int64_t* deoff = q.eoff_d.as<int64_t>(nb + 1);
      double* work = q.work.as<double>(6 * (size_t)q.eoff.back() + 6);
      double* eig = q.eig.as<double>((size_t)q.eoff.back() + 1);
      HIPCHECK(hipMemcpyAsync(dmoff, q.moff.data(), (nb + 1) * 8, hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(ddims, q.dims.data(), nb * 8, hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(deoff, q.eoff.data(), (nb + 1) * 8, hipMemcpyHostToDevice, s));
      launch_herm_eig(dmoff + first, ddims + first, (int)(nb - first), q.mats.ptr<double>(), work, eig,
                      deoff + first, s);
      HIPCHECK(hipGetLastError());
      const int64_t e0 = q.eoff[(size_t)first];
      HIPCHECK(hipMemcpyAsync(q.eig_host.data() + e0, eig + e0, (size_t)(q.eoff.back() - e0) * 8,
                              hipMemcpyDeviceToHost, s));
      sync(ctx);
      for (size_t i = (size_t)e0; i < q.eig_host.size(); ++i)
        if (std::isnan(q.eig_host[i])) throw Error(FTLM_ERR_DEVICE, "exact eigensolver did not converge");
      q.solved = nb;
    }
    q.has_result = true;
  });
}

int ftlm_exact_queue_result(ftlm_ctx* ctx, int64_t index, double* eig, int64_t* count) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    const ExactQueue& q = ctx->exq;
    if (index < 0 || index >= (int64_t)q.dims.size()) throw Error(FTLM_ERR_ARG, "queue index out of range");
    if (index >= q.solved) throw Error(FTLM_ERR_STATE, "exact queue entry has not been run");
    const int64_t a = q.eoff[(size_t)index], b = q.eoff[(size_t)index + 1];
    if (eig) std::copy(q.eig_host.begin() + a, q.eig_host.begin() + b, eig);
    if (count) *count = b - a;
  });
}

int ftlm_set_twist(ftlm_ctx* ctx, double phix, double phiy) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    if (!std::isfinite(phix) || !std::isfinite(phiy)) throw Error(FTLM_ERR_ARG, "twist must be finite");
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->model.phix = phix;
    ctx->model.phiy = phiy;
    set_amplitudes(ctx);
    HIPCHECK(hipMemcpy(ctx->g_dev.p, &ctx->g, sizeof(LatticeConst), hipMemcpyHostToDevice));
    ctx->cur.mx = ctx->cur.my = -1;
    ctx->prepared_R = 0;
    ++ctx->block_gen;
  });
}

int ftlm_pack_begin(ftlm_ctx* ctx, int R, int steps, uint64_t budget_bytes) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    check_R(R);
    if (steps < 1) throw Error(FTLM_ERR_ARG, "steps must be positive");
    HIPCHECK(hipSetDevice(ctx->device));
    PackState& p = ctx->pack;
    p.open = true;
    p.R = R;
    p.steps = steps;
    p.budget = budget_bytes ? budget_bytes : kPackDefaultBudget;
    p.arena.reset();
    p.entries.clear();
    p.cost.clear();
    p.alg_bytes.clear();
    p.starts.clear();
    p.max_dim = 0;
    p.has_result = false;
  });
}

int ftlm_pack_add(ftlm_ctx* ctx, uint64_t block_seed, int64_t* index) {
  bool added = true;
  const int rc = guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    HIPCHECK(hipSetDevice(ctx->device));
    added = pack_add(ctx, block_seed, index);
  });
  if (rc != FTLM_OK) return rc;
  return added ? FTLM_OK : FTLM_PACK_FULL;
}

int ftlm_pack_run(ftlm_ctx* ctx, int width, int wg_per_cu, int method) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    HIPCHECK(hipSetDevice(ctx->device));
    (void)pack_run(ctx, width, wg_per_cu, method);
    if (ctx->pack.slots.bytes > kPackSlotLimit) {  // leave the memory to solo blocks
      sync(ctx);
      ctx->pack.slots.release();
    }
  });
}

int ftlm_pack_result(ftlm_ctx* ctx, int64_t index, double* eig, double* w, int64_t* count) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    const PackState& p = ctx->pack;
    if (!p.has_result) throw Error(FTLM_ERR_STATE, "pack has not been run");
    if (index < 0 || index >= (int64_t)p.entries.size()) throw Error(FTLM_ERR_ARG, "pack index out of range");
    int64_t k = 0;
    for (int r = 0; r < p.R; ++r) {
      const int64_t col = index * p.R + r;
      for (int j = 0; j < p.m[col]; ++j) {
        if (eig) eig[k] = p.eig[(size_t)col * p.steps + j];
        if (w) w[k] = p.w[(size_t)col * p.steps + j];
        ++k;
      }
    }
    if (count) *count = k;
  });
}

int ftlm_pack_tune(ftlm_ctx* ctx, int tune_steps, int* best_width, int* best_wg_per_cu, double* best_ms,
                   int* ncand, int* cand_width, int* cand_wg, double* cand_ms) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    HIPCHECK(hipSetDevice(ctx->device));
    PackState& p = ctx->pack;
    if (!p.open || p.entries.empty()) throw Error(FTLM_ERR_STATE, "empty pack");
    const int cap = ncand ? *ncand : 0;
    // A large pack is timed on a sample of its entries (every stride-th in
    // decreasing cost), as long as the widest column groups still give 4
    // items per workgroup of the largest grid: each candidate's timing then
    // stays a small share of the run (eight candidates of two steps over the
    // whole pack cost a fifth of an 80-step run).
    const int64_t groups = (int64_t)p.entries.size() * ((p.R + 63) / 64);
    const int64_t need = (int64_t)4 * ctx->cus * 2;
    const int stride = (int)std::max<int64_t>(1, std::min<int64_t>(groups / need, 1 << 20));
    int n = 0, bw = 0, bg = 0;
    double bms = 0.0;
    // widths from narrow to wide, stopping at the first width that does not
    // beat the best so far: the time has been unimodal in W on every pack
    // measured (the 1,212 blocks of one 3x4 twist at R = 64: 377, 373, 395,
    // 550 ms for W = 8 to 64)
    for (int W : {8, 16, 32, 64}) {
      if (W > 16 && W >= 2 * p.R) continue;
      if (bw != 0 && bw < W / 2) break;
      for (int wg : {1, 2}) {  // 4 per CU measured no better (throughput-bound)
        const double ms = pack_time(ctx, W, wg, tune_steps > 0 ? tune_steps : 4, stride);
        if (n < cap) {
          if (cand_width) cand_width[n] = W;
          if (cand_wg) cand_wg[n] = wg;
          if (cand_ms) cand_ms[n] = ms;
        }
        ++n;
        if (bw == 0 || ms < bms) {
          bw = W;
          bg = wg;
          bms = ms;
        }
      }
    }
    p.best_w = bw;
    p.best_wg = bg;
    if (best_width) *best_width = bw;
    if (best_wg_per_cu) *best_wg_per_cu = bg;
    if (best_ms) *best_ms = bms;
    if (ncand) *ncand = std::min(n, cap);
  });
}

int ftlm_thermo(ftlm_ctx* ctx, int64_t nblk, const int64_t* off, const double* eig, const double* w,
                const int32_t* particles, const double* prefactor, int nbeta, const double* betas, int nmu,
                const double* mu, double* out) {
  return guarded([&] {
    if (!ctx || !off || !particles || !prefactor || !betas || !mu || !out) throw Error(FTLM_ERR_ARG, "null argument");
    if (nblk < 0 || nbeta <= 0 || nmu <= 0) throw Error(FTLM_ERR_ARG, "empty beta or mu list");
    if ((int64_t)nbeta * nmu > (1 << 24)) throw Error(FTLM_ERR_LIMIT, "too many (beta, mu) pairs");
    const int64_t nv = nblk ? off[nblk] : 0;
    if (nv < 0 || (nv > 0 && (!eig || !w))) throw Error(FTLM_ERR_ARG, "null spectrum");
    HIPCHECK(hipSetDevice(ctx->device));
    // per-value weight W = prefactor * w and particle number; per-mu emin as
    // the reference's running minimum ends (min over blocks of min E - mu N)
    std::vector<double> W((size_t)nv);
    std::vector<uint8_t> N((size_t)nv);
    std::vector<double> emin((size_t)nmu, std::numeric_limits<double>::infinity());
    for (int64_t b = 0; b < nblk; ++b) {
      if (off[b + 1] < off[b]) throw Error(FTLM_ERR_ARG, "block offsets must be non-decreasing");
      if (particles[b] < 0 || particles[b] > 255) throw Error(FTLM_ERR_ARG, "particle number out of range");
      if (off[b + 1] == off[b]) continue;
      double bmin = std::numeric_limits<double>::infinity();
      for (int64_t i = off[b]; i < off[b + 1]; ++i) {
        W[(size_t)i] = prefactor[b] * w[i];
        N[(size_t)i] = (uint8_t)particles[b];
        bmin = std::min(bmin, eig[i]);
      }
      const double Nb = (double)particles[b];
      for (int k = 0; k < nmu; ++k) emin[(size_t)k] = std::min(emin[(size_t)k], bmin - mu[k] * Nb);
    }
    const int np = nbeta * nmu;
    const int64_t nwg = thermo_workgroups(nv);
    hipStream_t s = ctx->stream;
    double* de = ctx->th_e.as<double>((size_t)std::max<int64_t>(nv, 1));
    double* dw = ctx->th_w.as<double>((size_t)std::max<int64_t>(nv, 1));
    uint8_t* dn = ctx->th_n.as<uint8_t>((size_t)std::max<int64_t>(nv, 1));
    double* db = ctx->th_beta.as<double>((size_t)nbeta);
    double* dm = ctx->th_mu.as<double>((size_t)nmu);
    double* dmin = ctx->th_emin.as<double>((size_t)nmu);
    double* dp = ctx->th_part.as<double>((size_t)std::max<int64_t>(nwg, 1) * np * 3);
    double* dout = ctx->th_out.as<double>((size_t)np * 4);
    if (nv) {
      HIPCHECK(hipMemcpyAsync(de, eig, (size_t)nv * 8, hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(dw, W.data(), (size_t)nv * 8, hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(dn, N.data(), (size_t)nv, hipMemcpyHostToDevice, s));
    }
    HIPCHECK(hipMemcpyAsync(db, betas, (size_t)nbeta * 8, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(dm, mu, (size_t)nmu * 8, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(dmin, emin.data(), (size_t)nmu * 8, hipMemcpyHostToDevice, s));
    launch_thermo(de, dw, dn, nv, db, nbeta, dm, nmu, dmin, dp, dout, s);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(out, dout, (size_t)np * 4 * 8, hipMemcpyDeviceToHost, s));
    sync(ctx);
  });
}

int ftlm_thermo_noalloc(ftlm_ctx* ctx, int64_t nv, const double* d_eig, const double* d_weight,
                        const uint8_t* d_particles, int nbeta, const double* d_betas, int nmu, const double* d_mu,
                        double* d_out, void* stream) {
  return guarded([&] {
    if (!ctx || !d_betas || !d_mu || !d_out) throw Error(FTLM_ERR_ARG, "null argument");
    if (nv < 0 || (nv > 0 && (!d_eig || !d_weight || !d_particles))) throw Error(FTLM_ERR_ARG, "null spectrum");
    if (nbeta <= 0 || nmu <= 0) throw Error(FTLM_ERR_ARG, "empty beta or mu list");
    if ((int64_t)nbeta * nmu > (1 << 24) ||
        thermo_noalloc_doubles(nv, nbeta, nmu) > kThermoScratchDoubles)
      throw Error(FTLM_ERR_LIMIT,
                  "ftlm_thermo_noalloc: 3 * ceil(nv / 4096) * nbeta * nmu partial sums exceed its fixed 32 MiB "
                  "scratch (use ftlm_thermo, or fewer values per call)");
    double* scr = thermo_scratch(ctx->device);
    if (!scr) throw Error(FTLM_ERR_STATE, "scratch of the device not set up");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    launch_thermo_noalloc(d_eig, d_weight, d_particles, nv, d_betas, nbeta, d_mu, nmu, scr, d_out, s);
    HIPCHECK(hipGetLastError());
    if (!stream) HIPCHECK(hipStreamSynchronize(s));
  });
}

int ftlm_memprobe(ftlm_ctx* ctx, int kind, int64_t rows, int width, int k, int reps, double* ms_per_launch,
                  double* bytes_per_launch) {
  return guarded([&] {
    if (!ctx || !ms_per_launch || !bytes_per_launch) throw Error(FTLM_ERR_ARG, "null argument");
    if (kind < 0 || kind > 4)
      throw Error(FTLM_ERR_ARG, "kind must be 0 (read), 1 (copy), 2 (gather), 3 (windowed gather) or 4 (mixed gather)");
    if (rows < 1 || width < 1 || width > 256 || reps < 1 || (kind >= 2 && k < 1))
      throw Error(FTLM_ERR_ARG, "rows, reps and k must be positive, width in [1, 256]");
    HIPCHECK(hipSetDevice(ctx->device));
    const size_t bytes = (size_t)rows * width * 16;
    DevBuf x, y;
    double* dx = x.as<double>(2 * (size_t)rows * width);
    double* dy = y.as<double>(2 * (size_t)rows * width);
    launch_zero(dx, 2 * rows * (int64_t)width, ctx->stream);
    launch_memprobe(kind, rows, width, k, ctx->cus, dx, dy, ctx->stream);  // warm-up
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < reps; ++i) launch_memprobe(kind, rows, width, k, ctx->cus, dx, dy, ctx->stream);
    HIPCHECK(hipEventRecord(e1, ctx->stream));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_launch = ms / reps;
    *bytes_per_launch = kind == 0 ? (double)bytes : kind == 1 ? 2.0 * bytes : (double)bytes * (k + 1);  // gathers: k rows in, 1 out
  });
}

int ftlm_pack_info(ftlm_ctx* ctx, int64_t* entries, uint64_t* bytes) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    if (entries) *entries = (int64_t)ctx->pack.entries.size();
    if (bytes) *bytes = (uint64_t)ctx->pack.arena.used;
  });
}

int ftlm_set_apply_config(ftlm_ctx* ctx, int path, int W) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    if (path < 0 || path > 2) throw Error(FTLM_ERR_ARG, "path must be 0 (automatic), 1 (hop list) or 2 (on the fly)");
    if (W < -1) throw Error(FTLM_ERR_ARG, "tile width must be -1 (automatic), 0 (untiled) or positive");
    if (W > (path == 2 ? kMaxTileOnTheFly : kMaxTileHopList))
      throw Error(FTLM_ERR_LIMIT, path == 2 ? "on-the-fly column tiles hold at most 128 vectors"
                                            : "hop-list column tiles hold at most 256 vectors");
    ctx->apply_path = path;
    ctx->apply_w = W;
    ctx->apply_path_r1 = path == 0 ? 2 : path;
    ctx->plan_key_dim = -1;
  });
}

int ftlm_get_apply_config(ftlm_ctx* ctx, int R, int* path, int* W) {
  return guarded([&] {
    if (!ctx || !path || !W) throw Error(FTLM_ERR_ARG, "null argument");
    need_block(ctx);
    const ApplyCfg a = resolve_cfg(ctx, R, ctx->apply_path, ctx->apply_w);
    *path = a.path;
    *W = a.w;
  });
}

int ftlm_set_vector_placement(ftlm_ctx* ctx, int candidates, uint64_t min_bytes) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    if (candidates != 0 && (candidates < 3 || candidates > 16))
      throw Error(FTLM_ERR_ARG, "placement candidates must be 0 (off) or 3 to 16");
    ctx->place_candidates = candidates;
    ctx->place_min_bytes = min_bytes;
    if (candidates == 0) ctx->place_pending = false;
  });
}

int ftlm_vector_placement_report(ftlm_ctx* ctx, int* n, double* ms, int* pick) {
  return guarded([&] {
    if (!ctx || !n) throw Error(FTLM_ERR_ARG, "null argument");
    const int have = (int)ctx->place_ms.size();
    if (ms)
      for (int i = 0; i < std::min(*n, have); ++i) ms[i] = ctx->place_ms[i];
    if (pick)
      for (int i = 0; i < (int)ctx->place_pick.size(); ++i) pick[i] = ctx->place_pick[i];
    *n = have;
  });
}

int ftlm_autotune(ftlm_ctx* ctx, int R, int* best_path, int* best_w, double* best_ms, int* ncand, int* cand_path,
                  int* cand_w, double* cand_ms) {
  return guarded([&] {
    if (!ctx) throw Error(FTLM_ERR_ARG, "null context");
    need_block(ctx);
    check_batch_R(R);
    const int64_t n = ctx->cur.dim;
    if (n == 0) throw Error(FTLM_ERR_ARG, "empty block");
    HIPCHECK(hipSetDevice(ctx->device));
    ensure_vectors(ctx, R);
    std::vector<ApplyCfg> cands = {{1, 0}};
    for (int w : {64, 32, 16, 8})
      if (w < R) cands.push_back({1, w});
    cands.push_back({2, 0});
    if (R == 1) cands = {{1, 0}, {2, 0}};
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    double bms = 1e300;
    ApplyCfg best = cands[0];
    int k = 0;
    for (const ApplyCfg& cf : cands) {
      const double* x = ctx->vec[0].ptr<double>();
      double* y = ctx->vec[1].ptr<double>();
      do_apply(ctx, cf, R, x, y, nullptr, nullptr, nullptr);  // warm (and build the hop list)
      HIPCHECK(hipEventRecord(e0, ctx->stream));
      for (int rep = 0; rep < 2; ++rep) do_apply(ctx, cf, R, x, y, nullptr, nullptr, nullptr);
      HIPCHECK(hipEventRecord(e1, ctx->stream));
      HIPCHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
      const double per = ms / 2.0;
      if (cand_w && cand_ms && cand_path && ncand && k < *ncand) {
        cand_path[k] = cf.path;
        cand_w[k] = cf.w;
        cand_ms[k] = per;
      }
      ++k;
      if (per < bms) {
        bms = per;
        best = cf;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ncand) *ncand = k;
    if (R == 1) {
      ctx->apply_path_r1 = best.path;
    } else {
      ctx->apply_path = best.path;
      ctx->apply_w = best.w;
    }
    ctx->plan_key_dim = -1;
    if (best_path) *best_path = best.path;
    if (best_w) *best_w = best.w;
    if (best_ms) *best_ms = bms;
  });
}

}  // extern "C"
