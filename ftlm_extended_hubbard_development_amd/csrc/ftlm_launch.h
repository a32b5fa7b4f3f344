// ftlm_launch.h -- host-side launchers of the FTLM kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "ftlm_types.h"

namespace ftlm {

// Message returned by ftlm_last_error() (shared by every C ABI source file).
void set_last_error(const char* msg);

// ---- basis (ftlm_basis.hip) ----
// Per up configuration: orbit minimum, shifts reaching it, their sign.
void launch_up_table(const LatticeConst* g, const uint64_t* binom, int nup, int64_t n_upc, uint8_t* is_canon,
                     int64_t* min_rank, uint8_t* tfirst, int8_t* tsign, uint64_t* amask, hipStream_t s);
// Segmented exclusive compaction: out[s][i] = flag ? rank among set flags of
// segment s : -1 ; counts[s] = number of set flags. One workgroup/segment.
void launch_segscan(const uint8_t* flags, int64_t seg_len, int nseg, int32_t* out, int64_t* counts,
                    hipStream_t s);
// Exclusive prefix sum of n int64 values into out[0..n] (out[n] = total).
void launch_prefix64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s);
// The same exclusive scan (out[n] = total) over many workgroups for long
// arrays; tmp holds prefix64_tmp_words(n) int64.
int64_t prefix64_tmp_words(int64_t n);
void launch_prefix64_tiled(const int64_t* in, int64_t n, int64_t* out, int64_t* tmp, hipStream_t s);
void launch_canon_fill(int64_t n_upc, int nup, int L, const uint64_t* binom, const uint8_t* is_canon,
                       const int32_t* cidx_of_rank, const int64_t* min_rank, const uint64_t* amask,
                       int32_t* upc_canon, uint64_t* can_bits, uint64_t* can_stab, uint8_t* nt_flag,
                       hipStream_t s);
void launch_nt_fill(int n_canon, const int32_t* nt_index, int32_t* can_nt, int32_t* nt_to_c, hipStream_t s);
void launch_nt_active(const LatticeConst* g, const uint64_t* binom, int ndn, int64_t n_dnc, int n_nt,
                      const int32_t* nt_to_c, const uint64_t* can_bits, const uint64_t* can_stab,
                      const double* phase_k, uint8_t* act, double* dnf, hipStream_t s);
void launch_nt_list(int n_nt, int64_t n_dnc, const int32_t* nt_loc, const int64_t* listoff, int32_t* list,
                    hipStream_t s);
void launch_row_counts(int n_canon, int64_t n_dnc, const int32_t* can_nt, const int64_t* nt_count,
                       int64_t* counts, hipStream_t s);
// Rows -> (up, down, degeneracy, norm factor) for export.
void launch_block_reps(const LatticeConst* g, BlockView bv, uint64_t* up, uint64_t* dn, int32_t* deg,
                       double* dnf, hipStream_t s);

// ---- H apply (ftlm_apply.hip) ----
// y = H x over R column vectors stored row-major [dim][R] (complex double).
// When dot != nullptr also accumulates per-column Re<x, y> partial sums into
// dot[grid][R] and reduces them in the last workgroup into dot_out[R].
struct ApplyPlan {
  int R;
  int lpr;      // lanes per row
  int vpl;      // vectors per lane
  int rpt;      // rows per tile
  int nc;       // hop candidate slots per row
  int grid;     // workgroups (persistent)
  int L;        // sites
  size_t lds;   // dynamic LDS bytes
  // column tiling (W < R): tile width, tiles per row, work items
  int W = 0, J = 1, n_items = 0;
  int64_t ntiles = 0;
  const int64_t* item_tiles = nullptr;  // device: first tile of each item
};
ApplyPlan plan_apply(const LatticeConst& g, int nup, int ndn, int64_t dim, int R, int cus);
// Column-tiled plan: item_tiles (host) receives the first tile of every
// (group, column tile) item; the caller uploads it and sets p.item_tiles.
ApplyPlan plan_apply_tiled(const LatticeConst& g, int nup, int ndn, const int64_t* offs_host, int n_canon, int R,
                           int W, int cus, std::vector<int64_t>& item_tiles);
void launch_apply(const ApplyPlan& p, const LatticeConst* g, BlockView bv, const double* x, double* y,
                  double* partials, unsigned* counter, double* dot_out, hipStream_t s);
void launch_apply_r1(const LatticeConst* g, int L, BlockView bv, const double* x, double* y, int cus,
                     double* partials, unsigned* counter, double* dot_out, hipStream_t s);
void launch_zero(double* p, int64_t n, hipStream_t s);
// Hop lists (built once per block, reused by every apply).
struct HopList {
  const int64_t* row_ptr = nullptr;  // dim + 1
  const int32_t* col = nullptr;
  const uint32_t* code = nullptr;
  const double* spec = nullptr;      // complex coefficients of special hops
  const double* diag = nullptr;      // diagonal energies
  const double* ptab = nullptr;      // [8][L] complex amp * phase
};
void launch_hops_count(const LatticeConst* g, BlockView bv, int64_t* cnt, int64_t* scnt, hipStream_t s);
void launch_hops_fill(const LatticeConst* g, BlockView bv, const int64_t* row_ptr, const int64_t* srow_ptr,
                      int32_t* hcol, uint32_t* hcode, double* spec, double* diag, hipStream_t s);
void launch_apply_hl(const HopList& h, int64_t dim, int R, int W, const int64_t* offs, int n_canon, int cus,
                     const double* x, double* y, double* partials, unsigned* counter, double* dot_out, hipStream_t s);
// Column-tiled hop-list gather for tiles of W <= 64 columns: 64 / W rows per
// wavefront, lane = (row in the wavefront, column in the tile). A task is a
// run of 4 * 64 / W consecutive rows of one canonical-up group and one
// column tile; tasks are ordered (group, tile, rows) and cut into one
// contiguous range per XCD, so the workgroups of an XCD gather the down-spin
// hops of one group's tile slab (rows * W * 16 B) from its L2.
struct TileTask {
  int64_t row0;
  int32_t nrows;
  int32_t tile;
};
int tile_task_rows(int W);
void launch_apply_hl_tiles(const HopList& h, int R, int W, const TileTask* tasks, int64_t ntask, int cus,
                           const double* x, double* y, hipStream_t s);
// CSR export: counts per row (pass 1) and entries (pass 2).
void launch_csr_count(const LatticeConst* g, BlockView bv, int all_dirs, int32_t* counts, hipStream_t s);
void launch_csr_fill(const LatticeConst* g, BlockView bv, int all_dirs, const int64_t* row_ptr, int32_t* col,
                     double* val, double* diag, hipStream_t s);
// Dense column-major Hermitised matrices of a batch of small blocks.
void launch_dense_build(const LatticeConst* g, const BlockView* bvs, int nblk, const int64_t* moff,
                        double* mats, int64_t max_dim, hipStream_t s);

// ---- Lanczos (ftlm_lanczos.hip) ----
struct LanczosState {
  double* alpha;   // [steps][R]
  double* beta;    // [steps][R]
  double* scale;   // [R] 1/b of the current vector (tree mode)
  double* sprev;   // [R]
  double* bprev;   // [R]
  double* dot;     // [R] reduction output
  double* nrm;     // [R] reduction output
  int32_t* done;   // [R]
  int32_t* mcount; // [R]
};
void launch_lanczos_init(int R, LanczosState st, hipStream_t s);
void launch_dot_serial(int64_t dim, int R, const double* x, const double* y, double* out, hipStream_t s);
void launch_norm_serial(int64_t dim, int R, const double* x, double* out, hipStream_t s);
void launch_recur_serial(int64_t dim, int R, const double* w, const double* q, double* qp, const double* alpha_k,
                         const double* bprev, const int32_t* done, hipStream_t s);
void launch_finish_serial(int R, int k, int kdim, LanczosState st, hipStream_t s);
void launch_normalize(int64_t dim, int R, double* q, const double* beta_k, const int32_t* done, int k,
                      const int32_t* mcount, hipStream_t s);
void launch_recur_tree(int64_t dim, int R, const double* y, const double* x, double* xp, int k, int kdim,
                       LanczosState st, double* partials, unsigned* counter, int grid, hipStream_t s);
// per-column Re<x, y>, tree order (per workgroup, then workgroups in index
// order in the last workgroup)
void launch_dot_tree(int64_t dim, int R, const double* x, const double* y, double* partials, unsigned* counter,
                     double* out, int grid, hipStream_t s);
void launch_transpose_in(int64_t dim, int R, int r0, int cnt, const double* src, double* x, hipStream_t s);
void launch_ritz_ql(int R, int steps, const double* alpha, const double* beta, const int32_t* mcount, int wdiv,
                    double* eig, double* w, double* scratch, int32_t* iscratch, int32_t* fail, hipStream_t s);
void launch_ritz(int R, int steps, const double* alpha, const double* beta, const int32_t* mcount, int wdiv,
                 double* eig, double* w, double* scratch, hipStream_t s);

// ---- packed Lanczos (ftlm_pack.hip) ----
// One block of a pack: its hop list, diagonal, coefficient tables and
// sample-major start vectors [R][dim]; its R columns are col0..col0+R-1 of
// the pack's alpha/beta/mcount arrays.
struct PackEntry {
  const int64_t* row_ptr;
  const int32_t* col;
  const uint32_t* code;
  const double2* spec;
  const double* diag;
  const double2* ptab;
  const double2* start;
  int64_t dim;
  int32_t kdim;  // min(steps, dim)
  int32_t col0;
};
// A work item: columns c0..c0+w-1 of one entry.
struct PackItem {
  int32_t entry, c0, w, pad;
};
struct PackArgs {
  const PackEntry* entries;
  const PackItem* items;      // in the order they are taken
  const int32_t* n_items;     // device scalar (read by the kernel: graph replays see new packs)
  unsigned* queue;            // next item (reset by launch_pack_reset)
  int32_t* fail;              // QL non-convergence count (reset likewise)
  double2* slots;             // per workgroup: 3 vectors of slot_stride
  int64_t slot_stride;        // >= max dim * W
  int32_t rcap;               // column capacity (row stride of alpha/beta)
  int32_t kmax;               // > 0: stop after kmax steps (tuning runs)
  int32_t tree;               // 0: column sums in row order (bit-identical to the
                              // reference); 1: fused per-thread partials reduced
                              // in a fixed order (tree reductions)
  double* alpha;              // [steps][rcap]
  double* beta;               // [steps][rcap]
  int32_t* mcount;            // [rcap]
};
void launch_pack_reset(const PackArgs& a, int rcap, hipStream_t s);
void launch_pack_lanczos(const PackArgs& a, int W, int grid, hipStream_t s);

// ---- seeded start vectors on the device (ftlm_rng.hip) ----
// Samples r0..r0+cnt-1 of block seed bseed, dim rows each, normalised, into
// output columns col0..col0+cnt-1 of out: layout 0 [dim][R] (column c),
// layout 1 [R][dim] (row c); norms: cnt doubles of scratch.
void launch_start_vectors(uint64_t bseed, int r0, int col0, int cnt, int R, int64_t dim, int layout, double* out,
                          double* norms, hipStream_t s);

// ---- thermodynamic sums (ftlm_thermo.hip) ----
// part: thermo_workgroups(nv) * nb * nm * 3 doubles; out: nb * nm * 4
// (emin, z, n1, n2), pair p = beta index * nm + mu index.
int64_t thermo_workgroups(int64_t nv);
// allocation-free variant: a fixed device array (kThermoScratchDoubles, one
// copy per GPU) holds the partials and the per-mu minima, which are computed
// on the device. thermo_scratch_setup resolves the array's device address
// with the device current (once per device); thermo_scratch returns it
// (nullptr before setup).
constexpr int64_t kThermoScratchDoubles = (int64_t)4 << 20;  // 32 MiB
constexpr int kMaxDevices = 64;
int thermo_scratch_setup(int device);
double* thermo_scratch(int device);
int64_t thermo_noalloc_doubles(int64_t nv, int nb, int nm);
void launch_thermo_noalloc(const double* E, const double* W, const uint8_t* N, int64_t nv, const double* betas,
                           int nb, const double* mus, int nm, double* scr, double* out, hipStream_t s);
void launch_thermo(const double* E, const double* W, const uint8_t* N, int64_t nv, const double* betas, int nb,
                   const double* mus, int nm, const double* emin, double* part, double* out, hipStream_t s);

// ---- memory-system probes (ftlm_probe.hip) ----
// kind 0 stream read, 1 stream copy, 2 random-row gather of k rows per
// output row; rows x width complex doubles in x (and y).
void launch_memprobe(int kind, int64_t rows, int width, int k, int cus, const double* x, double* y, hipStream_t s);

// ---- exact (ftlm_exact.hip) ----
void launch_herm_eig(const int64_t* moff, const int64_t* dims, int nblk, double* mats, double* work,
                     double* eig_out, const int64_t* eoff, hipStream_t s);

}  // namespace ftlm
