// ftlm_types.h -- plain structs shared by the host orchestration and the
// gfx950 kernels of the FTLM engine.
//
// Basis layout (one (N_up, N_down) sector, one momentum block k):
//   * up configurations are indexed by their colex rank among the C(L,N_up)
//     N_up-bit masks (= ascending integer order, the order Gosper's hack in
//     main.cpp:462-484 produces);
//   * a "canonical up" is an up mask that is the minimum of its translation
//     orbit. A representative (lexicographic minimum of its (up, down) orbit,
//     the parent order of main.cpp:643-651) always has a canonical up, and
//     the block's rows are ordered by (canonical up, down) exactly like the
//     reference's active parents;
//   * canonical ups whose stabiliser is trivial own all C(L,N_down) downs,
//     each with norm factor 1: row = offs[c] + rank(down);
//   * the few canonical ups with a non-trivial stabiliser ("nt" ups) carry a
//     per-down table: local row index (or -1 when the down is not the
//     stabiliser-minimum or the state is inactive at this k) and norm factor.
#pragma once
#include <stdint.h>

#define FTLM_MAX_SITES 63
#define FTLM_MAX_LX 63

struct LatticeConst {
  int lx, ly, L;
  int ndir;                 // number of hopping directions used by the apply
  int8_t dirs[8];           // those directions (ascending)
  uint8_t nb[FTLM_MAX_SITES][8];
  int nbonds;
  uint8_t bonds[2 * FTLM_MAX_SITES][2];
  uint64_t full;            // mask of the L site bits
  uint64_t col_ge[FTLM_MAX_LX + 1];  // sites with column x >= k
  uint64_t rows_lt[FTLM_MAX_LX + 1]; // sites with row y < k
  double U, V;
};

// Device view of the current block (all pointers are device pointers).
struct BlockView {
  int nup, ndn;
  int n_canon;              // canonical ups in the sector
  int64_t n_dnc;            // C(L, N_down)
  int64_t dim;              // rows of this block
  // per up configuration (indexed by colex rank)
  const int32_t* upc_canon;   // canonical index of the orbit minimum
  const uint8_t* upc_tfirst;  // first shift i with T_i(u) = orbit minimum
  const int8_t* upc_tsign;    // fermion sign of T_tfirst on u
  const uint64_t* upc_amask;  // all shifts i with T_i(u) = orbit minimum
  // per canonical up
  const uint64_t* can_bits;
  const uint64_t* can_stab;   // shifts fixing it (bit 0 always set)
  const int32_t* can_nt;      // index into the nt tables or -1
  const int64_t* offs;        // first row of each canonical up (n_canon+1)
  // per nt up x per down rank
  const int32_t* nt_loc;      // local row index or -1
  const double* nt_dnf;       // norm factor of active representatives
  const int32_t* nt_list;     // down ranks of the active rows, per nt up
  const int64_t* nt_listoff;  // start of each nt up's list
  const uint64_t* binom;      // binom[k * 64 + p] = C(p, k), k <= 63
  const double* ptab;         // [8][L] complex: amp(dir) * momentum phase(shift)
};

// Counters of the last Lanczos run (device-event timings in milliseconds).
struct RunStats {
  double ms_apply;     // summed duration of all H-apply launches
  double ms_total;     // whole run from first to last kernel
  int64_t n_apply;     // number of H-apply launches
  double bytes_apply;  // algorithmic bytes of those launches
};
