// ftlm_basis.hip -- on-device construction of (N_up, N_down, k) block bases.
//
// Replaces the serial hash-map basis of build_sector_basis /
// build_momentum_block (main.cpp:622-805) with closed-form tables:
//   representative (u, d) <=> u is the minimum of its up orbit and d is the
//   minimum of its orbit under the stabiliser of u,
// so a block is enumerated by ranks instead of by visiting all C(L,Nu)C(L,Nd)
// states. Row order, degeneracies and norm factors equal the reference's
// (checked bit for bit against reference-built blocks by the GPU tests).
#include "ftlm_device.h"
#include "ftlm_launch.h"

namespace ftlm {

namespace {

__global__ void k_up_table(const LatticeConst* __restrict__ gp, const uint64_t* __restrict__ binom, int nup,
                           int64_t n_upc, uint8_t* is_canon, int64_t* min_rank, uint8_t* tfirst, int8_t* tsg,
                           uint64_t* amask) {
  const LatticeConst& g = *gp;
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n_upc) return;
  const uint64_t u = colex_unrank(r, nup, g.L, binom);
  uint64_t best = u;
  uint64_t mask = 1;  // shift 0 maps u to itself
  for (int i = 1; i < g.L; ++i) {
    const uint64_t t = translate(u, i, g);
    if (t < best) {
      best = t;
      mask = 1ull << i;
    } else if (t == best) {
      mask |= 1ull << i;
    }
  }
  const int i0 = ctz(mask);
  is_canon[r] = best == u;
  min_rank[r] = colex_rank(best, binom);
  tfirst[r] = (uint8_t)i0;
  tsg[r] = (int8_t)tsign(u, i0, g);
  amask[r] = mask;
}

// One workgroup per segment; 1024 threads scan a segment in chunks.
__global__ void __launch_bounds__(1024) k_segscan(const uint8_t* __restrict__ flags, int64_t seg_len,
                                                  int32_t* __restrict__ out, int64_t* __restrict__ counts) {
  __shared__ int64_t part[1024];
  const int seg = blockIdx.x;
  const uint8_t* f = flags + (int64_t)seg * seg_len;
  int32_t* o = out + (int64_t)seg * seg_len;
  const int tid = threadIdx.x;
  const int64_t per = (seg_len + 1023) / 1024;
  const int64_t b0 = tid * per, b1 = (b0 + per < seg_len) ? b0 + per : seg_len;
  int64_t cnt = 0;
  for (int64_t i = b0; i < b1; ++i) cnt += f[i] ? 1 : 0;
  part[tid] = cnt;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int64_t v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int64_t run = part[tid] - cnt;
  for (int64_t i = b0; i < b1; ++i) {
    if (f[i]) o[i] = (int32_t)(run++);
    else o[i] = -1;
  }
  if (tid == 1023) counts[seg] = part[1023];
}

__global__ void __launch_bounds__(1024) k_prefix64(const int64_t* __restrict__ in, int64_t n,
                                                   int64_t* __restrict__ out) {
  __shared__ int64_t part[1024];
  const int tid = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t b0 = tid * per, b1 = (b0 + per < n) ? b0 + per : n;
  int64_t s = 0;
  for (int64_t i = b0; i < b1; ++i) s += in[i];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int64_t v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int64_t run = part[tid] - s;
  for (int64_t i = b0; i < b1; ++i) {
    out[i] = run;
    run += in[i];
  }
  if (tid == 1023) out[n] = part[1023];
}

// Multi-workgroup exclusive scan for long arrays: tiles of 4096 elements.
// Pass 1 sums each tile (coalesced), the tile sums are scanned by
// k_prefix64, pass 3 scans each tile locally (16 consecutive elements per
// thread, then the thread totals in LDS) and adds the tile's offset.
constexpr int kScanTile = 4096;

__global__ void __launch_bounds__(256) k_tile_sums(const int64_t* __restrict__ in, int64_t n,
                                                   int64_t* __restrict__ sums) {
  __shared__ int64_t red[256];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t s = 0;
  for (int j = 0; j < kScanTile / 256; ++j) {
    const int64_t i = base + j * 256 + threadIdx.x;
    if (i < n) s += in[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_tile_scan(const int64_t* __restrict__ in, int64_t n,
                                                   const int64_t* __restrict__ tile_off, int64_t* __restrict__ out) {
  __shared__ int64_t part[256];
  constexpr int kPer = kScanTile / 256;
  const int tid = threadIdx.x;
  const int64_t b0 = (int64_t)blockIdx.x * kScanTile + (int64_t)tid * kPer;
  int64_t v[kPer];
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    v[j] = b0 + j < n ? in[b0 + j] : 0;
    s += v[j];
  }
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int64_t a = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += a;
    __syncthreads();
  }
  int64_t run = tile_off[blockIdx.x] + part[tid] - s;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    if (b0 + j < n) out[b0 + j] = run;
    run += v[j];
  }
  if (b0 < n && n <= b0 + kPer) out[n] = run;  // the thread holding the last element
}

__global__ void k_canon_fill(int64_t n_upc, int nup, int L, const uint64_t* __restrict__ binom,
                             const uint8_t* __restrict__ is_canon, const int32_t* __restrict__ cidx,
                             const int64_t* __restrict__ min_rank, const uint64_t* __restrict__ amask,
                             int32_t* upc_canon, uint64_t* can_bits, uint64_t* can_stab, uint8_t* nt_flag) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n_upc) return;
  upc_canon[r] = cidx[min_rank[r]];
  if (is_canon[r]) {
    const int c = cidx[r];
    can_bits[c] = colex_unrank(r, nup, L, binom);
    can_stab[c] = amask[r];
    nt_flag[c] = amask[r] != 1ull;
  }
}

__global__ void k_nt_fill(int n_canon, const int32_t* __restrict__ nt_index, int32_t* can_nt, int32_t* nt_to_c) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_canon) return;
  const int x = nt_index[c];
  can_nt[c] = x;
  if (x >= 0) nt_to_c[x] = c;
}

// |z| for the orbit character sums. In an active block the sum is |G| times
// a unit root up to rounding, so the smaller component is below 2^-27 of the
// larger and hi*sqrt(1 + r^2) returns hi exactly, which is also the correctly
// rounded hypot the reference gets from std::abs (main.cpp:767).
__device__ __forceinline__ double cabs_ref(double re, double im) {
  const double a = fabs(re), b = fabs(im);
  const double hi = a > b ? a : b, lo = a > b ? b : a;
  if (hi == 0.0) return 0.0;
  const double r = lo / hi;
  return hi * sqrt(1.0 + r * r);
}

__global__ void k_nt_active(const LatticeConst* __restrict__ gp, const uint64_t* __restrict__ binom, int ndn,
                            int64_t n_dnc, int n_nt, const int32_t* __restrict__ nt_to_c,
                            const uint64_t* __restrict__ can_bits, const uint64_t* __restrict__ can_stab,
                            const double* __restrict__ phase_k, uint8_t* act, double* dnf) {
  const LatticeConst& g = *gp;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_nt * n_dnc) return;
  const int x = (int)(t / n_dnc);
  const int64_t b = t - (int64_t)x * n_dnc;
  const int c = nt_to_c[x];
  const uint64_t u = can_bits[c];
  const uint64_t st = can_stab[c];
  const uint64_t d = colex_unrank(b, ndn, g.L, binom);
  // representative iff d is the minimum of its orbit under stab(u)
  uint64_t m = st & ~1ull;
  bool rep = true;
  uint64_t fix = 1ull;
  while (m) {
    const int i = ctz(m);
    m &= m - 1;
    const uint64_t td = translate(d, i, g);
    if (td < d) {
      rep = false;
      break;
    }
    if (td == d) fix |= 1ull << i;
  }
  double dv = 0.0;
  bool on = false;
  if (rep) {
    // sum over the state's stabiliser in shift order (main.cpp:753-765)
    double sr = 0.0, si = 0.0;
    uint64_t f = fix;
    while (f) {
      const int i = ctz(f);
      f &= f - 1;
      const double sg = (double)(tsign(u, i, g) * tsign(d, i, g));
      sr = sr + sg * phase_k[2 * i];
      si = si + sg * phase_k[2 * i + 1];
    }
    dv = cabs_ref(sr, si) / sqrt((double)popc(fix));
    on = dv > 1e-12;
  }
  act[t] = on;
  dnf[t] = dv;
}

__global__ void k_nt_list(int n_nt, int64_t n_dnc, const int32_t* __restrict__ nt_loc,
                          const int64_t* __restrict__ listoff, int32_t* list) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_nt * n_dnc) return;
  const int x = (int)(t / n_dnc);
  const int64_t b = t - (int64_t)x * n_dnc;
  const int32_t loc = nt_loc[t];
  if (loc >= 0) list[listoff[x] + loc] = (int32_t)b;
}

__global__ void k_row_counts(int n_canon, int64_t n_dnc, const int32_t* __restrict__ can_nt,
                             const int64_t* __restrict__ nt_count, int64_t* counts) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_canon) return;
  const int x = can_nt[c];
  counts[c] = x < 0 ? n_dnc : nt_count[x];
}

__global__ void k_block_reps(const LatticeConst* __restrict__ gp, BlockView bv, uint64_t* up, uint64_t* dn,
                             int32_t* deg, double* dnf) {
  const LatticeConst& g = *gp;
  const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= bv.dim) return;
  const RowState s = row_state(bv, g, row);
  up[row] = s.up;
  dn[row] = s.dn;
  dnf[row] = s.dnf;
  int dg = 1;
  const uint64_t st = bv.can_stab[s.c];
  if (st != 1ull) {
    dg = 0;
    uint64_t m = st;
    while (m) {
      const int i = ctz(m);
      m &= m - 1;
      dg += translate(s.dn, i, g) == s.dn;
    }
  }
  deg[row] = dg;
}

inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void launch_up_table(const LatticeConst* g, const uint64_t* binom, int nup, int64_t n_upc, uint8_t* is_canon,
                     int64_t* min_rank, uint8_t* tfirst, int8_t* tsign_out, uint64_t* amask, hipStream_t s) {
  k_up_table<<<nblocks(n_upc, 256), 256, 0, s>>>(g, binom, nup, n_upc, is_canon, min_rank, tfirst, tsign_out,
                                                 amask);
}

void launch_segscan(const uint8_t* flags, int64_t seg_len, int nseg, int32_t* out, int64_t* counts,
                    hipStream_t s) {
  if (nseg <= 0 || seg_len <= 0) return;
  k_segscan<<<nseg, 1024, 0, s>>>(flags, seg_len, out, counts);
}

void launch_prefix64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s) {
  k_prefix64<<<1, 1024, 0, s>>>(in, n, out);
}

int64_t prefix64_tmp_words(int64_t n) { return 2 * ((n + kScanTile - 1) / kScanTile) + 1; }

void launch_prefix64_tiled(const int64_t* in, int64_t n, int64_t* out, int64_t* tmp, hipStream_t s) {
  if (n <= kScanTile) {
    k_prefix64<<<1, 1024, 0, s>>>(in, n, out);
    return;
  }
  const int64_t nt = (n + kScanTile - 1) / kScanTile;
  // tmp[0..nt): tile sums, then their exclusive scan in tmp[nt..2nt] (+ total)
  k_tile_sums<<<(unsigned)nt, 256, 0, s>>>(in, n, tmp);
  k_prefix64<<<1, 1024, 0, s>>>(tmp, nt, tmp + nt);
  k_tile_scan<<<(unsigned)nt, 256, 0, s>>>(in, n, tmp + nt, out);
}

void launch_canon_fill(int64_t n_upc, int nup, int L, const uint64_t* binom, const uint8_t* is_canon,
                       const int32_t* cidx_of_rank, const int64_t* min_rank, const uint64_t* amask,
                       int32_t* upc_canon, uint64_t* can_bits, uint64_t* can_stab, uint8_t* nt_flag,
                       hipStream_t s) {
  k_canon_fill<<<nblocks(n_upc, 256), 256, 0, s>>>(n_upc, nup, L, binom, is_canon, cidx_of_rank, min_rank, amask,
                                                   upc_canon, can_bits, can_stab, nt_flag);
}

void launch_nt_fill(int n_canon, const int32_t* nt_index, int32_t* can_nt, int32_t* nt_to_c, hipStream_t s) {
  k_nt_fill<<<nblocks(n_canon, 256), 256, 0, s>>>(n_canon, nt_index, can_nt, nt_to_c);
}

void launch_nt_active(const LatticeConst* g, const uint64_t* binom, int ndn, int64_t n_dnc, int n_nt,
                      const int32_t* nt_to_c, const uint64_t* can_bits, const uint64_t* can_stab,
                      const double* phase_k, uint8_t* act, double* dnf, hipStream_t s) {
  const int64_t n = (int64_t)n_nt * n_dnc;
  if (n == 0) return;
  k_nt_active<<<nblocks(n, 256), 256, 0, s>>>(g, binom, ndn, n_dnc, n_nt, nt_to_c, can_bits, can_stab, phase_k,
                                              act, dnf);
}

void launch_nt_list(int n_nt, int64_t n_dnc, const int32_t* nt_loc, const int64_t* listoff, int32_t* list,
                    hipStream_t s) {
  const int64_t n = (int64_t)n_nt * n_dnc;
  if (n == 0) return;
  k_nt_list<<<nblocks(n, 256), 256, 0, s>>>(n_nt, n_dnc, nt_loc, listoff, list);
}

void launch_row_counts(int n_canon, int64_t n_dnc, const int32_t* can_nt, const int64_t* nt_count,
                       int64_t* counts, hipStream_t s) {
  k_row_counts<<<nblocks(n_canon, 256), 256, 0, s>>>(n_canon, n_dnc, can_nt, nt_count, counts);
}

void launch_block_reps(const LatticeConst* g, BlockView bv, uint64_t* up, uint64_t* dn, int32_t* deg,
                       double* dnf, hipStream_t s) {
  if (bv.dim == 0) return;
  k_block_reps<<<nblocks(bv.dim, 256), 256, 0, s>>>(g, bv, up, dn, deg, dnf);
}

}  // namespace ftlm
