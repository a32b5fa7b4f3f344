// ftlm_device.h -- device-side helpers of the FTLM engine (gfx950).
//
// All floating-point code that must round like the reference is compiled
// with -ffp-contract=off: complex products are (ac - bd, ad + bc) with
// separately rounded products, as g++ -O3 without FMA evaluates
// std::complex<double> arithmetic in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ftlm_types.h"

namespace ftlm {

struct cd {
  double re, im;
};

__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd cscale(double s, cd a) { return {s * a.re, s * a.im}; }
// conj(a) * b in the reference's rounding: (ar*br + ai*bi, ar*bi - ai*br)
__device__ __forceinline__ cd cconjmul(cd a, cd b) { return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }

__device__ __forceinline__ int popc(uint64_t x) { return __popcll(x); }
__device__ __forceinline__ int ctz(uint64_t x) { return __ffsll((unsigned long long)x) - 1; }

// colex rank of an n-bit mask among all masks with the same popcount:
// sum over set bits (ascending position p_t, t = 0..n-1) of C(p_t, t+1).
__device__ __forceinline__ int64_t colex_rank(uint64_t bits, const uint64_t* binom) {
  int64_t r = 0;
  int t = 1;
  while (bits) {
    const int p = ctz(bits);
    r += (int64_t)binom[t * 64 + p];
    ++t;
    bits &= bits - 1;
  }
  return r;
}

// inverse of colex_rank for masks with n bits among L sites
__device__ __forceinline__ uint64_t colex_unrank(int64_t r, int n, int L, const uint64_t* binom) {
  uint64_t bits = 0;
  int p = L - 1;
  for (int t = n; t >= 1; --t) {
    while ((int64_t)binom[t * 64 + p] > r) --p;
    bits |= 1ull << p;
    r -= (int64_t)binom[t * 64 + p];
    --p;
  }
  return bits;
}

// Translation by shift i = dy*lx + dx: site (x, y) -> ((x+dx)%lx, (y+dy)%ly)
// (main.cpp:579-620).
__device__ __forceinline__ uint64_t translate(uint64_t b, int i, const LatticeConst& g) {
  const int dx = i % g.lx, dy = i / g.lx;
  if (dx) b = ((b << dx) & g.col_ge[dx]) | ((b >> (g.lx - dx)) & ~g.col_ge[dx] & g.full);
  if (dy) {
    const int s = g.lx * dy;
    b = ((b << s) | (b >> (g.L - s))) & g.full;
  }
  return b;
}

// Fermion sign of translating the occupied set b by shift i: the parity of
// inversions of the translated site list (main.cpp:598-607). Same-row pairs
// invert when exactly one of them wraps in x; pairs in different rows invert
// when exactly one wraps in y.
__device__ __forceinline__ int tsign(uint64_t b, int i, const LatticeConst& g) {
  const int dx = i % g.lx, dy = i / g.lx;
  int par = 0;
  if (dx) {
    const uint64_t wrap = b & g.col_ge[g.lx - dx];  // columns x >= lx-dx wrap
    const uint64_t stay = b & ~g.col_ge[g.lx - dx];
    const uint64_t row = (g.lx == 64) ? ~0ull : ((1ull << g.lx) - 1);
    for (int y = 0; y < g.ly; ++y) {
      const uint64_t m = row << (y * g.lx);
      par ^= popc(wrap & m) & popc(stay & m) & 1;
    }
  }
  if (dy) {
    const int lo = popc(b & g.rows_lt[g.ly - dy]);
    const int hi = popc(b) - lo;
    par ^= lo & hi & 1;
  }
  return par ? -1 : 1;
}

__device__ __forceinline__ int inv_shift(int i, const LatticeConst& g) {
  const int dx = i % g.lx, dy = i / g.lx;
  return ((g.ly - dy) % g.ly) * g.lx + (g.lx - dx) % g.lx;
}

// (-1)^(occupied sites strictly between a and b) (main.cpp:503-512)
__device__ __forceinline__ int hop_fsign(uint64_t bits, int a, int b) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  if (hi - lo <= 1) return 1;
  const uint64_t m = ((1ull << (hi - lo - 1)) - 1) << (lo + 1);
  return (popc(bits & m) & 1) ? -1 : 1;
}

// U*doublons + V*sum_bonds n_i n_j (main.cpp:486-501)
__device__ __forceinline__ double diag_energy(uint64_t up, uint64_t dn, const LatticeConst& g) {
  const int dbl = popc(up & dn);
  int nn = 0;
  for (int k = 0; k < g.nbonds; ++k) {
    const int a = g.bonds[k][0], b = g.bonds[k][1];
    const int na = (int)((up >> a) & 1) + (int)((dn >> a) & 1);
    const int nb = (int)((up >> b) & 1) + (int)((dn >> b) & 1);
    nn += na * nb;
  }
  return g.U * (double)dbl + g.V * (double)nn;
}

// Row -> (canonical up index, up bits, down bits, norm factor).
struct RowState {
  int c;
  uint64_t up, dn;
  double dnf;
};

// Row state when its canonical up group c is already known.
__device__ __forceinline__ RowState row_state_in(const BlockView& bv, const LatticeConst& g, int c, int64_t row) {
  RowState s;
  s.c = c;
  s.up = bv.can_bits[c];
  const int64_t local = row - bv.offs[c];
  const int x = bv.can_nt[c];
  int64_t drank;
  if (x < 0) {
    drank = local;
    s.dnf = 1.0;
  } else {
    drank = bv.nt_list[bv.nt_listoff[x] + local];
    s.dnf = bv.nt_dnf[(int64_t)x * bv.n_dnc + drank];
  }
  s.dn = colex_unrank(drank, bv.ndn, g.L, bv.binom);
  return s;
}

__device__ __forceinline__ RowState row_state(const BlockView& bv, const LatticeConst& g, int64_t row) {
  // last c with offs[c] <= row
  int lo = 0, hi = bv.n_canon - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bv.offs[mid] <= row) lo = mid;
    else hi = mid - 1;
  }
  RowState s;
  s.c = lo;
  s.up = bv.can_bits[lo];
  const int64_t local = row - bv.offs[lo];
  const int x = bv.can_nt[lo];
  int64_t drank;
  if (x < 0) {
    drank = local;
    s.dnf = 1.0;
  } else {
    drank = bv.nt_list[bv.nt_listoff[x] + local];
    s.dnf = bv.nt_dnf[(int64_t)x * bv.n_dnc + drank];
  }
  s.dn = colex_unrank(drank, bv.ndn, g.L, bv.binom);
  return s;
}

// One hop of the row (up, dn) in block view bv: electron of `spin` moves
// from site `from` to the empty site `to` along direction `dir`. Returns
// false when the target state is inactive in this block (main.cpp:791-793).
// Otherwise gives the target row and the coefficient
//   amp(dir) * fsign * sign * phase[shift] * (dnf[target] / dnf[row])
// (main.cpp:683-716, 794-799) with the reference's rounding: the table
// ptab[dir][shift] holds amp(dir) (*) phase[shift] computed on the host by the
// same std::complex expression, signs are exact negations and a norm ratio
// of exactly 1 is skipped as the exact no-op it is.
__device__ __forceinline__ bool hop_target_ex(const BlockView& bv, const LatticeConst& g, int c_row, uint64_t up,
                                              uint64_t dn, double dnf_row, int spin, int from, int to, int dir,
                                              int64_t* col, cd* coef, uint32_t* code, bool* special) {
  const uint64_t bits = spin ? dn : up;
  int sgn = hop_fsign(bits, from, to);
  const uint64_t moved = bits ^ (1ull << from) ^ (1ull << to);
  int c;
  uint64_t tdn;
  int jsel;
  if (spin == 0) {
    const int64_t r = colex_rank(moved, bv.binom);
    c = bv.upc_canon[r];
    const uint64_t am = bv.upc_amask[r];
    if ((am & (am - 1)) == 0) {
      const int i = bv.upc_tfirst[r];
      tdn = translate(dn, i, g);
      sgn *= bv.upc_tsign[r] * tsign(dn, i, g);
      jsel = inv_shift(i, g);
    } else {
      // orbit minimum reached by several shifts: lexicographic minimum over
      // the down parts, ties broken by the smallest relation shift (the
      // first emplace of main.cpp:669 wins)
      uint64_t m = am;
      uint64_t best = ~0ull;
      int bj = 1 << 30, bi = 0;
      while (m) {
        const int i = ctz(m);
        m &= m - 1;
        const uint64_t t = translate(dn, i, g);
        const int j = inv_shift(i, g);
        if (t < best || (t == best && j < bj)) {
          best = t;
          bj = j;
          bi = i;
        }
      }
      tdn = best;
      jsel = bj;
      sgn *= tsign(moved, bi, g) * tsign(dn, bi, g);
    }
  } else {
    c = c_row;
    const uint64_t st = bv.can_stab[c];
    if (st == 1ull) {
      tdn = moved;
      jsel = 0;
    } else {
      uint64_t m = st;
      uint64_t best = ~0ull;
      int bj = 1 << 30, bi = 0;
      while (m) {
        const int i = ctz(m);
        m &= m - 1;
        const uint64_t t = translate(moved, i, g);
        const int j = inv_shift(i, g);
        if (t < best || (t == best && j < bj)) {
          best = t;
          bj = j;
          bi = i;
        }
      }
      tdn = best;
      jsel = bj;
      sgn *= tsign(up, bi, g) * tsign(moved, bi, g);
    }
  }
  const int x = bv.can_nt[c];
  const int64_t b = colex_rank(tdn, bv.binom);
  double dnf_t = 1.0;
  if (x < 0) {
    *col = bv.offs[c] + b;
  } else {
    const int64_t loc = bv.nt_loc[(int64_t)x * bv.n_dnc + b];
    if (loc < 0) return false;
    *col = bv.offs[c] + loc;
    dnf_t = bv.nt_dnf[(int64_t)x * bv.n_dnc + b];
  }
  const double* pt = bv.ptab + 2 * (dir * g.L + jsel);
  cd v = {pt[0], pt[1]};
  const bool ratio_needed = dnf_t != 1.0 || dnf_row != 1.0;
  if (ratio_needed) {
    const double ratio = dnf_t / dnf_row;
    v = {v.re * ratio, v.im * ratio};
  }
  if (sgn < 0) v = {-v.re, -v.im};
  *coef = v;
  if (code) *code = (uint32_t)(dir * g.L + jsel) | (sgn < 0 ? (1u << 31) : 0u);
  if (special) *special = ratio_needed;
  return true;
}

// Column sums of per-workgroup partials (partials[b * R + r], b < nb), run
// by the last workgroup of a grid (256 threads) once every workgroup has
// written its row. Columns are taken 256 at a time; the 256 / cols threads of
// a column each add a strided share of the workgroups into eight independent
// accumulators (so the loads overlap), and the shares are then added in
// thread order: a fixed order, the same sums run to run. use(r, sum) is
// called once per column, by one thread.
template <typename F>
__device__ __forceinline__ void last_block_column_sums(const double* __restrict__ partials, unsigned nb, int R,
                                                       F&& use) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  for (int c0 = 0; c0 < R; c0 += 256) {
    const int cols = R - c0 < 256 ? R - c0 : 256;
    const int segs = 256 / cols;
    const int q = t / cols, r = t % cols;
    double acc = 0.0;
    if (q < segs) {
      double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      unsigned b = (unsigned)q;
      const unsigned step = (unsigned)segs;
      for (; b + 7 * step < nb; b += 8 * step) {
#pragma unroll
        for (int u = 0; u < 8; ++u)
          a[u] += __builtin_nontemporal_load(&partials[(int64_t)(b + u * step) * R + c0 + r]);
      }
      for (; b < nb; b += step) a[0] += __builtin_nontemporal_load(&partials[(int64_t)b * R + c0 + r]);
      acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
    __syncthreads();  // the previous chunk's readers are done with sh
    if (q < segs) sh[q * cols + r] = acc;
    __syncthreads();
    if (t < cols) {
      double sum = 0.0;
      for (int j = 0; j < segs; ++j) sum += sh[j * cols + t];
      use(c0 + t, sum);
    }
  }
}

__device__ __forceinline__ bool hop_target(const BlockView& bv, const LatticeConst& g, int c_row, uint64_t up,
                                           uint64_t dn, double dnf_row, int spin, int from, int to, int dir,
                                           int64_t* col, cd* coef) {
  return hop_target_ex(bv, g, c_row, up, dn, dnf_row, spin, from, to, dir, col, coef, nullptr, nullptr);
}

}  // namespace ftlm
