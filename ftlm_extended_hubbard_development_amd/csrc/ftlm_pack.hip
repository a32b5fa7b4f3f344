// ftlm_pack.hip -- packed Lanczos over many small blocks in one persistent
// launch.
//
// A pack holds blocks of any (N_up, N_down, k) sector and twist, each with
// its own hop list, start vectors and R samples. Work items are (block,
// group of W columns); a persistent grid of workgroups takes items from a
// queue ordered by decreasing cost, and one workgroup runs every Lanczos step
// of its item (main.cpp:1026-1057) in a private slot of three [dim][W]
// vectors. The reductions of a column (alpha dot, norm) are summed row by
// row in the reference's order (main.cpp:807-821) by the column's own
// thread, so alpha/beta are bit-identical to the reference whatever W, the
// grid size or the item order. No grid-wide synchronisation: each workgroup
// exits once the queue is drained.
#include "ftlm_device.h"
#include "ftlm_launch.h"

namespace ftlm {

namespace {

constexpr int kPackThreads = 512;
constexpr int kG = 4;  // hops per gather group in the apply loop
constexpr int kPackWaves = kPackThreads / 64;
constexpr uint32_t kNegP = 1u << 31, kSpecialP = 1u << 30;

__device__ __forceinline__ cd pack_coef(uint32_t code, const double2* __restrict__ ptab,
                                        const double2* __restrict__ spec) {
  if (code & kSpecialP) {
    const double2 v = spec[code & (kSpecialP - 1)];
    return {v.x, v.y};
  }
  const double2 v = ptab[code & (kSpecialP - 1)];
  return (code & kNegP) ? cd{-v.x, -v.y} : cd{v.x, v.y};
}

__global__ void k_pack_reset(PackArgs a, int rcap) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i == 0) {
    a.queue[0] = 0u;
    a.fail[0] = 0;
  }
  for (int64_t j = i; j < rcap; j += (int64_t)gridDim.x * blockDim.x) a.mcount[j] = 0;
}

// serial column sums in the order of k_dot_serial / k_norm_serial
__device__ __forceinline__ double col_dot(const double2* __restrict__ x, const double2* __restrict__ y, int64_t dim,
                                          int W, int c) {
  double s = 0.0;
  int64_t i = 0;
  for (; i + 8 <= dim; i += 8) {
    double2 a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = x[(i + u) * W + c];
      b[u] = y[(i + u) * W + c];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) s = s + (a[u].x * b[u].x + a[u].y * b[u].y);
  }
  for (; i < dim; ++i) {
    const double2 a = x[i * W + c], b = y[i * W + c];
    s = s + (a.x * b.x + a.y * b.y);
  }
  return s;
}

__device__ __forceinline__ double col_norm2(const double2* __restrict__ x, int64_t dim, int W, int c) {
  double s = 0.0;
  int64_t i = 0;
  for (; i + 8 <= dim; i += 8) {
    double2 a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = x[(i + u) * W + c];
#pragma unroll
    for (int u = 0; u < 8; ++u) s = s + (a[u].x * a[u].x + a[u].y * a[u].y);
  }
  for (; i < dim; ++i) {
    const double2 a = x[i * W + c];
    s = s + (a.x * a.x + a.y * a.y);
  }
  return s;
}

// Reductions: with a.tree == 0 one thread per column sums its rows in row
// order (bit-identical to the reference). With a.tree == 1 the alpha dot is
// accumulated inside the apply loop and the norm inside the update loop,
// per thread (thread p of a column owns rows p, p + NP, ...), and the NP
// partials of a column are added in index order: two vector passes fewer
// per step and no serial phase; a fixed order, so runs repeat exactly.
template <int W, bool TREE>
__global__ void __launch_bounds__(kPackThreads, 4) k_pack_lanczos(PackArgs a) {
  constexpr int RPW = 64 / W;             // rows per wave pass
  constexpr int NP = kPackWaves * RPW;    // threads per column
  __shared__ int s_item;
  __shared__ double s_alpha[W], s_b[W], s_bprev[W];
  __shared__ double s_red[NP][W];
  __shared__ int s_done[W];
  __shared__ int s_live;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int c = lane & (W - 1);
  const int rs = lane / W;
  constexpr bool tree = TREE;  // a.tree, fixed per instantiation
  double2* x0 = a.slots + (int64_t)blockIdx.x * 3 * a.slot_stride;
  double2* x1 = x0 + a.slot_stride;
  double2* y = x1 + a.slot_stride;
  const int n_items = a.n_items[0];
  for (;;) {
    if (tid == 0) s_item = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);  // uniform: item and entry fields load into scalar registers
    __syncthreads();
    if (it >= n_items) break;
    const PackItem item = a.items[it];
    const PackEntry e = a.entries[item.entry];
    const int64_t dim = e.dim;
    const int w = item.w;
    const int col = e.col0 + item.c0;  // first global column of the item
    // start columns (sample-major [R][dim]) -> x0; x1 = 0 (q_prev)
    for (int64_t t = tid; t < dim * W; t += kPackThreads) {
      const int64_t i = t / W;
      const int cc = (int)(t - i * W);
      if (cc < w) {
        x0[t] = e.start[(int64_t)(item.c0 + cc) * dim + i];
        x1[t] = make_double2(0.0, 0.0);
      }
    }
    if (tid < W) {
      s_bprev[tid] = 0.0;
      s_done[tid] = tid < w ? 0 : 1;
    }
    __syncthreads();
    const int kdim = a.kmax > 0 && a.kmax < e.kdim ? a.kmax : e.kdim;
    const int p = wave * RPW + rs;  // this thread's rows: p, p + NP, ... (column c)
    for (int k = 0; k < kdim; ++k) {
      // y = H x0 (main.cpp:833-845: diagonal first, then hops in order)
      const bool act = !s_done[c];
      double dpart = 0.0;
      for (int64_t i = p; i < dim; i += NP) {
        if (!act) continue;
        const double dg = e.diag[i];
        const double2 xs = x0[i * W + c];
        cd acc = {0.0 + dg * xs.x, 0.0 + dg * xs.y};
        const int64_t e0 = e.row_ptr[i], e1 = e.row_ptr[i + 1];
        int64_t h = e0;
        // hops go in groups of kG (the last one partial): the next group's
        // hop entries load while this group's vector rows are in flight; the
        // sums keep the hop order
        int32_t ncl[kG];
        uint32_t ncode[kG];
#pragma unroll
        for (int q = 0; q < kG; ++q) {
          ncl[q] = h + q < e1 ? e.col[h + q] : 0;
          ncode[q] = h + q < e1 ? e.code[h + q] : 0u;
        }
        for (; h < e1; h += kG) {
          const int n = e1 - h < kG ? (int)(e1 - h) : kG;
          double2 xv[kG];
          uint32_t code[kG];
#pragma unroll
          for (int q = 0; q < kG; ++q) {
            xv[q] = q < n ? x0[(int64_t)ncl[q] * W + c] : make_double2(0.0, 0.0);
            code[q] = ncode[q];
          }
#pragma unroll
          for (int q = 0; q < kG; ++q) {
            const int64_t hn = h + kG + q;
            ncl[q] = hn < e1 ? e.col[hn] : 0;
            ncode[q] = hn < e1 ? e.code[hn] : 0u;
          }
          cd cf[kG];
#pragma unroll
          for (int q = 0; q < kG; ++q) cf[q] = q < n ? pack_coef(code[q], e.ptab, e.spec) : cd{0.0, 0.0};
#pragma unroll
          for (int q = 0; q < kG; ++q)
            if (q < n) acc = cadd(acc, cmul(cf[q], {xv[q].x, xv[q].y}));
        }
        y[i * W + c] = make_double2(acc.re, acc.im);
        if (tree) dpart = dpart + (xs.x * acc.re + xs.y * acc.im);
      }
      if (tree) s_red[p][c] = dpart;
      __syncthreads();
      // alpha = Re <x0, y> (main.cpp:1037): serial over rows, or the partials
      if (tid < W && !s_done[tid]) {
        if (tree) {
          double s = 0.0;
          for (int q = 0; q < NP; ++q) s = s + s_red[q][tid];
          s_alpha[tid] = s;
        } else {
          s_alpha[tid] = col_dot(x0, y, dim, W, tid);
        }
      }
      __syncthreads();
      // x1 = (y - alpha x0) - bprev x1 (main.cpp:1039-1041); thread (p, c)
      // again owns rows p, p + NP, ...
      double npart = 0.0;
      for (int64_t t = tid; t < dim * W; t += kPackThreads) {
        const int cc = (int)(t & (W - 1));
        if (s_done[cc]) continue;
        const double al = s_alpha[cc], b = s_bprev[cc];
        const double2 hw = y[t], hq = x0[t], hp = x1[t];
        const double2 v = make_double2((hw.x - al * hq.x) - b * hp.x, (hw.y - al * hq.y) - b * hp.y);
        x1[t] = v;
        if (tree) npart = npart + (v.x * v.x + v.y * v.y);
      }
      if (tree) s_red[p][c] = npart;
      __syncthreads();
      // beta = sqrt(|x1|^2) and the stopping rule (main.cpp:1043-1056)
      if (tid < W && !s_done[tid]) {
        const int64_t gc = col + tid;
        double nrm;
        if (tree) {
          nrm = 0.0;
          for (int q = 0; q < NP; ++q) nrm = nrm + s_red[q][tid];
        } else {
          nrm = col_norm2(x1, dim, W, tid);
        }
        a.alpha[(int64_t)k * a.rcap + gc] = s_alpha[tid];
        a.mcount[gc] = k + 1;
        const double b = sqrt(nrm);
        s_b[tid] = b;
        if (b < 1e-12 || k + 1 == kdim) {
          s_done[tid] = 2;  // finished this step: not normalised
        } else {
          a.beta[(int64_t)k * a.rcap + gc] = b;
          s_bprev[tid] = b;
        }
      }
      if (tid == 0) s_live = 0;
      __syncthreads();
      // x1 /= beta (main.cpp:1053-1055), then (x0, x1) <- (x1, x0)
      for (int64_t t = tid; t < dim * W; t += kPackThreads) {
        const int cc = (int)(t & (W - 1));
        if (s_done[cc]) continue;
        const double b = s_b[cc];
        const double2 v = x1[t];
        x1[t] = make_double2(v.x / b, v.y / b);
      }
      if (tid < W) {
        if (s_done[tid] == 2) s_done[tid] = 1;
        if (!s_done[tid]) atomicAdd(&s_live, 1);
      }
      __syncthreads();
      double2* tmp = x0;
      x0 = x1;
      x1 = tmp;
      if (s_live == 0) break;
      __syncthreads();
    }
    // restore the slot roles for the next item
    x0 = a.slots + (int64_t)blockIdx.x * 3 * a.slot_stride;
    x1 = x0 + a.slot_stride;
    __syncthreads();
  }
}

}  // namespace

void launch_pack_reset(const PackArgs& a, int rcap, hipStream_t s) {
  k_pack_reset<<<(unsigned)std::max<int64_t>(1, std::min<int64_t>((rcap + 255) / 256, 1024)), 256, 0, s>>>(a, rcap);
}

template <int W, bool TREE>
void launch_pack_w(const PackArgs& a, int grid, hipStream_t s) {
  k_pack_lanczos<W, TREE><<<grid, kPackThreads, 0, s>>>(a);
}

void launch_pack_lanczos(const PackArgs& a, int W, int grid, hipStream_t s) {
  switch (W) {
    case 4: a.tree ? launch_pack_w<4, true>(a, grid, s) : launch_pack_w<4, false>(a, grid, s); break;
    case 8: a.tree ? launch_pack_w<8, true>(a, grid, s) : launch_pack_w<8, false>(a, grid, s); break;
    case 16: a.tree ? launch_pack_w<16, true>(a, grid, s) : launch_pack_w<16, false>(a, grid, s); break;
    case 32: a.tree ? launch_pack_w<32, true>(a, grid, s) : launch_pack_w<32, false>(a, grid, s); break;
    default: a.tree ? launch_pack_w<64, true>(a, grid, s) : launch_pack_w<64, false>(a, grid, s); break;
  }
}

}  // namespace ftlm
