// ftlm_lanczos.hip -- batched Lanczos recurrence and tridiagonal solve.
//
// All R random vectors of a block advance together (main.cpp:1011-1057 runs
// them one by one). Two reduction orders are provided:
//   serial -- dot products and norms are summed row by row in the
//             reference's order (main.cpp:807-821), one thread per column;
//             with the bit-exact H apply this reproduces the reference's
//             alpha/beta and Ritz values bit for bit. Cost grows with dim
//             (one latency-bound pass per reduction), so it suits blocks up
//             to ~1e5 rows.
//   tree   -- the alpha dot is fused into the H apply and the norm into the
//             recurrence, each summed per workgroup and then over
//             workgroups in index order (deterministic, run to run). The
//             vectors are kept unnormalised with per-column scale factors,
//             so a step is two kernels: apply and recurrence.
#include "ftlm_device.h"
#include "ftlm_launch.h"

namespace ftlm {

namespace {

inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// rows each thread of the one-column-per-thread recurrence keeps in flight
constexpr int kRecurRows = 4;

__global__ void k_lanczos_init(int R, LanczosState st) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  st.scale[r] = 1.0;
  st.sprev[r] = 0.0;
  st.bprev[r] = 0.0;
  st.done[r] = 0;
  st.mcount[r] = 0;
}

// out[r] = Re sum_i conj(x[i][r]) y[i][r], summed serially over i
__global__ void k_dot_serial(int64_t dim, int R, const double2* __restrict__ x, const double2* __restrict__ y,
                             double* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double s = 0.0;
  int64_t i = 0;
  for (; i + 8 <= dim; i += 8) {
    double2 a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = x[(i + u) * R + r];
      b[u] = y[(i + u) * R + r];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) s = s + (a[u].x * b[u].x + a[u].y * b[u].y);
  }
  for (; i < dim; ++i) {
    const double2 a = x[i * R + r], b = y[i * R + r];
    s = s + (a.x * b.x + a.y * b.y);
  }
  out[r] = s;
}

// out[r] = sum_i |x[i][r]|^2, summed serially over i (std::norm = re*re + im*im)
__global__ void k_norm_serial(int64_t dim, int R, const double2* __restrict__ x, double* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double s = 0.0;
  int64_t i = 0;
  for (; i + 8 <= dim; i += 8) {
    double2 a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = x[(i + u) * R + r];
#pragma unroll
    for (int u = 0; u < 8; ++u) s = s + (a[u].x * a[u].x + a[u].y * a[u].y);
  }
  for (; i < dim; ++i) {
    const double2 a = x[i * R + r];
    s = s + (a.x * a.x + a.y * a.y);
  }
  out[r] = s;
}

// qp = (w - alpha q) - bprev qp  (main.cpp:1039-1041)
__global__ void k_recur_serial(int64_t n, int R, const double2* __restrict__ w, const double2* __restrict__ q,
                               double2* __restrict__ qp, const double* __restrict__ alpha,
                               const double* __restrict__ bprev, const int32_t* __restrict__ done) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % R);
    if (done[r]) continue;
    const double a = alpha[r], b = bprev[r];
    const double2 hw = w[e], hq = q[e], hp = qp[e];
    const double re = (hw.x - a * hq.x) - b * hp.x;
    const double im = (hw.y - a * hq.y) - b * hp.y;
    qp[e] = make_double2(re, im);
  }
}

// alpha/beta bookkeeping and the stopping rule of main.cpp:1043-1056
__global__ void k_finish_serial(int R, int k, int kdim, LanczosState st) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R || st.done[r]) return;
  st.alpha[(int64_t)k * R + r] = st.dot[r];
  st.mcount[r] = k + 1;
  const double b = sqrt(st.nrm[r]);
  if (b < 1e-12 || k + 1 == kdim) {
    st.done[r] = 1;
    return;
  }
  st.beta[(int64_t)k * R + r] = b;
  st.bprev[r] = b;
}

// q /= beta_k for columns still running (main.cpp:1053-1055)
__global__ void k_normalize(int64_t n, int R, double2* __restrict__ q, const double* __restrict__ beta_k,
                            const int32_t* __restrict__ done) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % R);
    if (done[r]) continue;
    const double b = beta_k[r];
    const double2 v = q[e];
    q[e] = make_double2(v.x / b, v.y / b);
  }
}

// Tree-order recurrence on unnormalised vectors. With q = s x, Hq = s y:
//   x_next = s y - (alpha s) x - (bprev sprev) xp,  alpha = s^2 Re<x, y>
// Column ownership is fixed per thread: for R <= 256 (CPT = 1) thread t owns
// column t % R of rows t / R (+ multiples of 256/R); for larger R (CPT = 4,
// R <= 1024) it owns columns t, t+256, ... of every row. Each thread keeps U
// rows in flight and adds its rows' |x_next|^2 in increasing row order.
// Per-column norm partials are summed per workgroup in thread order and over
// workgroups by the last workgroup, which also applies the stopping rule
// (main.cpp:1043-1056) and updates the scale factors. The instantiation per
// column count keeps the R <= 256 kernel at a few dozen registers (8 waves
// per SIMD) instead of the four-column kernel's ~190.
template <int CPT, int U>
__global__ void __launch_bounds__(256) k_recur_tree(int64_t dim, int R, const double2* __restrict__ y,
                                                    const double2* __restrict__ x, double2* __restrict__ xp,
                                                    int k, int kdim, LanczosState st, double* partials,
                                                    unsigned* counter) {
  __shared__ double sh[CPT][256];
  const int tid = threadIdx.x;
  const int rpi = CPT == 1 ? 256 / R : 1;   // rows per iteration
  const int rsub = CPT == 1 ? tid / R : 0;
  const int r0 = CPT == 1 ? tid % R : tid;
  double acc[CPT], s_[CPT], c2_[CPT], c3_[CPT];
  bool run[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int r = r0 + 256 * j;
    acc[j] = 0.0;
    s_[j] = c2_[j] = c3_[j] = 0.0;
    run[j] = rsub < rpi && r < R && !st.done[r];
    if (run[j]) {
      const double sc = st.scale[r];
      s_[j] = sc;
      c2_[j] = (sc * sc * st.dot[r]) * sc;
      c3_[j] = st.bprev[r] * st.sprev[r];
    }
  }
  const int64_t stride = (int64_t)gridDim.x * rpi;
  for (int64_t row = (int64_t)blockIdx.x * rpi + rsub; row < dim; row += U * stride) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      if (!run[j]) continue;
      double2 vy[U], vx[U], vp[U];
      int64_t e[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        e[u] = (row + u * stride) * R + r0 + 256 * j;
        if (row + u * stride < dim) {
          vy[u] = y[e[u]];
          vx[u] = x[e[u]];
          vp[u] = xp[e[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (row + u * stride >= dim) continue;
        const double re = (s_[j] * vy[u].x - c2_[j] * vx[u].x) - c3_[j] * vp[u].x;
        const double im = (s_[j] * vy[u].y - c2_[j] * vx[u].y) - c3_[j] * vp[u].y;
        xp[e[u]] = make_double2(re, im);
        acc[j] += re * re + im * im;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < CPT; ++j) sh[j][tid] = acc[j];
  __syncthreads();
  for (int r = tid; r < R; r += 256) {
    double sum = 0.0;
    if (CPT == 1) {
      for (int q = r; q < rpi * R; q += R) sum += sh[0][q];
    } else {
      sum = sh[r / 256][r % 256];
    }
    partials[(int64_t)blockIdx.x * R + r] = sum;
  }
  __threadfence();
  __syncthreads();
  __shared__ unsigned last;
  if (tid == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  last_block_column_sums(partials, gridDim.x, R, [&](int r, double s) {
    if (st.done[r]) return;
    const double sc = st.scale[r];
    st.alpha[(int64_t)k * R + r] = sc * sc * st.dot[r];
    st.mcount[r] = k + 1;
    const double b = sqrt(s);
    if (b < 1e-12 || k + 1 == kdim) {
      st.done[r] = 1;
      st.scale[r] = 0.0;
      return;
    }
    st.beta[(int64_t)k * R + r] = b;
    st.sprev[r] = sc;
    st.scale[r] = 1.0 / b;
    st.bprev[r] = b;
  });
  if (tid == 0) *counter = 0u;
}

// Per-column Re<x, y> with the column ownership of k_recur_tree.
template <int CPT>
__global__ void __launch_bounds__(256) k_dot_tree(int64_t dim, int R, const double2* __restrict__ x,
                                                  const double2* __restrict__ y, double* partials, unsigned* counter,
                                                  double* out) {
  __shared__ double sh[CPT][256];
  const int tid = threadIdx.x;
  const int rpi = CPT == 1 ? 256 / R : 1;
  const int rsub = CPT == 1 ? tid / R : 0;
  const int r0 = CPT == 1 ? tid % R : tid;
  double acc[CPT];
  bool run[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    acc[j] = 0.0;
    run[j] = rsub < rpi && r0 + 256 * j < R;
  }
  const int64_t stride = (int64_t)gridDim.x * rpi;
  for (int64_t row = (int64_t)blockIdx.x * rpi + rsub; row < dim; row += 4 * stride) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      if (!run[j]) continue;
      double2 a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (row + u * stride < dim) {
          const int64_t e = (row + u * stride) * R + r0 + 256 * j;
          a[u] = x[e];
          b[u] = y[e];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (row + u * stride < dim) acc[j] += a[u].x * b[u].x + a[u].y * b[u].y;
    }
  }
#pragma unroll
  for (int j = 0; j < CPT; ++j) sh[j][tid] = acc[j];
  __syncthreads();
  for (int r = tid; r < R; r += 256) {
    double sum = 0.0;
    if (CPT == 1) {
      for (int q = r; q < rpi * R; q += R) sum += sh[0][q];
    } else {
      sum = sh[r / 256][r % 256];
    }
    partials[(int64_t)blockIdx.x * R + r] = sum;
  }
  __threadfence();
  __syncthreads();
  __shared__ unsigned last;
  if (tid == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  last_block_column_sums(partials, gridDim.x, R, [&](int r, double s) { out[r] = s; });
  if (tid == 0) *counter = 0u;
}

// src: cnt vectors, each dim contiguous complex; x: [dim][R] at columns r0..
__global__ void k_transpose_in(int64_t dim, int R, int r0, int cnt, const double2* __restrict__ src,
                               double2* __restrict__ x) {
  __shared__ double2 tile[32][33];
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 256 threads: 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j;
    const int64_t i = i0 + tx;
    if (c < cnt && i < dim) tile[j][tx] = src[(int64_t)c * dim + i];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int64_t i = i0 + j;
    const int c = c0 + tx;
    if (c < cnt && i < dim) x[i * R + r0 + c] = tile[tx][j];
  }
}

// ---- classical Jacobi on the Lanczos tridiagonal (main.cpp:913-994) -------
// One 64-lane workgroup per column. The symmetric matrix lives as its packed
// upper triangle (row-major, diagonal included) in LDS or, for large m, in a
// global scratch area. The pivot is the largest |a_ij| (i < j) with the
// reference's tie rule (first in row-major order); rotations update the
// packed entries with the reference's expression order, and only the first
// row of the eigenvector matrix is accumulated (the weights need v[0][j]).
__device__ __forceinline__ int64_t pk(int i, int j, int n) {
  // i <= j
  return (int64_t)i * n - ((int64_t)i * (i - 1)) / 2 + (j - i);
}

// ---- ordering of the Ritz values --------------------------------------
// The reference orders each sample's Ritz values with std::sort over an
// index array (main.cpp:1076-1081). std::sort is not stable, so with
// exactly tied eigenvalues (converged ghost copies) the weights' order
// depends on the algorithm. To keep records byte-identical, the same
// introsort is run here on one lane: median-of-three quicksort partitions
// down to runs of 16, a heapsort fallback after 2*floor(log2 n) levels, then
// a final insertion sort.
struct ByValue {
  const double* v;
  __device__ bool operator()(int a, int b) const { return v[a] < v[b]; }
};

__device__ void lin_insert(int* a, int last, const ByValue& lt) {
  const int val = a[last];
  int next = last - 1;
  while (lt(val, a[next])) {
    a[last] = a[next];
    last = next;
    --next;
  }
  a[last] = val;
}

__device__ void insertion_sort(int* a, int lo, int hi, const ByValue& lt) {
  if (lo == hi) return;
  for (int i = lo + 1; i < hi; ++i) {
    if (lt(a[i], a[lo])) {
      const int val = a[i];
      for (int j = i; j > lo; --j) a[j] = a[j - 1];
      a[lo] = val;
    } else {
      lin_insert(a, i, lt);
    }
  }
}

__device__ void sift_down(int* a, int first, int hole, int len, int value, const ByValue& lt) {
  const int top = hole;
  int child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (lt(a[first + child], a[first + child - 1])) --child;
    a[first + hole] = a[first + child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    a[first + hole] = a[first + child - 1];
    hole = child - 1;
  }
  int parent = (hole - 1) / 2;
  while (hole > top && lt(a[first + parent], value)) {
    a[first + hole] = a[first + parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  a[first + hole] = value;
}

__device__ void heap_sort_range(int* a, int first, int last, const ByValue& lt) {
  const int len = last - first;
  if (len >= 2) {
    for (int parent = (len - 2) / 2;; --parent) {
      sift_down(a, first, parent, len, a[first + parent], lt);
      if (parent == 0) break;
    }
  }
  while (last - first > 1) {
    --last;
    const int value = a[last];
    a[last] = a[first];
    sift_down(a, first, 0, last - first, value, lt);
  }
}

__device__ void swap_at(int* a, int i, int j) {
  const int t = a[i];
  a[i] = a[j];
  a[j] = t;
}

__device__ int partition_pivot(int* a, int first, int last, const ByValue& lt) {
  const int mid = first + (last - first) / 2;
  const int x = first + 1, y = mid, z = last - 1;
  if (lt(a[x], a[y])) {
    if (lt(a[y], a[z])) swap_at(a, first, y);
    else if (lt(a[x], a[z])) swap_at(a, first, z);
    else swap_at(a, first, x);
  } else if (lt(a[x], a[z])) {
    swap_at(a, first, x);
  } else if (lt(a[y], a[z])) {
    swap_at(a, first, z);
  } else {
    swap_at(a, first, y);
  }
  int lo = first + 1, hi = last;
  while (true) {
    while (lt(a[lo], a[first])) ++lo;
    --hi;
    while (lt(a[first], a[hi])) --hi;
    if (!(lo < hi)) return lo;
    swap_at(a, lo, hi);
    ++lo;
  }
}

__device__ void intro_sort(int* a, int n, const ByValue& lt) {
  if (n <= 1) return;
  int lg = 0;
  while ((1 << (lg + 1)) <= n) ++lg;
  // explicit stack of (first, last, depth); sub-ranges are disjoint, so the
  // order in which they are finished does not change the result
  int sf[64], sl[64], sd[64];
  int top = 0;
  sf[0] = 0;
  sl[0] = n;
  sd[0] = 2 * lg;
  ++top;
  while (top > 0) {
    --top;
    int first = sf[top], last = sl[top], depth = sd[top];
    while (last - first > 16) {
      if (depth == 0) {
        heap_sort_range(a, first, last, lt);
        break;
      }
      --depth;
      const int cut = partition_pivot(a, first, last, lt);
      sf[top] = cut;
      sl[top] = last;
      sd[top] = depth;
      ++top;
      last = cut;
    }
  }
  if (n > 16) {
    insertion_sort(a, 0, 16, lt);
    for (int i = 16; i < n; ++i) lin_insert(a, i, lt);
  } else {
    insertion_sort(a, 0, n, lt);
  }
}

__global__ void __launch_bounds__(64) k_ritz(int R, int steps, const double* __restrict__ alpha,
                                             const double* __restrict__ beta, const int32_t* __restrict__ mcount,
                                             int wdiv, double* __restrict__ eig_out, double* __restrict__ w_out,
                                             double* __restrict__ scratch, int use_lds) {
  extern __shared__ double lds[];
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = mcount[r];
  const int64_t np = (int64_t)steps * (steps + 1) / 2;
  double* A = use_lds ? lds : scratch + (int64_t)r * (np + 3 * steps);
  double* v0 = A + np;
  double* ev = v0 + steps;
  const int64_t nn = (int64_t)n * (n + 1) / 2;
  for (int64_t e = lane; e < nn; e += 64) A[e] = 0.0;
  __syncthreads();
  for (int i = lane; i < n; i += 64) {
    A[pk(i, i, n)] = alpha[(int64_t)i * R + r];
    if (i + 1 < n) A[pk(i, i + 1, n)] = beta[(int64_t)i * R + r];
    v0[i] = i == 0 ? 1.0 : 0.0;
  }
  __syncthreads();
  if (n >= 2) {
    const int64_t maxit = 50ll * n * n;
    for (int64_t it = 0; it < maxit; ++it) {
      // pivot search: lane scans columns j = i+1+lane, +64, ... of each row
      double best = 0.0;
      int bp = 0, bq = 1;
      bool found = false;
      for (int i = 0; i < n - 1; ++i) {
        const int64_t base = pk(i, i, n);
        for (int j = i + 1 + lane; j < n; j += 64) {
          const double v = fabs(A[base + (j - i)]);
          if (v > best) {
            best = v;
            bp = i;
            bq = j;
            found = true;
          }
        }
      }
      // wave argmax; ties -> smallest (row, column)
      for (int off = 32; off >= 1; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int op = __shfl_xor(bp, off), oq = __shfl_xor(bq, off);
        const int of = __shfl_xor((int)found, off);
        bool take = false;
        if (of) {
          if (!found) take = true;
          else if (ob > best) take = true;
          else if (ob == best && (op < bp || (op == bp && oq < bq))) take = true;
        }
        if (take) {
          best = ob;
          bp = op;
          bq = oq;
          found = true;
        }
      }
      if (!found) {
        bp = 0;
        bq = 1;
        best = 0.0;
      }
      if (best < 1e-12) break;
      const int p = bp, q = bq;
      const double app = A[pk(p, p, n)], aqq = A[pk(q, q, n)], apq = A[pk(p, q, n)];
      const double tau = (aqq - app) / (2.0 * apq);
      const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
      const double c = 1.0 / sqrt(1.0 + t * t);
      const double s = t * c;
      __syncthreads();
      for (int kk = lane; kk < n; kk += 64) {
        if (kk == p || kk == q) continue;
        const int64_t ip = kk < p ? pk(kk, p, n) : pk(p, kk, n);
        const int64_t iq = kk < q ? pk(kk, q, n) : pk(q, kk, n);
        const double akp = A[ip], akq = A[iq];
        A[ip] = c * akp - s * akq;
        A[iq] = s * akp + c * akq;
      }
      if (lane == 0) {
        A[pk(p, p, n)] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
        A[pk(q, q, n)] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
        A[pk(p, q, n)] = 0.0;
        const double vp = v0[p], vq = v0[q];
        v0[p] = c * vp - s * vq;
        v0[q] = s * vp + c * vq;
      }
      __syncthreads();
    }
  }
  int* order = reinterpret_cast<int*>(ev + steps);
  for (int i = lane; i < n; i += 64) {
    ev[i] = A[pk(i, i, n)];
    order[i] = i;
  }
  __syncthreads();
  if (lane == 0) intro_sort(order, n, ByValue{ev});
  __syncthreads();
  for (int k = lane; k < n; k += 64) {
    const int o = order[k];
    eig_out[(int64_t)r * steps + k] = ev[o];
    w_out[(int64_t)r * steps + k] = (v0[o] * v0[o]) / (double)wdiv;
  }
}

// ---- implicit QL on the Lanczos tridiagonal (tree-order runs) ------------
// One thread per column: symmetric tridiagonal QL with Wilkinson shifts
// (the EISPACK tql2 scheme the reference's Fortran ancestor uses), carrying
// only the first row of the eigenvector matrix, which is all the weights
// need. O(m^2) per column instead of the Jacobi sweep's O(m^4); eigenvalues
// agree with Jacobi's to rounding. Ordering uses the same introsort.
__device__ __forceinline__ double hyp(double a, double b) {
  a = fabs(a);
  b = fabs(b);
  const double hi = a > b ? a : b, lo = a > b ? b : a;
  if (hi == 0.0) return 0.0;
  const double r = lo / hi;
  return hi * sqrt(1.0 + r * r);
}

__global__ void k_ritz_ql(int R, int steps, const double* __restrict__ alpha, const double* __restrict__ beta,
                          const int32_t* __restrict__ mcount, int wdiv, double* __restrict__ eig_out,
                          double* __restrict__ w_out, double* __restrict__ scratch, int32_t* __restrict__ iscratch,
                          int32_t* __restrict__ fail) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const int n = mcount[r];
  double* d = scratch + (int64_t)r * 3 * steps;
  double* e = d + steps;
  double* z = e + steps;
  int* order = iscratch + (int64_t)r * steps;
  for (int i = 0; i < n; ++i) {
    d[i] = alpha[(int64_t)i * R + r];
    e[i] = i + 1 < n ? beta[(int64_t)i * R + r] : 0.0;
    z[i] = i == 0 ? 1.0 : 0.0;
  }
  // splitting: relative to the neighbouring diagonals, or below eps ||T||
  // (lets clusters of Ritz values near zero converge)
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i) tnorm = fmax(tnorm, fabs(d[i]) + fabs(e[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0));
  const double floor_e = 2.220446049250313e-16 * tnorm;
  bool ok = true;
  for (int l = 0; l < n && ok; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 2.220446049250313e-16 * dd || fabs(e[m]) <= floor_e) break;
      }
      if (m != l) {
        if (++iter > 60) {
          ok = false;
          break;
        }
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double rr = hyp(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? rr : -rr));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          const double f = s * e[i];
          const double b = c * e[i];
          rr = hyp(f, g);
          e[i + 1] = rr;
          if (rr == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / rr;
          c = g / rr;
          g = d[i + 1] - p;
          rr = (d[i] - g) * s + 2.0 * c * b;
          p = s * rr;
          d[i + 1] = g + p;
          g = c * rr - b;
          const double zf = z[i + 1];
          z[i + 1] = s * z[i] + c * zf;
          z[i] = c * z[i] - s * zf;
        }
        if (rr == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  if (!ok) atomicAdd(fail, 1);
  for (int i = 0; i < n; ++i) order[i] = i;
  intro_sort(order, n, ByValue{d});
  for (int k = 0; k < n; ++k) {
    const int o = order[k];
    eig_out[(int64_t)r * steps + k] = d[o];
    w_out[(int64_t)r * steps + k] = (z[o] * z[o]) / (double)wdiv;
  }
}

}  // namespace

void launch_lanczos_init(int R, LanczosState st, hipStream_t s) {
  k_lanczos_init<<<nblocks(R, 256), 256, 0, s>>>(R, st);
}

void launch_dot_serial(int64_t dim, int R, const double* x, const double* y, double* out, hipStream_t s) {
  k_dot_serial<<<nblocks(R, 64), 64, 0, s>>>(dim, R, reinterpret_cast<const double2*>(x),
                                             reinterpret_cast<const double2*>(y), out);
}

void launch_norm_serial(int64_t dim, int R, const double* x, double* out, hipStream_t s) {
  k_norm_serial<<<nblocks(R, 64), 64, 0, s>>>(dim, R, reinterpret_cast<const double2*>(x), out);
}

static unsigned stream_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void launch_recur_serial(int64_t dim, int R, const double* w, const double* q, double* qp, const double* alpha_k,
                         const double* bprev, const int32_t* done, hipStream_t s) {
  const int64_t n = dim * R;
  k_recur_serial<<<stream_grid(n), 256, 0, s>>>(n, R, reinterpret_cast<const double2*>(w),
                                                reinterpret_cast<const double2*>(q), reinterpret_cast<double2*>(qp),
                                                alpha_k, bprev, done);
}

void launch_finish_serial(int R, int k, int kdim, LanczosState st, hipStream_t s) {
  k_finish_serial<<<nblocks(R, 256), 256, 0, s>>>(R, k, kdim, st);
}

void launch_normalize(int64_t dim, int R, double* q, const double* beta_k, const int32_t* done, int k,
                      const int32_t* mcount, hipStream_t s) {
  (void)k;
  (void)mcount;
  const int64_t n = dim * R;
  k_normalize<<<stream_grid(n), 256, 0, s>>>(n, R, reinterpret_cast<double2*>(q), beta_k, done);
}

void launch_recur_tree(int64_t dim, int R, const double* y, const double* x, double* xp, int k, int kdim,
                       LanczosState st, double* partials, unsigned* counter, int grid, hipStream_t s) {
  const double2* yy = reinterpret_cast<const double2*>(y);
  const double2* xx = reinterpret_cast<const double2*>(x);
  double2* pp = reinterpret_cast<double2*>(xp);
  if (R <= 256)
    k_recur_tree<1, kRecurRows><<<grid, 256, 0, s>>>(dim, R, yy, xx, pp, k, kdim, st, partials, counter);
  else
    k_recur_tree<4, 4><<<grid, 256, 0, s>>>(dim, R, yy, xx, pp, k, kdim, st, partials, counter);
}

void launch_dot_tree(int64_t dim, int R, const double* x, const double* y, double* partials, unsigned* counter,
                     double* out, int grid, hipStream_t s) {
  if (R <= 256)
    k_dot_tree<1><<<grid, 256, 0, s>>>(dim, R, reinterpret_cast<const double2*>(x),
                                       reinterpret_cast<const double2*>(y), partials, counter, out);
  else
    k_dot_tree<4><<<grid, 256, 0, s>>>(dim, R, reinterpret_cast<const double2*>(x),
                                       reinterpret_cast<const double2*>(y), partials, counter, out);
}

void launch_transpose_in(int64_t dim, int R, int r0, int cnt, const double* src, double* x, hipStream_t s) {
  dim3 grid(nblocks(dim, 32), nblocks(cnt, 32));
  k_transpose_in<<<grid, 256, 0, s>>>(dim, R, r0, cnt, reinterpret_cast<const double2*>(src),
                                      reinterpret_cast<double2*>(x));
}

void launch_ritz_ql(int R, int steps, const double* alpha, const double* beta, const int32_t* mcount, int wdiv,
                    double* eig, double* w, double* scratch, int32_t* iscratch, int32_t* fail, hipStream_t s) {
  k_ritz_ql<<<nblocks(R, 64), 64, 0, s>>>(R, steps, alpha, beta, mcount, wdiv, eig, w, scratch, iscratch, fail);
}

void launch_ritz(int R, int steps, const double* alpha, const double* beta, const int32_t* mcount, int wdiv,
                 double* eig, double* w, double* scratch, hipStream_t s) {
  const size_t words = (size_t)steps * (steps + 1) / 2 + 3 * (size_t)steps;
  const size_t bytes = words * 8;
  const int use_lds = bytes <= 60 * 1024;
  k_ritz<<<R, 64, use_lds ? bytes : 0, s>>>(R, steps, alpha, beta, mcount, wdiv, eig, w, scratch, use_lds);
}

}  // namespace ftlm
