// ftlm_rng.hip -- seeded Gaussian start vectors generated on the device.
//
// The reference draws sample r's start vector on the host (main.cpp:1015-1025):
//   std::mt19937_64 rng(mix_u64(block_seed ^ r));
//   std::normal_distribution<double> normal(0, 1);
//   q[i] = Complex(normal(rng), normal(rng));   // g++: real part = 2nd draw
//   normalize(q);
// This kernel produces the same stream on the GPU, one workgroup per sample:
//
// * mt19937_64 (the C++ standard's parameters n = 312, m = 156, r = 31,
//   a = 0xb5026f5aa96619e9, tempering (29, 0x5555555555555555), (17,
//   0x71d67fffeda60000), (37, 0xfff7eee000000000), 43; seeding multiplier
//   6364136223846793005). One state refill updates word k from words k, k+1
//   and k+156 (mod 312) in increasing k, so it splits into three parallel
//   phases: k < 156 (all inputs old), 156 <= k < 311 (word k-156 new, k+1 old)
//   and k = 311 (words 0 and 155 new). The state is double-buffered in LDS
//   (a refill reads one copy and writes the other), so a refill costs three
//   barriers.
// * canonical doubles as libstdc++'s generate_canonical<double, 53> computes
//   them from a 64-bit engine: (double)z / 2^64, and 1 - 2^-53 if that rounds
//   to 1.
// * the Marsaglia polar method in libstdc++'s order: an attempt takes two
//   consecutive draws, x = 2 u1 - 1, y = 2 u2 - 1, r2 = x x + y y, rejected when
//   r2 > 1 or r2 == 0; an accepted attempt gives y * mult first and x * mult
//   (saved) second, mult = sqrt(-2 log(r2) / r2). Element i of the vector is
//   the i-th accepted attempt: real = x * mult, imaginary = y * mult. Attempts
//   never straddle a refill (312 draws = 156 attempts), so each refill's
//   accepted attempts are placed by a prefix count (wavefront ballots and the
//   counts of the preceding wavefronts).
// * normalisation: each element divided by the column's 2-norm.
//
// The arithmetic is the same IEEE double sequence (built with
// -ffp-contract=off); the one function that may round differently from the
// host's is log (the device's versus glibc's), so an element can differ in
// its last bit from the host stream where the two logs differ, and the
// norm is summed in tree order. Runs that must reproduce the reference bit
// for bit use the host stream.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "ftlm_launch.h"

namespace ftlm {
namespace {

constexpr int kN = 312;
constexpr int kM = 156;
constexpr uint64_t kMatrixA = 0xb5026f5aa96619e9ull;
constexpr uint64_t kUpper = 0xffffffff80000000ull;
constexpr uint64_t kLower = 0x000000007fffffffull;
constexpr int kThreads = 320;  // 312 state words, 5 wavefronts

__device__ __host__ inline uint64_t mix64_dev(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

__device__ inline uint64_t twist(uint64_t cur, uint64_t next, uint64_t far) {
  const uint64_t y = (cur & kUpper) | (next & kLower);
  return far ^ (y >> 1) ^ ((y & 1ull) ? kMatrixA : 0ull);
}

__device__ inline uint64_t temper(uint64_t z) {
  z ^= (z >> 29) & 0x5555555555555555ull;
  z ^= (z << 17) & 0x71d67fffeda60000ull;
  z ^= (z << 37) & 0xfff7eee000000000ull;
  z ^= z >> 43;
  return z;
}

__device__ inline double canonical(uint64_t z) {
  const double r = (double)z * 5.421010862427522170037e-20;  // 2^-64, exact scaling
  return r >= 1.0 ? 0.99999999999999988898 : r;              // nextafter(1, 0)
}

// out element (i, r): layout 0 [dim][R] (stride R), layout 1 [R][dim]
// sample r0 + b goes to output column col0 + b
__global__ __launch_bounds__(kThreads) void k_start_vectors(uint64_t bseed, int r0, int col0, int R, int64_t dim,
                                                           int layout, double* __restrict__ out,
                                                           double* __restrict__ norms) {
  constexpr int kWaves = kThreads / 64;
  __shared__ uint64_t mt[2][kN];
  __shared__ int wcnt[kWaves];
  __shared__ double red[kThreads];
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int r = r0 + blockIdx.x;
  // seeding (sequential recurrence, one thread)
  if (t == 0) {
    uint64_t s = mix64_dev(bseed ^ (uint64_t)r);
    mt[0][0] = s;
    for (int i = 1; i < kN; ++i) {
      s = 6364136223846793005ull * (s ^ (s >> 62)) + (uint64_t)i;
      mt[0][i] = s;
    }
  }
  __syncthreads();
  int cur = 0;  // mt[cur] holds the state
  const int c = col0 + blockIdx.x;
  const int64_t stride = layout == 0 ? (int64_t)R : 1;
  double* col = layout == 0 ? out + 2 * (int64_t)c : out + 2 * (int64_t)c * dim;
  int64_t base = 0;
  double ssq = 0.0;
  while (base < dim) {
    // refill from mt[cur] into mt[cur ^ 1]: phase 1 (k < 156), phase 2
    // (156 <= k < 311), phase 3 (k = 311)
    const uint64_t* o = mt[cur];
    uint64_t* nw = mt[cur ^ 1];
    if (t < kN - kM) nw[t] = twist(o[t], o[t + 1], o[t + kM]);
    __syncthreads();
    if (t >= kN - kM && t < kN - 1) nw[t] = twist(o[t], o[t + 1], nw[t - (kN - kM)]);
    __syncthreads();
    if (t == 0) nw[kN - 1] = twist(o[kN - 1], nw[0], nw[kM - 1]);
    __syncthreads();
    cur ^= 1;
    // attempt t (t < 156) takes draws 2t, 2t+1 of this refill
    int acc = 0;
    double x = 0.0, y = 0.0, r2 = 0.0;
    if (t < kN / 2) {
      x = 2.0 * canonical(temper(nw[2 * t])) - 1.0;
      y = 2.0 * canonical(temper(nw[2 * t + 1])) - 1.0;
      r2 = x * x + y * y;
      acc = (r2 > 1.0 || r2 == 0.0) ? 0 : 1;
    }
    // rank of each accepted attempt: accepted attempts of the earlier
    // wavefronts plus those of lower lanes in its own (attempt order = t)
    const unsigned long long bal = __ballot(acc);
    if (lane == 0) wcnt[wv] = __popcll(bal);
    __syncthreads();
    int before = __popcll(bal & ((1ull << lane) - 1ull));
    int total = 0;
    for (int w = 0; w < kWaves; ++w) {
      if (w < wv) before += wcnt[w];
      total += wcnt[w];
    }
    if (acc) {
      const int64_t i = base + before;
      if (i < dim) {
        const double mult = sqrt(-2.0 * log(r2) / r2);
        const double re = x * mult;
        const double im = y * mult;
        col[2 * i * stride] = re;
        col[2 * i * stride + 1] = im;
        ssq += re * re + im * im;
      }
    }
    base += total;
    // the next refill writes the other state copy and, after three
    // barriers, wcnt: no barrier needed here
  }
  // column norm: thread partials, then a fixed-order tree
  red[t] = ssq;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int i = 0; i < kThreads; ++i) s += red[i];
    norms[blockIdx.x] = sqrt(s);
  }
}

__global__ void k_normalize_starts(int col0, int R, int cnt, int64_t dim, int layout, double* __restrict__ out,
                                   const double* __restrict__ norms) {
  const int64_t n = dim * cnt;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e % cnt);  // column within the chunk
    const int64_t i = e / cnt;
    const double nrm = norms[j];
    double* p = layout == 0 ? out + 2 * (i * R + col0 + j) : out + 2 * ((int64_t)(col0 + j) * dim + i);
    p[0] = p[0] / nrm;
    p[1] = p[1] / nrm;
  }
}

}  // namespace

void launch_start_vectors(uint64_t bseed, int r0, int col0, int cnt, int R, int64_t dim, int layout, double* out,
                          double* norms, hipStream_t s) {
  if (cnt <= 0 || dim <= 0) return;
  hipLaunchKernelGGL(k_start_vectors, dim3(cnt), dim3(kThreads), 0, s, bseed, r0, col0, R, dim, layout, out, norms);
  const int64_t n = dim * cnt;
  const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_normalize_starts, dim3(grid), dim3(256), 0, s, col0, R, cnt, dim, layout, out, norms);
}

}  // namespace ftlm
