// ftlm_thermo.hip -- grand-canonical sums of all blocks' spectra for every
// (beta, mu) pair in one FP64 launch.
//
// The reference accumulates block by block and mu by mu on the host
// (accumulate_block_thermo, main.cpp:1172-1219): per mu a running minimum
// emin of E - mu N, rescaled sums when it drops, and per Ritz value (or exact
// eigenvalue) wt = prefactor * w * exp(-beta (E - mu N - emin)), z += wt,
// n1 += N wt, n2 += N^2 wt. Here the final emin is known up front -- it is
// the minimum over blocks of (min_j E_j) - mu N, since E -> E - c is
// monotone in floating point -- so the rescaling disappears and every value
// contributes once per pair with the same expression:
//   sh = E - mu * N - emin;  wt = W * exp(-beta * sh);  W = prefactor * w.
// The results equal the reference's up to the rounding of the summation
// order (no rescale products), i.e. to ~1e-15 relative.
//
// Work split: each wavefront owns kV consecutive strided values per lane
// (value v0 + 64 j + lane), held in registers while it loops over all pairs;
// per pair, lane sums in index order, a butterfly over the wavefront, the four
// wavefronts of a workgroup in order, then the workgroups in index order in
// a second kernel: a fixed order, so results are identical run to run.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "ftlm_launch.h"

namespace ftlm {

// Device scratch of the allocation-free entry point (ftlm_thermo_noalloc):
// workgroup partials and the per-mu minima live in this fixed array, one
// copy per GPU, held until the process ends. Host code uses the address
// hipGetSymbolAddress returns for the device (thermo_scratch_setup), never
// the array's name.
static __device__ __align__(256) double scratch[kThermoScratchDoubles];

namespace {

constexpr int kV = 16;              // values per lane
constexpr int kWave = 64 * kV;      // values per wavefront
constexpr int kWg = 4 * kWave;      // values per workgroup
constexpr int kPB = 64;             // pairs per LDS batch

__global__ __launch_bounds__(256) void k_thermo(const double* __restrict__ E, const double* __restrict__ W,
                                                const uint8_t* __restrict__ N, int64_t nv,
                                                const double* __restrict__ betas, int nb,
                                                const double* __restrict__ mus, int nm,
                                                const double* __restrict__ emin, double* __restrict__ part) {
  __shared__ double red[4][kPB][3];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t v0 = (int64_t)blockIdx.x * kWg + (int64_t)wave * kWave + lane;
  double e[kV], w[kV], n[kV];
  int cnt = 0;
  for (int j = 0; j < kV; ++j) {
    const int64_t i = v0 + (int64_t)j * 64;
    if (i < nv) {
      e[j] = E[i];
      w[j] = W[i];
      n[j] = (double)N[i];
      cnt = j + 1;
    } else {
      e[j] = 0.0;
      w[j] = 0.0;
      n[j] = 0.0;
    }
  }
  const int np = nb * nm;
  for (int p0 = 0; p0 < np; p0 += kPB) {
    const int pn = np - p0 < kPB ? np - p0 : kPB;
    for (int q = 0; q < pn; ++q) {
      const int p = p0 + q;
      const double beta = betas[p / nm];
      const double mu = mus[p % nm];
      const double em = emin[p % nm];
      double z = 0.0, n1 = 0.0, n2 = 0.0;
      for (int j = 0; j < kV; ++j) {
        if (j < cnt) {
          const double sh = e[j] - mu * n[j] - em;
          const double wt = w[j] * exp(-beta * sh);
          z += wt;
          n1 += n[j] * wt;
          n2 += n[j] * n[j] * wt;
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        z += __shfl_xor(z, o, 64);
        n1 += __shfl_xor(n1, o, 64);
        n2 += __shfl_xor(n2, o, 64);
      }
      if (lane == 0) {
        red[wave][q][0] = z;
        red[wave][q][1] = n1;
        red[wave][q][2] = n2;
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < pn * 3; t += 256) {
      const int q = t / 3;
      const int c = t % 3;
      part[((int64_t)blockIdx.x * np + p0 + q) * 3 + c] = ((red[0][q][c] + red[1][q][c]) + red[2][q][c]) + red[3][q][c];
    }
    __syncthreads();
  }
}

// part_min[g * nm + k] = min over the workgroup's values of E - mu_k N (the
// same tiling as k_thermo). E - c is monotone in E, so the minimum equals
// min_blocks (min_j E_j) - mu N_b computed on the host by ftlm_thermo.
__global__ __launch_bounds__(256) void k_emin_part(const double* __restrict__ E, const uint8_t* __restrict__ N,
                                                   int64_t nv, const double* __restrict__ mus, int nm,
                                                   double* __restrict__ part_min) {
  __shared__ double red[256];
  const int64_t v0 = (int64_t)blockIdx.x * kWg + threadIdx.x;
  double e[kV], n[kV];
  for (int j = 0; j < kV; ++j) {
    const int64_t i = v0 + (int64_t)j * 256;
    e[j] = i < nv ? E[i] : __builtin_inf();
    n[j] = i < nv ? (double)N[i] : 0.0;
  }
  for (int k = 0; k < nm; ++k) {
    const double mu = mus[k];
    double m = __builtin_inf();
    for (int j = 0; j < kV; ++j) m = fmin(m, e[j] - mu * n[j]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + o]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part_min[(int64_t)blockIdx.x * nm + k] = red[0];
    __syncthreads();
  }
}

__global__ void k_emin_final(const double* __restrict__ part_min, int nwg, int nm, double* __restrict__ emin) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nm) return;
  double m = __builtin_inf();
  for (int g = 0; g < nwg; ++g) m = fmin(m, part_min[(int64_t)g * nm + k]);
  emin[k] = m;
}

// out[p] = (emin, z, n1, n2) summing the workgroup partials in index order.
__global__ void k_thermo_reduce(const double* __restrict__ part, int nwg, int np, int nm,
                                const double* __restrict__ emin, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * 3) return;
  const int p = t / 3;
  const int c = t % 3;
  double s = 0.0;
  for (int g = 0; g < nwg; ++g) s += part[((int64_t)g * np + p) * 3 + c];
  out[(int64_t)p * 4 + 1 + c] = s;
  if (c == 0) out[(int64_t)p * 4] = emin[p % nm];
}

}  // namespace

int64_t thermo_workgroups(int64_t nv) { return (nv + kWg - 1) / kWg; }

namespace {
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};
}  // namespace

int thermo_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return (int)hipSuccess;
  double* p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return (int)e;
  g_scratch_dev[device] = p;
  return (int)hipSuccess;
}

double* thermo_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_dev[device];
}

int64_t thermo_noalloc_doubles(int64_t nv, int nb, int nm) {
  const int64_t nwg = thermo_workgroups(nv);
  return std::max<int64_t>(nwg, 1) * nb * nm * 3 + std::max<int64_t>(nwg, 1) * nm + nm;
}

void launch_thermo_noalloc(const double* E, const double* W, const uint8_t* N, int64_t nv, const double* betas,
                           int nb, const double* mus, int nm, double* scr, double* out, hipStream_t s) {
  const int64_t nwg = thermo_workgroups(nv);
  double* part = scr;
  double* part_min = part + std::max<int64_t>(nwg, 1) * nb * nm * 3;
  double* emin = part_min + std::max<int64_t>(nwg, 1) * nm;
  if (nwg > 0) {
    hipLaunchKernelGGL(k_emin_part, dim3((unsigned)nwg), dim3(256), 0, s, E, N, nv, mus, nm, part_min);
    hipLaunchKernelGGL(k_emin_final, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, part_min, (int)nwg, nm,
                       emin);
  } else {
    hipLaunchKernelGGL(k_emin_final, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, part_min, 0, nm, emin);
  }
  launch_thermo(E, W, N, nv, betas, nb, mus, nm, emin, part, out, s);
}

void launch_thermo(const double* E, const double* W, const uint8_t* N, int64_t nv, const double* betas, int nb,
                   const double* mus, int nm, const double* emin, double* part, double* out, hipStream_t s) {
  const int64_t nwg = thermo_workgroups(nv);
  const int np = nb * nm;
  if (nwg > 0) hipLaunchKernelGGL(k_thermo, dim3((unsigned)nwg), dim3(256), 0, s, E, W, N, nv, betas, nb, mus, nm,
                                  emin, part);
  hipLaunchKernelGGL(k_thermo_reduce, dim3((unsigned)((np * 3 + 255) / 256)), dim3(256), 0, s, part, (int)nwg, np,
                     nm, emin, out);
}

}  // namespace ftlm
