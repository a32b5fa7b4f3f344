// ftlm_probe.hip -- memory-system probes with the H apply's access shapes.
//
// They state what the HBM path of this GPU delivers for the batched apply's
// traffic without its hop metadata and arithmetic:
//   kind 0  stream read: each wavefront sums whole rows of `width` complex
//           doubles (one 16-byte load per lane per 64 columns);
//   kind 1  stream copy of the same rows;
//   kind 2  random-row gather: each output row adds `k` source rows chosen by
//           a fixed hash of (row, j) and writes the sum -- one wavefront per
//           row with the apply's lane layout, k rows in flight in groups of
//           four, the XCD-contiguous row ranges of the apply;
//   kind 3  the same gather with every source row drawn inside the output
//           row's window of kProbeWindow rows (12,870 = C(16,8): one
//           canonical-up group of the half-filled 4x4 block, 26 MB at 128
//           columns, where the apply's down-spin hops land);
//   kind 4  the same gather with sources alternately in the window and over
//           the whole vector (the apply's mix of down- and up-spin hops).
// The sums are written so that no load is dead code.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ftlm_launch.h"

namespace ftlm {
namespace {

constexpr int kProbeXcd = 8;
constexpr int64_t kProbeWindow = 12870;

__device__ __forceinline__ uint64_t probe_hash(uint64_t z) {
  z ^= z >> 33;
  z *= 0xff51afd7ed558ccdULL;
  z ^= z >> 33;
  z *= 0xc4ceb9fe1a85ec53ULL;
  z ^= z >> 33;
  return z;
}

// first row, end and step of this wavefront (XCD-contiguous ranges)
__device__ __forceinline__ void probe_rows(int64_t rows, int64_t* first, int64_t* end, int64_t* step) {
  const int wave = threadIdx.x >> 6;
  int64_t u0 = 0, u1 = rows, ustep = (int64_t)gridDim.x * 4, uf = (int64_t)blockIdx.x * 4 + wave;
  if (gridDim.x >= kProbeXcd) {
    const int xcd = blockIdx.x % kProbeXcd;
    const int l = blockIdx.x / kProbeXcd;
    const int nwg = (gridDim.x - xcd + kProbeXcd - 1) / kProbeXcd;
    const int64_t chunk = (rows + kProbeXcd - 1) / kProbeXcd;
    u0 = xcd * chunk;
    u1 = u0 + chunk < rows ? u0 + chunk : rows;
    uf = u0 + (int64_t)l * 4 + wave;
    ustep = (int64_t)nwg * 4;
  }
  *first = uf;
  *end = u1;
  *step = ustep;
}

template <int VPL>
__global__ void __launch_bounds__(256) k_probe_stream(int64_t rows, int width, int copy,
                                                      const double2* __restrict__ x, double2* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  int64_t r, end, step;
  probe_rows(rows, &r, &end, &step);
  double acc = 0.0;
  // four rows in flight per wavefront
  for (; r < end; r += 4 * step) {
    double2 t[4][VPL];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int v = 0; v < VPL; ++v) {
        const int c = lane + 64 * v;
        const int64_t rr = r + q * step;
        t[q][v] = (rr < end && c < width) ? x[rr * width + c] : make_double2(0.0, 0.0);
      }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int v = 0; v < VPL; ++v) {
        const int c = lane + 64 * v;
        const int64_t rr = r + q * step;
        if (rr < end && c < width) {
          if (copy) y[rr * width + c] = t[q][v];
          else acc += t[q][v].x + t[q][v].y;
        }
      }
  }
  if (!copy && acc == 1234.5) y[0] = make_double2(acc, acc);  // keeps the loads live
}

template <int VPL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VPL <= 2 ? 8 : 4)))
k_probe_gather(int64_t rows, int width, int k, int mode, const double2* __restrict__ x, double2* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  int64_t r, end, step;
  probe_rows(rows, &r, &end, &step);
  for (; r < end; r += step) {
    // the row's window [w0, w0 + wl), inside [0, rows) since w0 <= r < rows
    const int64_t w0 = r / kProbeWindow * kProbeWindow;
    const int64_t wl = rows - w0 < kProbeWindow ? rows - w0 : kProbeWindow;
    double2 acc[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) acc[v] = make_double2(0.0, 0.0);
    for (int j = 0; j < k; j += 4) {
      int64_t src[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint64_t h = probe_hash((uint64_t)r * 131 + (uint64_t)(j + q));
        const bool local = mode == 1 || (mode == 2 && ((j + q) & 1) == 0);
        src[q] = local ? w0 + (int64_t)(h % (uint64_t)wl) : (int64_t)(h % (uint64_t)rows);
      }
      double2 t[4][VPL];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
          const int c = lane + 64 * v;
          t[q][v] = (j + q < k && c < width) ? x[src[q] * width + c] : make_double2(0.0, 0.0);
        }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
          acc[v].x += t[q][v].x;
          acc[v].y += t[q][v].y;
        }
    }
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
      const int c = lane + 64 * v;
      if (c < width) y[r * width + c] = acc[v];
    }
  }
}

template <int VPL>
void probe_launch(int kind, int64_t rows, int width, int k, int grid, const double2* x, double2* y, hipStream_t s) {
  if (kind >= 2)
    k_probe_gather<VPL><<<grid, 256, 0, s>>>(rows, width, k, kind - 2, x, y);
  else
    k_probe_stream<VPL><<<grid, 256, 0, s>>>(rows, width, kind == 1 ? 1 : 0, x, y);
}

}  // namespace

void launch_memprobe(int kind, int64_t rows, int width, int k, int cus, const double* x, double* y, hipStream_t s) {
  const double2* xx = reinterpret_cast<const double2*>(x);
  double2* yy = reinterpret_cast<double2*>(y);
  int64_t grid = (rows + 3) / 4;
  if (grid > (int64_t)cus * 8) grid = (int64_t)cus * 8;
  if (grid < 1) grid = 1;
  const int vpl = (width + 63) / 64;
  if (vpl <= 1) probe_launch<1>(kind, rows, width, k, (int)grid, xx, yy, s);
  else if (vpl == 2) probe_launch<2>(kind, rows, width, k, (int)grid, xx, yy, s);
  else probe_launch<4>(kind, rows, width, k, (int)grid, xx, yy, s);
}

}  // namespace ftlm
