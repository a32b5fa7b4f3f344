// ftlm_exact.hip -- batched FP64 Hermitian eigenvalues for small blocks.
//
// Replaces LAPACK zheev('N','U') in diagonalize_block_exact
// (main.cpp:888-910): one workgroup per block reduces the Hermitised dense
// matrix (column-major, A[i + j*n]) to real symmetric tridiagonal form with
// Householder reflections, then one lane runs implicit QL with Wilkinson
// shifts on (d, e). Eigenvalues are returned ascending, like zheev's.
// Both steps are backward stable, so the eigenvalues agree with zheev's to
// a few ulps of ||A||. Blocks of kBigN rows or more run in a 1024-thread
// workgroup (a second launch over the same batch), the others in 256 threads:
// a large block's reduction is bound by how many of its matrix rows are in
// flight at once.
#include "ftlm_device.h"
#include "ftlm_launch.h"

namespace ftlm {

namespace {

constexpr int kBigN = 512;

template <int kT>
__device__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if (tid < off) red[tid] += red[tid + off];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

__device__ double hypot_safe(double a, double b) {
  a = fabs(a);
  b = fabs(b);
  const double hi = a > b ? a : b, lo = a > b ? b : a;
  if (hi == 0.0) return 0.0;
  const double r = lo / hi;
  return hi * sqrt(1.0 + r * r);
}

// Implicit QL with Wilkinson shifts on a symmetric tridiagonal matrix
// (d: diagonal, e[i]: coupling of i and i+1). Returns false if an eigenvalue
// needed more than 60 sweeps. A coupling splits the matrix when it is below
// eps times its two diagonal neighbours or below eps times the matrix norm:
// the second test (a perturbation of eps ||T||, the size of the reduction's
// own rounding) is what lets clusters of eigenvalues near zero converge,
// where the relative test alone waits for couplings below rounding noise.
__device__ bool tridiag_ql(double* d, double* e, int n) {
  if (n <= 1) return true;
  e[n - 1] = 0.0;
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i)
    tnorm = fmax(tnorm, fabs(d[i]) + fabs(e[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0));
  const double floor_e = 2.220446049250313e-16 * tnorm;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 2.220446049250313e-16 * dd || fabs(e[m]) <= floor_e) break;
      }
      if (m != l) {
        if (++iter > 60) return false;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot_safe(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          const double f = s * e[i];
          const double b = c * e[i];
          r = hypot_safe(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  return true;
}

// kT threads per block; the launch for kT handles the blocks of its size
// class (big: n >= kBigN) and skips the others.
template <int kT>
__global__ void __launch_bounds__(kT) k_herm_eig(const int64_t* __restrict__ moff, const int64_t* __restrict__ dims,
                                                 double* mats, double* work, double* __restrict__ eig_out,
                                                 const int64_t* __restrict__ eoff) {
  // the Householder vector v and p = A22 v in LDS when they fit (always for
  // the small class), else in the block's work area
  constexpr int kLdsN = kT > 256 ? 2048 : kBigN;
  __shared__ double red[kT];
  __shared__ double shv[2 * kLdsN], shp[2 * kLdsN];
  const int b = blockIdx.x;
  const int64_t n = dims[b];
  if (n == 0 || (n >= kBigN) != (kT > 256)) return;
  double* A = mats + 2 * moff[b];
  // work: per block v (n complex), p (n complex), d (n), e (n) = 6n doubles
  double* wk = work + 6 * eoff[b];
  double* V = n <= kLdsN ? shv : wk;          // 2n
  double* P = n <= kLdsN ? shp : wk + 2 * n;  // 2n
  double* D = wk + 4 * n;   // n
  double* E = wk + 5 * n;   // n
  const int tid = threadIdx.x;
  auto a = [&](int64_t i, int64_t j) -> double* { return A + 2 * (i + j * n); };
  for (int64_t k = 0; k + 2 < n; ++k) {
    double part = 0.0;
    for (int64_t i = k + 1 + tid; i < n; i += kT) {
      const double* z = a(i, k);
      part += z[0] * z[0] + z[1] * z[1];
    }
    const double xn = sqrt(block_sum<kT>(part, red));
    if (xn == 0.0) continue;
    const double x0r = a(k + 1, k)[0], x0i = a(k + 1, k)[1];
    const double ax0 = hypot_safe(x0r, x0i);
    const double phr = ax0 > 0.0 ? x0r / ax0 : 1.0, phi = ax0 > 0.0 ? x0i / ax0 : 0.0;
    const double alr = -phr * xn, ali = -phi * xn;
    part = 0.0;
    for (int64_t i = k + 1 + tid; i < n; i += kT) {
      double vr = a(i, k)[0], vi = a(i, k)[1];
      if (i == k + 1) {
        vr -= alr;
        vi -= ali;
      }
      V[2 * i] = vr;
      V[2 * i + 1] = vi;
      part += vr * vr + vi * vi;
    }
    const double vn = sqrt(block_sum<kT>(part, red));
    if (vn == 0.0) continue;
    for (int64_t i = k + 1 + tid; i < n; i += kT) {
      V[2 * i] /= vn;
      V[2 * i + 1] /= vn;
    }
    __syncthreads();
    // p = A22 v
    double kr = 0.0;
    for (int64_t i = k + 1 + tid; i < n; i += kT) {
      double sr = 0.0, si = 0.0;
      for (int64_t j = k + 1; j < n; ++j) {
        const double* z = a(i, j);
        const double vr = V[2 * j], vi = V[2 * j + 1];
        sr += z[0] * vr - z[1] * vi;
        si += z[0] * vi + z[1] * vr;
      }
      P[2 * i] = sr;
      P[2 * i + 1] = si;
      kr += V[2 * i] * sr + V[2 * i + 1] * si;  // Re(conj(v_i) p_i)
    }
    const double K = block_sum<kT>(kr, red);
    for (int64_t i = k + 1 + tid; i < n; i += kT) {
      P[2 * i] -= K * V[2 * i];
      P[2 * i + 1] -= K * V[2 * i + 1];
    }
    __syncthreads();
    // A22 -= 2 (v w^H + w v^H)
    const int64_t m = n - k - 1;
    for (int64_t t = tid; t < m * m; t += kT) {
      const int64_t i = k + 1 + t % m, j = k + 1 + t / m;
      const double vir = V[2 * i], vii = V[2 * i + 1], vjr = V[2 * j], vji = V[2 * j + 1];
      const double wir = P[2 * i], wii = P[2 * i + 1], wjr = P[2 * j], wji = P[2 * j + 1];
      // v_i conj(w_j) + w_i conj(v_j)
      const double re = (vir * wjr + vii * wji) + (wir * vjr + wii * vji);
      const double im = (vii * wjr - vir * wji) + (wii * vjr - wir * vji);
      double* z = a(i, j);
      z[0] -= 2.0 * re;
      z[1] -= 2.0 * im;
    }
    __syncthreads();
    if (tid == 0) {
      a(k + 1, k)[0] = alr;
      a(k + 1, k)[1] = ali;
    }
    __syncthreads();
  }
  for (int64_t i = tid; i < n; i += kT) {
    D[i] = a(i, i)[0];
    E[i] = i + 1 < n ? hypot_safe(a(i + 1, i)[0], a(i + 1, i)[1]) : 0.0;
  }
  __syncthreads();
  if (tid == 0) {
    const bool ok = tridiag_ql(D, E, (int)n);
    // ascending insertion sort
    for (int64_t i = 1; i < n; ++i) {
      const double x = D[i];
      int64_t j = i - 1;
      while (j >= 0 && D[j] > x) {
        D[j + 1] = D[j];
        --j;
      }
      D[j + 1] = x;
    }
    for (int64_t i = 0; i < n; ++i) eig_out[eoff[b] + i] = ok ? D[i] : __builtin_nan("");
  }
}

}  // namespace

void launch_herm_eig(const int64_t* moff, const int64_t* dims, int nblk, double* mats, double* work,
                     double* eig_out, const int64_t* eoff, hipStream_t s) {
  if (nblk <= 0) return;
  k_herm_eig<256><<<nblk, 256, 0, s>>>(moff, dims, mats, work, eig_out, eoff);
  k_herm_eig<1024><<<nblk, 1024, 0, s>>>(moff, dims, mats, work, eig_out, eoff);
}

}  // namespace ftlm
