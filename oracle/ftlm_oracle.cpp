// ftlm_oracle.cpp -- CPU restatement of the reference FTLM n(mu) path.
//
// TEST INFRASTRUCTURE ONLY. This file is the correctness oracle: only the
// tests, the smoke check and the CPU-baseline leg of the benchmark may load
// or run it, and only as a checker. The product (the HIP library and the
// ftlm_n_vs_mu host executable) never links, loads or calls it.
//
// Every function restates one piece of the reference algorithm in
// src/main.cpp (PeizhiMai/FTLM_extended_Hubbard_development) and cites the
// lines it follows. Floating-point expressions keep the reference's
// evaluation order, so with g++ -O3 (no -march, no FMA contraction, as the
// reference's CMakeLists.txt:35-37 builds) the FTLM results are bit-identical
// to the reference's. Exact diagonalisation is the one deliberate departure:
// the reference calls LAPACK zheev (main.cpp:888-910); this oracle uses its
// own Householder tridiagonalisation + implicit QL (orc_herm_eigvals), whose
// eigenvalues agree with any backward-stable Hermitian solver to ~1e-14.
//
// How each function is pinned (besides bit-equality with the reference
// compiled from source into oracle/_ref, where that builds) is listed in the
// tests: Burnside block dimensions, brute-force Fock-space spectra, free
// fermions, the atomic limit, the mt19937_64 known answer, Toeplitz
// tridiagonal closed forms and numpy eigvalsh.

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <numeric>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#if defined(_OPENMP)
#include <omp.h>
#endif

namespace orc {

using cplx = std::complex<double>;
using u64 = std::uint64_t;

constexpr double kTwoPi = 2.0 * 3.141592653589793238462643383279502884;

struct Model {
  int lx = 2, ly = 2;
  double tx = 1.0, ty = 1.0, tp = 0.0, phix = 1.0, phiy = 1.0, u = 8.0, v = 0.0;
};

[[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }

// ---- lattice: main.cpp:429-460 --------------------------------------------
// neighbours in direction order +x, -x, +y, -y, +x+y, -x+y, +x-y, -x-y;
// bonds are the ordered pairs (s, s+x) and (s, s+y), sorted and deduplicated
// as ordered pairs (so a 2-wide direction contributes both (a,b) and (b,a)).
struct Lattice {
  int lx = 0, ly = 0, L = 0;
  std::vector<std::array<int, 8>> nb;
  std::vector<std::pair<int, int>> bonds;
};

Lattice make_lattice(int lx, int ly) {
  Lattice g;
  g.lx = lx;
  g.ly = ly;
  g.L = lx * ly;
  g.nb.resize(g.L);
  auto at = [lx, ly](int x, int y) { return ((y + ly) % ly) * lx + ((x + lx) % lx); };
  for (int y = 0; y < ly; ++y) {
    for (int x = 0; x < lx; ++x) {
      const int s = at(x, y);
      g.nb[s] = {at(x + 1, y), at(x - 1, y), at(x, y + 1), at(x, y - 1),
                 at(x + 1, y + 1), at(x - 1, y + 1), at(x + 1, y - 1), at(x - 1, y - 1)};
      g.bonds.emplace_back(s, at(x + 1, y));
      g.bonds.emplace_back(s, at(x, y + 1));
    }
  }
  std::sort(g.bonds.begin(), g.bonds.end());
  g.bonds.erase(std::unique(g.bonds.begin(), g.bonds.end()), g.bonds.end());
  return g;
}

// ---- n-particle configurations in ascending order: main.cpp:462-484 -------
std::vector<u64> combos(int L, int n) {
  std::vector<u64> out;
  if (n < 0 || n > L) return out;
  if (n == 0) {
    out.push_back(0);
    return out;
  }
  u64 s = (u64{1} << n) - 1;
  const u64 end = u64{1} << L;
  while (s < end) {
    out.push_back(s);
    const u64 low = s & (~s + 1);
    const u64 ripple = s + low;
    if (ripple == 0) break;
    s = (((ripple ^ s) >> 2) / low) | ripple;
  }
  return out;
}

// ---- U*doublons + V*sum_bonds n_i n_j: main.cpp:486-501 -------------------
double diag_energy(u64 up, u64 dn, const Lattice& g, double U, double V) {
  const int dbl = __builtin_popcountll(up & dn);
  int nn = 0;
  for (const auto& b : g.bonds) {
    const int ni = int((up >> b.first) & 1) + int((dn >> b.first) & 1);
    const int nj = int((up >> b.second) & 1) + int((dn >> b.second) & 1);
    nn += ni * nj;
  }
  return U * double(dbl) + V * double(nn);
}

// ---- (-1)^(occupied strictly between): main.cpp:503-512 -------------------
double hop_sign(u64 bits, int a, int b) {
  const int lo = std::min(a, b), hi = std::max(a, b);
  if (hi - lo <= 1) return 1.0;
  const u64 mask = ((u64{1} << (hi - lo - 1)) - 1) << (lo + 1);
  return (__builtin_popcountll(bits & mask) & 1) ? -1.0 : 1.0;
}

// ---- SplitMix64 finaliser and block seed: main.cpp:514-533 ----------------
u64 mix(u64 z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

u64 block_seed(u64 seed, int nup, int ndn, int mx, int my) {
  u64 s = mix(seed);
  s ^= mix(u64(nup + 1));
  s ^= mix(u64(ndn + 17));
  s ^= mix(u64(mx + 257));
  s ^= mix(u64(my + 65537));
  return mix(s);
}

// ---- twist phase per hopping direction: main.cpp:544-577 ------------------
cplx twist(const Model& m, int dir) {
  const double ax = kTwoPi * m.phix / double(m.lx);
  const double ay = kTwoPi * m.phiy / double(m.ly);
  double a = 0.0;
  switch (dir) {
    case 0: a = ax; break;
    case 1: a = -ax; break;
    case 2: a = ay; break;
    case 3: a = -ay; break;
    case 4: a = ax + ay; break;
    case 5: a = -ax + ay; break;
    case 6: a = ax - ay; break;
    case 7: a = -ax - ay; break;
    default: fail("bad direction");
  }
  return cplx(std::cos(a), std::sin(a));
}

double hop_t(const Model& m, int dir) { return dir < 2 ? m.tx : (dir < 4 ? m.ty : m.tp); }

// ---- translation of one spin species and its fermion sign: main.cpp:579-613
// The sign is the parity of inversions of the translated site list taken in
// ascending order of the original sites.
std::pair<u64, int> shift_spin(u64 bits, int dx, int dy, int lx, int ly) {
  int moved[64];
  int cnt = 0;
  for (int s = 0; s < lx * ly; ++s) {
    if ((bits >> s) & 1) moved[cnt++] = ((s / lx + dy) % ly) * lx + (s % lx + dx) % lx;
  }
  int inv = 0;
  u64 out = 0;
  for (int i = 0; i < cnt; ++i) {
    out |= u64{1} << moved[i];
    for (int j = i + 1; j < cnt; ++j) inv += moved[i] > moved[j];
  }
  return {out, (inv & 1) ? -1 : 1};
}

// both species; shift index j = dy*lx + dx (main.cpp:614-620, 634-638)
void shift_state(u64 up, u64 dn, int j, int lx, int ly, u64& up2, u64& dn2, int& sign) {
  const int dx = j % lx, dy = j / lx;
  const auto a = shift_spin(up, dx, dy, lx, ly);
  const auto b = shift_spin(dn, dx, dy, lx, ly);
  up2 = a.first;
  dn2 = b.first;
  sign = a.second * b.second;
}

// ---- sector basis: main.cpp:622-721 ----------------------------------------
struct Hop {
  int parent, shift, sign;
  cplx amp;
};
struct Parent {
  u64 up, dn;
  int deg;
  double diag;
  std::vector<Hop> hops;
};
struct Sector {
  int nup = 0, ndn = 0;
  size_t full_dim = 0;
  std::vector<Parent> parents;
};

struct KeyHash {
  size_t operator()(const std::pair<u64, u64>& k) const {
    return std::hash<u64>{}(k.first * 0x9e3779b97f4a7c15ULL ^ (k.second + 0x632be59bd9b4e019ULL));
  }
};
struct Rel {
  int parent, shift, sign;
};

Sector make_sector(const Lattice& g, const Model& m, int nup, int ndn) {
  Sector sec;
  sec.nup = nup;
  sec.ndn = ndn;
  const auto ups = combos(g.L, nup);
  const auto dns = combos(g.L, ndn);
  sec.full_dim = ups.size() * dns.size();
  std::unordered_map<std::pair<u64, u64>, Rel, KeyHash> rel;
  rel.reserve(sec.full_dim);
  // States are visited in (up ascending, down ascending) order, so the first
  // unseen member of an orbit is its lexicographic minimum and becomes the
  // parent. Every orbit member records the FIRST shift reaching it.
  for (u64 up : ups) {
    for (u64 dn : dns) {
      const auto key = std::make_pair(up, dn);
      if (rel.count(key)) continue;
      const int pid = int(sec.parents.size());
      Parent p;
      p.up = up;
      p.dn = dn;
      p.diag = diag_energy(up, dn, g, m.u, m.v);
      rel.emplace(key, Rel{pid, 0, 1});
      int deg = 0;
      for (int j = 0; j < g.L; ++j) {
        u64 u2, d2;
        int sg;
        shift_state(up, dn, j, g.lx, g.ly, u2, d2, sg);
        if (u2 == up && d2 == dn) {
          ++deg;
        } else {
          rel.emplace(std::make_pair(u2, d2), Rel{pid, j, sg});
        }
      }
      p.deg = deg;
      sec.parents.push_back(std::move(p));
    }
  }
  // Hops of each parent: up electrons (ascending site) x 8 directions, then
  // down electrons; zero-amplitude hops are kept like the reference does.
  for (Parent& p : sec.parents) {
    for (int spin = 0; spin < 2; ++spin) {
      const u64 bits = spin == 0 ? p.up : p.dn;
      for (int from = 0; from < g.L; ++from) {
        if (!((bits >> from) & 1)) continue;
        for (int dir = 0; dir < 8; ++dir) {
          const int to = g.nb[from][dir];
          if ((bits >> to) & 1) continue;
          const cplx amp = -hop_t(m, dir) * hop_sign(bits, from, to) * twist(m, dir);
          const u64 moved = bits ^ (u64{1} << from) ^ (u64{1} << to);
          const auto key = spin == 0 ? std::make_pair(moved, p.dn) : std::make_pair(p.up, moved);
          const auto it = rel.find(key);
          if (it == rel.end()) fail("hop target missing from relation table");
          p.hops.push_back(Hop{it->second.parent, it->second.shift, it->second.sign, amp});
        }
      }
    }
  }
  return sec;
}

// ---- momentum block (CSR): main.cpp:723-805 --------------------------------
struct Block {
  int mx = 0, my = 0, particles = 0;
  size_t dim = 0;
  std::vector<u64> up, dn;        // representatives of active rows
  std::vector<int> deg;           // their degeneracy
  std::vector<double> dnf;        // their norm factor
  std::vector<double> diag;
  std::vector<int> row_ptr, col;
  std::vector<cplx> val;
};

std::vector<cplx> momentum_phases(const Lattice& g, int mx, int my) {
  std::vector<cplx> ph(g.L);
  for (int j = 0; j < g.L; ++j) {
    const int dx = j % g.lx, dy = j / g.lx;
    const double arg = kTwoPi * (double(mx * dx) / double(g.lx) + double(my * dy) / double(g.ly));
    ph[j] = cplx(std::cos(arg), std::sin(arg));
  }
  return ph;
}

Block make_block(const Sector& sec, const Lattice& g, int mx, int my) {
  const auto ph = momentum_phases(g, mx, my);
  const size_t np = sec.parents.size();
  std::vector<double> dnf(np, 0.0);
  std::vector<int> idx(np, -1);
  int next = 0;
  for (size_t p = 0; p < np; ++p) {
    const Parent& P = sec.parents[p];
    cplx sum = 0.0;
    for (int j = 0; j < g.L; ++j) {
      u64 u2, d2;
      int sg;
      shift_state(P.up, P.dn, j, g.lx, g.ly, u2, d2, sg);
      if (u2 == P.up && d2 == P.dn) sum += double(sg) * ph[j];
    }
    dnf[p] = std::abs(sum) / std::sqrt(double(P.deg));
    if (dnf[p] > 1e-12) idx[p] = next++;
  }
  Block b;
  b.mx = mx;
  b.my = my;
  b.particles = sec.nup + sec.ndn;
  b.dim = size_t(next);
  if (b.dim == 0) return b;
  b.diag.assign(b.dim, 0.0);
  b.row_ptr.push_back(0);
  for (size_t p = 0; p < np; ++p) {
    const int src = idx[p];
    if (src < 0) continue;
    const Parent& P = sec.parents[p];
    b.up.push_back(P.up);
    b.dn.push_back(P.dn);
    b.deg.push_back(P.deg);
    b.dnf.push_back(dnf[p]);
    b.diag[src] = P.diag;
    for (const Hop& h : P.hops) {
      const int tgt = idx[h.parent];
      if (tgt < 0) continue;
      b.col.push_back(tgt);
      b.val.push_back(h.amp * double(h.sign) * ph[h.shift] * (dnf[h.parent] / dnf[p]));
    }
    b.row_ptr.push_back(int(b.col.size()));
  }
  return b;
}

// ---- H apply and BLAS-1: main.cpp:807-845 ----------------------------------
void apply(const Block& b, const cplx* x, cplx* y) {
  for (size_t i = 0; i < b.dim; ++i) {
    cplx acc = 0.0;
    acc += b.diag[i] * x[i];
    for (int e = b.row_ptr[i]; e < b.row_ptr[i + 1]; ++e) acc += b.val[e] * x[b.col[e]];
    y[i] = acc;
  }
}

// order 0: serial sums (the reference). order 1: sums in blocks of 64 rows
// combined pairwise, a stand-in for a GPU tree reduction, used only to
// measure how far a different (equally valid) rounding order moves results.
double sum_norm2(const cplx* x, size_t n, int order) {
  if (order == 0) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += std::norm(x[i]);
    return s;
  }
  std::vector<double> part((n + 63) / 64, 0.0);
  for (size_t i = 0; i < n; ++i) part[i / 64] += std::norm(x[i]);
  while (part.size() > 1) {
    std::vector<double> nx((part.size() + 1) / 2, 0.0);
    for (size_t i = 0; i < part.size(); ++i) nx[i / 2] += part[i];
    part.swap(nx);
  }
  return part.empty() ? 0.0 : part[0];
}

cplx sum_dot(const cplx* a, const cplx* b, size_t n, int order) {
  if (order == 0) {
    cplx s = 0.0;
    for (size_t i = 0; i < n; ++i) s += std::conj(a[i]) * b[i];
    return s;
  }
  std::vector<cplx> part((n + 63) / 64, cplx(0.0));
  for (size_t i = 0; i < n; ++i) part[i / 64] += std::conj(a[i]) * b[i];
  while (part.size() > 1) {
    std::vector<cplx> nx((part.size() + 1) / 2, cplx(0.0));
    for (size_t i = 0; i < part.size(); ++i) nx[i / 2] += part[i];
    part.swap(nx);
  }
  return part.empty() ? cplx(0.0) : part[0];
}

// ---- seeded unit-norm complex Gaussian start vector: main.cpp:1015-1025 ----
// g++ evaluates Complex(normal(rng), normal(rng)) right to left, so the real
// part is the SECOND draw and the imaginary part the first.
std::vector<cplx> start_vector(u64 bseed, int sample, size_t n) {
  std::mt19937_64 rng(mix(bseed ^ u64(sample)));
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<cplx> q(n);
  for (size_t i = 0; i < n; ++i) {
    const double first = nd(rng);
    const double second = nd(rng);
    q[i] = cplx(second, first);
  }
  const double nrm = std::sqrt(sum_norm2(q.data(), n, 0));
  if (nrm <= 0.0) fail("Attempted to normalize a zero vector.");
  for (auto& z : q) z /= nrm;
  return q;
}

// ---- Lanczos recurrence without reorthogonalisation: main.cpp:1027-1057 ----
// Returns the alpha (m entries) and beta (m-1 entries) of one sample.
void lanczos(const Block& b, std::vector<cplx> q, int steps, int order,
             std::vector<double>& alpha, std::vector<double>& beta) {
  const size_t n = b.dim;
  const int kdim = std::min<int>(steps, int(n));
  std::vector<cplx> qp(n, 0.0), qn(n), hq(n);
  alpha.clear();
  beta.clear();
  double bprev = 0.0;
  for (int k = 0; k < kdim; ++k) {
    apply(b, q.data(), hq.data());
    const double a = sum_dot(q.data(), hq.data(), n, order).real();
    for (size_t i = 0; i < n; ++i) qn[i] = hq[i] - a * q[i] - bprev * qp[i];
    alpha.push_back(a);
    const double bb = std::sqrt(sum_norm2(qn.data(), n, order));
    if (bb < 1e-12 || k + 1 == kdim) break;
    beta.push_back(bb);
    qp.swap(q);
    q.swap(qn);
    for (auto& z : q) z /= bb;
    bprev = bb;
  }
}

// ---- classical Jacobi with largest-pivot search: main.cpp:913-994 ----------
// a is a dense symmetric n x n row-major matrix (destroyed). Returns the
// diagonal after convergence and the first row of the accumulated rotations.
void jacobi(int n, std::vector<double>& a, std::vector<double>& eig, std::vector<double>& v0) {
  eig.assign(n, 0.0);
  v0.assign(n, 0.0);
  if (n == 0) return;
  if (n == 1) {
    eig[0] = a[0];
    v0[0] = 1.0;
    return;
  }
  std::vector<double> vec(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) vec[size_t(i) * n + i] = 1.0;
  auto A = [&](int i, int j) -> double& { return a[size_t(i) * n + j]; };
  auto Vv = [&](int i, int j) -> double& { return vec[size_t(i) * n + j]; };
  for (int it = 0; it < 50 * n * n; ++it) {
    int p = 0, q = 1;
    double big = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        const double x = std::abs(A(i, j));
        if (x > big) {
          big = x;
          p = i;
          q = j;
        }
      }
    if (big < 1e-12) break;
    const double app = A(p, p), aqq = A(q, q), apq = A(p, q);
    const double tau = (aqq - app) / (2.0 * apq);
    const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
    const double c = 1.0 / std::sqrt(1.0 + t * t);
    const double s = t * c;
    for (int k = 0; k < n; ++k) {
      if (k == p || k == q) continue;
      const double akp = A(k, p), akq = A(k, q);
      A(k, p) = c * akp - s * akq;
      A(p, k) = A(k, p);
      A(k, q) = s * akp + c * akq;
      A(q, k) = A(k, q);
    }
    A(p, p) = c * c * app - 2.0 * s * c * apq + s * s * aqq;
    A(q, q) = s * s * app + 2.0 * s * c * apq + c * c * aqq;
    A(p, q) = 0.0;
    A(q, p) = 0.0;
    for (int k = 0; k < n; ++k) {
      const double vkp = Vv(k, p), vkq = Vv(k, q);
      Vv(k, p) = c * vkp - s * vkq;
      Vv(k, q) = s * vkp + c * vkq;
    }
  }
  for (int i = 0; i < n; ++i) {
    eig[i] = A(i, i);
    v0[i] = Vv(0, i);
  }
}

// Ritz values (ascending) and first-component weights v1^2 / R of one
// tridiagonal matrix: main.cpp:1059-1094.
void ritz(const std::vector<double>& alpha, const std::vector<double>& beta, int R,
          std::vector<double>& eig_out, std::vector<double>& w_out) {
  const int m = int(alpha.size());
  std::vector<double> t(size_t(m) * m, 0.0);
  for (int i = 0; i < m; ++i) {
    t[size_t(i) * m + i] = alpha[i];
    if (i + 1 < m) {
      t[size_t(i) * m + i + 1] = beta[i];
      t[size_t(i + 1) * m + i] = beta[i];
    }
  }
  std::vector<double> eig, v0;
  jacobi(m, t, eig, v0);
  std::vector<int> ord(m);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int l, int r) { return eig[l] < eig[r]; });
  eig_out.clear();
  w_out.clear();
  for (int k = 0; k < m; ++k) {
    const int o = ord[k];
    eig_out.push_back(eig[o]);
    w_out.push_back((v0[o] * v0[o]) / double(R));
  }
}

// ---- one FTLM block: main.cpp:996-1113 -------------------------------------
void ftlm_block(const Block& b, int R, int steps, u64 bseed, int order, int threads,
                std::vector<double>& eig, std::vector<double>& w) {
  std::vector<std::vector<double>> E(R), W(R);
#if defined(_OPENMP)
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : omp_get_max_threads())
#endif
  for (int r = 0; r < R; ++r) {
    std::vector<double> al, be;
    lanczos(b, start_vector(bseed, r, b.dim), steps, order, al, be);
    ritz(al, be, R, E[r], W[r]);
  }
  (void)threads;
  eig.clear();
  w.clear();
  for (int r = 0; r < R; ++r) {
    eig.insert(eig.end(), E[r].begin(), E[r].end());
    w.insert(w.end(), W[r].begin(), W[r].end());
  }
}

// ---- Hermitian eigenvalues (own method; the reference uses zheev) ----------
// a: n x n complex, column-major (a[i + j*n] = A_ij), Hermitian. Householder
// reduction to real symmetric tridiagonal form, then implicit QL with
// Wilkinson shifts. Eigenvalues ascending.
void tridiag_ql(std::vector<double>& d, std::vector<double>& e) {
  const int n = int(d.size());
  if (n <= 1) return;
  e.resize(n, 0.0);
  e[n - 1] = 0.0;
  // a coupling splits the matrix below eps times its diagonal neighbours or
  // below eps ||T|| (the second lets clusters of eigenvalues near zero
  // converge; a perturbation of the reduction's own size)
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i)
    tnorm = std::max(tnorm, std::abs(d[i]) + std::abs(e[i]) + (i > 0 ? std::abs(e[i - 1]) : 0.0));
  const double floor_e = std::numeric_limits<double>::epsilon() * tnorm;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int mm;
    do {
      for (mm = l; mm < n - 1; ++mm) {
        const double dd = std::abs(d[mm]) + std::abs(d[mm + 1]);
        if (std::abs(e[mm]) <= std::numeric_limits<double>::epsilon() * dd || std::abs(e[mm]) <= floor_e) break;
      }
      if (mm != l) {
        if (++iter > 60) fail("QL iteration did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[mm] - d[l] + e[l] / (g + (g >= 0.0 ? std::abs(r) : -std::abs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = mm - 1; i >= l; --i) {
          double f = s * e[i];
          const double bb = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[mm] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * bb;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - bb;
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[mm] = 0.0;
      }
    } while (mm != l);
  }
  std::sort(d.begin(), d.end());
}

std::vector<double> herm_eigvals(int n, std::vector<cplx> a) {
  if (n == 0) return {};
  auto A = [&](int i, int j) -> cplx& { return a[size_t(i) + size_t(j) * n]; };
  std::vector<double> d(n), e(n, 0.0);
  std::vector<cplx> v(n), p(n);
  for (int k = 0; k + 2 < n; ++k) {
    // reflect column k below the diagonal onto its first entry
    double xn = 0.0;
    for (int i = k + 1; i < n; ++i) xn += std::norm(A(i, k));
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    const cplx x0 = A(k + 1, k);
    const double ax0 = std::abs(x0);
    const cplx ph = ax0 > 0.0 ? x0 / ax0 : cplx(1.0);
    const cplx alpha = -ph * xn;
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) {
      v[i] = A(i, k) - (i == k + 1 ? alpha : cplx(0.0));
      vn += std::norm(v[i]);
    }
    vn = std::sqrt(vn);
    if (vn == 0.0) continue;
    for (int i = k + 1; i < n; ++i) v[i] /= vn;
    // p = A22 v ; K = v^H p ; w = p - K v ; A22 -= 2 (v w^H + w v^H)
    for (int i = k + 1; i < n; ++i) {
      cplx s = 0.0;
      for (int j = k + 1; j < n; ++j) s += A(i, j) * v[j];
      p[i] = s;
    }
    cplx K = 0.0;
    for (int i = k + 1; i < n; ++i) K += std::conj(v[i]) * p[i];
    for (int i = k + 1; i < n; ++i) p[i] -= K.real() * v[i];
    for (int j = k + 1; j < n; ++j)
      for (int i = k + 1; i < n; ++i)
        A(i, j) -= 2.0 * (v[i] * std::conj(p[j]) + p[i] * std::conj(v[j]));
    A(k + 1, k) = alpha;
    A(k, k + 1) = std::conj(alpha);
    for (int i = k + 2; i < n; ++i) {
      A(i, k) = 0.0;
      A(k, i) = 0.0;
    }
  }
  for (int i = 0; i < n; ++i) d[i] = A(i, i).real();
  for (int i = 0; i + 1 < n; ++i) e[i] = std::abs(A(i + 1, i));
  tridiag_ql(d, e);
  return d;
}

// dense Hermitised block, then eigenvalues: main.cpp:847-911 (n==1 shortcut
// takes the real part of self-edges, main.cpp:852-860).
std::vector<double> exact_block(const Block& b) {
  const int n = int(b.dim);
  if (n == 0) return {};
  if (n == 1) {
    double ev = b.diag[0];
    for (int e = b.row_ptr[0]; e < b.row_ptr[1]; ++e)
      if (b.col[e] == 0) ev += b.val[e].real();
    return {ev};
  }
  std::vector<cplx> a(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) {
    a[size_t(i) + size_t(i) * n] = b.diag[i];
    for (int e = b.row_ptr[i]; e < b.row_ptr[i + 1]; ++e) a[size_t(i) + size_t(b.col[e]) * n] += b.val[e];
  }
  for (int j = 0; j < n; ++j) {
    a[size_t(j) + size_t(j) * n] = cplx(a[size_t(j) + size_t(j) * n].real(), 0.0);
    for (int i = 0; i < j; ++i) {
      const cplx avg = 0.5 * (a[size_t(i) + size_t(j) * n] + std::conj(a[size_t(j) + size_t(i) * n]));
      a[size_t(i) + size_t(j) * n] = avg;
      a[size_t(j) + size_t(i) * n] = std::conj(avg);
    }
  }
  return herm_eigvals(n, std::move(a));
}

// ---- streaming log-sum-exp accumulation: main.cpp:1172-1219 ---------------
void accumulate(const double* eig, const double* w, size_t cnt, double pref, int particles, double beta,
                const std::vector<double>& mu, std::vector<double>& emin, std::vector<double>& z,
                std::vector<double>& n1, std::vector<double>& n2) {
  const double N = double(particles);
  for (size_t k = 0; k < mu.size(); ++k) {
    double bmin = std::numeric_limits<double>::infinity();
    for (size_t i = 0; i < cnt; ++i) bmin = std::min(bmin, eig[i] - mu[k] * N);
    if (!std::isfinite(bmin)) continue;
    if (!std::isfinite(emin[k])) {
      emin[k] = bmin;
    } else if (bmin < emin[k]) {
      const double sc = std::exp(-beta * (emin[k] - bmin));
      z[k] *= sc;
      n1[k] *= sc;
      n2[k] *= sc;
      emin[k] = bmin;
    }
    for (size_t i = 0; i < cnt; ++i) {
      const double sh = eig[i] - mu[k] * N - emin[k];
      const double wt = pref * (w ? w[i] : 1.0) * std::exp(-beta * sh);
      z[k] += wt;
      n1[k] += N * wt;
      n2[k] += N * N * wt;
    }
  }
}

// ---- mu grid: main.cpp:1400-1407 -------------------------------------------
std::vector<double> linspace(double a, double b, int cnt) {
  std::vector<double> v(cnt);
  const double st = (b - a) / double(cnt - 1);
  for (int i = 0; i < cnt; ++i) v[i] = a + st * double(i);
  return v;
}

// ---- finalisation of one (beta, mu) point: main.cpp:1770-1788, 1805-1823 --
struct Point {
  double n, cc, kappa, z, logz;
};
Point finish(double z, double n1, double n2, double emin, double beta, int L) {
  Point p;
  p.z = z;
  p.logz = std::log(z) - beta * emin;
  const double mean = n1 / z;
  const double mean2 = n2 / z;
  const double cc = (mean2 - mean * mean) / double(L);
  p.n = mean / double(L);
  p.cc = std::max(0.0, cc);
  p.kappa = beta * p.cc;
  return p;
}

// ---- command-line driver: main.cpp:193-368 (flags), 1459-1839 (loop) -------
struct Opts {
  Model m;
  double beta = 2.0;
  std::vector<double> betas;
  double mu_min = -4.0, mu_max = 10.0;
  int mu_count = 61, samples = 5, steps = 80, threads = 0;
  size_t threshold = 256;
  u64 seed = 12345;
  size_t max_sector = 2000000;
  std::string output = "n_vs_mu.csv", checkpoint;
  int order = 0;
};

template <typename T>
T parse_as(const std::string& s) {
  std::istringstream in(s);
  T v{};
  in >> v;
  if (!in || !in.eof()) fail("Failed to parse value: " + s);
  return v;
}

Opts parse(int argc, char** argv) {
  Opts o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() {
      if (i + 1 >= argc) fail("Missing value for " + a);
      return std::string(argv[++i]);
    };
    if (a == "--lx") o.m.lx = parse_as<int>(val());
    else if (a == "--ly") o.m.ly = parse_as<int>(val());
    else if (a == "--tx") o.m.tx = parse_as<double>(val());
    else if (a == "--ty") o.m.ty = parse_as<double>(val());
    else if (a == "--tp") o.m.tp = parse_as<double>(val());
    else if (a == "--phix") o.m.phix = parse_as<double>(val());
    else if (a == "--phiy") o.m.phiy = parse_as<double>(val());
    else if (a == "--u") o.m.u = parse_as<double>(val());
    else if (a == "--v") o.m.v = parse_as<double>(val());
    else if (a == "--beta") o.beta = parse_as<double>(val());
    else if (a == "--beta-list") {
      std::istringstream in(val());
      std::string item;
      while (std::getline(in, item, ',')) o.betas.push_back(parse_as<double>(item));
    } else if (a == "--mu-min") o.mu_min = parse_as<double>(val());
    else if (a == "--mu-max") o.mu_max = parse_as<double>(val());
    else if (a == "--mu-count") o.mu_count = parse_as<int>(val());
    else if (a == "--samples") o.samples = parse_as<int>(val());
    else if (a == "--lanczos-steps") o.steps = parse_as<int>(val());
    else if (a == "--threads") o.threads = parse_as<int>(val());
    else if (a == "--exact-block-threshold") o.threshold = parse_as<size_t>(val());
    else if (a == "--seed") o.seed = parse_as<u64>(val());
    else if (a == "--max-sector-dim") o.max_sector = parse_as<size_t>(val());
    else if (a == "--checkpoint") o.checkpoint = val();
    else if (a == "--output") o.output = val();
    else if (a == "--sum-order") o.order = parse_as<int>(val());
    else fail("Unknown argument: " + a);
  }
  if (o.betas.empty()) o.betas.push_back(o.beta);
  o.beta = o.betas.front();
  return o;
}

// checkpoint record writer: main.cpp:1363-1398 (format), 1279-1295 (header)
void write_record(const std::string& path, int nup, int ndn, int mx, int my, int particles, u64 dim,
                  double pref, bool exact, const std::vector<double>& eig, const std::vector<double>& w) {
  if (path.empty()) return;
  {
    std::ifstream in(path, std::ios::binary);
    if (!in || in.peek() == std::ifstream::traits_type::eof()) {
      std::ofstream out(path, std::ios::binary);
      out.write("FTLMCP1\n", 8);
    }
  }
  std::ofstream out(path, std::ios::binary | std::ios::app);
  out.write("BLKREC1\n", 8);
  const int32_t hdr[5] = {nup, ndn, mx, my, particles};
  out.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
  out.write(reinterpret_cast<const char*>(&dim), 8);
  out.write(reinterpret_cast<const char*>(&pref), 8);
  const uint8_t ex = exact ? 1 : 0;
  out.write(reinterpret_cast<const char*>(&ex), 1);
  const u64 ne = eig.size(), nw = w.size();
  out.write(reinterpret_cast<const char*>(&ne), 8);
  out.write(reinterpret_cast<const char*>(&nw), 8);
  out.write(reinterpret_cast<const char*>(eig.data()), std::streamsize(8 * ne));
  out.write(reinterpret_cast<const char*>(w.data()), std::streamsize(8 * nw));
}

int run_driver(int argc, char** argv, bool ed_only) {
  const Opts o = parse(argc, argv);
  const Lattice g = make_lattice(o.m.lx, o.m.ly);
  const auto mu = linspace(o.mu_min, o.mu_max, o.mu_count);
  struct Stored {
    std::vector<double> eig, w;
    double pref;
    int particles;
  };
  std::vector<Stored> st;
  for (int nup = 0; nup <= g.L; ++nup) {
    for (int ndn = 0; ndn <= g.L; ++ndn) {
      const Sector sec = make_sector(g, o.m, nup, ndn);
      if (sec.full_dim == 0) continue;
      if (sec.full_dim > o.max_sector) fail("sector exceeds --max-sector-dim");
      size_t active = 0;
      for (int my = 0; my < g.ly; ++my) {
        for (int mx = 0; mx < g.lx; ++mx) {
          const Block b = make_block(sec, g, mx, my);
          active += b.dim;
          if (b.dim == 0) continue;
          Stored s;
          s.particles = b.particles;
          const bool ex = ed_only || b.dim <= o.threshold;
          if (ex) {
            s.eig = exact_block(b);
            s.pref = 1.0;
          } else {
            ftlm_block(b, o.samples, o.steps, block_seed(o.seed, nup, ndn, mx, my), o.order, o.threads, s.eig, s.w);
            s.pref = double(b.dim);
          }
          if (!ed_only)
            write_record(o.checkpoint, nup, ndn, mx, my, b.particles, b.dim, s.pref, ex, s.eig, s.w);
          st.push_back(std::move(s));
        }
      }
      std::cout << "  sector Nup=" << nup << " Ndown=" << ndn << " full_dim=" << sec.full_dim
                << " parents=" << sec.parents.size() << " active_k_dim_sum=" << active << "\n";
    }
  }
  const bool multi = o.betas.size() > 1 && !ed_only;
  std::ofstream out(o.output);
  if (!out) fail("Failed to open output file: " + o.output);
  out << (multi ? "beta," : "") << "mu,n,charge_correlation,compressibility,partition_like,log_partition\n";
  out << std::setprecision(15);
  const size_t nb = ed_only ? 1 : o.betas.size();
  for (size_t ib = 0; ib < nb; ++ib) {
    const double beta = o.betas[ib];
    std::vector<double> emin(mu.size(), std::numeric_limits<double>::infinity()), z(mu.size(), 0.0),
        n1(mu.size(), 0.0), n2(mu.size(), 0.0);
    for (const Stored& s : st)
      accumulate(s.eig.data(), s.w.empty() ? nullptr : s.w.data(), s.eig.size(), s.pref, s.particles, beta, mu,
                 emin, z, n1, n2);
    for (size_t k = 0; k < mu.size(); ++k) {
      const Point p = finish(z[k], n1[k], n2[k], emin[k], beta, g.L);
      if (multi) out << beta << ",";
      out << mu[k] << "," << p.n << "," << p.cc << "," << p.kappa << "," << p.z << "," << p.logz << "\n";
    }
  }
  std::cout << "Wrote " << o.output << "\n";
  return 0;
}

}  // namespace orc

// ============================================================================
// C API (loaded with ctypes by the tests and the CPU-baseline leg of bench.py)
// ============================================================================
extern "C" {

struct orc_model {
  int lx, ly;
  double tx, ty, tp, phix, phiy, u, v;
};

static orc::Model to_model(const orc_model* p) {
  orc::Model m;
  m.lx = p->lx;
  m.ly = p->ly;
  m.tx = p->tx;
  m.ty = p->ty;
  m.tp = p->tp;
  m.phix = p->phix;
  m.phiy = p->phiy;
  m.u = p->u;
  m.v = p->v;
  return m;
}

struct orc_block_h {
  orc::Model m;
  orc::Lattice g;
  orc::Block b;
  size_t full_dim = 0, parents = 0;
};

void orc_lattice(int lx, int ly, int* nb, int* bonds, int* nbonds) {
  const orc::Lattice g = orc::make_lattice(lx, ly);
  for (int s = 0; s < g.L; ++s)
    for (int d = 0; d < 8; ++d) nb[s * 8 + d] = g.nb[s][d];
  *nbonds = int(g.bonds.size());
  for (size_t i = 0; i < g.bonds.size(); ++i) {
    bonds[2 * i] = g.bonds[i].first;
    bonds[2 * i + 1] = g.bonds[i].second;
  }
}

unsigned long long orc_mix_u64(unsigned long long z) { return orc::mix(z); }
// n-th output (1-based) of std::mt19937_64 seeded with `seed` (known-answer
// pin of the engine behind the start vectors).
unsigned long long orc_mt_nth(unsigned long long seed, long long n) {
  std::mt19937_64 rng(seed);
  unsigned long long v = 0;
  for (long long i = 0; i < n; ++i) v = rng();
  return v;
}
unsigned long long orc_block_seed(unsigned long long seed, int nup, int ndn, int mx, int my) {
  return orc::block_seed(seed, nup, ndn, mx, my);
}
void orc_twist_phases(const orc_model* p, double* out16) {
  const orc::Model m = to_model(p);
  for (int d = 0; d < 8; ++d) {
    const auto z = orc::twist(m, d);
    out16[2 * d] = z.real();
    out16[2 * d + 1] = z.imag();
  }
}
double orc_diag_energy(const orc_model* p, unsigned long long up, unsigned long long dn) {
  const orc::Model m = to_model(p);
  return orc::diag_energy(up, dn, orc::make_lattice(m.lx, m.ly), m.u, m.v);
}
double orc_hop_sign(unsigned long long bits, int a, int b) { return orc::hop_sign(bits, a, b); }
void orc_shift_state(unsigned long long up, unsigned long long dn, int j, int lx, int ly,
                     unsigned long long* up2, unsigned long long* dn2, int* sign) {
  orc::u64 a, b;
  orc::shift_state(up, dn, j, lx, ly, a, b, *sign);
  *up2 = a;
  *dn2 = b;
}
long long orc_combos(int L, int n, unsigned long long* out) {
  const auto v = orc::combos(L, n);
  if (out) std::copy(v.begin(), v.end(), out);
  return (long long)v.size();
}

void* orc_block_new(const orc_model* p, int nup, int ndn, int mx, int my) {
  try {
    auto* h = new orc_block_h;
    h->m = to_model(p);
    h->g = orc::make_lattice(h->m.lx, h->m.ly);
    const orc::Sector sec = orc::make_sector(h->g, h->m, nup, ndn);
    h->full_dim = sec.full_dim;
    h->parents = sec.parents.size();
    h->b = orc::make_block(sec, h->g, mx, my);
    return h;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "orc_block_new: %s\n", e.what());
    return nullptr;
  }
}
void orc_block_free(void* h) { delete static_cast<orc_block_h*>(h); }
long long orc_block_dim(void* h) { return (long long)static_cast<orc_block_h*>(h)->b.dim; }
long long orc_block_nnz(void* h) { return (long long)static_cast<orc_block_h*>(h)->b.col.size(); }
long long orc_block_full_dim(void* h) { return (long long)static_cast<orc_block_h*>(h)->full_dim; }
long long orc_block_parents(void* h) { return (long long)static_cast<orc_block_h*>(h)->parents; }
int orc_block_particles(void* h) { return static_cast<orc_block_h*>(h)->b.particles; }

void orc_block_reps(void* hv, unsigned long long* up, unsigned long long* dn, int* deg, double* dnf) {
  const auto& b = static_cast<orc_block_h*>(hv)->b;
  for (size_t i = 0; i < b.dim; ++i) {
    up[i] = b.up[i];
    dn[i] = b.dn[i];
    deg[i] = b.deg[i];
    dnf[i] = b.dnf[i];
  }
}
void orc_block_csr(void* hv, int* row_ptr, int* col, double* val, double* diag) {
  const auto& b = static_cast<orc_block_h*>(hv)->b;
  if (b.dim == 0) return;
  std::copy(b.row_ptr.begin(), b.row_ptr.end(), row_ptr);
  std::copy(b.col.begin(), b.col.end(), col);
  for (size_t e = 0; e < b.val.size(); ++e) {
    val[2 * e] = b.val[e].real();
    val[2 * e + 1] = b.val[e].imag();
  }
  std::copy(b.diag.begin(), b.diag.end(), diag);
}
// x, y: R interleaved complex vectors in row-major [dim][R] layout
void orc_block_apply(void* hv, int R, const double* x, double* y) {
  const auto& b = static_cast<orc_block_h*>(hv)->b;
  std::vector<orc::cplx> xi(b.dim), yi(b.dim);
  for (int r = 0; r < R; ++r) {
    for (size_t i = 0; i < b.dim; ++i) xi[i] = orc::cplx(x[2 * (i * R + r)], x[2 * (i * R + r) + 1]);
    orc::apply(b, xi.data(), yi.data());
    for (size_t i = 0; i < b.dim; ++i) {
      y[2 * (i * R + r)] = yi[i].real();
      y[2 * (i * R + r) + 1] = yi[i].imag();
    }
  }
}
void orc_start_vector(unsigned long long bseed, int sample, long long n, double* out) {
  const auto q = orc::start_vector(bseed, sample, size_t(n));
  for (long long i = 0; i < n; ++i) {
    out[2 * i] = q[i].real();
    out[2 * i + 1] = q[i].imag();
  }
}
// One sample's alpha/beta from an explicit start vector (or the seeded one
// when start == NULL). Returns m (alpha count); beta receives m-1 values.
int orc_lanczos(void* hv, const double* start, unsigned long long bseed, int sample, int steps, int order,
                double* alpha, double* beta) {
  const auto& b = static_cast<orc_block_h*>(hv)->b;
  std::vector<orc::cplx> q;
  if (start) {
    q.resize(b.dim);
    for (size_t i = 0; i < b.dim; ++i) q[i] = orc::cplx(start[2 * i], start[2 * i + 1]);
  } else {
    q = orc::start_vector(bseed, sample, b.dim);
  }
  std::vector<double> al, be;
  orc::lanczos(b, q, steps, order, al, be);
  std::copy(al.begin(), al.end(), alpha);
  std::copy(be.begin(), be.end(), beta);
  return int(al.size());
}
void orc_ritz(int m, const double* alpha, const double* beta, int R, double* eig, double* w) {
  std::vector<double> al(alpha, alpha + m), be(beta, beta + (m > 0 ? m - 1 : 0)), e, ww;
  orc::ritz(al, be, R, e, ww);
  std::copy(e.begin(), e.end(), eig);
  std::copy(ww.begin(), ww.end(), w);
}
// Dense symmetric Jacobi (row-major n x n, destroyed): eigenvalues in the
// order Jacobi leaves them and the first row of the eigenvector matrix.
void orc_jacobi(int n, double* a, double* eig, double* v0) {
  std::vector<double> A(a, a + size_t(n) * n), e, v;
  orc::jacobi(n, A, e, v);
  std::copy(e.begin(), e.end(), eig);
  std::copy(v.begin(), v.end(), v0);
  std::copy(A.begin(), A.end(), a);
}
long long orc_ftlm_block(void* hv, int R, int steps, unsigned long long bseed, int order, int threads, double* eig,
                         double* w) {
  const auto& b = static_cast<orc_block_h*>(hv)->b;
  std::vector<double> e, ww;
  orc::ftlm_block(b, R, steps, bseed, order, threads, e, ww);
  if (eig) std::copy(e.begin(), e.end(), eig);
  if (w) std::copy(ww.begin(), ww.end(), w);
  return (long long)e.size();
}
// a solver failure fills the output with NaN (no exception crosses the C
// boundary), so a checker sees a mismatch instead of an abort
void orc_exact_block(void* hv, double* eig) {
  const auto& b = static_cast<orc_block_h*>(hv)->b;
  try {
    const auto e = orc::exact_block(b);
    std::copy(e.begin(), e.end(), eig);
  } catch (const std::exception&) {
    std::fill(eig, eig + b.dim, std::numeric_limits<double>::quiet_NaN());
  }
}
void orc_herm_eigvals(int n, const double* a, double* w) {
  std::vector<orc::cplx> A(size_t(n) * n);
  for (size_t i = 0; i < A.size(); ++i) A[i] = orc::cplx(a[2 * i], a[2 * i + 1]);
  try {
    const auto d = orc::herm_eigvals(n, std::move(A));
    std::copy(d.begin(), d.end(), w);
  } catch (const std::exception&) {
    std::fill(w, w + n, std::numeric_limits<double>::quiet_NaN());
  }
}
void orc_accumulate(const double* eig, const double* w, long long cnt, double pref, int particles, double beta,
                    const double* mu, int nmu, double* emin, double* z, double* n1, double* n2) {
  std::vector<double> M(mu, mu + nmu), E(emin, emin + nmu), Z(z, z + nmu), A(n1, n1 + nmu), B(n2, n2 + nmu);
  orc::accumulate(eig, w, size_t(cnt), pref, particles, beta, M, E, Z, A, B);
  std::copy(E.begin(), E.end(), emin);
  std::copy(Z.begin(), Z.end(), z);
  std::copy(A.begin(), A.end(), n1);
  std::copy(B.begin(), B.end(), n2);
}
void orc_linspace(double a, double b, int cnt, double* out) {
  const auto v = orc::linspace(a, b, cnt);
  std::copy(v.begin(), v.end(), out);
}
// Full driver (FTLM executable flags; ed != 0 mirrors the ED executable).
int orc_main(int argc, char** argv, int ed) {
  try {
    return orc::run_driver(argc, argv, ed != 0);
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << "\n";
    return 1;
  }
}

}  // extern "C"

#if defined(ORC_CLI)
int main(int argc, char** argv) {
  bool ed = false;
  std::vector<char*> args;
  for (int i = 0; i < argc; ++i) {
    if (std::string(argv[i]) == "--ed") {
      ed = true;
      continue;
    }
    args.push_back(argv[i]);
  }
  return orc_main(int(args.size()), args.data(), ed ? 1 : 0);
}
#endif
