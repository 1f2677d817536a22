// Spectral open boundary conditions on the device: Beyn's contour-integral
// method for the polynomial eigenvalue problem of a lead layer,
// (sum_n a_n w^n) v = 0 with n = -p..p, the reflected-mode selection and the
// surface Green's function built from the reflected modes, batched over the
// energies of a call.
//
// Follows src/qttools/nevp/beyn.py:14-379 (quadrature on two circles, the
// moments P_0 = q_0 Y and P_1 = q_1 Y, SVD projection with the rank cut
// s > s_max m_0 eps, reduced eigenproblem) and
// src/qttools/boundary_conditions/obc/spectral.py:221-629 (unit-norm modes,
// reflected-mode masks, x_ii from the Moore-Penrose pseudo-inverse of the
// reflected modes, fixed-point refinement), for block_sections = 1.
//
// The O(b^3) work runs on the existing kernels: every operator inverse
// T(z_k)^-1 on the batched Gauss-Jordan kernel, every product on the grouped
// FP64 MFMA GEMM (the quadrature sums accumulate in its epilogue).  The
// kernels of this file are the small per-energy parts:
//   * beyn_probe_kernel: the probe block Y (uniform [0,1) real and imaginary
//     parts from a counter-based hash of (seed, energy, row, column));
//   * operator_kernel: T(z) = sum_n z^n a_n for a chunk of quadrature points;
//   * jacobi_svd_kernel: one-sided (Hestenes) Jacobi SVD of P_0 (b x m_0),
//     one workgroup per energy, columns rotated in pairs by a round-robin
//     tournament (one pair per wave at a time);
//   * hqr_eig_kernel: eigenpairs of the reduced m x m matrix (m <= 64), one
//     wave per energy: Householder reduction to Hessenberg form, complex
//     single-shift QR with Wilkinson shifts and deflation to the Schur form,
//     eigenvectors by back substitution on the triangular factor;
//   * mode_kernel: residual, group velocity and decay of each mode and the
//     reflected-mode mask, one wave per (energy, mode).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rgf.h"
#include "qx_common.h"
#include "rgf_driver.h"

using namespace qx;

namespace {

constexpr double2 ONE = {1.0, 0.0};
constexpr double2 MONE = {-1.0, 0.0};
constexpr int kMaxModes = 64;      // m_0 limit of the reduced eigenproblem (one wave)
constexpr int kMaxCoef = 7;        // polynomial terms n = -3..3
constexpr int kPtsPerLaunch = 16;  // quadrature points assembled per operator launch

__device__ inline double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }
__device__ inline double cabs(double2 a) { return hypot(a.x, a.y); }
__device__ inline double2 cscale(double2 a, double s) { return cmk(a.x * s, a.y * s); }
__device__ inline double2 cdiv(double2 a, double2 b) { return cmul(a, crecip(b)); }
__device__ inline double2 csqrt_(double2 a) {
  const double r = cabs(a);
  if (r == 0.0) return cmk(0.0, 0.0);
  double re = sqrt(0.5 * (r + fabs(a.x)));
  double im = a.y / (2.0 * re);
  if (a.x < 0.0) {
    const double t = fabs(im);
    im = copysign(re, a.y);
    re = t;
  }
  return cmk(re, im);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}

__device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ inline double unit(uint64_t seed, uint64_t idx) {
  return (double)(mix64(seed ^ mix64(idx)) >> 11) * 0x1.0p-53;
}

// Y[e][r][c], e counted from e_base (so chunks draw what one call would).
__global__ void beyn_probe_kernel(double2* __restrict__ Y, int64_t count, uint64_t seed,
                                  int64_t e_base) {
  const int e = blockIdx.y;
  const uint64_t base = (uint64_t)(e_base + e) * (uint64_t)count;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = 2 * (base + (uint64_t)i);
    Y[(int64_t)e * count + i] = cmk(unit(seed, k), unit(seed, k + 1));
  }
}

struct OpArgs {
  const double2* a[kMaxCoef];
  int64_t sA;                            // energy stride of every coefficient block
  double2 zp[kPtsPerLaunch][kMaxCoef];   // z_k^n for the launch's points
  int ncoef, npts, nE;
  int64_t dd;
  double2* out;                          // [point][energy][d*d]
};

__global__ void __launch_bounds__(256) operator_kernel(OpArgs P) {
  const int q = blockIdx.y;  // point * nE + energy
  const int pt = q / P.nE, e = q % P.nE;
  double2* o = P.out + (int64_t)q * P.dd;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.dd;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = cmk(0.0, 0.0);
    for (int n = 0; n < P.ncoef; ++n) v = cadd(v, cmul(P.zp[pt][n], P.a[n][(int64_t)e * P.sA + i]));
    o[i] = v;
  }
}

// One-sided Jacobi SVD of P_0 (d x m, row-major, energy stride d*m); A is a
// column-major scratch copy (m x d per energy).  Outputs, compacted to the
// r singular values above s_max * m * eps in decreasing order (zero beyond r):
// U (d x m, unit columns), V (m x m), s (m), rank[e] = r.
__global__ void __launch_bounds__(256)
jacobi_svd_kernel(const double2* __restrict__ P0, double2* __restrict__ A, int d, int m,
                  double2* __restrict__ U, double2* __restrict__ V, double* __restrict__ s_out,
                  int* __restrict__ rank, int max_sweeps) {
  __shared__ double2 Vl[kMaxModes][kMaxModes + 1];  // column-major: Vl[col][row]
  __shared__ double off_w[4];
  __shared__ double sv[kMaxModes];
  __shared__ int perm[kMaxModes];
  __shared__ int r_sh;
  const int e = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double2* P = P0 + (int64_t)e * d * m;
  double2* Ae = A + (int64_t)e * d * m;
  double fro = 0.0;
  for (int64_t i = tid; i < (int64_t)d * m; i += blockDim.x) {
    const int r = (int)(i / m), c = (int)(i % m);
    const double2 x = P[i];
    Ae[(int64_t)c * d + r] = x;
    fro += x.x * x.x + x.y * x.y;
  }
  for (int i = tid; i < m * m; i += blockDim.x) {
    const int c = i / m, r = i % m;
    Vl[c][r] = cmk(r == c ? 1.0 : 0.0, 0.0);
  }
  fro = wave_sum(fro);
  if (lane == 0) off_w[wave] = fro;
  __syncthreads();
  // columns below eps ||P_0||_F stay below the rank cut (s_max m eps >=
  // sqrt(m) eps ||P_0||_F): they take part in no rotation
  const double floor2 = DBL_EPSILON * DBL_EPSILON * (off_w[0] + off_w[1] + off_w[2] + off_w[3]);
  __syncthreads();
  const int mp = m + (m & 1);
  const double tol = sqrt((double)d) * DBL_EPSILON;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off = 0.0;
    for (int t = 0; t < mp - 1; ++t) {
      for (int k = wave; k < mp / 2; k += 4) {
        const int pa = k, pb = mp - 1 - k;
        const int p = pa == 0 ? 0 : ((pa - 1 + t) % (mp - 1)) + 1;
        const int q = ((pb - 1 + t) % (mp - 1)) + 1;
        if (p >= m || q >= m) continue;
        double2* ap = Ae + (int64_t)p * d;
        double2* aq = Ae + (int64_t)q * d;
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int r = lane; r < d; r += 64) {
          const double2 x = ap[r], y = aq[r];
          al += x.x * x.x + x.y * x.y;
          be += y.x * y.x + y.y * y.y;
          gr += x.x * y.x + x.y * y.y;  // conj(x) y
          gi += x.x * y.y - x.y * y.x;
        }
        al = wave_sum(al);
        be = wave_sum(be);
        gr = wave_sum(gr);
        gi = wave_sum(gi);
        const double g = hypot(gr, gi);
        if (al <= floor2 || be <= floor2 || g == 0.0) continue;
        const double rel = g / (sqrt(al) * sqrt(be));
        off = fmax(off, rel);
        if (rel <= tol) continue;
        // a_q <- e^{-i phi} a_q makes conj(a_p) a_q real (= g); then a real
        // rotation by t = tan(theta) with t^2 g + t (be - al) - g = 0
        const double zeta = (be - al) / (2.0 * g);
        const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + tt * tt), sn = c * tt;
        const double2 ph = cmk(gr / g, -gi / g);  // e^{-i phi}
        for (int r = lane; r < d; r += 64) {
          const double2 x = ap[r];
          const double2 y = cmul(ph, aq[r]);
          ap[r] = cmk(c * x.x - sn * y.x, c * x.y - sn * y.y);
          aq[r] = cmk(sn * x.x + c * y.x, sn * x.y + c * y.y);
        }
        for (int r = lane; r < m; r += 64) {
          const double2 x = Vl[p][r];
          const double2 y = cmul(ph, Vl[q][r]);
          Vl[p][r] = cmk(c * x.x - sn * y.x, c * x.y - sn * y.y);
          Vl[q][r] = cmk(sn * x.x + c * y.x, sn * x.y + c * y.y);
        }
      }
      __syncthreads();
    }
    if (lane == 0) off_w[wave] = off;
    __syncthreads();
    const double offall = fmax(fmax(off_w[0], off_w[1]), fmax(off_w[2], off_w[3]));
    __syncthreads();
    if (offall <= tol) break;
  }
  // singular values (column norms), one wave per column
  for (int c = wave; c < m; c += 4) {
    double n2 = 0.0;
    for (int r = lane; r < d; r += 64) {
      const double2 x = Ae[(int64_t)c * d + r];
      n2 += x.x * x.x + x.y * x.y;
    }
    n2 = wave_sum(n2);
    if (lane == 0) sv[c] = sqrt(n2);
  }
  __syncthreads();
  if (tid == 0) {
    double smax = 0.0;
    for (int c = 0; c < m; ++c) {
      perm[c] = c;
      smax = fmax(smax, sv[c]);
    }
    for (int i = 0; i < m; ++i) {  // selection sort, decreasing
      int best = i;
      for (int j = i + 1; j < m; ++j)
        if (sv[perm[j]] > sv[perm[best]]) best = j;
      const int t = perm[i];
      perm[i] = perm[best];
      perm[best] = t;
    }
    const double cut = smax * m * DBL_EPSILON;
    int r = 0;
    while (r < m && sv[perm[r]] > cut) ++r;
    r_sh = r;
    rank[e] = r;
  }
  __syncthreads();
  const int r = r_sh;
  double2* Ue = U + (int64_t)e * d * m;
  for (int64_t i = tid; i < (int64_t)d * m; i += blockDim.x) {
    const int row = (int)(i / m), k = (int)(i % m);
    double2 v = cmk(0.0, 0.0);
    if (k < r) {
      const int c = perm[k];
      v = cscale(Ae[(int64_t)c * d + row], 1.0 / sv[c]);
    }
    Ue[i] = v;
  }
  double2* Ve = V + (int64_t)e * m * m;
  for (int i = tid; i < m * m; i += blockDim.x) {
    const int row = i / m, k = i % m;
    Ve[i] = k < r ? Vl[perm[k]][row] : cmk(0.0, 0.0);
  }
  for (int k = tid; k < m; k += blockDim.x) s_out[(int64_t)e * m + k] = k < r ? sv[perm[k]] : 0.0;
}

// Givens rotation G = [[c, s], [-conj(s), c]] with G [x; y] = [rho; 0].
__device__ inline void givens(double2 x, double2 y, double* c, double2* s, double2* rho) {
  const double ax = cabs(x), ay = cabs(y);
  if (ay == 0.0) {
    *c = 1.0;
    *s = cmk(0.0, 0.0);
    *rho = x;
    return;
  }
  if (ax == 0.0) {
    *c = 0.0;
    *s = cmk(y.x / ay, -y.y / ay);
    *rho = cmk(ay, 0.0);
    return;
  }
  const double nrm = hypot(ax, ay);
  const double2 ph = cmk(x.x / ax, x.y / ax);
  *c = ax / nrm;
  *s = cscale(cmul(ph, cmk(y.x, -y.y)), 1.0 / nrm);
  *rho = cscale(ph, nrm);
}

// Eigenpairs of the reduced matrix a = (U^H P_1 V)[:r, :r] / s (columns
// scaled by 1/s_j), one wave per energy.  w[e][k] = eigenvalue (0 for
// k >= r), W[e] (m x m, row-major) holds the eigenvectors in its first r
// columns (zero elsewhere); X is scratch (m x m per energy).  A QR iteration
// that does not converge leaves w = 0 and W = 0 for that energy and records
// (tag, energy) in the status word, like a singular pivot.
__global__ void __launch_bounds__(64)
hqr_eig_kernel(const double2* __restrict__ Ared, const double* __restrict__ sv,
               const int* __restrict__ rank, int m, double2* __restrict__ w_out,
               double2* __restrict__ W, double2* __restrict__ X, int* __restrict__ info, int tag,
               int64_t e_base) {
  __shared__ double2 H[kMaxModes][kMaxModes + 1];
  __shared__ double2 Z[kMaxModes][kMaxModes + 1];
  __shared__ double2 vh[kMaxModes];
  const int e = blockIdx.x, t = threadIdx.x;
  const int r = rank[e];
  const double2* Ae = Ared + (int64_t)e * m * m;
  const double* se = sv + (int64_t)e * m;
  for (int i = 0; i < r; ++i) {
    if (t < r) {
      H[i][t] = cscale(Ae[(int64_t)i * m + t], 1.0 / se[t]);
      Z[i][t] = cmk(i == t ? 1.0 : 0.0, 0.0);
    }
  }
  __syncthreads();
  // Householder reduction to upper Hessenberg form, Z accumulating the reflectors
  for (int k = 0; k + 2 < r; ++k) {
    const double2 xk = t > k && t < r ? H[t][k] : cmk(0.0, 0.0);
    const double tail = wave_sum(t > k + 1 && t < r ? xk.x * xk.x + xk.y * xk.y : 0.0);
    if (tail == 0.0) continue;  // uniform: already in Hessenberg form in this column
    const double2 x0 = H[k + 1][k];
    const double alpha = sqrt(tail + x0.x * x0.x + x0.y * x0.y);
    const double ax0 = cabs(x0);
    const double2 ph = ax0 > 0.0 ? cmk(x0.x / ax0, x0.y / ax0) : cmk(1.0, 0.0);
    if (t > k && t < r) vh[t] = t == k + 1 ? cadd(x0, cscale(ph, alpha)) : xk;
    __syncthreads();
    double vn = 0.0;
    for (int i = k + 1; i < r; ++i) vn += vh[i].x * vh[i].x + vh[i].y * vh[i].y;
    const double tau = 2.0 / vn;
    // left: rows k+1.., lane = column j >= k
    if (t >= k && t < r) {
      double2 sj = cmk(0.0, 0.0);
      for (int i = k + 1; i < r; ++i) sj = cadd(sj, cmul(cconj(vh[i]), H[i][t]));
      sj = cscale(sj, tau);
      for (int i = k + 1; i < r; ++i) H[i][t] = csub(H[i][t], cmul(vh[i], sj));
    }
    __syncthreads();
    // right: columns k+1.., lane = row i (H and Z)
    if (t < r) {
      double2 ti = cmk(0.0, 0.0), zi = cmk(0.0, 0.0);
      for (int j = k + 1; j < r; ++j) {
        ti = cadd(ti, cmul(H[t][j], vh[j]));
        zi = cadd(zi, cmul(Z[t][j], vh[j]));
      }
      ti = cscale(ti, tau);
      zi = cscale(zi, tau);
      for (int j = k + 1; j < r; ++j) {
        H[t][j] = csub(H[t][j], cmul(ti, cconj(vh[j])));
        Z[t][j] = csub(Z[t][j], cmul(zi, cconj(vh[j])));
      }
    }
    __syncthreads();
    if (t > k + 1 && t < r) H[t][k] = cmk(0.0, 0.0);
    __syncthreads();
  }
  // complex single-shift QR to the Schur form (full rows and columns updated)
  int ihi = r - 1, iter = 0, total = 0, failed = 0;
  while (ihi >= 0) {
    int l = 0;
    for (int k = ihi; k > 0; --k) {
      const double h = cabs1(H[k][k - 1]);
      if (h <= DBL_EPSILON * (cabs1(H[k - 1][k - 1]) + cabs1(H[k][k])) || h < DBL_MIN) {
        l = k;
        break;
      }
    }
    __syncthreads();
    if (l > 0 && t == 0) H[l][l - 1] = cmk(0.0, 0.0);
    if (l == ihi) {
      --ihi;
      iter = 0;
      __syncthreads();
      continue;
    }
    if (++total > 40 * r) {
      failed = 1;
      break;
    }
    ++iter;
    double2 mu;
    if (iter % 11 == 10) {
      // exceptional shift
      mu = cadd(H[ihi][ihi], cmk(0.75 * fabs(H[ihi][ihi - 1].x), 0.0));
    } else {
      const double2 a = H[ihi - 1][ihi - 1], b = H[ihi - 1][ihi];
      const double2 c = H[ihi][ihi - 1], d = H[ihi][ihi];
      const double2 half = cscale(csub(a, d), 0.5);
      const double2 disc = csqrt_(cadd(cmul(half, half), cmul(b, c)));
      const double2 mean = cscale(cadd(a, d), 0.5);
      const double2 m1 = cadd(mean, disc), m2 = csub(mean, disc);
      mu = cabs1(csub(m1, d)) <= cabs1(csub(m2, d)) ? m1 : m2;
    }
    double2 x = csub(H[l][l], mu), y = H[l + 1][l];
    for (int k = l; k < ihi; ++k) {
      if (k > l) {
        x = H[k][k - 1];
        y = H[k + 1][k - 1];
      }
      double c;
      double2 s, rho;
      givens(x, y, &c, &s, &rho);
      __syncthreads();
      if (k > l && t == 0) {
        H[k][k - 1] = rho;
        H[k + 1][k - 1] = cmk(0.0, 0.0);
      }
      const int j0 = k > l ? k : l;
      if (t >= j0 && t < r) {
        const double2 h1 = H[k][t], h2 = H[k + 1][t];
        H[k][t] = cadd(cscale(h1, c), cmul(s, h2));
        H[k + 1][t] = csub(cscale(h2, c), cmul(cconj(s), h1));
      }
      __syncthreads();
      const int i1 = min(k + 2, ihi);
      if (t <= i1) {
        const double2 h1 = H[t][k], h2 = H[t][k + 1];
        H[t][k] = cadd(cscale(h1, c), cmul(cconj(s), h2));
        H[t][k + 1] = csub(cscale(h2, c), cmul(s, h1));
      }
      if (t < r) {
        const double2 z1 = Z[t][k], z2 = Z[t][k + 1];
        Z[t][k] = cadd(cscale(z1, c), cmul(cconj(s), z2));
        Z[t][k + 1] = csub(cscale(z2, c), cmul(s, z1));
      }
      __syncthreads();
    }
  }
  __syncthreads();
  if (t == 0 && failed && atomicCAS(info, 0, tag) == 0) {
    info[1] = (int)(e_base + e);
    info[2] = -1;
  }
  double2* we = w_out + (int64_t)e * m;
  if (t < m) we[t] = t < r && !failed ? H[t][t] : cmk(0.0, 0.0);
  // eigenvectors of the triangular factor: lane k solves (T - t_kk) x = 0, x_k = 1
  double2* Xe = X + (int64_t)e * m * m;
  if (t < r && !failed) {
    const int k = t;
    const double2 lk = H[k][k];
    const double smin = fmax(DBL_EPSILON * cabs1(lk), DBL_MIN);
    Xe[(int64_t)k * m + k] = cmk(1.0, 0.0);
    for (int i = k - 1; i >= 0; --i) {
      double2 sum = cmk(0.0, 0.0);
      for (int j = i + 1; j <= k; ++j) sum = cadd(sum, cmul(H[i][j], Xe[(int64_t)j * m + k]));
      double2 den = csub(H[i][i], lk);
      if (cabs1(den) < smin) den = cmk(smin, 0.0);
      Xe[(int64_t)i * m + k] = cmk(-cdiv(sum, den).x, -cdiv(sum, den).y);
    }
  }
  // W = Z X (column k by lane k), zero outside the r x r block
  double2* We = W + (int64_t)e * m * m;
  if (t < m) {
    const int k = t;
    for (int i = 0; i < m; ++i) {
      double2 v = cmk(0.0, 0.0);
      if (k < r && i < r && !failed)
        for (int j = 0; j <= k; ++j) v = cadd(v, cmul(Z[i][j], Xe[(int64_t)j * m + k]));
      We[(int64_t)i * m + k] = v;
    }
  }
}

// Normalise the columns of vs (d x m) to unit 2-norm, one wave per (energy, column).
__global__ void __launch_bounds__(64) colnorm_kernel(double2* __restrict__ vs, int d, int m) {
  const int e = blockIdx.y, k = blockIdx.x, lane = threadIdx.x;
  double2* v = vs + (int64_t)e * d * m;
  double n2 = 0.0;
  for (int r = lane; r < d; r += 64) n2 += v[(int64_t)r * m + k].x * v[(int64_t)r * m + k].x +
                                          v[(int64_t)r * m + k].y * v[(int64_t)r * m + k].y;
  n2 = wave_sum(n2);
  if (n2 == 0.0) return;
  const double inv = 1.0 / sqrt(n2);
  for (int r = lane; r < d; r += 64) v[(int64_t)r * m + k] = cscale(v[(int64_t)r * m + k], inv);
}

struct ModeArgs {
  const double2* vs;          // d x m, unit columns
  const double2* Z[3];        // a_ji vs, a_ii vs, a_ij vs
  const double2* w;           // m eigenvalues
  const int* rank;
  int d, m;
  double min_decay, max_decay, min_propagation, residual_tolerance, eta_decay;
  int residual_normalization;
  double2* vrm;               // d x m: reflected modes, zero columns elsewhere
  double2* vrw;               // d x m: reflected modes / w
  double2* Dm;                // m x m: 1 on the diagonal of non-reflected modes (pre-zeroed)
  int* mask;                  // m per energy
  int* count;                 // reflected modes per energy (pre-zeroed)
};

// Reflected-mode mask of spectral.py:221-327 for mode k (blockIdx.x) of
// energy blockIdx.y; polynomial a_ji / w + a_ii + a_ij w.
__global__ void __launch_bounds__(64) mode_kernel(ModeArgs P) {
  const int k = blockIdx.x, e = blockIdx.y, lane = threadIdx.x;
  const int d = P.d, m = P.m;
  const int64_t off = (int64_t)e * d * m;
  const double2 w = P.w[(int64_t)e * m + k];
  bool refl = false;
  if (k < P.rank[e] && (w.x != 0.0 || w.y != 0.0)) {
    const double2 wi = crecip(w);
    double res = 0.0, vn = 0.0, d0r = 0.0, d0i = 0.0, d2r = 0.0, d2i = 0.0;
    for (int r = lane; r < d; r += 64) {
      const int64_t o = off + (int64_t)r * m + k;
      const double2 v = P.vs[o], zj = P.Z[0][o], zi = P.Z[1][o], zu = P.Z[2][o];
      const double2 p = cadd(cadd(cmul(zj, wi), zi), cmul(zu, w));
      res += p.x * p.x + p.y * p.y;
      vn += v.x * v.x + v.y * v.y;
      const double2 a0 = cmul(cconj(v), zj), a2 = cmul(cconj(v), zu);
      d0r += a0.x;
      d0i += a0.y;
      d2r += a2.x;
      d2i += a2.y;
    }
    res = wave_sum(res);
    vn = wave_sum(vn);
    d0r = wave_sum(d0r);
    d0i = wave_sum(d0i);
    d2r = wave_sum(d2r);
    d2i = wave_sum(d2i);
    double resid = sqrt(res) / sqrt(vn);
    if (P.residual_normalization) resid /= cabs(w);
    // dE/dk = -sum_n (i n) (v^H a_n v) w^n over n = -1, 0, 1
    const double2 t0 = cmul(cmul(cmk(0.0, -1.0), cmk(d0r, d0i)), wi);  // n = -1
    const double2 t2 = cmul(cmul(cmk(0.0, 1.0), cmk(d2r, d2i)), w);    // n = +1
    const double2 dE = cmk(-(t0.x + t2.x), -(t0.y + t2.y));
    // k = -i log(w): imaginary part -log|w|
    const double kim = -log(cabs(w));
    const double ratio = fabs(dE.x) / (fabs(dE.y) + DBL_EPSILON);
    const bool prop = fabs(kim) < P.min_decay && P.min_propagation < ratio && dE.x < 0.0;
    bool decay = kim < -P.min_decay;
    decay = decay || (P.min_propagation >= ratio && kim < -P.eta_decay);
    decay = decay && kim > -P.max_decay;
    refl = (prop || decay) && resid < P.residual_tolerance;
    if (resid != resid) refl = false;
  }
  if (lane == 0) {
    P.mask[(int64_t)e * m + k] = refl ? 1 : 0;
    if (refl) atomicAdd(P.count + e, 1);
    P.Dm[(int64_t)e * m * m + (int64_t)k * m + k] = cmk(refl ? 0.0 : 1.0, 0.0);
  }
  const double2 wi = refl ? crecip(w) : cmk(0.0, 0.0);
  for (int r = lane; r < d; r += 64) {
    const int64_t o = off + (int64_t)r * m + k;
    const double2 v = refl ? P.vs[o] : cmk(0.0, 0.0);
    P.vrm[o] = v;
    P.vrw[o] = cmul(v, wi);
  }
}

__global__ void zero_ints_kernel2(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

dim3 grid2(int64_t count, int ny) {
  return dim3((unsigned)std::min<int64_t>(1024, (count + 255) / 256), (unsigned)ny);
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Carve {
  char* base = nullptr;
  size_t bytes = 0;
  template <class T>
  T* take(size_t n) {
    T* p = base ? reinterpret_cast<T*>(base + bytes) : nullptr;
    bytes += al256(n * sizeof(T));
    return p;
  }
};

// Workspace of one energy chunk (E energies, np points per operator chunk).
struct Work {
  double2 *Y, *P0, *P1, *Acol, *U, *V, *T1, *Ared, *W, *X, *w, *vs, *Z[3], *vrm, *vrw, *Dm, *G,
      *Ginv, *W2, *Xij, *S, *Tm, *T, *Tinv;
  double* sv;
  int *rank, *count, *mask;
  char *aux_b, *aux_m, *aux_pts;
};

size_t carve(char* base, int d, int m, int E, int np, Work* w) {
  Carve c{base, 0};
  const size_t dm = (size_t)d * m, mm = (size_t)m * m, dd = (size_t)d * d;
  w->Y = c.take<double2>(dm * E);
  w->P0 = c.take<double2>(dm * E);
  w->P1 = c.take<double2>(dm * E);
  w->Acol = c.take<double2>(dm * E);
  w->U = c.take<double2>(dm * E);
  w->V = c.take<double2>(mm * E);
  w->T1 = c.take<double2>(dm * E);
  w->Ared = c.take<double2>(mm * E);
  w->W = c.take<double2>(mm * E);
  w->X = c.take<double2>(mm * E);
  w->w = c.take<double2>((size_t)m * E);
  w->vs = c.take<double2>(dm * E);
  for (int k = 0; k < 3; ++k) w->Z[k] = c.take<double2>(dm * E);
  w->vrm = c.take<double2>(dm * E);
  w->vrw = c.take<double2>(dm * E);
  w->Dm = c.take<double2>(mm * E);
  w->G = c.take<double2>(mm * E);
  w->Ginv = c.take<double2>(mm * E);
  w->W2 = c.take<double2>(dm * E);
  w->Xij = c.take<double2>(dd * E);
  w->S = c.take<double2>(dd * E);
  w->Tm = c.take<double2>(dd * E);
  w->sv = c.take<double>((size_t)m * E);
  w->rank = c.take<int>((size_t)E);
  w->count = c.take<int>((size_t)E);
  w->mask = c.take<int>((size_t)m * E);
  w->aux_b = c.take<char>(zinv_aux_bytes(d) * (size_t)E);
  w->aux_m = c.take<char>(zinv_aux_bytes(m) * (size_t)E);
  w->T = c.take<double2>(dd * E * np);
  w->Tinv = c.take<double2>(dd * E * np);
  w->aux_pts = c.take<char>(zinv_aux_bytes(d) * (size_t)E * np);
  return c.bytes;
}

int check_opts(rgf_ctx* ctx, int64_t d, int64_t nE, const rgf_beyn_opts* bo, int ncoef) {
  if (d < 1 || d > 1024) return ctx->fail(RGF_EUNSUPPORTED, "beyn: block size must be in [1, 1024]");
  if (nE < 0) return ctx->fail(RGF_EINVAL, "beyn: negative energy count");
  if (!bo) return ctx->fail(RGF_EINVAL, "beyn: options required");
  if (bo->m_0 < 1 || bo->m_0 > kMaxModes)
    return ctx->fail(RGF_EUNSUPPORTED, "beyn: m_0 must be in [1, 64]");
  if (bo->num_quad_points < 1) return ctx->fail(RGF_EINVAL, "beyn: num_quad_points must be >= 1");
  if (!(bo->r_o > 0.0) || !(bo->r_i > 0.0)) return ctx->fail(RGF_EINVAL, "beyn: radii must be positive");
  if (ncoef < 3 || ncoef > kMaxCoef || ncoef % 2 == 0)
    return ctx->fail(RGF_EUNSUPPORTED, "beyn: 3, 5 or 7 coefficient blocks");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  return RGF_OK;
}

// Beyn on one energy chunk: leaves w (m per energy, 0 beyond the rank) and
// vs (d x m, unit columns within the rank, zero beyond) in the workspace.
int beyn_chunk(Ctx* c, int d, int E, int64_t e_glob, const double2* const* coef, int ncoef,
               const rgf_beyn_opts* bo, int np, const Work& w) {
  const int m = bo->m_0;
  const int64_t dd = (int64_t)d * d, dm = (int64_t)d * m, mm = (int64_t)m * m;
  hipStream_t st = c->stream;
  int rc = c->hip(launch_fill_zero(w.P0, dm, dm, E, st), "zero P0");
  if (!rc) rc = c->hip(launch_fill_zero(w.P1, dm, dm, E, st), "zero P1");
  if (rc) return rc;
  hipLaunchKernelGGL(beyn_probe_kernel, grid2(dm, E), dim3(256), 0, st, w.Y, dm, bo->seed, e_glob);
  rc = c->hip(hipGetLastError(), "probe");
  if (rc) return rc;
  // quadrature nodes: outer circle (weight + w), inner circle (weight - w)
  const int nq = bo->num_quad_points;
  const int p = ncoef / 2;
  std::vector<double2> zs;
  std::vector<double> sg;
  for (int circle = 0; circle < 2; ++circle) {
    const double rad = circle == 0 ? bo->r_o : bo->r_i;
    for (int k = 0; k < nq; ++k) {
      const double ang = 2.0 * M_PI * ((k + 1) + 0.5) / nq;
      zs.push_back(double2{rad * std::cos(ang), rad * std::sin(ang)});
      sg.push_back(circle == 0 ? 1.0 : -1.0);
    }
  }
  auto hmul = [](double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; };
  auto hpow = [&](double2 z, int n) {
    double2 r{1.0, 0.0};
    const bool neg = n < 0;
    for (int i = 0; i < (neg ? -n : n); ++i) r = hmul(r, z);
    if (neg) {
      const double den = r.x * r.x + r.y * r.y;
      r = double2{r.x / den, -r.y / den};
    }
    return r;
  };
  const int npts = (int)zs.size();
  const double wq = 1.0 / nq;
  for (int p0 = 0; p0 < npts; p0 += np) {
    const int n = std::min(np, npts - p0);
    for (int l0 = 0; l0 < n; l0 += kPtsPerLaunch) {
      const int nl = std::min(kPtsPerLaunch, n - l0);
      OpArgs A{};
      for (int k = 0; k < ncoef; ++k) A.a[k] = coef[k];
      A.sA = dd;
      for (int i = 0; i < nl; ++i)
        for (int k = 0; k < ncoef; ++k) A.zp[i][k] = hpow(zs[p0 + l0 + i], k - p);
      A.ncoef = ncoef;
      A.npts = nl;
      A.nE = E;
      A.dd = dd;
      A.out = w.T + (int64_t)l0 * E * dd;
      hipLaunchKernelGGL(operator_kernel, grid2(dd, nl * E), dim3(256), 0, st, A);
      rc = c->hip(hipGetLastError(), "operator");
      if (rc) return rc;
    }
    rc = c->inv(w.T, dd, d, n * E, Mat{w.Tinv, dd, d}, w.aux_pts, 901, e_glob);
    if (rc) return rc;
    // P_0 += sum w z T^-1 Y, P_1 += sum w z^2 T^-1 Y, two points per launch
    for (int i = 0; i < n; i += 2) {
      PB b0(d, m), b1(d, m);
      for (int j = i; j < std::min(i + 2, n); ++j) {
        const double2 z = zs[p0 + j];
        const double s = sg[p0 + j] * wq;
        const double2 c0{s * z.x, s * z.y};
        const double2 c1 = hmul(c0, z);
        const Mat Ti{w.Tinv + (int64_t)j * E * dd, dd, d};
        b0.term(Ti, kOpN, Mat{w.Y, dm, m}, kOpN, d, c0);
        b1.term(Ti, kOpN, Mat{w.Y, dm, m}, kOpN, d, c1);
      }
      std::vector<ZProblem> probs = {b0.add(Mat{w.P0, dm, m}, ONE).out(Mat{w.P0, dm, m}).done(),
                                     b1.add(Mat{w.P1, dm, m}, ONE).out(Mat{w.P1, dm, m}).done()};
      rc = c->gemm(probs, E, d);
      if (rc) return rc;
    }
  }
  // SVD projection and the reduced eigenproblem
  hipLaunchKernelGGL(jacobi_svd_kernel, dim3(E), dim3(256), 0, st, w.P0, w.Acol, d, m, w.U, w.V, w.sv,
                     w.rank, 60);
  rc = c->hip(hipGetLastError(), "jacobi svd");
  if (rc) return rc;
  rc = c->gemm({PB(d, m).term(Mat{w.P1, dm, m}, kOpN, Mat{w.V, mm, m}, kOpN, m).out(Mat{w.T1, dm, m}).done()},
               E, d);
  if (!rc)
    rc = c->gemm({PB(m, m).term(Mat{w.U, dm, m}, kOpC, Mat{w.T1, dm, m}, kOpN, d).out(Mat{w.Ared, mm, m}).done()},
                 E, m);
  if (rc) return rc;
  hipLaunchKernelGGL(hqr_eig_kernel, dim3(E), dim3(64), 0, st, w.Ared, w.sv, w.rank, m, w.w, w.W, w.X,
                     c->info_dev, 905, e_glob);
  rc = c->hip(hipGetLastError(), "eig");
  if (rc) return rc;
  rc = c->gemm({PB(d, m).term(Mat{w.U, dm, m}, kOpN, Mat{w.W, mm, m}, kOpN, m).out(Mat{w.vs, dm, m}).done()},
               E, d);
  if (rc) return rc;
  hipLaunchKernelGGL(colnorm_kernel, dim3(m, E), dim3(64), 0, st, w.vs, d, m);
  return c->hip(hipGetLastError(), "normalise modes");
}

int64_t pick_chunk(Ctx* c, int d, int m, int64_t nE, int np, size_t* bytes) {
  Work w;
  const size_t one = carve(nullptr, d, m, 1, np, &w);
  size_t budget = c->ws_owned || !c->ws ? c->ws_budget : c->ws_bytes;
  int64_t E = (int64_t)(budget / std::max<size_t>(one, 1));
  E = std::max<int64_t>(1, std::min<int64_t>(E, nE));
  E = std::min<int64_t>(E, 65535 / np);
  if (c->max_chunk > 0) E = std::min<int64_t>(E, c->max_chunk);
  *bytes = carve(nullptr, d, m, (int)E, np, &w);
  return E;
}

}  // namespace

extern "C" {

int rgf_beyn_z(rgf_ctx* ctx, int64_t d, int64_t n_energies, int32_t n_coef, const void* const* a_xx,
               const rgf_beyn_opts* opts, void* ws, void* vs, int32_t* n_valid) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  int rc = check_opts(ctx, d, n_energies, opts, n_coef);
  if (rc) return rc;
  if (n_energies == 0) return RGF_OK;
  if (!a_xx || !ws || !vs) return ctx->fail(RGF_EINVAL, "rgf_beyn_z: null buffer");
  for (int k = 0; k < n_coef; ++k)
    if (!a_xx[k]) return ctx->fail(RGF_EINVAL, "rgf_beyn_z: null coefficient block");
  const int m = opts->m_0;
  const int np = (int)std::min<int64_t>(2 * opts->num_quad_points, 16);
  size_t bytes = 0;
  const int64_t chunk = pick_chunk(ctx, (int)d, m, n_energies, np, &bytes);
  char* wsp = nullptr;
  rc = ctx->ensure_workspace(bytes, &wsp);
  if (rc) return rc;
  Work w;
  carve(wsp, (int)d, m, (int)chunk, np, &w);
  rc = ctx->reset_info();
  if (rc) return rc;
  const int64_t dd = d * d, dm = d * m;
  for (int64_t e0 = 0; e0 < n_energies; e0 += chunk) {
    const int E = (int)std::min<int64_t>(chunk, n_energies - e0);
    const double2* coef[kMaxCoef];
    for (int k = 0; k < n_coef; ++k) coef[k] = static_cast<const double2*>(a_xx[k]) + e0 * dd;
    rc = beyn_chunk(ctx, (int)d, E, e0, coef, n_coef, opts, np, w);
    if (!rc)
      rc = ctx->hip(hipMemcpyAsync(static_cast<double2*>(ws) + e0 * m, w.w, (size_t)E * m * sizeof(double2),
                                   hipMemcpyDeviceToDevice, ctx->stream), "copy w");
    if (!rc)
      rc = ctx->hip(hipMemcpyAsync(static_cast<double2*>(vs) + e0 * dm, w.vs, (size_t)E * dm * sizeof(double2),
                                   hipMemcpyDeviceToDevice, ctx->stream), "copy vs");
    if (!rc && n_valid)
      rc = ctx->hip(hipMemcpyAsync(n_valid + e0, w.rank, (size_t)E * sizeof(int), hipMemcpyDeviceToDevice,
                                   ctx->stream), "copy rank");
    if (rc) return rc;
  }
  return RGF_OK;
  GUARD_END
}

int rgf_spectral_obc_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, const void* a_ii, const void* a_ij,
                       const void* a_ji, const rgf_beyn_opts* beyn, const rgf_spectral_opts* opts,
                       void* x_ii, void* eigenvalues, int32_t* n_reflected) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  int rc = check_opts(ctx, b, n_energies, beyn, 3);
  if (rc) return rc;
  if (!opts) return ctx->fail(RGF_EINVAL, "rgf_spectral_obc_z: options required");
  if (opts->num_ref_iterations < 1)
    return ctx->fail(RGF_EINVAL, "Number of refinement iterations must be at least 1.");
  if (n_energies == 0) return RGF_OK;
  if (!a_ii || !a_ij || !a_ji || !x_ii) return ctx->fail(RGF_EINVAL, "rgf_spectral_obc_z: null buffer");
  const int d = (int)b, m = beyn->m_0;
  const int np = (int)std::min<int64_t>(2 * beyn->num_quad_points, 16);
  size_t bytes = 0;
  const int64_t chunk = pick_chunk(ctx, d, m, n_energies, np, &bytes);
  char* wsp = nullptr;
  rc = ctx->ensure_workspace(bytes, &wsp);
  if (rc) return rc;
  Work w;
  carve(wsp, d, m, (int)chunk, np, &w);
  rc = ctx->reset_info();
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const int64_t dd = (int64_t)d * d, dm = (int64_t)d * m, mm = (int64_t)m * m;
  for (int64_t e0 = 0; e0 < n_energies; e0 += chunk) {
    const int E = (int)std::min<int64_t>(chunk, n_energies - e0);
    const double2* Aii = static_cast<const double2*>(a_ii) + e0 * dd;
    const double2* Aij = static_cast<const double2*>(a_ij) + e0 * dd;
    const double2* Aji = static_cast<const double2*>(a_ji) + e0 * dd;
    const double2* coef[3] = {Aji, Aii, Aij};
    rc = beyn_chunk(ctx, d, E, e0, coef, 3, beyn, np, w);
    if (rc) return rc;
    // mode products a_x vs, then the reflected-mode mask
    auto cm = [](const double2* p, int64_t s, int ld) { return Mat{const_cast<double2*>(p), s, ld}; };
    rc = ctx->gemm({PB(d, m).term(cm(Aji, dd, d), kOpN, Mat{w.vs, dm, m}, kOpN, d).out(Mat{w.Z[0], dm, m}).done(),
                    PB(d, m).term(cm(Aii, dd, d), kOpN, Mat{w.vs, dm, m}, kOpN, d).out(Mat{w.Z[1], dm, m}).done(),
                    PB(d, m).term(cm(Aij, dd, d), kOpN, Mat{w.vs, dm, m}, kOpN, d).out(Mat{w.Z[2], dm, m}).done()},
                   E, d);
    if (!rc) rc = ctx->hip(launch_fill_zero(w.Dm, mm, mm, E, st), "zero Dm");
    if (rc) return rc;
    int* count = n_reflected ? n_reflected + e0 : w.count;  // scratch when not requested
    hipLaunchKernelGGL(zero_ints_kernel2, dim3(1), dim3(256), 0, st, count, E);
    ModeArgs M{};
    M.vs = w.vs;
    for (int k = 0; k < 3; ++k) M.Z[k] = w.Z[k];
    M.w = w.w;
    M.rank = w.rank;
    M.d = d;
    M.m = m;
    M.min_decay = opts->min_decay;
    M.max_decay = opts->max_decay;
    M.min_propagation = opts->min_propagation;
    M.residual_tolerance = opts->residual_tolerance;
    M.eta_decay = opts->eta_decay;
    M.residual_normalization = opts->residual_normalization;
    M.vrm = w.vrm;
    M.vrw = w.vrw;
    M.Dm = w.Dm;
    M.mask = w.mask;
    M.count = count;
    hipLaunchKernelGGL(mode_kernel, dim3(m, E), dim3(64), 0, st, M);
    rc = ctx->hip(hipGetLastError(), "mode mask");
    if (rc) return rc;
    if (eigenvalues) {
      rc = ctx->hip(hipMemcpyAsync(static_cast<double2*>(eigenvalues) + e0 * m, w.w,
                                   (size_t)E * m * sizeof(double2), hipMemcpyDeviceToDevice, st),
                    "copy eigenvalues");
      if (rc) return rc;
    }
    // x_ii a_ij = (vr / w) inv(vr^H vr) vr^H; non-reflected columns are zero
    // and carry a unit diagonal in the Gram matrix, which leaves the inverse
    // on the reflected block unchanged
    rc = ctx->gemm({PB(m, m).term(Mat{w.vrm, dm, m}, kOpC, Mat{w.vrm, dm, m}, kOpN, d)
                        .add(Mat{w.Dm, mm, m}, ONE).out(Mat{w.G, mm, m}).done()},
                   E, m);
    if (!rc) rc = ctx->inv(w.G, mm, m, E, Mat{w.Ginv, mm, m}, w.aux_m, 902, e0);
    if (!rc)
      rc = ctx->gemm({PB(m, d).term(Mat{w.Ginv, mm, m}, kOpN, Mat{w.vrm, dm, m}, kOpC, m).out(Mat{w.W2, dm, d}).done()},
                     E, m);
    if (!rc)
      rc = ctx->gemm({PB(d, d).term(Mat{w.vrw, dm, m}, kOpN, Mat{w.W2, dm, d}, kOpN, m).out(Mat{w.Xij, dd, d}).done()},
                     E, d);
    // x = inv(a_ii + a_ji x_ii a_ij)
    if (!rc)
      rc = ctx->gemm({PB(d, d).term(cm(Aji, dd, d), kOpN, Mat{w.Xij, dd, d}, kOpN, d)
                          .add(cm(Aii, dd, d), ONE).out(Mat{w.S, dd, d}).done()},
                     E, d);
    double2* X = static_cast<double2*>(x_ii) + e0 * dd;
    if (!rc) rc = ctx->inv(w.S, dd, d, E, Mat{X, dd, d}, w.aux_b, 903, e0);
    // refinement: x = inv(a_ii - (a_ji x) a_ij), num_ref_iterations times
    for (int it = 0; !rc && it < opts->num_ref_iterations; ++it) {
      rc = ctx->gemm({PB(d, d).term(cm(Aji, dd, d), kOpN, Mat{X, dd, d}, kOpN, d).out(Mat{w.Tm, dd, d}).done()},
                     E, d);
      if (!rc)
        rc = ctx->gemm({PB(d, d).term(Mat{w.Tm, dd, d}, kOpN, cm(Aij, dd, d), kOpN, d, MONE)
                            .add(cm(Aii, dd, d), ONE).out(Mat{w.S, dd, d}).done()},
                       E, d);
      if (!rc) rc = ctx->inv(w.S, dd, d, E, Mat{X, dd, d}, w.aux_b, 904, e0);
    }
    if (rc) return rc;
  }
  return RGF_OK;
  GUARD_END
}

}  // extern "C"
