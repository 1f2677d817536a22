// Open boundary conditions on the device: the Sancho-Rubio surface Green's
// function, one fixed-point refinement step with its residuals, and the
// contact self-energies (retarded, lesser, greater), each batched over all
// energies of a call.
//
// Follows src/qttools/boundary_conditions/obc/sancho_rubio.py:33-85 (the
// decimation loop and its stopping rule), obc.py:102-202 (fixed point and
// residuals) and src/quatrex/electron/solver.py:486-495 (contact
// self-energies).  Products run on the grouped FP64 MFMA GEMM, inverses on the
// batched Gauss-Jordan kernel; the kernels here are the O(b^2) reductions and
// the gamma / occupancy scaling.
#include <algorithm>
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
constexpr int kParts = 16;  // partial sums per energy in the norm reductions

// part[(e*kParts + p)*2 + {0,1}] = partial sums over elements p, p+kParts*256, ...
// mode 0: (|x|+|y|)^2 and 0;  mode 1: |x-y|^2 and |x|^2.
__global__ void __launch_bounds__(256)
pair_norm_kernel(const double2* __restrict__ X, const double2* __restrict__ Y, int64_t sE,
                 int64_t count, int mode, double* __restrict__ part) {
  const int e = blockIdx.y;
  const double2* x = X + (int64_t)e * sE;
  const double2* y = Y + (int64_t)e * sE;
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = x[i], b = y[i];
    if (mode == 0) {
      const double t = hypot(a.x, a.y) + hypot(b.x, b.y);
      s0 += t * t;
    } else {
      const double dr = a.x - b.x, di = a.y - b.y;
      s0 += dr * dr + di * di;
      s1 += a.x * a.x + a.y * a.y;
    }
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* o = part + ((int64_t)e * gridDim.x + blockIdx.x) * 2;
    o[0] = red[0][0];
    o[1] = red[1][0];
  }
}

// mode 0: out0[e] = sqrt(sum s0) (the norm of |alpha| + |beta|);
// mode 1: out0[e] = ||x - y||_F (absolute residual), out1[e] = out0[e] / ||x||_F.
__global__ void norm_finish_kernel(const double* __restrict__ part, int nparts, int nE, int mode,
                                   double* __restrict__ out0, double* __restrict__ out1) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  double s0 = 0.0, s1 = 0.0;
  for (int p = 0; p < nparts; ++p) {
    s0 += part[((int64_t)e * nparts + p) * 2];
    s1 += part[((int64_t)e * nparts + p) * 2 + 1];
  }
  const double a = sqrt(s0);
  out0[e] = a;
  if (mode == 1) out1[e] = a / sqrt(s1);
}

// lesser = i f gamma, greater = i (f - 1) gamma with gamma = i (S - S^H):
// both are -c (S - S^H) with c = f resp. f - 1 (the reference's products by i
// are exact rotations, so these are its values bit for bit).  One 32 x 32
// output tile per workgroup; the mirrored input tile is staged in LDS.
__global__ void __launch_bounds__(256)
contact_gamma_kernel(const double2* __restrict__ sig, int64_t sS, const double* __restrict__ occ,
                     int b, double2* __restrict__ lesser, double2* __restrict__ greater,
                     int64_t sO) {
  __shared__ double2 tm[32][33];
  const int T = (b + 31) / 32;
  const int ti = blockIdx.x / T, tj = blockIdx.x % T;
  const int e = blockIdx.y;
  const double2* S = sig + (int64_t)e * sS;
  const int x = threadIdx.x & 31, y0 = threadIdx.x >> 5;
  for (int y = y0; y < 32; y += 8) {
    const int r = tj * 32 + y, c = ti * 32 + x;
    tm[y][x] = (r < b && c < b) ? S[(int64_t)r * b + c] : cmk(0.0, 0.0);
  }
  __syncthreads();
  const double f = occ[e];
  const double g = f - 1.0;
  for (int y = y0; y < 32; y += 8) {
    const int r = ti * 32 + y, c = tj * 32 + x;
    if (r < b && c < b) {
      const double2 s = S[(int64_t)r * b + c];
      const double2 t = tm[x][y];  // S[c][r]
      const double dr = s.x - t.x, di = s.y + t.y;
      const int64_t o = (int64_t)e * sO + (int64_t)r * b + c;
      if (lesser) lesser[o] = cmk(-(f * dr), -(f * di));
      if (greater) greater[o] = cmk(-(g * dr), -(g * di));
    }
  }
}

// out = x + i eta s per element; s == NULL stands for the identity (then
// only the diagonal moves: v.y += eta, which is what adding the product of
// 1j * eta and an identity matrix yields).
__global__ void shift_kernel(const double2* __restrict__ x, const double2* __restrict__ s,
                             double eta, int b, int64_t bb, double2* __restrict__ out) {
  const int e = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bb;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = x[(int64_t)e * bb + i];
    if (s) {
      const double2 w = s[(int64_t)e * bb + i];
      v = cmk(v.x + (0.0 * w.x - eta * w.y), v.y + (0.0 * w.y + eta * w.x));
    } else if (i / b == i % b) {
      v.y += eta;
    }
    out[(int64_t)e * bb + i] = v;
  }
}

// eps <- eps + x and s <- eps (the inversion copy), per element of a chunk
__global__ void eps_merge_kernel(double2* __restrict__ eps, const double2* __restrict__ x,
                                 double2* __restrict__ s, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = cadd(eps[i], x[i]);
    eps[i] = v;
    s[i] = v;
  }
}

hipError_t launch_eps_merge(double2* eps, const double2* x, double2* s, int64_t count,
                            hipStream_t st) {
  if (count == 0) return hipSuccess;
  const int gx = (int)std::min<int64_t>(4096, (count + 255) / 256);
  hipLaunchKernelGGL(eps_merge_kernel, dim3(gx), dim3(256), 0, st, eps, x, s, count);
  return hipGetLastError();
}

hipError_t launch_shift(const double2* x, const double2* s, double eta, int b, int nE,
                        double2* out, hipStream_t st) {
  if (nE == 0) return hipSuccess;
  const int64_t bb = (int64_t)b * b;
  const int gx = (int)std::min<int64_t>(1024, (bb + 255) / 256);
  hipLaunchKernelGGL(shift_kernel, dim3(gx, nE), dim3(256), 0, st, x, s, eta, b, bb, out);
  return hipGetLastError();
}

hipError_t launch_pair_norms(const double2* X, const double2* Y, int64_t sE, int64_t count,
                             int nE, int mode, double* part, double* out0, double* out1,
                             hipStream_t s) {
  if (nE == 0) return hipSuccess;
  hipLaunchKernelGGL(pair_norm_kernel, dim3(kParts, nE), dim3(256), 0, s, X, Y, sE, count, mode,
                     part);
  hipLaunchKernelGGL(norm_finish_kernel, dim3((nE + 127) / 128), dim3(128), 0, s, part, kParts, nE,
                     mode, out0, out1);
  return hipGetLastError();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Carves consecutive 256-byte aligned pieces out of one workspace; `bytes`
// of a layout run with base == nullptr is the exact size to allocate.
struct Carve {
  char* base = nullptr;
  size_t bytes = 0;
  template <class T>
  T* take(size_t n) {
    T* p = base ? reinterpret_cast<T*>(base + bytes) : nullptr;
    bytes += align256(n * sizeof(T));
    return p;
  }
};

// energies per launch: the norm and gamma kernels put energies on grid y
constexpr int64_t kMaxLaunchEnergies = 65535;

int check_square(rgf_ctx* ctx, int64_t b, int64_t nE, const char* who) {
  if (b < 1 || b > 1024) return ctx->fail(RGF_EUNSUPPORTED, std::string(who) + ": b must be in [1, 1024]");
  if (nE < 0) return ctx->fail(RGF_EINVAL, std::string(who) + ": negative energy count");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  return RGF_OK;
}

}  // namespace

extern "C" {

int rgf_sancho_rubio_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, const void* a_ii,
                       const void* a_ij, const void* a_ji, int32_t max_iterations,
                       double convergence_tol, void* x_ii, int32_t* iterations, double* delta) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  int rc = check_square(ctx, b, n_energies, "rgf_sancho_rubio_z");
  if (rc) return rc;
  if (max_iterations < 0) return ctx->fail(RGF_EINVAL, "rgf_sancho_rubio_z: max_iterations < 0");
  if (iterations) *iterations = 0;
  if (delta) *delta = NAN;
  if (n_energies == 0) return RGF_OK;
  if (!a_ii || !a_ij || !a_ji || !x_ii) return ctx->fail(RGF_EINVAL, "rgf_sancho_rubio_z: null buffer");
  const int n = (int)b;
  const int64_t nE = n_energies;
  const int64_t bb = b * b;
  const size_t mat = (size_t)bb * sizeof(double2);
  // resident for all energies: eps, eps_s, the inversion copy of eps, alpha
  // and beta (two generations each), the per-energy deltas and partial sums;
  // per chunk of energies: the inverse, alpha g, beta g, the inversion aux
  struct Layout {
    double2 *Eps, *EpsS, *S, *Al[2], *Be[2], *G, *P1, *P2;
    double *dE, *part;
    char* aux;
  };
  auto layout = [&](char* base, int64_t chunk, Layout* L) {
    Carve c{base, 0};
    L->Eps = c.take<double2>(bb * nE);
    L->EpsS = c.take<double2>(bb * nE);
    L->S = c.take<double2>(bb * nE);
    for (int k = 0; k < 2; ++k) L->Al[k] = c.take<double2>(bb * nE);
    for (int k = 0; k < 2; ++k) L->Be[k] = c.take<double2>(bb * nE);
    L->dE = c.take<double>(nE);
    L->part = c.take<double>((size_t)nE * kParts * 2);
    L->G = c.take<double2>(bb * chunk);
    L->P1 = c.take<double2>(bb * chunk);
    L->P2 = c.take<double2>(bb * chunk);
    L->aux = c.take<char>(qx::zinv_aux_bytes(n) * chunk);
    return c.bytes;
  };
  Layout L;
  const size_t resident = layout(nullptr, 0, &L);
  const size_t per_chunk = layout(nullptr, 1, &L) - resident + 4 * 256;
  size_t budget = ctx->ws_owned || !ctx->ws ? ctx->ws_budget : ctx->ws_bytes;
  int64_t chunk = budget > resident + per_chunk ? (int64_t)((budget - resident) / per_chunk) : 1;
  chunk = std::max<int64_t>(1, std::min<int64_t>({chunk, nE, kMaxLaunchEnergies}));
  if (ctx->max_chunk > 0) chunk = std::min<int64_t>(chunk, ctx->max_chunk);
  char* ws = nullptr;
  rc = ctx->ensure_workspace(layout(nullptr, chunk, &L), &ws);
  if (rc) return rc;
  layout(ws, chunk, &L);
  double2 *Eps = L.Eps, *EpsS = L.EpsS, *S = L.S, *G = L.G, *P1 = L.P1, *P2 = L.P2;
  double2* Al[2] = {L.Al[0], L.Al[1]};
  double2* Be[2] = {L.Be[0], L.Be[1]};
  double *dE = L.dE, *part = L.part;
  char* aux = L.aux;

  hipStream_t st = ctx->stream;
  const size_t all = mat * nE;
  rc = ctx->hip(hipMemcpyAsync(Eps, a_ii, all, hipMemcpyDeviceToDevice, st), "copy a_ii");
  if (!rc) rc = ctx->hip(hipMemcpyAsync(EpsS, a_ii, all, hipMemcpyDeviceToDevice, st), "copy a_ii");
  if (!rc) rc = ctx->hip(hipMemcpyAsync(S, a_ii, all, hipMemcpyDeviceToDevice, st), "copy a_ii");
  if (!rc) rc = ctx->hip(hipMemcpyAsync(Al[0], a_ji, all, hipMemcpyDeviceToDevice, st), "copy a_ji");
  if (!rc) rc = ctx->hip(hipMemcpyAsync(Be[0], a_ij, all, hipMemcpyDeviceToDevice, st), "copy a_ij");
  if (!rc) rc = ctx->reset_info();
  if (rc) return rc;

  std::vector<double> host_delta(nE);
  auto M = [&](double2* p, int64_t e0) { return Mat{p + e0 * bb, bb, n}; };
  int cur = 0;
  int done = 0;
  double dmax = NAN;
  for (int it = 0; it < max_iterations; ++it) {
    const int nxt = cur ^ 1;
    for (int64_t e0 = 0; e0 < nE; e0 += chunk) {
      const int E = (int)std::min<int64_t>(chunk, nE - e0);
      // inverse = inv(epsilon)  (S holds a copy of epsilon; inverted in place)
      rc = ctx->inv(S + e0 * bb, bb, n, E, Mat{G, bb, n}, aux, 3, e0);
      if (rc) return rc;
      // alpha g, beta g
      std::vector<ZProblem> p1 = {
          PB(n, n).term(M(Al[cur], e0), kOpN, Mat{G, bb, n}, kOpN, n).out(Mat{P1, bb, n}).done(),
          PB(n, n).term(M(Be[cur], e0), kOpN, Mat{G, bb, n}, kOpN, n).out(Mat{P2, bb, n}).done()};
      rc = ctx->gemm(p1, E, n);
      if (rc) return rc;
      // epsilon_s -= alpha g beta, with -alpha g beta also written to G (free
      // now: the inverse was last read above); epsilon -= beta g alpha;
      // alpha <- alpha g alpha, beta <- beta g beta
      std::vector<ZProblem> p2 = {
          PB(n, n).term(Mat{P1, bb, n}, kOpN, M(Be[cur], e0), kOpN, n, MONE)
              .partial_here().add(M(EpsS, e0), ONE).out(M(EpsS, e0))
              .out2(kD2Partial, Mat{G, bb, n}).done(),
          PB(n, n).term(Mat{P2, bb, n}, kOpN, M(Al[cur], e0), kOpN, n, MONE)
              .add(M(Eps, e0), ONE).out(M(Eps, e0)).done(),
          PB(n, n).term(Mat{P1, bb, n}, kOpN, M(Al[cur], e0), kOpN, n).out(M(Al[nxt], e0)).done(),
          PB(n, n).term(Mat{P2, bb, n}, kOpN, M(Be[cur], e0), kOpN, n).out(M(Be[nxt], e0)).done()};
      rc = ctx->gemm(p2, E, n);
      if (rc) return rc;
      // epsilon -= alpha g beta; S <- epsilon for the next inverse
      rc = ctx->hip(launch_eps_merge(Eps + e0 * bb, G, S + e0 * bb, (int64_t)E * bb, st), "eps merge");
      if (rc) return rc;
    }
    cur = nxt;
    // delta = max_e ||(|alpha| + |beta|)||_F / 2, read on the host (the stopping rule)
    for (int64_t e0 = 0; e0 < nE; e0 += kMaxLaunchEnergies) {
      const int E = (int)std::min<int64_t>(kMaxLaunchEnergies, nE - e0);
      rc = ctx->hip(launch_pair_norms(Al[cur] + e0 * bb, Be[cur] + e0 * bb, bb, bb, E, 0,
                                      part + e0 * kParts * 2, dE + e0, nullptr, st),
                    "sancho-rubio delta");
      if (rc) return rc;
    }
    rc = ctx->hip(hipMemcpyAsync(host_delta.data(), dE, nE * sizeof(double), hipMemcpyDeviceToHost, st),
                  "read delta");
    if (!rc) rc = ctx->hip(hipStreamSynchronize(st), "sancho-rubio sync");
    if (rc) return rc;
    dmax = -INFINITY;
    for (double d : host_delta) {
      if (std::isnan(d)) { dmax = NAN; break; }
      dmax = std::max(dmax, d);
    }
    dmax /= 2;
    done = it + 1;
    if (dmax < convergence_tol) break;
  }
  // x_ii = inv(epsilon_s)
  rc = ctx->hip(hipMemcpyAsync(S, EpsS, all, hipMemcpyDeviceToDevice, st), "copy eps_s");
  if (rc) return rc;
  double2* X = static_cast<double2*>(x_ii);
  for (int64_t e0 = 0; e0 < nE; e0 += chunk) {
    const int E = (int)std::min<int64_t>(chunk, nE - e0);
    rc = ctx->inv(S + e0 * bb, bb, n, E, Mat{X + e0 * bb, bb, n}, aux, 3, e0);
    if (rc) return rc;
  }
  if (iterations) *iterations = done;
  if (delta) *delta = dmax;
  return RGF_OK;
  GUARD_END
}

int rgf_obc_fixed_point_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, const void* a_ii,
                          const void* a_ij, const void* a_ji, const void* x, void* x_ref,
                          double* abs_residuals, double* rel_residuals) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  int rc = check_square(ctx, b, n_energies, "rgf_obc_fixed_point_z");
  if (rc) return rc;
  if (n_energies == 0) return RGF_OK;
  if (!a_ii || !a_ij || !a_ji || !x || !x_ref)
    return ctx->fail(RGF_EINVAL, "rgf_obc_fixed_point_z: null buffer");
  if ((abs_residuals == nullptr) != (rel_residuals == nullptr))
    return ctx->fail(RGF_EINVAL, "rgf_obc_fixed_point_z: give both residual arrays or neither");
  if (x_ref == x) return ctx->fail(RGF_EINVAL, "rgf_obc_fixed_point_z: x_ref must not alias x");
  const int n = (int)b;
  const int64_t bb = b * b;
  const size_t mat = (size_t)bb * sizeof(double2);
  auto layout = [&](char* base, int64_t chunk, double2** T, double2** S, char** aux, double** part) {
    Carve c{base, 0};
    *T = c.take<double2>(bb * chunk);
    *S = c.take<double2>(bb * chunk);
    *aux = c.take<char>(qx::zinv_aux_bytes(n) * chunk);
    *part = c.take<double>((size_t)chunk * kParts * 2);
    return c.bytes;
  };
  double2 *T = nullptr, *S = nullptr;
  char* aux = nullptr;
  double* part = nullptr;
  const size_t per = layout(nullptr, 1, &T, &S, &aux, &part);
  size_t budget = ctx->ws_owned || !ctx->ws ? ctx->ws_budget : ctx->ws_bytes;
  int64_t chunk = std::max<int64_t>(1, (int64_t)(budget / per));
  chunk = std::min<int64_t>({chunk, n_energies, kMaxLaunchEnergies});
  if (ctx->max_chunk > 0) chunk = std::min<int64_t>(chunk, ctx->max_chunk);
  char* ws = nullptr;
  rc = ctx->ensure_workspace(layout(nullptr, chunk, &T, &S, &aux, &part), &ws);
  if (rc) return rc;
  layout(ws, chunk, &T, &S, &aux, &part);
  rc = ctx->reset_info();
  if (rc) return rc;
  const double2* Aii = static_cast<const double2*>(a_ii);
  const double2* Aij = static_cast<const double2*>(a_ij);
  const double2* Aji = static_cast<const double2*>(a_ji);
  const double2* Xin = static_cast<const double2*>(x);
  double2* Xout = static_cast<double2*>(x_ref);
  auto C = [&](const double2* p, int64_t e0) { return Mat{const_cast<double2*>(p) + e0 * bb, bb, n}; };
  for (int64_t e0 = 0; e0 < n_energies; e0 += chunk) {
    const int E = (int)std::min<int64_t>(chunk, n_energies - e0);
    // x_ref = inv(a_ii - a_ji x a_ij)
    rc = ctx->gemm({PB(n, n).term(C(Aji, e0), kOpN, C(Xin, e0), kOpN, n).out(Mat{T, bb, n}).done()}, E, n);
    if (!rc)
      rc = ctx->gemm({PB(n, n).term(Mat{T, bb, n}, kOpN, C(Aij, e0), kOpN, n, MONE)
                          .add(C(Aii, e0), ONE).out(Mat{S, bb, n}).done()},
                     E, n);
    if (!rc) rc = ctx->inv(S, bb, n, E, Mat{Xout + e0 * bb, bb, n}, aux, 4, e0);
    if (!rc && abs_residuals)
      rc = ctx->hip(launch_pair_norms(Xout + e0 * bb, Xin + e0 * bb, bb, bb, E, 1, part,
                                      abs_residuals + e0, rel_residuals + e0, ctx->stream),
                    "residual norms");
    if (rc) return rc;
  }
  return RGF_OK;
  GUARD_END
}

int rgf_contact_obc_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, const void* m_10,
                      const void* g_00, const void* m_01, const double* occupancies, void* sigma,
                      void* lesser, void* greater) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  int rc = check_square(ctx, b, n_energies, "rgf_contact_obc_z");
  if (rc) return rc;
  if (n_energies == 0) return RGF_OK;
  if (!m_10 || !g_00 || !m_01 || !sigma) return ctx->fail(RGF_EINVAL, "rgf_contact_obc_z: null buffer");
  if ((lesser || greater) && !occupancies)
    return ctx->fail(RGF_EINVAL, "rgf_contact_obc_z: lesser/greater need the occupancies");
  const int n = (int)b;
  const int64_t bb = b * b;
  const size_t mat = (size_t)bb * sizeof(double2);
  size_t budget = ctx->ws_owned || !ctx->ws ? ctx->ws_budget : ctx->ws_bytes;
  int64_t chunk = std::max<int64_t>(1, (int64_t)(budget / (mat + 256)));
  chunk = std::min<int64_t>({chunk, n_energies, kMaxLaunchEnergies});
  if (ctx->max_chunk > 0) chunk = std::min<int64_t>(chunk, ctx->max_chunk);
  char* ws = nullptr;
  rc = ctx->ensure_workspace(align256(mat * (size_t)chunk), &ws);
  if (rc) return rc;
  double2* T = reinterpret_cast<double2*>(ws);
  const double2* M10 = static_cast<const double2*>(m_10);
  const double2* G00 = static_cast<const double2*>(g_00);
  const double2* M01 = static_cast<const double2*>(m_01);
  double2* Sg = static_cast<double2*>(sigma);
  double2* Le = static_cast<double2*>(lesser);
  double2* Gr = static_cast<double2*>(greater);
  auto C = [&](const double2* p, int64_t e0) { return Mat{const_cast<double2*>(p) + e0 * bb, bb, n}; };
  const int Tn = (n + 31) / 32;
  for (int64_t e0 = 0; e0 < n_energies; e0 += chunk) {
    const int E = (int)std::min<int64_t>(chunk, n_energies - e0);
    // sigma = (m_10 g_00) m_01
    rc = ctx->gemm({PB(n, n).term(C(M10, e0), kOpN, C(G00, e0), kOpN, n).out(Mat{T, bb, n}).done()}, E, n);
    if (!rc)
      rc = ctx->gemm({PB(n, n).term(Mat{T, bb, n}, kOpN, C(M01, e0), kOpN, n).out(C(Sg, e0)).done()}, E, n);
    if (rc) return rc;
    if (Le || Gr) {
      hipLaunchKernelGGL(contact_gamma_kernel, dim3(Tn * Tn, E), dim3(256), 0, ctx->stream,
                         Sg + e0 * bb, bb, occupancies + e0, n, Le ? Le + e0 * bb : nullptr,
                         Gr ? Gr + e0 * bb : nullptr, bb);
      rc = ctx->hip(hipGetLastError(), "contact gamma");
      if (rc) return rc;
    }
  }
  return RGF_OK;
  GUARD_END
}

int rgf_boundary_system_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, double eta_obc,
                          const void* m_00, const void* m_01, const void* m_10, const void* s_00,
                          const void* s_01, const void* s_10, void* a_ii, void* a_ij, void* a_ji) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  int rc = check_square(ctx, b, n_energies, "rgf_boundary_system_z");
  if (rc) return rc;
  if (n_energies == 0) return RGF_OK;
  if (!m_00 || !a_ii) return ctx->fail(RGF_EINVAL, "rgf_boundary_system_z: null buffer");
  if ((a_ij && !m_01) || (a_ji && !m_10))
    return ctx->fail(RGF_EINVAL, "rgf_boundary_system_z: an output needs its input block");
  const int n = (int)b;
  const int64_t bb = b * b;
  struct Job { const void* m; const void* s; void* a; bool identity; };
  const Job jobs[3] = {{m_00, s_00, a_ii, true}, {m_01, s_01, a_ij, false}, {m_10, s_10, a_ji, false}};
  for (const Job& j : jobs) {
    if (!j.a) continue;
    for (int64_t e0 = 0; e0 < n_energies; e0 += kMaxLaunchEnergies) {
      const int E = (int)std::min<int64_t>(kMaxLaunchEnergies, n_energies - e0);
      const double2* m = static_cast<const double2*>(j.m) + e0 * bb;
      const double2* sv = j.s ? static_cast<const double2*>(j.s) + e0 * bb : nullptr;
      double2* a = static_cast<double2*>(j.a) + e0 * bb;
      if (!sv && !j.identity) {
        // no overlap: the off-diagonal coupling is copied unchanged
        if (a != m)
          rc = ctx->hip(hipMemcpyAsync(a, m, (size_t)E * bb * sizeof(double2), hipMemcpyDeviceToDevice,
                                       ctx->stream), "copy coupling");
      } else {
        rc = ctx->hip(launch_shift(m, sv, eta_obc, n, E, a, ctx->stream), "boundary shift");
      }
      if (rc) return rc;
    }
  }
  return RGF_OK;
  GUARD_END
}

}  // extern "C"
