// System-matrix assembly of the electron solver on the tridiagonal blocks,
// per energy e with z_e = E_e + i eta:
//   M = z_e S - P - (Sigma^R_herm + (Sigma^> - Sigma^<) / 2 + H)
// where S is the overlap (identity when absent) and the potential term P is
// S_ab (v_a + v_b) / 2 with an overlap, diag(v) without one.  Follows
// src/quatrex/electron/solver.py:668-709 (_assemble_system_matrix) with
// _add_overlap (:541-565), _apply_potential (:567-609) and
// _subtract_hamiltonian_and_self_energy (:611-665), in that operation order.
//
// One thread per element of one block for kEPT energies: grid (elements of
// the largest block, block, energy group); the block table holds the packed
// offset, the shape and the global row / column offsets of each of the 3N - 2
// tridiagonal blocks.
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/rgf.h"
#include "qx_common.h"
#include "rgf_driver.h"

using namespace qx;

namespace {

struct BlkInfo {
  int64_t off;     // packed element offset within its region (diag / upper / lower)
  int32_t rows, cols;
  int32_t row0, col0;  // global row / column of element (0, 0)
  int32_t region;      // 0 diag, 1 upper, 2 lower
  int32_t pad_;
};

struct AsmArgs {
  const BlkInfo* blk;
  int64_t rsz[3];                       // region sizes (energy strides of the packed outputs)
  const double2* z;
  const double2* h[3];
  int64_t h_stride[3];                  // 0: one H for every energy
  const double2* s[3];                  // overlap, or null (identity)
  const double* v;                      // potential, or null
  const double2* sr[3];                 // Sigma^R_herm, or null
  const double2* sl[3];                 // Sigma^<, or null
  const double2* sg[3];                 // Sigma^>, or null
  double2* a[3];
};

__device__ inline double2 ld(const double2* p, int64_t i) { return p ? p[i] : cmk(0.0, 0.0); }

// Energies per thread: the element's index math and its shared inputs
// (H when shared, overlap, potential) are loaded once for kEPT energies.
constexpr int kEPT = 8;

__global__ void __launch_bounds__(256) assemble_system_kernel(AsmArgs P, int n_energies) {
  const BlkInfo B = P.blk[blockIdx.y];
  const int e0 = blockIdx.z * kEPT;
  const int ne = min(kEPT, n_energies - e0);
  const int n = B.rows * B.cols;  // at most 1024 x 1024
  const int g = B.region;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int r = k / B.cols, c = k - r * B.cols;
    const int64_t i = B.off + k;
    const int ra = B.row0 + r, cb = B.col0 + c;
    // the energy-independent parts of z S - P
    const bool ovl = P.s[g] != nullptr;
    const double2 s = ovl ? P.s[g][i] : cmk(0.0, 0.0);
    const double pv = (!ovl && P.v && ra == cb) ? P.v[ra] : 0.0;  // diag(v) without an overlap
    const double2 hs = P.h_stride[g] ? cmk(0.0, 0.0) : P.h[g][i];
    for (int u = 0; u < ne; ++u) {
      const int e = e0 + u;
      const int64_t se = (int64_t)e * P.rsz[g];
      const double2 z = P.z[e];
      double2 m;
      if (ovl) {
        m = cmul(z, s);
        if (P.v) {
          const double2 p1 = cmk(s.x * P.v[ra], s.y * P.v[ra]);
          const double2 p2 = cmk(s.x * P.v[cb], s.y * P.v[cb]);
          m = csub(m, cmk(0.5 * (p1.x + p2.x), 0.5 * (p1.y + p2.y)));
        }
      } else {
        m = (ra == cb) ? z : cmk(0.0, 0.0);
        if (ra == cb) m = cmk(m.x - pv, m.y);
      }
      // - (Sigma^R_herm + (Sigma^> - Sigma^<) / 2 + H)
      const double2 dl = csub(ld(P.sg[g], se + i), ld(P.sl[g], se + i));
      double2 t = cadd(ld(P.sr[g], se + i), cmk(0.5 * dl.x, 0.5 * dl.y));
      t = cadd(t, P.h_stride[g] ? P.h[g][(int64_t)e * P.h_stride[g] + i] : hs);
      P.a[g][se + i] = csub(m, t);
    }
  }
}

}  // namespace

extern "C" int rgf_assemble_system_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                                     const rgf_assembly* in, rgf_bt* a) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  qx::Geo g;
  int rc = g.init(geom, &ctx->last_error);
  if (rc) return rc;
  if (n_energies < 0) return ctx->fail(RGF_EINVAL, "n_energies < 0");
  if (n_energies == 0) return RGF_OK;
  if (n_energies > 65535) return ctx->fail(RGF_EUNSUPPORTED, "assemble: at most 65535 energies per call");
  const bool multi = g.N > 1;
  if (!in || !in->z || !in->hamiltonian.diag || !a || !a->diag ||
      (multi && (!in->hamiltonian.upper || !in->hamiltonian.lower || !a->upper || !a->lower)))
    return ctx->fail(RGF_EINVAL, "assemble: z, H and the output buffers are required");
  auto full = [&](const rgf_bt_const* t) {
    return !t || (t->diag && (!multi || (t->upper && t->lower)));
  };
  if (!full(in->overlap) || !full(in->sigma_retarded_herm) || !full(in->sigma_lesser) ||
      !full(in->sigma_greater))
    return ctx->fail(RGF_EINVAL, "assemble: overlap and self-energies need diag, upper and lower");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  // block table: diagonal blocks, then upper, then lower
  std::vector<BlkInfo> blk;
  std::vector<int64_t> row(g.N + 1, 0);
  for (int i = 0; i < g.N; ++i) row[i + 1] = row[i] + g.bs[i];
  if (row[g.N] > 0x7fffffff) return ctx->fail(RGF_EUNSUPPORTED, "assemble: matrix too large");
  for (int i = 0; i < g.N; ++i)
    blk.push_back(BlkInfo{g.od[i], (int)g.bs[i], (int)g.bs[i], (int)row[i], (int)row[i], 0, 0});
  for (int i = 0; i + 1 < g.N; ++i) {
    blk.push_back(BlkInfo{g.ou[i], (int)g.bs[i], (int)g.bs[i + 1], (int)row[i], (int)row[i + 1], 1, 0});
    blk.push_back(BlkInfo{g.ou[i], (int)g.bs[i + 1], (int)g.bs[i], (int)row[i + 1], (int)row[i], 2, 0});
  }
  if (!ctx->tables_fixed) {
    rc = ctx->ensure_meta(blk.size() * sizeof(BlkInfo) + 256);
    if (rc) return rc;
    ctx->tables = static_cast<char*>(ctx->meta);
    ctx->tables_cap = ctx->meta_bytes;
  }
  ctx->tables_used = 0;
  void* dev = nullptr;
  rc = ctx->put_table(blk.data(), blk.size() * sizeof(BlkInfo), &dev);
  if (rc) return rc;
  AsmArgs P{};
  P.blk = static_cast<const BlkInfo*>(dev);
  P.rsz[0] = g.Dsz;
  P.rsz[1] = P.rsz[2] = g.Usz;
  P.z = static_cast<const double2*>(in->z);
  auto parts = [](const rgf_bt_const* t, const double2* (&out)[3]) {
    out[0] = t ? static_cast<const double2*>(t->diag) : nullptr;
    out[1] = t ? static_cast<const double2*>(t->upper) : nullptr;
    out[2] = t ? static_cast<const double2*>(t->lower) : nullptr;
  };
  parts(&in->hamiltonian, P.h);
  for (int k = 0; k < 3; ++k) P.h_stride[k] = in->hamiltonian_per_energy ? P.rsz[k] : 0;
  parts(in->overlap, P.s);
  P.v = in->potential;
  parts(in->sigma_retarded_herm, P.sr);
  parts(in->sigma_lesser, P.sl);
  parts(in->sigma_greater, P.sg);
  P.a[0] = static_cast<double2*>(a->diag);
  P.a[1] = static_cast<double2*>(a->upper);
  P.a[2] = static_cast<double2*>(a->lower);
  const int64_t bmax2 = g.bmax * g.bmax;
  const dim3 grid((unsigned)std::min<int64_t>(64, (bmax2 + 255) / 256), (unsigned)blk.size(),
                  (unsigned)((n_energies + kEPT - 1) / kEPT));
  hipLaunchKernelGGL(assemble_system_kernel, grid, dim3(256), 0, ctx->stream, P, (int)n_energies);
  return ctx->hip(hipGetLastError(), "assemble system");
  GUARD_END
}
