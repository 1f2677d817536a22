// Tridiagonal block densify / sparsify between block-sparse storage
// (DSDBCOO / DSDBCSR: data[e][nnz]) and the packed diag/upper/lower layout.
//
// The host builds, once per sparsity structure, a gather map per part (diag,
// upper, lower): entry k moves data[e][src[k]] to part[e][dst[k]] with the
// op flags below.  Every destination appears at most once, so the kernels
// need no atomics; the entries of a block are consecutive and in row-major
// order (block-sorted storage), so reads are coalesced and writes mostly so.
#include "qx_common.h"

namespace qx {

namespace {

// op flags: kMirror = the entry is the symmetric image of a stored one
// (op(v), densify only); kDiag = a stored diagonal element of a symmetric
// diagonal block, which densifies to (v + op(v)) / 2; op(v) = s * conj(v),
// or s * v with kNoConj, and s = -1 with kNegate.
constexpr uint8_t kMirror = 1, kDiagSym = 2, kNegate = 4, kNoConj = 8;

__device__ inline double2 sym_op(double2 v, uint8_t op) {
  double2 w = (op & kNoConj) ? v : cmk(v.x, -v.y);
  return (op & kNegate) ? cmk(-w.x, -w.y) : w;
}

__global__ void __launch_bounds__(256)
fill_zero_kernel(double2* p, int64_t count, int64_t stride) {
  const int e = blockIdx.y;
  double2* q = p + (int64_t)e * stride;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    q[i] = cmk(0.0, 0.0);
}

// Energies per thread: a map entry is loaded once and applied to kEPT
// energies (the map is shared by all energies of the stack).
constexpr int kEPT = 8;

__global__ void __launch_bounds__(256)
densify_kernel(const double2* __restrict__ data, int64_t sData, int64_t n,
               const int64_t* __restrict__ src, const int64_t* __restrict__ dst,
               const uint8_t* __restrict__ op, double2* __restrict__ out, int64_t sOut, int nE) {
  const int e0 = blockIdx.y * kEPT;
  const int ne = (nE - e0 < kEPT) ? (nE - e0) : kEPT;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = src[k], d = dst[k];
    const uint8_t f = op[k];
    for (int u = 0; u < ne; ++u) {
      const double2 v = data[(int64_t)(e0 + u) * sData + s];
      double2 w = v;
      if (f & kMirror) {
        w = sym_op(v, f);
      } else if (f & kDiagSym) {
        const double2 t = sym_op(v, f);
        w = cmk(0.5 * (v.x + t.x), 0.5 * (v.y + t.y));
      }
      out[(int64_t)(e0 + u) * sOut + d] = w;
    }
  }
}

__global__ void __launch_bounds__(256)
sparsify_kernel(const double2* __restrict__ in, int64_t sIn, int64_t n,
                const int64_t* __restrict__ src, const int64_t* __restrict__ dst,
                const uint8_t* __restrict__ op, double2* __restrict__ data, int64_t sData, int nE) {
  const int e0 = blockIdx.y * kEPT;
  const int ne = (nE - e0 < kEPT) ? (nE - e0) : kEPT;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (op[k] & kMirror) continue;
    const int64_t s = src[k], d = dst[k];
    for (int u = 0; u < ne; ++u)
      data[(int64_t)(e0 + u) * sData + s] = in[(int64_t)(e0 + u) * sIn + d];
  }
}

dim3 grid_for(int64_t n, int ny) {
  int64_t gx = (n + 255) / 256;
  if (gx > 4096) gx = 4096;
  if (gx < 1) gx = 1;
  return dim3((unsigned)gx, (unsigned)ny);
}

}  // namespace

hipError_t launch_fill_zero(double2* p, int64_t count, int64_t stride, int nE, hipStream_t s) {
  if (count <= 0 || nE <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_zero_kernel, grid_for(count, nE), dim3(256), 0, s, p, count, stride);
  return hipGetLastError();
}

hipError_t launch_densify(const double2* data, int64_t sData, int64_t n, const int64_t* src,
                          const int64_t* dst, const uint8_t* op, double2* out, int64_t sOut,
                          int nE, hipStream_t s) {
  if (n <= 0 || nE <= 0) return hipSuccess;
  hipLaunchKernelGGL(densify_kernel, grid_for(n, (nE + kEPT - 1) / kEPT), dim3(256), 0, s, data, sData,
                     n, src, dst, op, out, sOut, nE);
  return hipGetLastError();
}

hipError_t launch_sparsify(const double2* in, int64_t sIn, int64_t n, const int64_t* src,
                           const int64_t* dst, const uint8_t* op, double2* data, int64_t sData,
                           int nE, hipStream_t s) {
  if (n <= 0 || nE <= 0) return hipSuccess;
  hipLaunchKernelGGL(sparsify_kernel, grid_for(n, (nE + kEPT - 1) / kEPT), dim3(256), 0, s, in, sIn,
                     n, src, dst, op, data, sData, nE);
  return hipGetLastError();
}

}  // namespace qx
