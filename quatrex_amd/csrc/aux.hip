// Elementwise and reduction kernels around the RGF sweeps.
#include <cstring>

#include "qx_common.h"

namespace qx {

namespace {

// X <- (X - X^H)/2 in place, for nblk diagonal blocks (offset offs[i], size
// sizes[i]) of each energy.  A workgroup owns a pair of mirrored 32 x 32 tiles
// (ti <= tj): both are read row-coalesced into LDS, combined, and written
// back row-coalesced.
__global__ void __launch_bounds__(256)
project_kernel(double2* base, int64_t sE, const int64_t* offs, const int* sizes) {
  __shared__ double2 ta[32][33];
  __shared__ double2 tb[32][33];
  const int blk = blockIdx.y, e = blockIdx.z;
  const int b = sizes[blk];
  const int T = (b + 31) / 32;
  int p = blockIdx.x, ti = 0;
  while (ti < T && p >= T - ti) { p -= T - ti; ++ti; }
  if (ti >= T) return;
  const int tj = ti + p;
  double2* X = base + (int64_t)e * sE + offs[blk];
  const int x = threadIdx.x & 31, y0 = threadIdx.x >> 5;
  for (int y = y0; y < 32; y += 8) {
    const int r = ti * 32 + y, c = tj * 32 + x;
    ta[y][x] = (r < b && c < b) ? X[(int64_t)r * b + c] : cmk(0.0, 0.0);
    const int r2 = tj * 32 + y, c2 = ti * 32 + x;
    tb[y][x] = (r2 < b && c2 < b) ? X[(int64_t)r2 * b + c2] : cmk(0.0, 0.0);
  }
  __syncthreads();
  for (int y = y0; y < 32; y += 8) {
    const int r = ti * 32 + y, c = tj * 32 + x;
    if (r < b && c < b) {
      const double2 u = ta[y][x], l = tb[x][y];
      X[(int64_t)r * b + c] = cmk(0.5 * (u.x - l.x), 0.5 * (u.y + l.y));
    }
    if (ti != tj) {
      const int r2 = tj * 32 + y, c2 = ti * 32 + x;
      if (r2 < b && c2 < b) {
        const double2 u = tb[y][x], l = ta[x][y];
        X[(int64_t)r2 * b + c2] = cmk(0.5 * (u.x - l.x), 0.5 * (u.y + l.y));
      }
    }
  }
}

// out[e*sOut + slot] = sign * (tr(X1 Y1) - tr(X2 Y2)),  tr(XY) = sum_rc X[r][c] Y[c][r].
__global__ void __launch_bounds__(256)
trace_kernel(const TraceJob* jobs, double2* out, int64_t sOut) {
  const TraceJob J = jobs[blockIdx.x];
  const int e = blockIdx.y;
  const int b = J.b;
  const double2* X1 = J.X1 + (int64_t)e * J.sX1;
  const double2* Y1 = J.Y1 + (int64_t)e * J.sY1;
  const double2* X2 = J.X2 + (int64_t)e * J.sX2;
  const double2* Y2 = J.Y2 + (int64_t)e * J.sY2;
  double sr = 0.0, si = 0.0;
  for (int idx = threadIdx.x; idx < b * b; idx += blockDim.x) {
    const int r = idx / b, c = idx % b;
    const double2 p = cmul(X1[idx], Y1[(int64_t)c * b + r]);
    const double2 q = cmul(X2[idx], Y2[(int64_t)c * b + r]);
    sr += p.x - q.x;
    si += p.y - q.y;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(int64_t)e * sOut + J.slot] = cmk(J.sign * red[0][0], J.sign * red[1][0]);
}

// out[e*sOut + j] = ||X_j||_F for a list of matrices (rows x cols, ld).
__global__ void __launch_bounds__(256)
norm_kernel(const NormJob* jobs, double* out, int64_t sOut) {
  const NormJob J = jobs[blockIdx.x];
  const int e = blockIdx.y;
  const double2* X = J.X + (int64_t)e * J.sX;
  double acc = 0.0;
  for (int idx = threadIdx.x; idx < J.rows * J.cols; idx += blockDim.x) {
    const double2 v = X[(int64_t)(idx / J.cols) * J.ld + idx % J.cols];
    acc += v.x * v.x + v.y * v.y;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(int64_t)e * sOut + blockIdx.x] = sqrt(red[0]);
}

// One workgroup copies a by-value payload (kernel argument) to device memory.
__global__ void __launch_bounds__(256) store_words_kernel(uint64_t* dst, WordPayload p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = p.w[i];
}

__global__ void identity_kernel(double2* X, int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n * n) X[idx] = cmk((idx / n) == (idx % n) ? 1.0 : 0.0, 0.0);
}

// A = -H for every energy (one copy of H, energy stride 0 on input).
__global__ void neg_copy_kernel(const double2* __restrict__ h, double2* __restrict__ a, int64_t n,
                                int64_t sA) {
  const int e = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = h[i];
    a[(int64_t)e * sA + i] = cmk(-v.x, -v.y);
  }
}

// diag blocks: A_ii += z_e I - S_i  (S_i optional b_i x b_i, shared by energies)
__global__ void shift_diag_kernel(double2* a, int64_t sA, const double2* z, const int64_t* offs,
                                  const int* sizes, const double2* const* sig) {
  const int blk = blockIdx.y, e = blockIdx.z;
  const int b = sizes[blk];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= b * b) return;
  const int r = idx / b, c = idx % b;
  double2* A = a + (int64_t)e * sA + offs[blk];
  double2 v = A[idx];
  if (r == c) v = cadd(v, z[e]);
  const double2* S = sig[blk];
  if (S) v = csub(v, S[idx]);
  A[idx] = v;
}

__global__ void z2c_kernel(const double2* src, int64_t sSrc, float2* dst, int64_t sDst,
                           int64_t count) {
  const int e = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = src[(int64_t)e * sSrc + i];
    dst[(int64_t)e * sDst + i] = make_float2((float)v.x, (float)v.y);
  }
}

__global__ void c2z_kernel(const float2* src, int64_t sSrc, double2* dst, int64_t sDst,
                           int64_t count) {
  const int e = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = src[(int64_t)e * sSrc + i];
    dst[(int64_t)e * sDst + i] = cmk((double)v.x, (double)v.y);
  }
}

}  // namespace

hipError_t launch_project(double2* base, int64_t sE, const int64_t* offs_dev, const int* sizes_dev,
                          int nblk, int bmax, int nE, hipStream_t stream) {
  if (nblk == 0 || nE == 0) return hipSuccess;
  const int T = (bmax + 31) / 32;
  dim3 grid(T * (T + 1) / 2, nblk, nE);
  hipLaunchKernelGGL(project_kernel, grid, dim3(256), 0, stream, base, sE, offs_dev, sizes_dev);
  return hipGetLastError();
}

hipError_t launch_traces(const TraceJob* jobs_dev, int njobs, int nE, double2* out, int64_t sOut,
                         hipStream_t stream) {
  if (njobs == 0 || nE == 0) return hipSuccess;
  hipLaunchKernelGGL(trace_kernel, dim3(njobs, nE), dim3(256), 0, stream, jobs_dev, out, sOut);
  return hipGetLastError();
}

hipError_t launch_norms(const NormJob* jobs_dev, int njobs, int nE, double* out, int64_t sOut,
                        hipStream_t stream) {
  if (njobs == 0 || nE == 0) return hipSuccess;
  hipLaunchKernelGGL(norm_kernel, dim3(njobs, nE), dim3(256), 0, stream, jobs_dev, out, sOut);
  return hipGetLastError();
}

hipError_t launch_store_bytes(void* dst, const void* src, size_t bytes, hipStream_t stream) {
  const size_t words = (bytes + 7) / 8;
  const char* in = static_cast<const char*>(src);
  uint64_t* out = static_cast<uint64_t*>(dst);
  for (size_t w0 = 0; w0 < words; w0 += kPayloadWords) {
    const size_t n = words - w0 < (size_t)kPayloadWords ? words - w0 : (size_t)kPayloadWords;
    WordPayload p;
    const size_t off = w0 * 8;
    const size_t take = bytes - off < n * 8 ? bytes - off : n * 8;
    std::memcpy(p.w, in + off, take);
    if (take < n * 8) std::memset(reinterpret_cast<char*>(p.w) + take, 0, n * 8 - take);
    hipLaunchKernelGGL(store_words_kernel, dim3(1), dim3(256), 0, stream, out + w0, p, (int)n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_identity(double2* X, int n, hipStream_t stream) {
  hipLaunchKernelGGL(identity_kernel, dim3((n * n + 255) / 256), dim3(256), 0, stream, X, n);
  return hipGetLastError();
}

hipError_t launch_assemble(const double2* h_diag, const double2* h_upper, const double2* h_lower,
                           int64_t Dsz, int64_t Usz, int nE, const double2* z, const int64_t* offs,
                           const int* sizes, const double2* const* sig, int nblk, int bmax,
                           double2* a_diag, double2* a_upper, double2* a_lower, hipStream_t stream) {
  if (nE == 0) return hipSuccess;
  auto grid = [](int64_t n, int ne) {
    int64_t gx = (n + 255) / 256;
    return dim3((unsigned)(gx > 2048 ? 2048 : gx), (unsigned)ne);
  };
  hipLaunchKernelGGL(neg_copy_kernel, grid(Dsz, nE), dim3(256), 0, stream, h_diag, a_diag, Dsz, Dsz);
  if (Usz > 0) {
    hipLaunchKernelGGL(neg_copy_kernel, grid(Usz, nE), dim3(256), 0, stream, h_upper, a_upper, Usz, Usz);
    hipLaunchKernelGGL(neg_copy_kernel, grid(Usz, nE), dim3(256), 0, stream, h_lower, a_lower, Usz, Usz);
  }
  dim3 g2((bmax * bmax + 255) / 256, nblk, nE);
  hipLaunchKernelGGL(shift_diag_kernel, g2, dim3(256), 0, stream, a_diag, Dsz, z, offs, sizes, sig);
  return hipGetLastError();
}

hipError_t launch_z2c(const double2* src, int64_t sSrc, float2* dst, int64_t sDst, int64_t count,
                      int nE, hipStream_t stream) {
  if (count == 0 || nE == 0) return hipSuccess;
  int gx = (int)((count + 255) / 256);
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(z2c_kernel, dim3(gx, nE), dim3(256), 0, stream, src, sSrc, dst, sDst, count);
  return hipGetLastError();
}

hipError_t launch_c2z(const float2* src, int64_t sSrc, double2* dst, int64_t sDst, int64_t count,
                      int nE, hipStream_t stream) {
  if (count == 0 || nE == 0) return hipSuccess;
  int gx = (int)((count + 255) / 256);
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(c2z_kernel, dim3(gx, nE), dim3(256), 0, stream, src, sSrc, dst, sDst, count);
  return hipGetLastError();
}

}  // namespace qx
