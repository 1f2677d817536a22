// Batched complex FP64 matrix inverse: blocked in-place Gauss-Jordan
// elimination with partial (row) pivoting.
//
// For each panel of NB columns [k, k+nb):
//   1. zinv_panel_kernel (one workgroup per energy, one row per thread, the
//      panel row held in registers): unblocked GJE on the panel columns with
//      pivot search over rows >= c by wavefront shuffle max-reduction
//      (|re|+|im|, first index on ties, as LAPACK izamax).  Writes the
//      transformed panel T, Ahat = T - I_panel, the row-permutation record and
//      the list of (destination, source) rows moved by the panel's swaps.
//   2. zinv_swap_kernel: applies those row moves to the non-panel columns and
//      copies the permuted panel rows to Wp (nb x n).
//   3. Trailing update on MFMA (grouped GEMM, two column ranges):
//      S[:, nonpanel] += Ahat * Wp[:, nonpanel].
// After the last panel S holds (P A)^{-1}; zinv_finish_kernel writes
// A^{-1}[:, j] = S[:, rpinv[j]] to the output, where rp is the accumulated row
// permutation (row i of P A is row rp[i] of A).
#include "qx_common.h"

namespace qx {

namespace {

template <int NB>
struct PanelAux {
  // byte offsets inside one energy's aux area
  __host__ __device__ static size_t ahat(int) { return 0; }
  __host__ __device__ static size_t wp(int n) { return (size_t)n * NB * sizeof(double2); }
  __host__ __device__ static size_t rp(int n) { return wp(n) + (size_t)n * NB * sizeof(double2); }
  __host__ __device__ static size_t moves(int n) { return rp(n) + (size_t)n * sizeof(int); }
  __host__ __device__ static size_t bytes(int n) {
    size_t b = moves(n) + 4 * NB * sizeof(int);
    return (b + 255) & ~(size_t)255;
  }
};

inline int panel_nb(int n) { return n <= 512 ? 32 : 16; }

template <int NB, int MAXT>
__global__ void __launch_bounds__(MAXT)
zinv_panel_kernel(double2* S, int64_t sS, int n, int k, int nbe, char* aux, int64_t sAux,
                  int* info, int info_tag, int64_t e0) {
  const int e = blockIdx.x;
  double2* A = S + (int64_t)e * sS;
  char* ax = aux + (int64_t)e * sAux;
  double2* Ahat = reinterpret_cast<double2*>(ax + PanelAux<NB>::ahat(n));
  int* rp = reinterpret_cast<int*>(ax + PanelAux<NB>::rp(n));
  int* moves = reinterpret_cast<int*>(ax + PanelAux<NB>::moves(n));

  __shared__ double red_key[16];
  __shared__ int red_idx[16];
  __shared__ double2 piv_row[NB];
  __shared__ double2 swp[2][NB];
  __shared__ int s_piv[NB];
  __shared__ int src_of[1024];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  const int r = tid;
  const bool valid = r < n;
  const double2 zero = cmk(0.0, 0.0);

  double re[NB], im[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    double2 v = zero;
    if (valid && q < nbe) v = A[(int64_t)r * n + k + q];
    re[q] = v.x;
    im[q] = v.y;
  }

  // The column loop stays rolled; entry jj of the row is read and written
  // through unrolled selects so that the row lives in registers.
#pragma unroll 1
  for (int jj = 0; jj < nbe; ++jj) {
    const int c = k + jj;
    double pr = 0.0, pi = 0.0;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      pr = (q == jj) ? re[q] : pr;
      pi = (q == jj) ? im[q] : pi;
    }
    double key = (valid && r >= c) ? fabs(pr) + fabs(pi) : -1.0;
    int idx = r;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ok = __shfl_xor(key, off);
      const int oi = __shfl_xor(idx, off);
      if (ok > key || (ok == key && oi < idx)) { key = ok; idx = oi; }
    }
    if (lane == 0) { red_key[wave] = key; red_idx[wave] = idx; }
    __syncthreads();
    double bk = red_key[0];
    int p = red_idx[0];
    for (int w = 1; w < nw; ++w) {
      const double ok = red_key[w];
      const int oi = red_idx[w];
      if (ok > bk || (ok == bk && oi < p)) { bk = ok; p = oi; }
    }
    if (p != c) {  // uniform
      if (r == p) {
#pragma unroll
        for (int q = 0; q < NB; ++q) swp[0][q] = cmk(re[q], im[q]);
      }
      if (r == c) {
#pragma unroll
        for (int q = 0; q < NB; ++q) swp[1][q] = cmk(re[q], im[q]);
      }
      __syncthreads();
      if (r == p) {
#pragma unroll
        for (int q = 0; q < NB; ++q) { const double2 v = swp[1][q]; re[q] = v.x; im[q] = v.y; }
      }
      if (r == c) {
#pragma unroll
        for (int q = 0; q < NB; ++q) { const double2 v = swp[0][q]; re[q] = v.x; im[q] = v.y; }
      }
    }
    if (tid == 0) s_piv[jj] = p;
    if (r == c) {
      double dr = 0.0, di = 0.0;
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        dr = (q == jj) ? re[q] : dr;
        di = (q == jj) ? im[q] : di;
      }
      if (dr == 0.0 && di == 0.0) {
        if (atomicCAS(info, 0, info_tag) == 0) {
          info[1] = (int)(e0 + e);
          info[2] = c;
        }
      }
      const double2 inv = crecip(cmk(dr, di));
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const double xr = (q == jj) ? 1.0 : re[q];
        const double xi = (q == jj) ? 0.0 : im[q];
        re[q] = xr * inv.x - xi * inv.y;
        im[q] = xr * inv.y + xi * inv.x;
        piv_row[q] = cmk(re[q], im[q]);
      }
    }
    __syncthreads();
    if (valid && r != c) {
      double fr = 0.0, fi = 0.0;
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        fr = (q == jj) ? re[q] : fr;
        fi = (q == jj) ? im[q] : fi;
      }
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const double2 pv = piv_row[q];
        const double cr = (q == jj) ? 0.0 : re[q];
        const double ci = (q == jj) ? 0.0 : im[q];
        re[q] = cr - (fr * pv.x - fi * pv.y);
        im[q] = ci - (fr * pv.y + fi * pv.x);
      }
    }
  }

  if (valid) {
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      if (q < nbe) {
        A[(int64_t)r * n + k + q] = cmk(re[q], im[q]);
        Ahat[(int64_t)r * NB + q] = cmk(re[q] - (r == k + q ? 1.0 : 0.0), im[q]);
      }
    }
  }

  // Row moves of this panel: simulate the swaps on an index map.
  for (int t = tid; t < n; t += blockDim.x) src_of[t] = t;
  __syncthreads();
  if (tid == 0) {
    for (int jj = 0; jj < nbe; ++jj) {
      const int a = k + jj, b = s_piv[jj];
      const int tmp = src_of[a];
      src_of[a] = src_of[b];
      src_of[b] = tmp;
    }
  }
  __syncthreads();
  if (k == 0) {
    for (int t = tid; t < n; t += blockDim.x) rp[t] = src_of[t];
  } else {
    int v0 = 0, v1 = 0;
    if (tid < nbe) {
      v0 = rp[src_of[k + tid]];
      v1 = rp[src_of[s_piv[tid]]];
    }
    __syncthreads();
    if (tid < nbe) {
      rp[k + tid] = v0;
      rp[s_piv[tid]] = v1;
    }
  }
  if (tid < nbe) {
    moves[2 * tid] = k + tid;                  // destination
    moves[2 * tid + 1] = src_of[k + tid];      // source
    moves[2 * (NB + tid)] = s_piv[tid];
    moves[2 * (NB + tid) + 1] = src_of[s_piv[tid]];
  } else if (tid < NB) {
    // pad to 2*NB moves with copies of move 0 (same destination, same value)
    moves[2 * tid] = k;
    moves[2 * tid + 1] = src_of[k];
    moves[2 * (NB + tid)] = k;
    moves[2 * (NB + tid) + 1] = src_of[k];
  }
}

template <int NB>
__global__ void __launch_bounds__(64)
zinv_swap_kernel(double2* S, int64_t sS, int n, int k, int nbe, char* aux, int64_t sAux) {
  const int e = blockIdx.y;
  const int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= n || (c >= k && c < k + nbe)) return;
  double2* A = S + (int64_t)e * sS;
  char* ax = aux + (int64_t)e * sAux;
  double2* Wp = reinterpret_cast<double2*>(ax + PanelAux<NB>::wp(n));
  const int* moves = reinterpret_cast<const int*>(ax + PanelAux<NB>::moves(n));
  // moves holds 2*NB (destination, source) pairs: entries [0, nbe) are the
  // panel rows, [NB, NB+nbe) the pivot rows, the rest duplicates of pair 0.
  // All loads of a column precede its stores; values are staged in LDS.
  __shared__ double2 v[2 * NB][64];
  const int t = threadIdx.x;
#pragma unroll 8
  for (int q = 0; q < 2 * NB; ++q) v[q][t] = A[(int64_t)moves[2 * q + 1] * n + c];
#pragma unroll 8
  for (int q = 0; q < 2 * NB; ++q) A[(int64_t)moves[2 * q] * n + c] = v[q][t];
#pragma unroll 8
  for (int q = 0; q < NB; ++q) Wp[(int64_t)q * n + c] = v[q][t];
}

template <int NB>
__global__ void __launch_bounds__(256)
zinv_finish_kernel(const double2* S, int64_t sS, int n, const char* aux, int64_t sAux,
                   double2* Out, int64_t sOut, int ldo, int rows_per_wg) {
  __shared__ int rpinv[1024];
  const int e = blockIdx.y;
  const int* rp = reinterpret_cast<const int*>(aux + (int64_t)e * sAux + PanelAux<NB>::rp(n));
  for (int t = threadIdx.x; t < n; t += blockDim.x) rpinv[rp[t]] = t;
  __syncthreads();
  const double2* A = S + (int64_t)e * sS;
  double2* O = Out + (int64_t)e * sOut;
  const int r0 = blockIdx.x * rows_per_wg;
  const int total = rows_per_wg * n;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int r = r0 + idx / n, j = idx % n;
    if (r < n) O[(int64_t)r * ldo + j] = A[(int64_t)r * n + rpinv[j]];
  }
}

template <int NB, int MAXT>
hipError_t zinv_run(double2* S, int64_t sS, int n, int nE, double2* Out, int64_t sOut, int ldo,
                    void* aux, int* info, int info_tag, int64_t e0, int tile_cfg,
                    hipStream_t stream) {
  char* ax = static_cast<char*>(aux);
  const int64_t sAux = (int64_t)PanelAux<NB>::bytes(n);
  const int threads = ((n + 63) / 64) * 64;
  for (int k = 0; k < n; k += NB) {
    const int nbe = (n - k < NB) ? (n - k) : NB;
    hipLaunchKernelGGL((zinv_panel_kernel<NB, MAXT>), dim3(nE), dim3(threads), 0, stream, S, sS, n,
                       k, nbe, ax, sAux, info, info_tag, e0);
    if (nbe == n) break;  // single panel: nothing else to update
    hipLaunchKernelGGL((zinv_swap_kernel<NB>), dim3((n + 63) / 64, nE), dim3(64), 0, stream, S, sS,
                       n, k, nbe, ax, sAux);
    ZGroup g{};
    g.nE = nE;
    int np = 0;
    const int64_t sAuxC = sAux / (int64_t)sizeof(double2);
    double2* ahat = reinterpret_cast<double2*>(ax + PanelAux<NB>::ahat(n));
    double2* wp = reinterpret_cast<double2*>(ax + PanelAux<NB>::wp(n));
    auto add = [&](int c0, int nc) {
      if (nc <= 0) return;
      ZProblem& P = g.p[np++];
      P.m = n; P.n = nc; P.nterms = 1; P.naddends = 1; P.npre = 1; P.d2mode = kD2None;
      P.t[0] = ZTerm{ahat, wp + c0, sAuxC, sAuxC, NB, n, kOpN, kOpN, nbe, 0, cmk(1.0, 0.0)};
      P.ad[0] = ZAddend{S + c0, sS, n, 0, cmk(1.0, 0.0), cmk(0.0, 0.0)};
      P.D = S + c0; P.sD = sS; P.ldd = n;
    };
    add(0, k);
    add(k + nbe, n - k - nbe);
    g.nprob = np;
    hipError_t err = launch_zgemm_update(g, tile_cfg, stream);
    if (err != hipSuccess) return err;
  }
  const int rows = (n >= 256) ? 4 : ((1024 + n - 1) / n);
  hipLaunchKernelGGL((zinv_finish_kernel<NB>), dim3((n + rows - 1) / rows, nE), dim3(256), 0,
                     stream, S, sS, n, ax, sAux, Out, sOut, ldo, rows);
  return hipGetLastError();
}

}  // namespace

size_t zinv_aux_bytes(int n) {
  return panel_nb(n) == 32 ? PanelAux<32>::bytes(n) : PanelAux<16>::bytes(n);
}

hipError_t launch_zinv(double2* S, int64_t sS, int n, int nE, double2* Out, int64_t sOut, int ldo,
                       void* aux, int* info, int info_tag, int64_t e0, int tile_cfg,
                       hipStream_t stream) {
  if (n < 1 || n > 1024) return hipErrorInvalidValue;
  if (nE == 0) return hipSuccess;
  if (panel_nb(n) == 32)
    return zinv_run<32, 512>(S, sS, n, nE, Out, sOut, ldo, aux, info, info_tag, e0, tile_cfg,
                             stream);
  return zinv_run<16, 1024>(S, sS, n, nE, Out, sOut, ldo, aux, info, info_tag, e0, tile_cfg,
                            stream);
}

}  // namespace qx
