// Batched complex FP64 matrix inverse: blocked in-place Gauss-Jordan
// elimination with partial pivoting, rows never moved (implicit pivoting).
//
// Step c eliminates column c with pivot row pi(c), the row of largest
// |re|+|im| among the rows not yet used as pivots (LAPACK izamax's measure,
// first physical row on ties).  Nothing is swapped: after all steps the matrix
// holds M with A^{-1}[i][j] = M[pi(i)][pi^{-1}(j)].  Per panel of NB columns
// [k, k+nb):
//   1. zinv_panel_kernel (one workgroup per energy, one physical row per
//      thread, the row's panel part in registers): unblocked GJE on the panel
//      with the pivot search by wavefront shuffle max-reduction.  Writes the
//      transformed panel T in place and the pivot rows, moves the pivot rows'
//      non-panel columns to Wp (nb x n) and zeroes them in S.
//   2. Trailing update on MFMA (grouped GEMM, two column ranges):
//      S[:, nonpanel] += T * Wp[:, nonpanel].
//      This is the blocked GJE update S += (T - J) Wp with J[pi(k+q)][q] = 1:
//      J Wp restores the pivot rows, which the move to Wp left at zero.
// zinv_finish_kernel then writes Out[i][j] = S[pi(i)][pi^{-1}(j)].
#include "qx_common.h"

namespace qx {

namespace {

template <int NB>
struct PanelAux {
  // byte offsets inside one energy's aux area
  __host__ __device__ static size_t wp(int) { return 0; }
  // pivot row of each step (pi)
  __host__ __device__ static size_t piv(int n) { return wp(n) + (size_t)n * NB * sizeof(double2); }
  // step at which each row was the pivot (pi^{-1}; -1 while unused)
  __host__ __device__ static size_t step(int n) { return piv(n) + (size_t)n * sizeof(int); }
  __host__ __device__ static size_t bytes(int n) {
    size_t b = step(n) + (size_t)n * sizeof(int);
    return (b + 255) & ~(size_t)255;
  }
};

inline int panel_nb(int n) { return n <= 512 ? 32 : 16; }

// Shared state of one panel workgroup.
// Panel rows are moved between HBM and registers through LDS in chunks of
// kChunk columns, so that global accesses are row-contiguous.
constexpr int kChunk = 8;

template <int NB, int MAXT>
struct PanelShared {
  double red_key[16];
  int red_idx[16];
  double2 piv_row[NB];
  int s_piv[NB];
  double2 stage[MAXT][kChunk + 1];
};

template <int NB, int MAXT>
__global__ void __launch_bounds__(MAXT)
zinv_panel_kernel(double2* S, int64_t sS, int n, int k, int nbe, char* aux, int64_t sAux,
                  int* info, int info_tag, int64_t e0) {
  const int e = blockIdx.x;
  double2* A = S + (int64_t)e * sS;
  char* ax = aux + (int64_t)e * sAux;
  double2* Wp = reinterpret_cast<double2*>(ax + PanelAux<NB>::wp(n));
  int* piv = reinterpret_cast<int*>(ax + PanelAux<NB>::piv(n));
  int* step_of = reinterpret_cast<int*>(ax + PanelAux<NB>::step(n));

  __shared__ PanelShared<NB, MAXT> sh;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  const int r = tid;
  const bool valid = r < n;

  int step = -1;
  if (valid && k > 0) step = step_of[r];
  // The row's panel part lives in registers and is rotated by one position
  // after every step, so the current pivot column is always position 0 and
  // the (rolled) loop body indexes registers with constants only.
  double re[NB], im[NB];
  const int rows = blockDim.x;
#pragma unroll
  for (int q0 = 0; q0 < NB; q0 += kChunk) {
    // rows x kChunk tile, kChunk consecutive lanes per row segment
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int idx = tid + i * rows;
      const int rr = idx / kChunk, q = q0 + idx % kChunk;
      sh.stage[rr][idx % kChunk] = (rr < n && q < nbe) ? A[(int64_t)rr * n + k + q] : cmk(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const double2 v = sh.stage[r][q];
      re[q0 + q] = v.x;
      im[q0 + q] = v.y;
    }
    __syncthreads();
  }

#pragma unroll 1
  for (int jj = 0; jj < nbe; ++jj) {
    const int c = k + jj;
    double key = (valid && step < 0) ? fabs(re[0]) + fabs(im[0]) : -1.0;
    key = (key != key) ? HUGE_VAL : key;  // NaN: keep the order total
    // wavefront max of the key, then the lowest lane holding it
    double wmax = key;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) wmax = fmax(wmax, __shfl_xor(wmax, off));
    const unsigned long long hit = __ballot(key == wmax);
    if (lane == 0) {
      sh.red_key[wave] = wmax;
      sh.red_idx[wave] = wave * 64 + (int)__builtin_ctzll(hit);
    }
    __syncthreads();
    double bk = sh.red_key[0];
    int p = sh.red_idx[0];
    for (int w = 1; w < nw; ++w) {
      const double ok = sh.red_key[w];
      if (ok > bk) { bk = ok; p = sh.red_idx[w]; }  // waves in row order: first wins ties
    }
    p = (p < n) ? p : n - 1;  // memory safety only; a candidate row always wins
    const bool me = (r == p);
    if (me) {
      // the pivot thread publishes its scaled row (position 0: the pivot
      // column, which becomes 1/pivot)
      if (re[0] == 0.0 && im[0] == 0.0) {
        if (atomicCAS(info, 0, info_tag) == 0) {
          info[1] = (int)(e0 + e);
          info[2] = c;
        }
      }
      const double2 inv = crecip(cmk(re[0], im[0]));
      sh.piv_row[0] = inv;
#pragma unroll
      for (int q = 1; q < NB; ++q)
        sh.piv_row[q] = cmk(re[q] * inv.x - im[q] * inv.y, re[q] * inv.y + im[q] * inv.x);
      step = c;
    }
    if (tid == 0) sh.s_piv[jj] = p;
    __syncthreads();
    // new[q] = keep * row[q] - f * prow[q] (row[0] counts as 0), with
    // keep = 0, f = -1 for the pivot row (new = prow); then rotate by one.
    const double keep = me ? 0.0 : 1.0;
    const double fr = me ? -1.0 : re[0], fi = me ? 0.0 : im[0];
    const double2 p0 = sh.piv_row[0];
    const double n0r = -(fr * p0.x - fi * p0.y);
    const double n0i = -(fr * p0.y + fi * p0.x);
#pragma unroll
    for (int q0 = 1; q0 < NB; q0 += 8) {
      // groups of 8 pivot-row values: keeps the LDS reads from all being
      // hoisted at once
      asm volatile("" ::: "memory");
#pragma unroll
      for (int q = q0; q < q0 + 8 && q < NB; ++q) {
        const double2 pv = sh.piv_row[q];
        const double nr = keep * re[q] - (fr * pv.x - fi * pv.y);
        const double ni = keep * im[q] - (fr * pv.y + fi * pv.x);
        re[q - 1] = nr;
        im[q - 1] = ni;
      }
    }
    re[NB - 1] = n0r;
    im[NB - 1] = n0i;
  }

  // after nbe rotations position q holds panel column (q + nbe) mod NB
  if (valid) step_of[r] = step;
  if (tid < nbe) piv[k + tid] = sh.s_piv[tid];
#pragma unroll
  for (int q0 = 0; q0 < NB; q0 += kChunk) {
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      int col = q + nbe;
      col = (col >= NB) ? col - NB : col;
      if (col >= q0 && col < q0 + kChunk) sh.stage[r][col - q0] = cmk(re[q], im[q]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int idx = tid + i * rows;
      const int rr = idx / kChunk, q = q0 + idx % kChunk;
      if (rr < n && q < nbe) A[(int64_t)rr * n + k + q] = sh.stage[rr][idx % kChunk];
    }
    __syncthreads();
  }
  // move the pivot rows' non-panel columns to Wp, leaving zeros in S.  Every
  // thread owns one column (rows >= n); the values of 16 pivot rows are loaded
  // before any of them is stored, so the loads are not serialised behind the
  // stores (the pointers may alias as far as the compiler knows).
  if (nbe < n) {
    const int c = tid;
    const bool outside = c < n && (c < k || c >= k + nbe);
#pragma unroll 1
    for (int q0 = 0; q0 < nbe; q0 += 16) {
      double2 v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int q = q0 + u;
        v[u] = (outside && q < nbe) ? A[(int64_t)sh.s_piv[q] * n + c] : cmk(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int q = q0 + u;
        if (outside && q < nbe) {
          Wp[(int64_t)q * n + c] = v[u];
          A[(int64_t)sh.s_piv[q] * n + c] = cmk(0.0, 0.0);
        }
      }
    }
  }
}

// Rows staged per workgroup of the finish kernel: rows * n <= kFinishElems.
constexpr int kFinishElems = 2048;

template <int NB>
__global__ void __launch_bounds__(256)
zinv_finish_kernel(const double2* S, int64_t sS, int n, const char* aux, int64_t sAux,
                   double2* Out, int64_t sOut, int ldo, int rows_per_wg) {
  __shared__ int colsrc[1024];
  __shared__ double2 rows_lds[kFinishElems];
  const int e = blockIdx.y;
  const char* ax = aux + (int64_t)e * sAux;
  const int* piv = reinterpret_cast<const int*>(ax + PanelAux<NB>::piv(n));
  const int* step_of = reinterpret_cast<const int*>(ax + PanelAux<NB>::step(n));
  // (the clamps only matter for NaN input, where the pivot order is meaningless)
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int s = step_of[t];
    colsrc[t] = (s >= 0 && s < n) ? s : 0;
  }
  const double2* A = S + (int64_t)e * sS;
  double2* O = Out + (int64_t)e * sOut;
  const int r0 = blockIdx.x * rows_per_wg;
  const int total = rows_per_wg * n;
  // source rows pi(r) read whole and coalesced ...
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int r = r0 + idx / n, t = idx % n;
    if (r < n) {
      const int pr = piv[r];
      rows_lds[idx] = A[(int64_t)((pr >= 0 && pr < n) ? pr : 0) * n + t];
    }
  }
  __syncthreads();
  // ... and permuted by column inside LDS: Out[r][j] = S[pi(r)][pi^{-1}(j)]
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int rr = idx / n, j = idx % n, r = r0 + rr;
    if (r < n) O[(int64_t)r * ldo + j] = rows_lds[rr * n + colsrc[j]];
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
    ZGroup g{};
    g.nE = nE;
    int np = 0;
    const int64_t sAuxC = sAux / (int64_t)sizeof(double2);
    double2* wp = reinterpret_cast<double2*>(ax + PanelAux<NB>::wp(n));
    auto add = [&](int c0, int nc) {
      if (nc <= 0) return;
      ZProblem& P = g.p[np++];
      P.m = n; P.n = nc; P.nterms = 1; P.naddends = 1; P.npre = 1; P.d2mode = kD2None;
      P.t[0] = ZTerm{S + k, wp + c0, sS, sAuxC, n, n, kOpN, kOpN, nbe, 0, cmk(1.0, 0.0)};
      P.ad[0] = ZAddend{S + c0, sS, n, 0, cmk(1.0, 0.0), cmk(0.0, 0.0)};
      P.D = S + c0; P.sD = sS; P.ldd = n;
    };
    add(0, k);
    add(k + nbe, n - k - nbe);
    g.nprob = np;
    hipError_t err = launch_zgemm_update(g, tile_cfg, stream);
    if (err != hipSuccess) return err;
  }
  const int rows = kFinishElems / n;  // n <= 1024: at least 2
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
