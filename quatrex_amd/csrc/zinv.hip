// Batched complex FP64 matrix inverse: blocked in-place Gauss-Jordan
// elimination with partial pivoting, rows never moved (implicit pivoting).
//
// Step c eliminates column c with pivot row pi(c), the row of largest
// |re|+|im| among the rows not yet used as pivots (LAPACK izamax's measure,
// first physical row on ties).  Nothing is swapped: after all steps the matrix
// holds M with A^{-1}[i][j] = M[pi(i)][pi^{-1}(j)].  A panel of NB columns
// [k, k+nb) is eliminated by zinv_panel_kernel (one workgroup per energy, one
// physical row per thread, the row's panel part in registers; for matrices up
// to 32 rows zinv_panel2_kernel, two threads per row): unblocked GJE on the
// panel.  In zinv_panel_kernel a step takes one barrier: before it every
// wave publishes its best candidate row (DPP maximum of the key) and the
// key; after it every thread takes the winner and applies
//   row <- keep row + g prow   (keep = 1, g = -a_r / a_p; for the pivot row
//   keep = 0, g = 1 / a_p)
// with the unscaled pivot row prow, a_r the row's pivot-column value and a_p
// the pivot, so no thread scales the pivot row alone while the others wait.
// The panel kernels write the transformed panel T in place, move the pivot
// rows' non-panel columns to Wp (nb x n) and zero them in S.  The blocked
// GJE update is
//   S[:, nonpanel] += T * Wp[:, nonpanel],
// i.e. S += (T - J) Wp with J[pi(k+q)][q] = 1: J Wp restores the pivot rows,
// which the move to Wp left at zero.
//
// Each update reads and writes all of S (with short K it waits on memory more
// than on the matrix cores), so panels are taken in pairs (P1, P2) and the
// columns R outside both are updated once, with K = nb1 + nb2:
//   1. panel P1 -> T1, WA
//   2. S[:, P2] += T1 WA[:, P2]                   (P2 columns only)
//   3. panel P2 -> T2, WB; the move zeroes rows pi(P2) of T1 as well, and
//      WB[:, P1] holds those rows of T1
//   4. S[:, P1] += T2 WB[:, P1]  (T1 -> T1')
//   5. S[:, R] += T1' WA[:, R] + T2 WB[:, R]     (one pass over R)
// (2 and 4 are K = nb updates of nb columns, run on 32 x 32 tiles.)
// Row by row this is the sequence of the two single-panel updates: rows
// pi(P1) and pi(P2) are zero in S[:, R] before step 5, and rows pi(P2) of T1
// are zero there, so their P1 contribution enters only through WB.  WB's R
// columns were moved before P1's update reached them; the missing
// WB[:, P1] WA[:, R] enters step 5 through T2 WB[:, P1] WA[:, R], which
// T1' carries (step 4 before step 5).
// zinv_finish_kernel then writes Out[i][j] = S[pi(i)][pi^{-1}(j)].
#include <algorithm>

#include "qx_common.h"

namespace qx {

namespace {

template <int NB>
struct PanelAux {
  // byte offsets inside one energy's aux area: two pivot-row buffers (the
  // panels of a pair), then the pivot bookkeeping
  __host__ __device__ static size_t wp(int n, int slot) {
    return (size_t)slot * n * NB * sizeof(double2);
  }
  // pivot row of each step (pi)
  __host__ __device__ static size_t piv(int n) { return wp(n, 2); }
  // step at which each row was the pivot (pi^{-1}; -1 while unused)
  __host__ __device__ static size_t step(int n) { return piv(n) + (size_t)n * sizeof(int); }
  __host__ __device__ static size_t bytes(int n) {
    size_t b = step(n) + (size_t)n * sizeof(int);
    return (b + 255) & ~(size_t)255;
  }
};

inline int panel_nb(int n) { return n <= 512 ? 32 : 16; }

// Maximum over each row of 16 lanes, left in the row's last lane (row_shr
// 1, 2, 4, 8 on both halves of the double; a lane whose source lies before its
// row keeps its own value).
template <int CTRL>
__device__ inline double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)(b & 0xffffffffll), (int)(b & 0xffffffffll), CTRL,
                                             0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double row_max16_d(double v) {
  v = fmax(v, dpp_d<0x111>(v));
  v = fmax(v, dpp_d<0x112>(v));
  v = fmax(v, dpp_d<0x114>(v));
  v = fmax(v, dpp_d<0x118>(v));
  return v;
}

// Shared state of one panel workgroup.
// Panel rows are moved between HBM and registers through LDS in chunks of
// kChunk columns, so that global accesses are row-contiguous.  That staging
// area is dynamic LDS of blockDim.x x (kChunk + 1) elements, sized by the
// launch (one row per thread): a static one sized for the largest block
// would cap small matrices at two workgroups per CU.
constexpr int kChunk = 8;

template <int NB>
struct PanelShared {
  double red_key[16];
  int red_idx[16];
  double2 piv_row[NB];
  int s_piv[NB];
};
// The one-thread-per-row panel kernel: per wave its candidate pivot row and
// (key, row), both double-buffered by step.
template <int NB, int MAXW>
struct PanelShared1 {
  double2 cand[2][MAXW][NB];
  double2 red[2][16];
  int s_piv[NB];
};

inline size_t panel_stage_bytes(int threads) {
  return (size_t)threads * (kChunk + 1) * sizeof(double2);
}

template <int NB, int MAXT>
__global__ void __launch_bounds__(MAXT)
zinv_panel_kernel(double2* S, int64_t sS, int n, int k, int nbe, char* aux, int64_t sAux,
                  int wp_slot, int* info, int info_tag, int64_t e0) {
  const int e = blockIdx.x;
  double2* A = S + (int64_t)e * sS;
  char* ax = aux + (int64_t)e * sAux;
  double2* Wp = reinterpret_cast<double2*>(ax + PanelAux<NB>::wp(n, wp_slot));
  int* piv = reinterpret_cast<int*>(ax + PanelAux<NB>::piv(n));
  int* step_of = reinterpret_cast<int*>(ax + PanelAux<NB>::step(n));

  __shared__ PanelShared1<NB, MAXT / 64> sh;
  extern __shared__ __align__(16) double2 stage_raw[];
  auto stage = reinterpret_cast<double2(*)[kChunk + 1]>(stage_raw);

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
  // rows x kChunk tiles, kChunk consecutive lanes per row segment
  // (coalesced); thread tid moves column cc of rows rr0, rr0 + rstep, ...
  // (rows is a multiple of kChunk).  The next chunk's global loads are in
  // flight while the current one passes through LDS.
  const int cc = tid % kChunk, rr0 = tid / kChunk, rstep = rows / kChunk;
  double2 ld[2][kChunk];
  auto fetch = [&](int q0, double2* dst) {
    const bool col_ok = q0 + cc < nbe;
    const double2* src = A + (int64_t)rr0 * n + k + q0 + cc;
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int rr = rr0 + i * rstep;
      dst[i] = (rr < n && col_ok) ? src[(int64_t)i * rstep * n] : cmk(0.0, 0.0);
    }
  };
  fetch(0, ld[0]);
#pragma unroll
  for (int c = 0; c < NB / kChunk; ++c) {
    if (c + 1 < NB / kChunk) fetch((c + 1) * kChunk, ld[(c + 1) & 1]);
#pragma unroll
    for (int i = 0; i < kChunk; ++i) stage[rr0 + i * rstep][cc] = ld[c & 1][i];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const double2 v = stage[r][q];
      re[c * kChunk + q] = v.x;
      im[c * kChunk + q] = v.y;
    }
    __syncthreads();
  }

  // Wave maximum of the pivot key of the current pivot column (position 0)
  // and the lane of its row (lowest on ties; -1 when the wave has no
  // candidate row left): DPP maxima of the four 16-lane rows.
  auto wave_pick = [&](double& wmax, int& cl) {
    double key = (valid && step < 0) ? fabs(re[0]) + fabs(im[0]) : -1.0;
    key = (key != key) ? HUGE_VAL : key;  // NaN: keep the order total
    const double v = row_max16_d(key);
    wmax = fmax(fmax(readlane_d(v, 15), readlane_d(v, 31)),
                fmax(readlane_d(v, 47), readlane_d(v, 63)));
    cl = (int)__builtin_ctzll(__ballot(key == wmax));
  };
  // Every wave publishes its candidate row (unscaled) and (key, row) before
  // the step's one barrier, double-buffered by step; after it every thread
  // takes the winner (first wave on ties: rows ascend with the wave).
  auto publish = [&](int b, double wmax, int cl) {
    if (lane == cl) {
      // as 8-byte halves (ds_write2_b64 takes them from any registers)
      double* ps = reinterpret_cast<double*>(sh.cand[b][wave]);
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        ps[2 * q] = re[q];
        ps[2 * q + 1] = im[q];
      }
    }
    if (lane == 0) sh.red[b][wave] = cmk(wmax, (double)(wave * 64 + cl));
  };
  {
    double wmax;
    int cl;
    wave_pick(wmax, cl);
    publish(0, wmax, cl);
  }

#pragma unroll 1
  for (int jj = 0; jj < nbe; ++jj) {
    const int c = k + jj;
    const int b = jj & 1;
    __syncthreads();
    int p, wb;
    {
      // lane w < nw holds wave w's entry
      const double2 ent = sh.red[b][lane & 15];
      const double key = (lane < nw) ? ent.x : -2.0;
      const double v = row_max16_d(key);
      const double mx = readlane_d(v, 15);
      wb = (int)__builtin_ctzll(__ballot(key == mx) & 0xffffull);
      p = (int)readlane_d(ent.y, wb);
      p = (p < n) ? p : n - 1;  // memory safety only; a candidate row always wins
    }
    const double2* prow = sh.cand[b][wb];
    const double2 pa = prow[0];
    if (tid == 0) {
      sh.s_piv[jj] = p;
      if (pa.x == 0.0 && pa.y == 0.0 && atomicCAS(info, 0, info_tag) == 0) {
        info[1] = (int)(e0 + e);
        info[2] = c;
      }
    }
    const bool me = r == p;
    if (me) step = c;
    // new[q - 1] = keep row[q] + g prow[q] (q >= 1), new[NB - 1] = g, with
    // keep = 1, g = -a_r / a_p for the other rows and keep = 0, g = 1 / a_p
    // for the pivot row (a_r: the row's pivot-column value, a_p: the pivot);
    // pivot-row values in groups of four, the next group's reads issued
    // before the current group's update
    constexpr int G = 4, NG = (NB - 1 + G - 1) / G;
    double2 pv[2][G];
    auto load = [&](int grp, double2* dst) {
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const int q = 1 + G * grp + i;
        if (q < NB) dst[i] = prow[q];
      }
    };
    load(0, pv[0]);
    const double2 inv = pivot_recip_inline(pa);
    const double gr = me ? inv.x : -(re[0] * inv.x - im[0] * inv.y);
    const double gi = me ? inv.y : -(re[0] * inv.y + im[0] * inv.x);
    const double keep = me ? 0.0 : 1.0;
    double wmax = -1.0;
    int cl = 0;
#pragma unroll
    for (int grp = 0; grp < NG; ++grp) {
      if (grp + 1 < NG) load(grp + 1, pv[(grp + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const int q = 1 + G * grp + i;
        if (q < NB) {
          const double2 v = pv[grp & 1][i];
          re[q - 1] = fma(gr, v.x, fma(-gi, v.y, keep * re[q]));
          im[q - 1] = fma(gr, v.y, fma(gi, v.x, keep * im[q]));
        }
      }
      // the next step's pivot search (position 0 is final after the first
      // group) beside the second group's update
      if (grp == 1) wave_pick(wmax, cl);
      __builtin_amdgcn_sched_barrier(0);
    }
    re[NB - 1] = gr;  // the eliminated column rotates into the last position
    im[NB - 1] = gi;
    publish(b ^ 1, wmax, cl);
  }
  __syncthreads();

  // after nbe rotations position q holds panel column (q + nbe) mod NB
  if (valid) step_of[r] = step;
  if (tid < nbe) piv[k + tid] = sh.s_piv[tid];
#pragma unroll
  for (int q0 = 0; q0 < NB; q0 += kChunk) {
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      int col = q + nbe;
      col = (col >= NB) ? col - NB : col;
      if (col >= q0 && col < q0 + kChunk) stage[r][col - q0] = cmk(re[q], im[q]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int idx = tid + i * rows;
      const int rr = idx / kChunk, q = q0 + idx % kChunk;
      if (rr < n && q < nbe) A[(int64_t)rr * n + k + q] = stage[rr][idx % kChunk];
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

// The same elimination with two threads per row for 32-column panels: lanes l and l + 32 of a wave hold the two 16-column halves of
// row 32 * wave + (l & 31), so each thread updates 16 values per step instead
// of 32, twice as many waves hide the pivot search, and a 32 x 32 matrix fills
// its wave instead of half of it.  The halves exchange their position 0 once
// per step (the lower half needs the row's pivot-column value, the upper one
// the value that rotates into its last position).  Combined position
// P = 16 h + q rotates exactly as the single-thread row does.
template <int MAXT>
__global__ void __launch_bounds__(MAXT)
zinv_panel2_kernel(double2* S, int64_t sS, int n, int k, int nbe, char* aux, int64_t sAux,
                   int wp_slot, int* info, int info_tag, int64_t e0) {
  constexpr int NB = 32, H = 16;
  const int e = blockIdx.x;
  double2* A = S + (int64_t)e * sS;
  char* ax = aux + (int64_t)e * sAux;
  double2* Wp = reinterpret_cast<double2*>(ax + PanelAux<NB>::wp(n, wp_slot));
  int* piv = reinterpret_cast<int*>(ax + PanelAux<NB>::piv(n));
  int* step_of = reinterpret_cast<int*>(ax + PanelAux<NB>::step(n));

  __shared__ PanelShared<NB> sh;
  extern __shared__ __align__(16) double2 stage_raw[];
  auto stage = reinterpret_cast<double2(*)[kChunk + 1]>(stage_raw);  // [blockDim.x / 2] rows

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  const int h = lane >> 5;
  const int r = wave * 32 + (lane & 31);
  const bool valid = r < n;
  const int nthr = blockDim.x;

  int step = -1;
  if (valid && k > 0) step = step_of[r];
  double re[H], im[H];
#pragma unroll
  for (int q0 = 0; q0 < NB; q0 += kChunk) {
#pragma unroll
    for (int i = 0; i < kChunk / 2; ++i) {
      const int idx = tid + i * nthr;
      const int rr = idx / kChunk, q = q0 + idx % kChunk;
      stage[rr][idx % kChunk] = (rr < n && q < nbe) ? A[(int64_t)rr * n + k + q] : cmk(0.0, 0.0);
    }
    __syncthreads();
    if (h == q0 / H) {
#pragma unroll
      for (int q = 0; q < kChunk; ++q) {
        const double2 v = stage[r][q];
        re[q0 % H + q] = v.x;
        im[q0 % H + q] = v.y;
      }
    }
    __syncthreads();
  }

#pragma unroll 1
  for (int jj = 0; jj < nbe; ++jj) {
    const int c = k + jj;
    const double ore = __shfl_xor(re[0], 32), oim = __shfl_xor(im[0], 32);
    const double pre = h ? ore : re[0], pim = h ? oim : im[0];  // the row's pivot-column value
    double key = (valid && step < 0 && h == 0) ? fabs(pre) + fabs(pim) : -1.0;
    key = (key != key) ? HUGE_VAL : key;  // NaN: keep the order total
    double wmax = key;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) wmax = fmax(wmax, __shfl_xor(wmax, off));
    const unsigned long long hit = __ballot(key == wmax);
    if (lane == 0) {
      sh.red_key[wave] = wmax;
      sh.red_idx[wave] = wave * 32 + ((int)__builtin_ctzll(hit) & 31);
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
      if (h == 0 && pre == 0.0 && pim == 0.0) {
        if (atomicCAS(info, 0, info_tag) == 0) {
          info[1] = (int)(e0 + e);
          info[2] = c;
        }
      }
      const double2 inv = crecip(cmk(pre, pim));
      if (h == 0) sh.piv_row[0] = inv;
#pragma unroll
      for (int q = 0; q < H; ++q) {
        if (h == 0 && q == 0) continue;
        sh.piv_row[h * H + q] = cmk(re[q] * inv.x - im[q] * inv.y, re[q] * inv.y + im[q] * inv.x);
      }
      step = c;
    }
    if (tid == 0) sh.s_piv[jj] = p;
    __syncthreads();
    const double keep = me ? 0.0 : 1.0;
    const double fr = me ? -1.0 : pre, fi = me ? 0.0 : pim;
    // the value rotating into the last position: for the lower half the
    // partner's position 0 (panel column 16 + ...), for the upper half the
    // eliminated column itself
    double lr, li;
    if (h == 0) {
      const double2 pv = sh.piv_row[H];
      lr = keep * ore - (fr * pv.x - fi * pv.y);
      li = keep * oim - (fr * pv.y + fi * pv.x);
    } else {
      const double2 p0 = sh.piv_row[0];
      lr = -(fr * p0.x - fi * p0.y);
      li = -(fr * p0.y + fi * p0.x);
    }
    const int base = h * H;
#pragma unroll
    for (int q0 = 1; q0 < H; q0 += 8) {
      asm volatile("" ::: "memory");
#pragma unroll
      for (int q = q0; q < q0 + 8 && q < H; ++q) {
        const double2 pv = sh.piv_row[base + q];
        const double nr = keep * re[q] - (fr * pv.x - fi * pv.y);
        const double ni = keep * im[q] - (fr * pv.y + fi * pv.x);
        re[q - 1] = nr;
        im[q - 1] = ni;
      }
    }
    re[H - 1] = lr;
    im[H - 1] = li;
  }

  // after nbe rotations combined position P = 16 h + q holds panel column (P + nbe) mod NB
  if (valid && h == 0) step_of[r] = step;
  if (tid < nbe) piv[k + tid] = sh.s_piv[tid];
#pragma unroll
  for (int q0 = 0; q0 < NB; q0 += kChunk) {
#pragma unroll
    for (int q = 0; q < H; ++q) {
      int col = h * H + q + nbe;
      col = (col >= NB) ? col - NB : col;
      if (col >= q0 && col < q0 + kChunk) stage[r][col - q0] = cmk(re[q], im[q]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kChunk / 2; ++i) {
      const int idx = tid + i * nthr;
      const int rr = idx / kChunk, q = q0 + idx % kChunk;
      if (rr < n && q < nbe) A[(int64_t)rr * n + k + q] = stage[rr][idx % kChunk];
    }
    __syncthreads();
  }
  // move the pivot rows' non-panel columns to Wp (one column per thread), as
  // in the single-thread-per-row kernel
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

// Rows staged per workgroup of the finish kernel: rows * n <= kFinishElems,
// rows <= kFinishRows; kFinishThreads threads per workgroup.
constexpr int kFinishElems = 2048;
constexpr int kFinishRows = 256;
constexpr int kFinishThreads = 256;

template <int NB>
__global__ void __launch_bounds__(256)
zinv_finish_kernel(const double2* S, int64_t sS, int n, const char* aux, int64_t sAux,
                   double2* Out, int64_t sOut, int ldo, int rows_per_wg) {
  __shared__ int colsrc[1024];
  __shared__ int rowsrc[kFinishRows];
  __shared__ double2 rows_lds[kFinishElems];
  const int e = blockIdx.y;
  const char* ax = aux + (int64_t)e * sAux;
  const int* piv = reinterpret_cast<const int*>(ax + PanelAux<NB>::piv(n));
  const int* step_of = reinterpret_cast<const int*>(ax + PanelAux<NB>::step(n));
  const int r0 = blockIdx.x * rows_per_wg;
  // (the clamps only matter for NaN input, where the pivot order is meaningless)
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int s = step_of[t];
    colsrc[t] = (s >= 0 && s < n) ? s : 0;
  }
  for (int t = threadIdx.x; t < rows_per_wg; t += blockDim.x) {
    const int pr = (r0 + t < n) ? piv[r0 + t] : 0;
    rowsrc[t] = (pr >= 0 && pr < n) ? pr : 0;
  }
  __syncthreads();
  const double2* A = S + (int64_t)e * sS;
  double2* O = Out + (int64_t)e * sOut;
  const int total = rows_per_wg * n;
  // source rows pi(r) read whole and coalesced, all of a thread's loads in
  // flight before their LDS stores ...
  constexpr int kU = kFinishElems / kFinishThreads;
  double2 v[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int idx = threadIdx.x + u * kFinishThreads;
    const int rr = idx / n, t = idx - rr * n;
    // (indexes clamped so that even a speculated LDS read stays in its array)
    v[u] = (idx < total && r0 + rr < n) ? A[(int64_t)rowsrc[min(rr, kFinishRows - 1)] * n + t]
                                        : cmk(0.0, 0.0);
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int idx = threadIdx.x + u * kFinishThreads;
    if (idx < total) rows_lds[idx] = v[u];
  }
  __syncthreads();
  // ... and permuted by column inside LDS: Out[r][j] = S[pi(r)][pi^{-1}(j)]
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int idx = threadIdx.x + u * kFinishThreads;
    const int rr = idx / n, j = idx - rr * n, r = r0 + rr;
    if (idx < total && r < n) O[(int64_t)r * ldo + j] = rows_lds[min(rr * n + colsrc[j], kFinishElems - 1)];
  }
}

template <int NB, int MAXT>
hipError_t zinv_run(double2* S, int64_t sS, int n, int nE, double2* Out, int64_t sOut, int ldo,
                    void* aux, int* info, int info_tag, int64_t e0, int tile_cfg,
                    hipStream_t stream) {
  char* ax = static_cast<char*>(aux);
  const int64_t sAux = (int64_t)PanelAux<NB>::bytes(n);
  const int64_t sAuxC = sAux / (int64_t)sizeof(double2);
  double2* wa = reinterpret_cast<double2*>(ax + PanelAux<NB>::wp(n, 0));
  double2* wb = reinterpret_cast<double2*>(ax + PanelAux<NB>::wp(n, 1));
  // blocks up to 32: two threads per row, which fills the wave that one thread
  // per row would leave half idle (measured at 4096 matrices of 32 x 32: 0.196
  // against 0.275 ms).  Larger blocks keep one thread per row: with two, a
  // 256-row panel needs 8 waves of 152 VGPRs and one workgroup fills a CU
  // (3.92 against 3.50 ms for 512 matrices of 256 x 256).
  const bool two = NB == 32 && n <= 32;
  const int threads = two ? 2 * ((n + 31) / 32) * 32 : ((n + 63) / 64) * 64;
  auto panel = [&](int k, int nbe, int slot) {
    if constexpr (NB == 32) {
      if (two) {
        hipLaunchKernelGGL((zinv_panel2_kernel<512>), dim3(nE), dim3(threads),
                           panel_stage_bytes(threads / 2), stream, S, sS, n, k, nbe, ax, sAux,
                           slot, info, info_tag, e0);
        return;
      }
    }
    hipLaunchKernelGGL((zinv_panel_kernel<NB, MAXT>), dim3(nE), dim3(threads),
                       panel_stage_bytes(threads), stream, S, sS, n, k, nbe, ax, sAux, slot, info,
                       info_tag, e0);
  };
  // operands: columns c0.. of S (stride sS) or of a pivot-row buffer (stride sAuxC), ld n
  auto tS = [&](int c0, const double2* B, int k) {
    return ZTerm{S + c0, B, sS, sAuxC, n, n, kOpN, kOpN, k, 0, cmk(1.0, 0.0)};
  };
  // D (m x nc at D, per-energy stride sD, ld n) += sum of the terms
  auto add = [&](ZGroup& g, int m, int nc, double2* D, int64_t sD, ZTerm t0, const ZTerm* t1) {
    if (m <= 0 || nc <= 0) return;
    ZProblem& P = g.p[g.nprob++];
    P.m = m; P.n = nc; P.nterms = t1 ? 2 : 1; P.naddends = 1; P.npre = 1; P.d2mode = kD2None;
    P.t[0] = t0;
    if (t1) P.t[1] = *t1;
    P.ad[0] = ZAddend{D, sD, n, 0, cmk(1.0, 0.0), cmk(0.0, 0.0)};
    P.D = D; P.sD = sD; P.ldd = n;
  };
  auto run = [&](ZGroup& g) { return g.nprob ? launch_zgemm_update(g, tile_cfg, stream) : hipSuccess; };
  // the narrow K = NB updates of a pair's panel columns on full 32 x 32 tiles
  // (a 64 x 64 tile would be half empty; 512 inverses of 256 x 256: 3.16 ->
  // 3.03 ms, the 3M 32 x 32 tile 3.10 ms)
  auto run_narrow = [&](ZGroup& g) {
    return g.nprob ? launch_zgemm_update(g, kTile32x32, stream) : hipSuccess;
  };
  hipError_t err = hipSuccess;
  for (int k = 0; k < n;) {
    const int nb1 = (n - k < NB) ? (n - k) : NB;
    const int k2 = k + nb1;
    const int nb2 = (n - k2 < NB) ? (n - k2) : NB;
    panel(k, nb1, 0);
    if (nb1 == n) break;  // single panel: nothing else to update
    if (nb2 == 0) {       // last panel, unpaired: columns before it
      ZGroup g{};
      g.nE = nE;
      add(g, n, k, S, sS, tS(k, wa, nb1), nullptr);
      err = run(g);
      break;
    }
    const int k3 = k2 + nb2;  // R = [0, k) and [k3, n)
    {
      ZGroup g{};
      g.nE = nE;
      add(g, n, nb2, S + k2, sS, tS(k, wa + k2, nb1), nullptr);
      if ((err = run_narrow(g)) != hipSuccess) return err;
    }
    panel(k2, nb2, 1);
    {
      // S[:, P1] += T2 WB[:, P1]: T1 becomes T1' = T1 + T2 WB[:, P1], which
      // carries WB's missing P1 contribution into the pass over R below
      ZGroup g{};
      g.nE = nE;
      add(g, n, nb1, S + k, sS, tS(k2, wb + k, nb2), nullptr);
      if ((err = run_narrow(g)) != hipSuccess) return err;
    }
    {
      // S[:, R] += T1' WA[:, R] + T2 WB[:, R]
      ZGroup g{};
      g.nE = nE;
      const ZTerm l2 = tS(k2, wb, nb2), r2 = tS(k2, wb + k3, nb2);
      add(g, n, k, S, sS, tS(k, wa, nb1), &l2);
      add(g, n, n - k3, S + k3, sS, tS(k, wa + k3, nb1), &r2);
      if ((err = run(g)) != hipSuccess) return err;
    }
    k = k3;
  }
  if (err != hipSuccess) return err;
  const int rows = std::min(kFinishElems / n, kFinishRows);  // n <= 1024: at least 2
  hipLaunchKernelGGL((zinv_finish_kernel<NB>), dim3((n + rows - 1) / rows, nE), dim3(256), 0,
                     stream, S, sS, n, ax, sAux, Out, sOut, ldo, rows);
  return hipGetLastError();
}

}  // namespace

// Non-decreasing in n: a workspace sized for the largest block of a geometry
// also holds the need of every smaller block.  Blocks up to 512 use 32-column
// panels (about 1032 n bytes), larger ones 16-column panels (about 520 n), so
// for n > 512 the need of a 512-row block can exceed that of n itself.
size_t zinv_aux_bytes(int n) {
  if (panel_nb(n) == 32) return PanelAux<32>::bytes(n);
  return std::max(PanelAux<32>::bytes(512), PanelAux<16>::bytes(n));
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
