// Persistent, on-chip selected solve for geometries whose blocks are all at
// most 32 x 32 (the regime where the launch-per-product sweep is bound by
// HBM round trips of its intermediates and by launch count).
//
// One workgroup (4 waves) runs the whole forward and backward sweep of one
// energy point with every intermediate block in LDS: nine 32 x 32 complex128
// block slots (152 KB, one workgroup per CU).  Inputs are read from HBM when a
// step needs them and results are written once; the only blocks that travel
// back through memory are the left-connected g_i, g^<>_i written by the
// forward sweep and read by the backward one (they are outputs too) and,
// when the current is requested, the Sigma~ blocks.  Workgroups are persistent
// (about one per CU) and pull energy points from a device work queue that
// spans several problems (geometries), ordered by cost on the host; the
// triple products g W g^H run back to back with V = g W never leaving LDS.
//
// The recurrences are those of RGF.selected_solve (rgf.py:216-412); the
// backward lesser/greater blocks use B = g_i Sigma_ij G_jj^H, which serves
// both G^<_ij = -g^<_i R^H + B - U G^<_jj and, as B U^H = Z M^H, the
// T1 = K g^<_i - B U^H term (the same products as the reference, grouped
// to fit the nine slots).
//
// Product tiles: wave w computes the 16 x 16 output tile (w / 2, w % 2) with
// v_mfma_f64_16x16x4f64, complex as three real products (3M: ar br, ai bi,
// (ar + ai)(br + bi)) into three accumulators.  All operand fragments of a
// product are read from the LDS slots (row-major, 33-element row stride:
// conflict-free column reads) before its first MFMA, so the MFMA chain does
// not wait on LDS between k-steps.  The input blocks of a step are loaded
// from HBM together (one latency per step, not one per block), and the
// global addends of the W product are prefetched into registers.
// Blocks smaller than 32 are zero-padded; tiles that lie outside a
// product's valid rows / columns skip their MFMAs and the k loop stops at the
// valid inner size.  Inverted blocks carry an identity pad, which no product
// reads (k stops at the valid size and the other operand's pad is zero).
#include "qx_common.h"

namespace qx {

namespace {

constexpr int kB = 32;
constexpr int kLd = 33;
constexpr int kNS = 9;
constexpr int kThreads = 256;
typedef double2 Slot[kB][kLd];

struct SmallShared {
  Slot s[kNS];
  int piv[kB];     // pivot row of each elimination step
  int step_of[kB]; // step at which each row was the pivot (-1: not yet)
  double red[2][4];
  int job;
};

struct Acc {
  double4_t p1, p2, p3;  // ar br, ai bi, (ar + ai)(br + bi)
};

__device__ inline double4_t mf(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ inline Acc acc_zero() {
  Acc a;
  a.p1 = double4_t{0, 0, 0, 0};
  a.p2 = a.p1;
  a.p3 = a.p1;
  return a;
}

// acc += sgn * op(A) op(B) over the inner index range [0, k) for this wave's
// tile (tr, tc); op = plain or conjugate transpose (ah / bh).  Every
// fragment is loaded first (k-steps beyond the valid range read in-bounds
// pad and are zeroed by selection), then the MFMAs run back to back.
template <bool AH, bool BH>
__device__ inline void mma(Acc& c, const Slot& A, const Slot& B, int k, double sgn, int tr, int tc,
                           int lane) {
  const int l16 = lane & 15, kq = lane >> 4;
  const int row = tr * 16 + l16, col = tc * 16 + l16;
  const int nks = (k + 3) >> 2;
  double ax[kB / 4], ay[kB / 4], bx[kB / 4], by[kB / 4];
#pragma unroll
  for (int s = 0; s < kB / 4; ++s) {
    const int kk = s * 4 + kq;
    const double2 a = AH ? A[kk][row] : A[row][kk];
    const double2 b = BH ? B[col][kk] : B[kk][col];
    const bool on = s < nks;
    ax[s] = on ? sgn * a.x : 0.0;
    ay[s] = on ? (AH ? -sgn * a.y : sgn * a.y) : 0.0;
    bx[s] = on ? b.x : 0.0;
    by[s] = on ? (BH ? -b.y : b.y) : 0.0;
  }
#pragma unroll
  for (int s = 0; s < kB / 4; ++s) {
    c.p1 = mf(ax[s], bx[s], c.p1);
    c.p2 = mf(ay[s], by[s], c.p2);
    c.p3 = mf(ax[s] + ay[s], bx[s] + by[s], c.p3);
  }
}

// Register r of a lane holds tile element (4 r + lane / 16, lane % 16).
__device__ inline void tile_vals(const Acc& c, double2 (&v)[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = cmk(c.p1[r] - c.p2[r], c.p3[r] - c.p1[r] - c.p2[r]);
}

// Element of a global block (row-major, ld), zero outside rows x cols.
__device__ inline double2 gld(const double2* F, int ld, int rows, int cols, int r, int c) {
  return (F && r < rows && c < cols) ? F[(int64_t)r * ld + c] : cmk(0.0, 0.0);
}

// One product's epilogue ingredients: up to three addends
//   c1 * F[r][c] + c2 * conj(F[c][r])
// from an LDS slot (s) or a global block (g, ld, rows, cols).
struct Add {
  const Slot* s;
  const double2* g;
  int ld, rows, cols;
  double c1, c2;  // real coefficients suffice for every use here
  const double2* r;  // this thread's four tile elements, prefetched into registers
};
__device__ inline Add add_slot(const Slot& s, double c1, double c2 = 0.0) {
  return Add{&s, nullptr, 0, 0, 0, c1, c2, nullptr};
}
__device__ inline Add add_glob(const double2* g, int ld, int rows, int cols, double c1) {
  return Add{nullptr, g, ld, rows, cols, c1, 0.0, nullptr};
}
__device__ inline Add add_regs(const double2* r) { return Add{nullptr, nullptr, 0, 0, 0, 1.0, 0.0, r}; }
__device__ inline Add add_none() { return Add{nullptr, nullptr, 0, 0, 0, 0.0, 0.0, nullptr}; }

// Element (row, col) of this wave's tile that register r of `lane` holds.
__device__ inline int tile_row(int tr, int r, int lane) { return tr * 16 + 4 * r + (lane >> 4); }
__device__ inline int tile_col(int tc, int lane) { return tc * 16 + (lane & 15); }

__device__ inline double2 add_val(const Add& a, int r, int c) {
  double2 out = cmk(0.0, 0.0);
  if (a.s) {
    const Slot& s = *a.s;
    if (a.c1 != 0.0) out = cadd(out, cmk(a.c1 * s[r][c].x, a.c1 * s[r][c].y));
    if (a.c2 != 0.0) {
      const double2 t = s[c][r];
      out = cadd(out, cmk(a.c2 * t.x, -a.c2 * t.y));
    }
  } else if (a.g) {
    const double2 t = gld(a.g, a.ld, a.rows, a.cols, r, c);
    out = cmk(a.c1 * t.x, a.c1 * t.y);
  }
  return out;
}

// Zero-filled loads of global blocks (rows x cols, ld) into slots, with
// c2 * src2 (same shape; optional) added and `pad` on the diagonal beyond the
// valid square (the identity pad of a block about to be inverted).  The
// global loads of all NB blocks are issued before the barrier that protects
// the slots' previous readers, so the blocks share one memory latency; the
// caller synchronises before reading the slots.
struct Blk {
  Slot* s;
  const double2* src;
  int ld, rows, cols;
  const double2* src2;
  double c2, pad;
};
__device__ inline Blk blk(Slot& s, const double2* src, int ld, int rows, int cols,
                          const double2* src2 = nullptr, double c2 = 0.0, double pad = 0.0) {
  return Blk{&s, src, ld, rows, cols, src2, c2, pad};
}
template <int NB>
__device__ inline void load_blks(const Blk (&b)[NB]) {
  const int tid = threadIdx.x;
  double2 v[NB][kB * kB / kThreads];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
#pragma unroll
    for (int q = 0; q < kB * kB / kThreads; ++q) {
      const int idx = tid + q * kThreads;
      const int r = idx >> 5, c = idx & 31;
      double2 x = gld(b[i].src, b[i].ld, b[i].rows, b[i].cols, r, c);
      if (b[i].src2 && r < b[i].rows && c < b[i].cols) {
        const double2 w = b[i].src2[(int64_t)r * b[i].ld + c];
        x = cmk(x.x + b[i].c2 * w.x, x.y + b[i].c2 * w.y);
      }
      if (r == c && r >= b[i].rows) x = cmk(b[i].pad, 0.0);
      v[i][q] = x;
    }
  }
  __syncthreads();  // the slots' previous content may still be read
#pragma unroll
  for (int i = 0; i < NB; ++i) {
#pragma unroll
    for (int q = 0; q < kB * kB / kThreads; ++q) {
      const int idx = tid + q * kThreads;
      (*b[i].s)[idx >> 5][idx & 31] = v[i][q];
    }
  }
}
__device__ inline void load_blk(Slot& s, const double2* src, int ld, int rows, int cols,
                                const double2* src2 = nullptr, double c2 = 0.0, double pad = 0.0) {
  const Blk b[1] = {blk(s, src, ld, rows, cols, src2, c2, pad)};
  load_blks<1>(b);
}

// Global write of a slot's valid part.  mode 0: dst = scale * s;
// 1: dst[r][c] = scale * conj(s[c][r]) (dst is cols x rows of the slot);
// 2: dst = (s - s^H) / 2 (square).
// Mode 0 reads, per thread, exactly the elements the thread wrote as a
// product epilogue (wave tile mapping), so a product may overwrite the slot
// right after without a barrier; modes 1 and 2 read across threads.
__device__ inline void store_blk(double2* dst, int ld, int rows, int cols, const Slot& s, int mode,
                                 double scale = 1.0) {
  if (!dst) return;
  const int tid = threadIdx.x;
  if (mode == 0) {
    const int lane = tid & 63, wave = tid >> 6;
    const int tr = wave >> 1, tc = wave & 1, c = tile_col(tc, lane);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = tile_row(tr, q, lane);
      if (r < rows && c < cols) dst[(int64_t)r * ld + c] = cmk(scale * s[r][c].x, scale * s[r][c].y);
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < kB * kB / kThreads; ++q) {
    const int idx = tid + q * kThreads;
    const int r = idx >> 5, c = idx & 31;
    if (r < rows && c < cols) {
      double2 v;
      if (mode == 0) {
        v = cmk(scale * s[r][c].x, scale * s[r][c].y);
      } else if (mode == 1) {
        v = cmk(scale * s[c][r].x, -scale * s[c][r].y);
      } else {
        const double2 u = s[r][c], l = s[c][r];
        v = cmk(0.5 * (u.x - l.x), 0.5 * (u.y + l.y));
      }
      dst[(int64_t)r * ld + c] = v;
    }
  }
}

// In-place Gauss-Jordan inverse of the n x n valid part of a slot (identity
// pad beyond n), partial pivoting without moving rows: step c eliminates
// column c with the unused row of largest |re| + |im| (first on ties); the
// result is permuted back at the end, Out[i][j] = M[pi(i)][pi^-1(j)].
// Thread t owns row t / 8 and columns 4 (t % 8) .. 4 (t % 8) + 3.
__device__ void invert(SmallShared& sh, Slot& M, int n, int* info, int tag, int64_t energy) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int r = tid >> 3, c0 = (tid & 7) * 4;
  if (tid < kB) {
    sh.step_of[tid] = -1;
    sh.piv[tid] = tid;
  }
  __syncthreads();
#pragma unroll 1
  for (int c = 0; c < n; ++c) {
    // every wave finds the pivot itself (no broadcast needed): lanes < n
    // carry the candidates of column c
    double key = -1.0;
    if (lane < n && sh.step_of[lane] < 0) {
      const double2 v = M[lane][c];
      key = fabs(v.x) + fabs(v.y);
      key = (key != key) ? HUGE_VAL : key;
    }
    double wmax = key;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) wmax = fmax(wmax, __shfl_xor(wmax, off));
    const unsigned long long hit = __ballot(key == wmax);
    int p = (int)__builtin_ctzll(hit);
    p = p < n ? p : n - 1;
    const double2 pv = M[p][c];
    const double2 a = M[r][c];
    double2 prow[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) prow[j] = M[p][c0 + j];
    __syncthreads();  // every read of step c is done before any write
    if (tid == 0) {
      sh.step_of[p] = c;
      sh.piv[c] = p;
      if (pv.x == 0.0 && pv.y == 0.0 && atomicCAS(info, 0, tag) == 0) {
        info[1] = (int)energy;
        info[2] = c;
      }
    }
    const double2 inv = crecip(pv);
    if (r < n) {
      if (r == p) {
#pragma unroll
        for (int j = 0; j < 4; ++j) M[r][c0 + j] = (c0 + j == c) ? inv : cmul(prow[j], inv);
      } else {
        const double2 f = cmul(a, inv);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int cc = c0 + j;
          if (cc == c) M[r][cc] = cmk(-f.x, -f.y);
          else if (cc < n) M[r][cc] = csub(M[r][cc], cmul(f, prow[j]));
        }
      }
    }
    __syncthreads();
  }
  // permute back: Out[i][j] = M[pi(i)][pi^-1(j)]
  double2 v[4];
  const int sr = r < n ? sh.piv[r] : r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int cc = c0 + j;
    const int sc = cc < n ? sh.step_of[cc] : cc;
    v[j] = M[sr][sc];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 4; ++j) M[r][c0 + j] = v[j];
  __syncthreads();
}

// sign * (tr(X1 Y1) - tr(X2 Y2)) over the valid b x b parts; thread 0 gets it.
__device__ double2 trace2(SmallShared& sh, const Slot& X1, const Slot& Y1, const Slot& X2,
                          const Slot& Y2, int b) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int q = 0; q < kB * kB / kThreads; ++q) {
    const int idx = tid + q * kThreads;
    const int r = idx >> 5, c = idx & 31;
    if (r < b && c < b) {
      const double2 p = cmul(X1[r][c], Y1[c][r]);
      const double2 m = cmul(X2[r][c], Y2[c][r]);
      sr += p.x - m.x;
      si += p.y - m.y;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off);
    si += __shfl_xor(si, off);
  }
  if (lane == 0) {
    sh.red[0][wave] = sr;
    sh.red[1][wave] = si;
  }
  __syncthreads();
  double2 out = cmk(sh.red[0][0] + sh.red[0][1] + sh.red[0][2] + sh.red[0][3],
                    sh.red[1][0] + sh.red[1][1] + sh.red[1][2] + sh.red[1][3]);
  __syncthreads();
  return out;
}

// Uniform slot allocator (every thread holds the same mask).
struct Slots {
  unsigned used = 0;
  __device__ int take() {
    const int s = __builtin_ctz(~used);
    used |= 1u << s;
    return s;
  }
  __device__ void give(int s) { used &= ~(1u << s); }
};

}  // namespace

namespace {

// D = sum of products (the `terms` callback) + addends for the whole 32 x 32
// slot `dst`; the tile of wave w is (w / 2, w % 2) and tiles outside the
// valid m x n skip their MFMAs.  `partial`: the value after the first
// `partial_after` addends also goes there (padded 32 x 32 layout).  Each
// thread writes the elements of D it computed; `sync_first` waits for every
// wave before those writes, which is needed when D is an operand, a
// transposed addend, or a slot other threads read since the last barrier
// (a transposing or projecting store).  The trailing barrier publishes D.
template <typename F>
__device__ inline void product(SmallShared& sh, int dst, int m, int n, F&& terms, Add a0, Add a1,
                               Add a2, double2* partial = nullptr, int partial_after = 0,
                               bool sync_first = false) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tr = wave >> 1, tc = wave & 1;
  Acc acc = acc_zero();
  if (tr * 16 < m && tc * 16 < n) terms(acc, tr, tc, lane);
  double2 v[4];
  tile_vals(acc, v);
  const Add adds[3] = {a0, a1, a2};
  const int col = tile_col(tc, lane);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (partial && a == partial_after) {
      // the value before the remaining addends (Sigma~ of the current),
      // padded 32 x 32 layout
#pragma unroll
      for (int r = 0; r < 4; ++r) partial[tile_row(tr, r, lane) * kB + col] = v[r];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (adds[a].r) v[r] = cadd(v[r], adds[a].r[r]);
      else v[r] = cadd(v[r], add_val(adds[a], tile_row(tr, r, lane), col));
    }
  }
  if (sync_first) __syncthreads();
  Slot& D = sh.s[dst];
#pragma unroll
  for (int r = 0; r < 4; ++r) D[tile_row(tr, r, lane)][col] = v[r];
  __syncthreads();
}

__global__ void __launch_bounds__(kThreads, 1) sweep_small_kernel(SmallLaunch L) {
  __shared__ SmallShared sh;
  const int tid = threadIdx.x;
  double2* wg_scratch = L.scratch ? (double2*)(L.scratch + (int64_t)blockIdx.x * L.scratch_per_wg) : nullptr;
  for (;;) {
    if (tid == 0) sh.job = atomicAdd(L.counter, 1);
    __syncthreads();
    const int job = sh.job;
    __syncthreads();
    if (job >= L.total_jobs) break;
    int p = 0;
    while (p + 1 < L.nprob && job >= L.probs[p + 1].job0) ++p;
    const SmallProblem& P = L.probs[p];
    const int e = job - P.job0;
    const int64_t eg = P.e_base + e;
    const int N = P.N;
    const int* bs = P.bs;
    const int64_t* od = P.od;
    const int64_t* ou = P.ou;
    auto D_ = [&](const double2* base, int i) { return base ? base + (int64_t)e * P.Dsz + od[i] : nullptr; };
    auto U_ = [&](const double2* base, int i) { return base ? base + (int64_t)e * P.Usz + ou[i] : nullptr; };
    auto Dw = [&](double2* base, int i) { return base ? base + (int64_t)e * P.Dsz + od[i] : nullptr; };
    auto Uw = [&](double2* base, int i) { return base ? base + (int64_t)e * P.Usz + ou[i] : nullptr; };
    auto OB = [&](const double2* const* arr, int i) -> const double2* {
      if (!arr || !arr[i]) return nullptr;
      return arr[i] + (int64_t)e * bs[i] * bs[i];
    };
    // left-connected g_i: the G^R diagonal output, or this workgroup's scratch
    double2* gstore = P.GRd ? nullptr : wg_scratch;
    double2* sigt = wg_scratch ? wg_scratch + (int64_t)(P.GRd ? 0 : N) * kB * kB : nullptr;
    auto Gd = [&](int i) -> double2* { return P.GRd ? Dw(P.GRd, i) : gstore + (int64_t)i * kB * kB; };
    auto Gld = [&](int i) { return P.GRd ? bs[i] : kB; };
    auto SigT = [&](int k, int i) { return sigt + ((int64_t)k * N + i) * kB * kB; };
    const bool cur = P.cur != nullptr;
    Slots S;

    // ---- block 0 (rgf.py:216-239) ----
    int cg, cl, cgg;
    {
      const int b = bs[0];
      cg = S.take();
      const int ws[2] = {S.take(), S.take()};
      {
        const Blk blks[3] = {blk(sh.s[cg], D_(P.Ad, 0), b, b, b, OB(P.obr, 0), -1.0, 1.0),
                             blk(sh.s[ws[0]], D_(P.SLd, 0), b, b, b, OB(P.obl, 0), 1.0),
                             blk(sh.s[ws[1]], D_(P.SGd, 0), b, b, b, OB(P.obg, 0), 1.0)};
        load_blks<3>(blks);
      }
      __syncthreads();
      invert(sh, sh.s[cg], b, L.info, 1, eg);
      store_blk(Gd(0), Gld(0), b, b, sh.s[cg], 0);
      double2* OUT[2] = {Dw(P.GLd, 0), Dw(P.GGd, 0)};
      int res[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int w = ws[k];
        const int v = S.take();
        product(sh, v, b, b, [&](Acc& a, int tr, int tc, int lane) {
          mma<false, false>(a, sh.s[cg], sh.s[w], b, 1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        product(sh, w, b, b, [&](Acc& a, int tr, int tc, int lane) {
          mma<false, true>(a, sh.s[v], sh.s[cg], b, 1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        S.give(v);
        store_blk(OUT[k], b, b, b, sh.s[w], 0);  // unprojected: the carried g^<>_0
        res[k] = w;
      }
      cl = res[0];
      cgg = res[1];
    }

    // ---- forward sweep (rgf.py:241-302) ----
    for (int j = 1; j < N; ++j) {
      const int i = j - 1;
      const int bi = bs[i], bj = bs[j];
      // the step's input blocks in one load: A_ji, A_ij, A_jj - B^R_j,
      // Sigma^<_ij, Sigma^>_ij; the W addends Sigma^<>_jj + B^<>_j of this
      // thread's tile elements go to registers
      const int a = S.take(), t = S.take(), s = S.take(), sgl = S.take(), sgg = S.take();
      {
        const Blk b[5] = {blk(sh.s[a], U_(P.Al, i), bi, bj, bi),              // A_ji (bj x bi)
                          blk(sh.s[t], U_(P.Au, i), bj, bi, bj),              // A_ij (bi x bj)
                          blk(sh.s[s], D_(P.Ad, j), bj, bj, bj, OB(P.obr, j), -1.0, 1.0),
                          blk(sh.s[sgl], U_(P.SLu, i), bj, bi, bj),           // Sigma^<_ij
                          blk(sh.s[sgg], U_(P.SGu, i), bj, bi, bj)};          // Sigma^>_ij
        load_blks<5>(b);
      }
      double2 wadd[2][4];
      {
        const int lane = tid & 63, wave = tid >> 6, c = tile_col(wave & 1, lane);
        const double2* SD[2] = {D_(P.SLd, j), D_(P.SGd, j)};
        const double2* BO[2] = {OB(P.obl, j), OB(P.obg, j)};
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rr = tile_row(wave >> 1, r, lane);
            wadd[k][r] = cadd(gld(SD[k], bj, bj, bj, rr, c), gld(BO[k], bj, bj, bj, rr, c));
          }
      }
      __syncthreads();
      const int pp = S.take();
      product(sh, pp, bj, bi, [&](Acc& ac, int tr, int tc, int lane) {
        mma<false, false>(ac, sh.s[a], sh.s[cg], bi, 1.0, tr, tc, lane);
      }, add_none(), add_none(), add_none());
      S.give(cg);
      // S = A_jj - B^R_j - P A_ij (in place: each thread reads its own elements)
      product(sh, s, bj, bj, [&](Acc& ac, int tr, int tc, int lane) {
        mma<false, false>(ac, sh.s[pp], sh.s[t], bi, -1.0, tr, tc, lane);
      }, add_slot(sh.s[s], 1.0), add_none(), add_none());
      S.give(t);
      invert(sh, sh.s[s], bj, L.info, j + 1, eg);
      store_blk(Gd(j), Gld(j), bj, bj, sh.s[s], 0);
      // lesser, then greater: Y = A_ji g^<>_i ; Q = P Sigma_ij ;
      // W = Y A_ji^H + Q^H - Q [= Sigma~] + Sigma_jj + B_j ; g^<>_j = (g_j W) g_j^H
      const int SU[2] = {sgl, sgg};
      double2* OUT[2] = {Dw(P.GLd, j), Dw(P.GGd, j)};
      int carried[2] = {cl, cgg};
      int res[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int y = S.take();
        product(sh, y, bj, bi, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, false>(ac, sh.s[a], sh.s[carried[k]], bi, 1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        S.give(carried[k]);
        const int q = S.take();
        product(sh, q, bj, bj, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, false>(ac, sh.s[pp], sh.s[SU[k]], bi, 1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        S.give(SU[k]);
        const int w = S.take();
        product(sh, w, bj, bj, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, true>(ac, sh.s[y], sh.s[a], bi, 1.0, tr, tc, lane);
        }, add_slot(sh.s[q], -1.0, 1.0), add_regs(wadd[k]), add_none(), cur ? SigT(k, j) : nullptr, 1);
        S.give(y);
        S.give(q);
        const int v = S.take();
        product(sh, v, bj, bj, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, false>(ac, sh.s[s], sh.s[w], bj, 1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        product(sh, w, bj, bj, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, true>(ac, sh.s[v], sh.s[s], bj, 1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        S.give(v);
        store_blk(OUT[k], bj, bj, bj, sh.s[w], 0);  // unprojected left-connected g^<>_j
        res[k] = w;
      }
      S.give(a);
      S.give(pp);
      cg = s;
      cl = res[0];
      cgg = res[1];
    }

    // ---- last block: G_{N-1} = g_{N-1}; stored diagonal G^<> projected ----
    {
      const int n1 = N - 1, b = bs[n1];
      store_blk(Dw(P.GLd, n1), b, b, b, sh.s[cl], 2);
      store_blk(Dw(P.GGd, n1), b, b, b, sh.s[cgg], 2);
      if (cur) {
        const int x = S.take(), y = S.take();
        if (N > 1) {
          load_blk(sh.s[x], SigT(1, n1), kB, b, b);
          load_blk(sh.s[y], SigT(0, n1), kB, b, b);
          __syncthreads();
          const double2 jv = trace2(sh, sh.s[x], sh.s[cl], sh.s[cgg], sh.s[y], b);
          if (tid == 0) P.cur[(int64_t)e * (N + 1) + n1] = jv;
        }
        load_blk(sh.s[x], OB(P.obg, n1), b, b, b);
        load_blk(sh.s[y], OB(P.obl, n1), b, b, b);
        __syncthreads();
        const double2 jn = trace2(sh, sh.s[x], sh.s[cl], sh.s[cgg], sh.s[y], b);
        if (tid == 0) P.cur[(int64_t)e * (N + 1) + N] = cmk(-jn.x, -jn.y);
        if (N == 1) {
          // J_0 with the same (first = last) block
          if (tid == 0) P.cur[(int64_t)e * (N + 1)] = jn;
        }
        S.give(x);
        S.give(y);
      }
      S.give(cg);
    }

    // ---- backward sweep (rgf.py:315-412) ----
    int Xl = cl, Xg = cgg;
    for (int i = N - 2; i >= 0; --i) {
      const int j = i + 1;
      const int bi = bs[i], bj = bs[j];
      // the step's input blocks in one load: G_jj (retarded, full), g_i
      // (left-connected), Sigma^<_ij, Sigma^>_ij, A_ij, A_ji
      const int X = S.take(), g = S.take(), sgl = S.take(), sgg = S.take(), aij = S.take(),
                aji = S.take();
      {
        const Blk b[6] = {blk(sh.s[X], Gd(j), Gld(j), bj, bj), blk(sh.s[g], Gd(i), Gld(i), bi, bi),
                          blk(sh.s[sgl], U_(P.SLu, i), bj, bi, bj),
                          blk(sh.s[sgg], U_(P.SGu, i), bj, bi, bj),
                          blk(sh.s[aij], U_(P.Au, i), bj, bi, bj),
                          blk(sh.s[aji], U_(P.Al, i), bi, bj, bi)};
        load_blks<6>(b);
      }
      __syncthreads();
      // B^<> = (g_i Sigma^<>_ij) G_jj^H
      int Bk[2];
      const int SU[2] = {sgl, sgg};
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int z = S.take();
        product(sh, z, bi, bj, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, false>(ac, sh.s[g], sh.s[SU[k]], bi, 1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        S.give(SU[k]);
        Bk[k] = S.take();
        product(sh, Bk[k], bi, bj, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, true>(ac, sh.s[z], sh.s[X], bj, 1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        S.give(z);
      }
      // U = g_i A_ij ; R = G_jj A_ji
      const int u = S.take();
      product(sh, u, bi, bj, [&](Acc& ac, int tr, int tc, int lane) {
        mma<false, false>(ac, sh.s[g], sh.s[aij], bi, 1.0, tr, tc, lane);
      }, add_none(), add_none(), add_none());
      S.give(aij);
      const int rr = S.take();
      product(sh, rr, bj, bi, [&](Acc& ac, int tr, int tc, int lane) {
        mma<false, false>(ac, sh.s[X], sh.s[aji], bj, 1.0, tr, tc, lane);
      }, add_none(), add_none(), add_none());
      S.give(aji);
      const int t = S.take();
      // G^R_ij = -U G_jj (t is free now)
      if (P.GRu) {
        product(sh, t, bi, bj, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, false>(ac, sh.s[u], sh.s[X], bj, -1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        store_blk(Uw(P.GRu, i), bj, bi, bj, sh.s[t], 0);
      }
      S.give(X);
      // G^R_ji = -R g_i
      if (P.GRl) {
        product(sh, t, bj, bi, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, false>(ac, sh.s[rr], sh.s[g], bi, -1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        store_blk(Uw(P.GRl, i), bi, bj, bi, sh.s[t], 0);
      }
      // K = U R ; G_ii = g_i + K g_i
      const int kk = S.take();
      product(sh, kk, bi, bi, [&](Acc& ac, int tr, int tc, int lane) {
        mma<false, false>(ac, sh.s[u], sh.s[rr], bj, 1.0, tr, tc, lane);
      }, add_none(), add_none(), add_none());
      product(sh, t, bi, bi, [&](Acc& ac, int tr, int tc, int lane) {
        mma<false, false>(ac, sh.s[kk], sh.s[g], bi, 1.0, tr, tc, lane);
      }, add_slot(sh.s[g], 1.0), add_none(), add_none());
      store_blk(Gd(i), Gld(i), bi, bi, sh.s[t], 0);
      S.give(t);
      S.give(g);
      // lesser, then greater
      double2* OU[2] = {Uw(P.GLu, i), Uw(P.GGu, i)};
      double2* OL[2] = {Uw(P.GLl, i), Uw(P.GGl, i)};
      double2* OD[2] = {Dw(P.GLd, i), Dw(P.GGd, i)};
      int Xc[2] = {Xl, Xg};
      int res[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        // T2 = U G^<>_jj
        const int t2 = S.take();
        product(sh, t2, bi, bj, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, false>(ac, sh.s[u], sh.s[Xc[k]], bj, 1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none());
        S.give(Xc[k]);
        const int gl = S.take();
        load_blk(sh.s[gl], Dw(k == 0 ? P.GLd : P.GGd, i), bi, bi, bi);  // g^<>_i (unprojected)
        __syncthreads();
        // G^<>_ij = -g^<>_i R^H + B - T2 ; G^<>_ji = -(G^<>_ij)^H
        const int st = S.take();
        product(sh, st, bi, bj, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, true>(ac, sh.s[gl], sh.s[rr], bi, -1.0, tr, tc, lane);
        }, add_slot(sh.s[Bk[k]], 1.0), add_slot(sh.s[t2], -1.0), add_none());
        store_blk(OU[k], bj, bi, bj, sh.s[st], 0);
        store_blk(OL[k], bi, bj, bi, sh.s[st], 1, -1.0);
        // T1 = K g^<>_i - B U^H  (reuses st, which the transposing store read)
        product(sh, st, bi, bi, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, false>(ac, sh.s[kk], sh.s[gl], bi, 1.0, tr, tc, lane);
          mma<false, true>(ac, sh.s[Bk[k]], sh.s[u], bj, -1.0, tr, tc, lane);
        }, add_none(), add_none(), add_none(), nullptr, 0, true);
        S.give(Bk[k]);
        // G^<>_ii = g^<>_i + T2 U^H + T1 - T1^H  (in place over g^<>_i)
        product(sh, gl, bi, bi, [&](Acc& ac, int tr, int tc, int lane) {
          mma<false, true>(ac, sh.s[t2], sh.s[u], bj, 1.0, tr, tc, lane);
        }, add_slot(sh.s[gl], 1.0), add_slot(sh.s[st], 1.0, -1.0), add_none());
        S.give(t2);
        S.give(st);
        store_blk(OD[k], bi, bi, bi, sh.s[gl], 2);
        res[k] = gl;
      }
      S.give(u);
      S.give(rr);
      S.give(kk);
      Xl = res[0];
      Xg = res[1];
      if (cur) {
        const int x = S.take(), y = S.take();
        if (i >= 1) {
          load_blk(sh.s[x], SigT(1, i), kB, bi, bi);
          load_blk(sh.s[y], SigT(0, i), kB, bi, bi);
        } else {
          load_blk(sh.s[x], OB(P.obg, 0), bi, bi, bi);
          load_blk(sh.s[y], OB(P.obl, 0), bi, bi, bi);
        }
        __syncthreads();
        const double2 jv = trace2(sh, sh.s[x], sh.s[Xl], sh.s[Xg], sh.s[y], bi);
        if (tid == 0) P.cur[(int64_t)e * (N + 1) + i] = jv;
        S.give(x);
        S.give(y);
      }
    }
    __syncthreads();
  }
}

__global__ void reset_counter_kernel(int* c) {
  if (threadIdx.x == 0) c[0] = 0;
}

}  // namespace

size_t sweep_small_scratch_per_wg(int n_blocks, bool need_g, bool current) {
  const size_t blk = (size_t)kB * kB * sizeof(double2);
  return blk * ((need_g ? n_blocks : 0) + (current ? 2 * n_blocks : 0));
}

hipError_t launch_sweep_small(const SmallLaunch& L, int grid, hipStream_t stream) {
  if (L.total_jobs == 0) return hipSuccess;
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, L.counter);
  hipLaunchKernelGGL(sweep_small_kernel, dim3(grid), dim3(kThreads), 0, stream, L);
  return hipGetLastError();
}

}  // namespace qx
