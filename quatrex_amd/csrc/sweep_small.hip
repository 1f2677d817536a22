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
// Products run in rounds: a round is a set of products that do not depend
// on one another, wave w computes one 16 x 16 output tile of each of them
// (tile (w - p) mod 4 of product p, so that products smaller than 16 x 16
// spread over the waves), and the round ends with one barrier.  A round whose
// outputs overwrite one of its own operands waits for every wave once more
// before writing.  Outputs that are only stored (G^R_ij, G^R_ji, G_ii,
// G^<>_ij and their conjugate transposes) go from registers to HBM without
// an LDS slot.
//
// Tile products: v_mfma_f64_16x16x4f64, complex as three real products (3M:
// ar br, ai bi, (ar + ai)(br + bi)) into three accumulators; the sign of a
// term and the conjugations of op(A) / op(B) are the MFMA's negate-A
// modifier and the sign of the fragment sums, so no fragment is scaled.  The
// inner index always runs over the 32 padded rows: blocks smaller than 32 are
// zero-padded and inverted blocks carry an identity pad, and no valid output
// element sees a pad product (a pad column of one operand meets a zero pad row
// of the other, or an identity pad meets zeros), so only the pad region of an
// output can differ from zero (it is never stored).  With every block at most
// 16 the inner loop runs over 16 (the KS = 4 kernel).
#include <utility>

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
  int piv[kB];          // pivot row of each elimination step
  int step_of[kB];      // step at which each row was the pivot
  double red[2][4];
  int job;
};

struct Acc {
  double4_t p1, p2, p3;  // ar br, ai bi, (ar + ai)(br + bi)
};

template <bool NEG>
__device__ inline double4_t mf(double a, double b, double4_t c) {
  // f64 MFMA with the BLGP field used as the negate-A modifier
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, NEG ? 1 : 0);
}

__device__ inline Acc acc_zero() {
  Acc a;
  a.p1 = double4_t{0, 0, 0, 0};
  a.p2 = a.p1;
  a.p3 = a.p1;
  return a;
}

// acc += (NEG ? -1 : 1) op(A) op(B) over KS k-steps of 4 for this wave's tile
// (tr, tc); op = plain or conjugate transpose (AH / BH).  With op(A) =
// (ar, sa ai) and op(B) = (br, sb bi): P1 += ar br, P2 += sa sb ai bi,
// P3 += (ar + sa ai)(br + sb bi), each negated for NEG.  All fragments are read
// before the first MFMA, so the MFMA chain does not wait on LDS.
template <int KS, bool AH, bool BH, bool NEG>
__device__ inline void mma(Acc& c, const Slot& A, const Slot& B, int tr, int tc, int lane) {
  const int l16 = lane & 15, kq = lane >> 4;
  const int row = tr * 16 + l16, col = tc * 16 + l16;
  double2 a[KS], b[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int kk = s * 4 + kq;
    a[s] = AH ? A[kk][row] : A[row][kk];
    b[s] = BH ? B[col][kk] : B[kk][col];
  }
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const double sa = AH ? a[s].x - a[s].y : a[s].x + a[s].y;
    const double sb = BH ? b[s].x - b[s].y : b[s].x + b[s].y;
    c.p1 = mf<NEG>(a[s].x, b[s].x, c.p1);
    c.p2 = mf<NEG != (AH != BH)>(a[s].y, b[s].y, c.p2);
    c.p3 = mf<NEG>(sa, sb, c.p3);
  }
}

// Register r of a lane holds tile element (4 r + lane / 16, lane % 16).
__device__ inline void tile_vals(const Acc& c, double2 (&v)[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = cmk(c.p1[r] - c.p2[r], c.p3[r] - c.p1[r] - c.p2[r]);
}

__device__ inline int tile_row(int tr, int r, int lane) { return tr * 16 + 4 * r + (lane >> 4); }
__device__ inline int tile_col(int tc, int lane) { return tc * 16 + (lane & 15); }
// Tile of product p of a round computed by `wave`.
__device__ inline void tile_of(int p, int wave, int& tr, int& tc) {
  const int t = (wave - p) & 3;
  tr = t >> 1;
  tc = t & 1;
}

// Global-memory accesses through address-space-1 pointers: global (not
// flat) instructions, which count on the vector-memory counter only, so an
// LDS wait does not also wait for prefetches still in flight.
typedef __attribute__((address_space(1))) double gdouble;
__device__ inline double2 gget(const double2* p, int64_t i) {
  const gdouble* q = (const gdouble*)p + 2 * i;
  return cmk(q[0], q[1]);
}
__device__ inline void gput(double2* p, int64_t i, double2 v) {
  gdouble* q = (gdouble*)p + 2 * i;
  q[0] = v.x;
  q[1] = v.y;
}

// Element of a global block (row-major, ld), zero outside rows x cols.
__device__ inline double2 gld(const double2* F, int ld, int rows, int cols, int r, int c) {
  return (F && r < rows && c < cols) ? gget(F, (int64_t)r * ld + c) : cmk(0.0, 0.0);
}

// One product's epilogue ingredients: up to three addends
//   c1 * F[r][c] + c2 * conj(F[c][r])
// from an LDS slot (s), or this thread's four tile elements prefetched into
// registers (r).
struct Add {
  const Slot* s;
  double c1, c2;  // real coefficients suffice for every use here
  const double2* r;
};
__device__ inline Add add_slot(const Slot& s, double c1, double c2 = 0.0) { return Add{&s, c1, c2, nullptr}; }
__device__ inline Add add_regs(const double2* r) { return Add{nullptr, 1.0, 0.0, r}; }
__device__ inline Add add_none() { return Add{nullptr, 0.0, 0.0, nullptr}; }

__device__ inline double2 add_val(const Add& a, int q, int r, int c) {
  if (a.r) return a.r[q];
  double2 out = cmk(0.0, 0.0);
  if (a.s) {
    const Slot& s = *a.s;
    if (a.c1 != 0.0) out = cmk(a.c1 * s[r][c].x, a.c1 * s[r][c].y);
    if (a.c2 != 0.0) {
      const double2 t = s[c][r];
      out = cadd(out, cmk(a.c2 * t.x, -a.c2 * t.y));
    }
  }
  return out;
}

// Where a product's tile goes: an LDS slot (whole padded tile), a global
// block (valid rows x cols, row-major, scaled) and / or the conjugate
// transpose of the valid part into a second global block (gt[c][r]).
struct Dst {
  int slot = -1;
  double2* g = nullptr;
  int ld = 0, rows = 0, cols = 0;
  double gs = 1.0;
  double2* gt = nullptr;
  int ldt = 0;
  double ts = 1.0;
  __device__ Dst glob(double2* p, int ld_, int rows_, int cols_, double s = 1.0) const {
    Dst d = *this;
    d.g = p;
    d.ld = ld_;
    d.rows = rows_;
    d.cols = cols_;
    d.gs = s;
    return d;
  }
  __device__ Dst globT(double2* p, int ldt_, double s) const {
    Dst d = *this;
    d.gt = p;
    d.ldt = ldt_;
    d.ts = s;
    return d;
  }
};
__device__ inline Dst to_slot(int s) {
  Dst d;
  d.slot = s;
  return d;
}
__device__ inline Dst to_glob(double2* p, int ld, int rows, int cols, double s = 1.0) {
  return Dst{}.glob(p, ld, rows, cols, s);
}

// A product of a round: D (m x n valid) = terms + addends.  `partial`: the
// value after the first `partial_after` addends also goes there (padded
// 32 x 32 layout, global).  `on` = false skips the product (uniform).
template <typename T>
struct Prod {
  bool on;
  int m, n;
  T terms;  // void(Acc&, int tr, int tc, int lane)
  Dst d;
  Add a0, a1, a2;
  double2* partial;
  int partial_after;
};
template <typename T>
__device__ inline Prod<T> prod(int m, int n, T terms, Dst d, Add a0 = add_none(), Add a1 = add_none(),
                               Add a2 = add_none(), double2* partial = nullptr, int partial_after = 0,
                               bool on = true) {
  return Prod<T>{on, m, n, terms, d, a0, a1, a2, partial, partial_after};
}

template <int I, typename T>
__device__ inline void compute(const Prod<T>& p, double2 (&v)[4], int wave, int lane) {
  int tr, tc;
  tile_of(I, wave, tr, tc);
  Acc acc = acc_zero();
  if (p.on && tr * 16 < p.m && tc * 16 < p.n) p.terms(acc, tr, tc, lane);
  tile_vals(acc, v);
  if (!p.on) return;
  const Add adds[3] = {p.a0, p.a1, p.a2};
  const int col = tile_col(tc, lane);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (p.partial && a == p.partial_after) {
#pragma unroll
      for (int r = 0; r < 4; ++r) gput(p.partial, tile_row(tr, r, lane) * kB + col, v[r]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = cadd(v[r], add_val(adds[a], r, tile_row(tr, r, lane), col));
  }
}

template <int I, typename T>
__device__ inline void emit(SmallShared& sh, const Prod<T>& p, const double2 (&v)[4], int wave,
                            int lane) {
  if (!p.on) return;
  int tr, tc;
  tile_of(I, wave, tr, tc);
  const int col = tile_col(tc, lane);
  const Dst& d = p.d;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = tile_row(tr, r, lane);
    if (d.slot >= 0) sh.s[d.slot][row][col] = v[r];
    if (d.g && row < d.rows && col < d.cols)
      gput(d.g, (int64_t)row * d.ld + col, cmk(d.gs * v[r].x, d.gs * v[r].y));
    if (d.gt && row < d.rows && col < d.cols)
      gput(d.gt, (int64_t)col * d.ldt + row, cmk(d.ts * v[r].x, -d.ts * v[r].y));
  }
}

template <typename... P, size_t... I>
__device__ inline void round_impl(SmallShared& sh, bool sync_first, std::index_sequence<I...>,
                                  const P&... ps) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double2 v[sizeof...(P)][4];
  (compute<I>(ps, v[I], wave, lane), ...);
  if (sync_first) __syncthreads();
  (emit<I>(sh, ps, v[I], wave, lane), ...);
  __syncthreads();
}

// Runs the products of one round and publishes their outputs (trailing
// barrier).  `sync_first`: wait for every wave between computing and writing,
// needed when an output slot is an operand or slot addend of any product of
// the round (other than its own elements as an addend).
template <typename... P>
__device__ inline void run_round(SmallShared& sh, bool sync_first, const P&... ps) {
  round_impl(sh, sync_first, std::index_sequence_for<P...>{}, ps...);
}

// Zero-filled loads of global blocks (rows x cols, ld) into slots, with
// c2 * src2 (same shape; optional) added and `pad` on the diagonal beyond the
// valid square (the identity pad of a block about to be inverted).  Thread t
// moves elements t + 256 q (q < 4) of the padded 32 x 32 block.
struct Blk {
  int s;
  const double2* src;
  int ld, rows, cols;
  const double2* src2;
  double c2, pad;
};
__device__ inline Blk blk(int s, const double2* src, int ld, int rows, int cols,
                          const double2* src2 = nullptr, double c2 = 0.0, double pad = 0.0) {
  return Blk{s, src, ld, rows, cols, src2, c2, pad};
}
__device__ inline void fetch_blk(const Blk& b, double2 (&v)[4]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + q * kThreads;
    const int r = idx >> 5, c = idx & 31;
    double2 x = gld(b.src, b.ld, b.rows, b.cols, r, c);
    if (b.src2 && r < b.rows && c < b.cols) {
      const double2 w = gget(b.src2, (int64_t)r * b.ld + c);
      x = cmk(x.x + b.c2 * w.x, x.y + b.c2 * w.y);
    }
    if (r == c && r >= b.rows) x = cmk(b.pad, 0.0);
    v[q] = x;
  }
}
__device__ inline void commit_blk(SmallShared& sh, int s, const double2 (&v)[4]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = tid + q * kThreads;
    sh.s[s][idx >> 5][idx & 31] = v[q];
  }
}
// The global loads of all NB blocks are issued before the barrier that
// protects the slots' previous readers, so the blocks share one memory
// latency; the caller synchronises before reading the slots.
template <int NB>
__device__ inline void load_blks(SmallShared& sh, const Blk (&b)[NB]) {
  double2 v[NB][4];
#pragma unroll
  for (int i = 0; i < NB; ++i) fetch_blk(b[i], v[i]);
  __syncthreads();  // the slots' previous content may still be read
#pragma unroll
  for (int i = 0; i < NB; ++i) commit_blk(sh, b[i].s, v[i]);
}
__device__ inline void load_blk(SmallShared& sh, int s, const double2* src, int ld, int rows, int cols) {
  const Blk b[1] = {blk(s, src, ld, rows, cols)};
  load_blks<1>(sh, b);
}

// Global write of a slot's valid part after a barrier.  mode 0: dst = s;
// 2: dst = (s - s^H) / 2 (square).
__device__ inline void store_blk(double2* dst, int ld, int rows, int cols, const Slot& s, int mode) {
  if (!dst) return;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kB * kB / kThreads; ++q) {
    const int idx = tid + q * kThreads;
    const int r = idx >> 5, c = idx & 31;
    if (r < rows && c < cols) {
      double2 v = s[r][c];
      if (mode == 2) {
        const double2 l = s[c][r];
        v = cmk(0.5 * (v.x - l.x), 0.5 * (v.y + l.y));
      }
      gput(dst, (int64_t)r * ld + c, v);
    }
  }
}

// Pivot key of a candidate row: the high word of |re| + |im| (non-negative
// doubles order like their bit patterns; NaN above infinity), with its low six
// bits replaced by a marker bit (every candidate key is nonzero) and 31 - row
// (ties go to the lower row).
__device__ inline unsigned pivot_key(double2 v, int row) {
  const unsigned hi = (unsigned)(__double_as_longlong(fabs(v.x) + fabs(v.y)) >> 32);
  return (hi & ~63u) | 32u | (unsigned)(31 - row);
}

// Maximum over each row of 16 lanes, left in the row's last lane (row_shr
// 1, 2, 4, 8; a lane whose source lies before its row keeps its own value).
__device__ inline unsigned row_max16(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x111, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x112, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x114, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x118, 0xf, 0xf, false));
  return v;
}

// Value of lane `addr / 4` of this wave (per-lane address; ds_bpermute).
__device__ inline double bpermute_d(double x, int addr) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Pivot of the column whose values lanes of half hc hold in `v` (row l % 32):
// the unused row of largest key (DPP row maxima of the half's two 16-lane
// rows); marks it used and returns it with its value (every lane of the
// wave gets the same).
__device__ inline int pick_pivot(double2 v, int hc, int n, bool& used, double2& pv) {
  const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
  unsigned key = (h == hc && r < n && !used) ? pivot_key(v, r) : 0u;
  key = row_max16(key);
  const unsigned mx = max((unsigned)__builtin_amdgcn_readlane((int)key, 32 * hc + 15),
                          (unsigned)__builtin_amdgcn_readlane((int)key, 32 * hc + 31));
  int p = 31 - (int)(mx & 31u);
  p = p < n ? p : n - 1;
  if (r == p) used = true;
  pv = cmk(readlane_d(v.x, 32 * hc + p), readlane_d(v.y, 32 * hc + p));
  return p;
}

// Gauss-Jordan inverse of the n x n valid part of slot ms (identity pad
// beyond n) with partial pivoting and no row moves, by ONE wave with the
// matrix in registers (no barrier, so the other waves can work meanwhile):
// step c eliminates column c with the unused row of largest |re| + |im|
// (compared on the high word; first row on ties), and the result is permuted
// back at the end, Out[i][j] = M[pi(i)][pi^-1(j)].  Lane l holds row l % 32,
// columns 16 (l / 32) .. 16 (l / 32) + 15, rotated by one register position
// per step (for free: the update writes register k + 1's new value into
// register k), so that column c is always register 0 of the lanes of its
// half and the step loop stays rolled with constant register indices.
// Step c: every lane fetches its half of the pivot row from the pivot lane
// (ds_bpermute) and its row's column-c value from the other half if needed,
// and applies
//   row <- keep row - g prow,  keep = 1, g = m_rc / pivot  (keep = 0,
//   g = -1/pivot for the pivot row),
// then sets column c to -g.  The step is software-pipelined: register 1
// (column c + 1) is updated first and the search and reciprocal of step
// c + 1 run beside the update of the other fifteen registers (one wave has
// no other work to hide the search's dependent chain behind: 1630 against
// 1920 cycles per column measured).  Slot ts is scratch for the permutation
// back into slot ms.
__device__ __forceinline__ void invert_wave(SmallShared& sh, int ms, int ts, int n, int* info, int tag,
                                            int64_t energy) {
  const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
  double2 m[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = sh.s[ms][r][16 * h + k];
  bool used = false;
  double2 pv;
  int p = pick_pivot(m[0], 0, n, used, pv);
  double2 inv = pivot_recip(pv);
#pragma unroll 1
  for (int c = 0; c < n; ++c) {
    const int hc = c >> 4;
    if (lane == 0) {
      sh.piv[c] = p;
      sh.step_of[p] = c;
      if (pv.x == 0.0 && pv.y == 0.0 && atomicCAS(info, 0, tag) == 0) {
        info[1] = (int)energy;
        info[2] = c;
      }
    }
    const int addr = (32 * h + p) * 4;
    double2 pr[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) pr[k] = cmk(bpermute_d(m[k].x, addr), bpermute_d(m[k].y, addr));
    const double ox = __shfl_xor(m[0].x, 32), oy = __shfl_xor(m[0].y, 32);
    const bool me = r == p;
    const double2 a = (h == hc) ? m[0] : cmk(ox, oy);  // this row's column-c value
    const double2 f = cmul(a, inv);
    const double keep = me ? 0.0 : 1.0;
    const double2 g = me ? cmk(-inv.x, -inv.y) : f;
    double2 nw[16];
    // register 1 holds column c + 1 in the half that owns it
    nw[1] = cmk(keep * m[1].x - (g.x * pr[1].x - g.y * pr[1].y), keep * m[1].y - (g.x * pr[1].y + g.y * pr[1].x));
    int pn = p;
    double2 invn = inv, pvn = pv;
    if (c + 1 < n) {
      pn = pick_pivot(nw[1], (c + 1) >> 4, n, used, pvn);
      invn = pivot_recip(pvn);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k == 1) continue;
      nw[k] = cmk(keep * m[k].x - (g.x * pr[k].x - g.y * pr[k].y), keep * m[k].y - (g.x * pr[k].y + g.y * pr[k].x));
    }
    if (h == hc) nw[0] = cmk(-g.x, -g.y);
    // rotate by one position
#pragma unroll
    for (int k = 0; k < 15; ++k) m[k] = nw[k + 1];
    m[15] = nw[0];
    p = pn;
    pv = pvn;
    inv = invn;
  }
  // register k now holds column 16 h + ((k + n) mod 16); permute back
  // through slot ts: Out[i][j] = M[pi(i)][pi^-1(j)]
#pragma unroll
  for (int k = 0; k < 16; ++k) sh.s[ts][r][16 * h + ((k + n) & 15)] = m[k];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  asm volatile("" ::: "memory");
  const int sr = r < n ? sh.piv[r] : r;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int cc = 16 * h + k;
    const int sc = cc < n ? sh.step_of[cc] : cc;
    m[k] = sh.s[ts][sr][sc];
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) sh.s[ms][r][16 * h + k] = m[k];
}

// tr(X1 Y1) - tr(X2 Y2) over the valid b x b parts; every thread gets it.
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

// W = Y A^H + Q^H - Q [= Sigma~, to `partial`] + addends (this lane's
// elements of all four tiles, in registers), all four tiles by one wave.
template <int KS>
__device__ inline void w_product(SmallShared& sh, int y, int a, int q, int out, int m,
                                 const double2 (&add)[4][4], double2* partial, int lane) {
  const Add qa = add_slot(sh.s[q], -1.0, 1.0);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int tr = t >> 1, tc = t & 1;
    Acc acc = acc_zero();
    if (tr * 16 < m && tc * 16 < m) mma<KS, false, true, false>(acc, sh.s[y], sh.s[a], tr, tc, lane);
    double2 v[4];
    tile_vals(acc, v);
    const int col = tile_col(tc, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = tile_row(tr, r, lane);
      v[r] = cadd(v[r], add_val(qa, r, row, col));
      if (partial) gput(partial, row * kB + col, v[r]);
      sh.s[out][row][col] = cadd(v[r], add[t][r]);
    }
  }
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

// Product terms: T<AH, BH, NEG>(A, B) is acc += (-1)^NEG op(A) op(B).
#define QX_T(AH, BH, NEG, A_, B_)                                                   \
  [&](Acc& ac, int tr, int tc, int lane) {                                          \
    mma<KS, AH, BH, NEG>(ac, sh.s[A_], sh.s[B_], tr, tc, lane);                     \
  }
#define QX_T2(AH1, BH1, NEG1, A1, B1, AH2, BH2, NEG2, A2, B2)                       \
  [&](Acc& ac, int tr, int tc, int lane) {                                          \
    mma<KS, AH1, BH1, NEG1>(ac, sh.s[A1], sh.s[B1], tr, tc, lane);                  \
    mma<KS, AH2, BH2, NEG2>(ac, sh.s[A2], sh.s[B2], tr, tc, lane);                  \
  }

template <int KS>
__global__ void __launch_bounds__(kThreads, 1) sweep_small_kernel(SmallLaunch L) {
  __shared__ SmallShared sh;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double2* wg_scratch = L.scratch ? (double2*)(L.scratch + (int64_t)blockIdx.x * L.scratch_per_wg) : nullptr;
  for (;;) {
    if (tid == 0) sh.job = atomicAdd(L.counter, 1);
    __syncthreads();
    const int job = sh.job;
    __syncthreads();
    if (job >= L.total_jobs) break;
    int p = 0;
    // the problem table and its index tables are read through the scalar
    // (constant) path
    typedef __attribute__((address_space(4))) const SmallProblem cprob;
    const cprob* probs = (const cprob*)L.probs;
    while (p + 1 < L.nprob && job >= probs[p + 1].job0) ++p;
    const cprob& P = probs[p];
    const int e = job - P.job0;
    const int64_t eg = P.e_base + e;
    const int N = P.N;
    const __attribute__((address_space(4))) int* bs = (const __attribute__((address_space(4))) int*)P.bs;
    const __attribute__((address_space(4))) int64_t* od = (const __attribute__((address_space(4))) int64_t*)P.od;
    const __attribute__((address_space(4))) int64_t* ou = (const __attribute__((address_space(4))) int64_t*)P.ou;
    auto D_ = [&](const double2* base, int i) { return base ? base + (int64_t)e * P.Dsz + od[i] : nullptr; };
    auto U_ = [&](const double2* base, int i) { return base ? base + (int64_t)e * P.Usz + ou[i] : nullptr; };
    auto Dw = [&](double2* base, int i) { return base ? base + (int64_t)e * P.Dsz + od[i] : nullptr; };
    auto Uw = [&](double2* base, int i) { return base ? base + (int64_t)e * P.Usz + ou[i] : nullptr; };
    auto OB = [&](const double2* const* arr, int i) -> const double2* {
      if (!arr) return nullptr;
      typedef const double2* cptr;
      typedef __attribute__((address_space(4))) const cptr ccptr;
      const double2* b = ((const ccptr*)arr)[i];
      return b ? b + (int64_t)e * bs[i] * bs[i] : nullptr;
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
      const int w0 = S.take(), w1 = S.take();
      {
        const Blk blks[3] = {blk(cg, D_(P.Ad, 0), b, b, b, OB(P.obr, 0), -1.0, 1.0),
                             blk(w0, D_(P.SLd, 0), b, b, b, OB(P.obl, 0), 1.0),
                             blk(w1, D_(P.SGd, 0), b, b, b, OB(P.obg, 0), 1.0)};
        load_blks<3>(sh, blks);
      }
      const int vl = S.take(), vg = S.take();
      __syncthreads();
      if (wave == 0) invert_wave(sh, cg, vl, b, L.info, 1, eg);
      __syncthreads();
      store_blk(Gd(0), Gld(0), b, b, sh.s[cg], 0);
      // V = g_0 W_0 ; g^<>_0 = V g_0^H (stored unprojected: the carried value)
      run_round(sh, false, prod(b, b, QX_T(false, false, false, cg, w0), to_slot(vl)),
                prod(b, b, QX_T(false, false, false, cg, w1), to_slot(vg)));
      run_round(sh, false,
                prod(b, b, QX_T(false, true, false, vl, cg), to_slot(w0).glob(Dw(P.GLd, 0), b, b, b)),
                prod(b, b, QX_T(false, true, false, vg, cg), to_slot(w1).glob(Dw(P.GGd, 0), b, b, b)));
      S.give(vl);
      S.give(vg);
      cl = w0;
      cgg = w1;
    }

    // ---- forward sweep (rgf.py:241-302) ----
    for (int j = 1; j < N; ++j) {
      const int i = j - 1;
      const int bi = bs[i], bj = bs[j];
      // the step's input blocks in one load: A_ji, A_ij, A_jj - B^R_j,
      // Sigma^<_ij, Sigma^>_ij; the W addends Sigma^<_jj + B^<_j (wave 0) and
      // Sigma^>_jj + B^>_j (wave 1) of the lane's elements go to registers
      const int a = S.take(), t = S.take(), s = S.take(), sgl = S.take(), sgg = S.take();
      double2 wadd[4][4];
      {
        const double2* SD = wave == 0 ? D_(P.SLd, j) : D_(P.SGd, j);
        const double2* BO = wave == 0 ? OB(P.obl, j) : OB(P.obg, j);
        if (wave >= 2) SD = BO = nullptr;
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
          const int c = tile_col(tt & 1, lane);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rr = tile_row(tt >> 1, r, lane);
            wadd[tt][r] = cadd(gld(SD, bj, bj, bj, rr, c), gld(BO, bj, bj, bj, rr, c));
          }
        }
      const Blk b[5] = {blk(a, U_(P.Al, i), bi, bj, bi),              // A_ji (bj x bi)
                          blk(t, U_(P.Au, i), bj, bi, bj),              // A_ij (bi x bj)
                          blk(s, D_(P.Ad, j), bj, bj, bj, OB(P.obr, j), -1.0, 1.0),
                          blk(sgl, U_(P.SLu, i), bj, bi, bj),           // Sigma^<_ij
                          blk(sgg, U_(P.SGu, i), bj, bi, bj)};          // Sigma^>_ij
        load_blks<5>(sh, b);
      }
      const int pp = S.take();
      __syncthreads();
      // P = A_ji g_i ; Y^<> = A_ji g^<>_i over the carried g^<>_i
      run_round(sh, true, prod(bj, bi, QX_T(false, false, false, a, cg), to_slot(pp)),
                prod(bj, bi, QX_T(false, false, false, a, cl), to_slot(cl)),
                prod(bj, bi, QX_T(false, false, false, a, cgg), to_slot(cgg)));
      S.give(cg);
      const int yl = cl, yg = cgg;
      // S = A_jj - B^R_j - P A_ij (in place) ; Q^<> = P Sigma^<>_ij over Sigma_ij
      run_round(sh, true, prod(bj, bj, QX_T(false, false, true, pp, t), to_slot(s), add_slot(sh.s[s], 1.0)),
                prod(bj, bj, QX_T(false, false, false, pp, sgl), to_slot(sgl)),
                prod(bj, bj, QX_T(false, false, false, pp, sgg), to_slot(sgg)));
      S.give(t);
      const int ql = sgl, qg = sgg;
      // g_j = S^-1 by wave 3 while waves 0 and 1 form
      // W^<> = Y^<> A_ji^H + Q^<>^H - Q^<> [= Sigma~] + Sigma^<>_jj + B^<>_j
      const int wl = S.take(), wg = S.take();
      if (wave == 3)
        invert_wave(sh, s, pp, bj, L.info, j + 1, eg);
      else if (wave == 0)
        w_product<KS>(sh, yl, a, ql, wl, bj, wadd, cur ? SigT(0, j) : nullptr, lane);
      else if (wave == 1)
        w_product<KS>(sh, yg, a, qg, wg, bj, wadd, cur ? SigT(1, j) : nullptr, lane);
      __syncthreads();
      S.give(pp);
      store_blk(Gd(j), Gld(j), bj, bj, sh.s[s], 0);
      S.give(a);
      S.give(yl);
      S.give(yg);
      S.give(ql);
      S.give(qg);
      // V = g_j W ; g^<>_j = V g_j^H (into the W slots, stored unprojected)
      const int vl = S.take(), vg = S.take();
      run_round(sh, false, prod(bj, bj, QX_T(false, false, false, s, wl), to_slot(vl)),
                prod(bj, bj, QX_T(false, false, false, s, wg), to_slot(vg)));
      run_round(sh, false,
                prod(bj, bj, QX_T(false, true, false, vl, s), to_slot(wl).glob(Dw(P.GLd, j), bj, bj, bj)),
                prod(bj, bj, QX_T(false, true, false, vg, s), to_slot(wg).glob(Dw(P.GGd, j), bj, bj, bj)));
      S.give(vl);
      S.give(vg);
      cg = s;
      cl = wl;
      cgg = wg;
    }

    // ---- last block: G_{N-1} = g_{N-1}; stored diagonal G^<> projected ----
    {
      const int n1 = N - 1, b = bs[n1];
      store_blk(Dw(P.GLd, n1), b, b, b, sh.s[cl], 2);
      store_blk(Dw(P.GGd, n1), b, b, b, sh.s[cgg], 2);
      if (cur) {
        const int x = S.take(), y = S.take();
        if (N > 1) {
          const Blk xy[2] = {blk(x, SigT(1, n1), kB, b, b), blk(y, SigT(0, n1), kB, b, b)};
          load_blks<2>(sh, xy);
          __syncthreads();
          const double2 jv = trace2(sh, sh.s[x], sh.s[cl], sh.s[cgg], sh.s[y], b);
          if (tid == 0) gput(P.cur, (int64_t)e * (N + 1) + n1, jv);
        }
        const Blk xy[2] = {blk(x, OB(P.obg, n1), b, b, b), blk(y, OB(P.obl, n1), b, b, b)};
        load_blks<2>(sh, xy);
        __syncthreads();
        const double2 jn = trace2(sh, sh.s[x], sh.s[cl], sh.s[cgg], sh.s[y], b);
        if (tid == 0) gput(P.cur, (int64_t)e * (N + 1) + N, cmk(-jn.x, -jn.y));
        if (N == 1) {
          // J_0 with the same (first = last) block
          if (tid == 0) gput(P.cur, (int64_t)e * (N + 1), jn);
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
      // (left-connected), Sigma^<_ij, Sigma^>_ij, A_ij, A_ji; the left-connected
      // g^<>_i (unprojected) are fetched into registers one round before
      // slots free up for them
      const int X = S.take(), g = S.take(), sgl = S.take(), sgg = S.take(), aij = S.take(),
                aji = S.take();
      const Blk gb[2] = {blk(0, Dw(P.GLd, i), bi, bi, bi), blk(0, Dw(P.GGd, i), bi, bi, bi)};
      double2 gpre[4];
      {
        const Blk b[6] = {blk(X, Gd(j), Gld(j), bj, bj), blk(g, Gd(i), Gld(i), bi, bi),
                          blk(sgl, U_(P.SLu, i), bj, bi, bj), blk(sgg, U_(P.SGu, i), bj, bi, bj),
                          blk(aij, U_(P.Au, i), bj, bi, bj), blk(aji, U_(P.Al, i), bi, bj, bi)};
        load_blks<6>(sh, b);
      }
      __syncthreads();
      // Z^<> = g_i Sigma^<>_ij ; U = g_i A_ij ; R = G_jj A_ji (over their B operands)
      run_round(sh, true, prod(bi, bj, QX_T(false, false, false, g, sgl), to_slot(sgl)),
                prod(bi, bj, QX_T(false, false, false, g, sgg), to_slot(sgg)),
                prod(bi, bj, QX_T(false, false, false, g, aij), to_slot(aij)),
                prod(bj, bi, QX_T(false, false, false, X, aji), to_slot(aji)));
      const int zl = sgl, zg = sgg, u = aij, rr = aji;
      fetch_blk(gb[0], gpre);
      // B^<> = Z^<> G_jj^H ; G^R_ij = -U G_jj ; G^R_ji = -R g_i ; K = U R ;
      // T2^<> = U G^<>_jj (over Z, G_jj and the carried G^<>_jj)
      run_round(sh, true, prod(bi, bj, QX_T(false, true, false, zl, X), to_slot(zl)),
                prod(bi, bj, QX_T(false, true, false, zg, X), to_slot(zg)),
                prod(bi, bj, QX_T(false, false, true, u, X), to_glob(Uw(P.GRu, i), bj, bi, bj), add_none(),
                     add_none(), add_none(), nullptr, 0, P.GRu != nullptr),
                prod(bj, bi, QX_T(false, false, true, rr, g), to_glob(Uw(P.GRl, i), bi, bj, bi), add_none(),
                     add_none(), add_none(), nullptr, 0, P.GRl != nullptr),
                prod(bi, bi, QX_T(false, false, false, u, rr), to_slot(X)),
                prod(bi, bj, QX_T(false, false, false, u, Xl), to_slot(Xl)),
                prod(bi, bj, QX_T(false, false, false, u, Xg), to_slot(Xg)));
      const int bl = zl, bg = zg, kk = X, t2l = Xl, t2g = Xg;
      // g^<_i into the free slot
      const int gl = S.take();
      commit_blk(sh, gl, gpre);
      fetch_blk(gb[1], gpre);
      __syncthreads();
      // G_ii = g_i + K g_i ; G^<_ij = -g^<_i R^H + B^< - T2^< (and G^<_ji =
      // -(G^<_ij)^H) ; T1^< = K g^<_i - B^< U^H over g_i
      run_round(sh, true,
                prod(bi, bi, QX_T(false, false, false, kk, g), to_glob(Gd(i), Gld(i), bi, bi),
                     add_slot(sh.s[g], 1.0)),
                prod(bi, bj, QX_T(false, true, true, gl, rr),
                     to_glob(Uw(P.GLu, i), bj, bi, bj).globT(Uw(P.GLl, i), bi, -1.0), add_slot(sh.s[bl], 1.0),
                     add_slot(sh.s[t2l], -1.0)),
                prod(bi, bi, QX_T2(false, false, false, kk, gl, false, true, true, bl, u), to_slot(g)));
      const int t1l = g;
      S.give(bl);
      // g^>_i into the slot of B^<
      const int gg = S.take();
      commit_blk(sh, gg, gpre);
      __syncthreads();
      // G^>_ij, G^>_ji ; T1^> = K g^>_i - B^> U^H over K ;
      // G^<_ii = g^<_i + T2^< U^H + T1^< - T1^<^H (in place)
      run_round(sh, true,
                prod(bi, bj, QX_T(false, true, true, gg, rr),
                     to_glob(Uw(P.GGu, i), bj, bi, bj).globT(Uw(P.GGl, i), bi, -1.0), add_slot(sh.s[bg], 1.0),
                     add_slot(sh.s[t2g], -1.0)),
                prod(bi, bi, QX_T2(false, false, false, kk, gg, false, true, true, bg, u), to_slot(kk)),
                prod(bi, bi, QX_T(false, true, false, t2l, u), to_slot(gl), add_slot(sh.s[gl], 1.0),
                     add_slot(sh.s[t1l], 1.0, -1.0)));
      const int t1g = kk;
      // G^>_ii = g^>_i + T2^> U^H + T1^> - T1^>^H (in place)
      run_round(sh, false,
                prod(bi, bi, QX_T(false, true, false, t2g, u), to_slot(gg), add_slot(sh.s[gg], 1.0),
                     add_slot(sh.s[t1g], 1.0, -1.0)));
      store_blk(Dw(P.GLd, i), bi, bi, bi, sh.s[gl], 2);
      store_blk(Dw(P.GGd, i), bi, bi, bi, sh.s[gg], 2);
      S.give(t1g);
      S.give(t1l);
      S.give(bg);
      S.give(u);
      S.give(rr);
      S.give(t2l);
      S.give(t2g);
      Xl = gl;
      Xg = gg;
      if (cur) {
        const int x = S.take(), y = S.take();
        if (i >= 1) {
          const Blk xy[2] = {blk(x, SigT(1, i), kB, bi, bi), blk(y, SigT(0, i), kB, bi, bi)};
          load_blks<2>(sh, xy);
        } else {
          const Blk xy[2] = {blk(x, OB(P.obg, 0), bi, bi, bi), blk(y, OB(P.obl, 0), bi, bi, bi)};
          load_blks<2>(sh, xy);
        }
        __syncthreads();
        const double2 jv = trace2(sh, sh.s[x], sh.s[Xl], sh.s[Xg], sh.s[y], bi);
        if (tid == 0) gput(P.cur, (int64_t)e * (N + 1) + i, jv);
        S.give(x);
        S.give(y);
      }
    }
    __syncthreads();
  }
}

#undef QX_T
#undef QX_T2

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
  if (L.bmax <= 16)
    hipLaunchKernelGGL(sweep_small_kernel<4>, dim3(grid), dim3(kThreads), 0, stream, L);
  else
    hipLaunchKernelGGL(sweep_small_kernel<8>, dim3(grid), dim3(kThreads), 0, stream, L);
  return hipGetLastError();
}

}  // namespace qx
