// Grouped, energy-batched complex GEMM on CDNA4 FP64 matrix cores.
//
// Each launch carries up to kMaxProblems independent problems
//   D = sum_t alpha_t op(A_t) op(B_t) + sum_a (c1_a F_a + c2_a F_a^H)
// over nE energies.  Complex products are four real MFMAs
// (v_mfma_f64_16x16x4f64): re += ar*br - ai*bi, im += ar*bi + ai*br.
// Tiles of op(A) (TM x TK) and op(B) (TK x TN) are staged through LDS,
// k-major so that every MFMA fragment read is 16 contiguous complex values;
// conjugation and alpha are applied while staging, so the inner loop is the
// same for every op combination.  Global loads are issued for the next k-tile
// before the MFMAs of the current one (register double buffering, two LDS
// stages, one barrier per k-tile).
#include "qx_common.h"

namespace qx {

namespace {

constexpr int TK = 16;
constexpr int PAD = 1;

template <int TM, int TN>
struct TileShape {
  static constexpr int kWavesM = TM / 32;
  static constexpr int kWavesN = TN / 32;
  static constexpr int kThreads = 64 * kWavesM * kWavesN;
  static constexpr int kAPer = TM * TK / kThreads;
  static constexpr int kBPer = TK * TN / kThreads;
};

template <int TM, int TN>
__device__ inline void load_a(const ZTerm& t, int e, int m, int m0, int k0, int tid,
                              double2 (&ra)[TileShape<TM, TN>::kAPer]) {
  using S = TileShape<TM, TN>;
  const double2* A = t.A + (int64_t)e * t.sA;
  const double2 zero = cmk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < S::kAPer; ++i) {
    const int idx = tid + i * S::kThreads;
    int mm, kk;
    if (t.opA == kOpN) { mm = idx / TK; kk = idx % TK; }
    else { mm = idx % TM; kk = idx / TM; }
    const int gm = m0 + mm, gk = k0 + kk;
    double2 v = zero;
    if (gm < m && gk < t.k) {
      v = (t.opA == kOpN) ? A[(int64_t)gm * t.lda + gk] : A[(int64_t)gk * t.lda + gm];
    }
    ra[i] = v;
  }
}

template <int TM, int TN>
__device__ inline void store_a(const ZTerm& t, int tid, const double2 (&ra)[TileShape<TM, TN>::kAPer],
                               double2 (*As)[TM + PAD]) {
  using S = TileShape<TM, TN>;
  const bool cj = (t.opA == kOpC);
#pragma unroll
  for (int i = 0; i < S::kAPer; ++i) {
    const int idx = tid + i * S::kThreads;
    int mm, kk;
    if (!cj) { mm = idx / TK; kk = idx % TK; }
    else { mm = idx % TM; kk = idx / TM; }
    double2 v = ra[i];
    if (cj) v.y = -v.y;
    As[kk][mm] = cmul(t.alpha, v);
  }
}

template <int TM, int TN>
__device__ inline void load_b(const ZTerm& t, int e, int n, int n0, int k0, int tid,
                              double2 (&rb)[TileShape<TM, TN>::kBPer]) {
  using S = TileShape<TM, TN>;
  const double2* B = t.B + (int64_t)e * t.sB;
  const double2 zero = cmk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < S::kBPer; ++i) {
    const int idx = tid + i * S::kThreads;
    int kk, nn;
    if (t.opB == kOpN) { kk = idx / TN; nn = idx % TN; }
    else { kk = idx % TK; nn = idx / TK; }
    const int gk = k0 + kk, gn = n0 + nn;
    double2 v = zero;
    if (gn < n && gk < t.k) {
      v = (t.opB == kOpN) ? B[(int64_t)gk * t.ldb + gn] : B[(int64_t)gn * t.ldb + gk];
    }
    rb[i] = v;
  }
}

template <int TM, int TN>
__device__ inline void store_b(const ZTerm& t, int tid, const double2 (&rb)[TileShape<TM, TN>::kBPer],
                               double2 (*Bs)[TN + PAD]) {
  using S = TileShape<TM, TN>;
  const bool cj = (t.opB == kOpC);
#pragma unroll
  for (int i = 0; i < S::kBPer; ++i) {
    const int idx = tid + i * S::kThreads;
    int kk, nn;
    if (!cj) { kk = idx / TN; nn = idx % TN; }
    else { kk = idx % TK; nn = idx / TK; }
    double2 v = rb[i];
    if (cj) v.y = -v.y;
    Bs[kk][nn] = v;
  }
}

// kRole only separates kernel names in traces: 0 = sweep products,
// 1 = trailing updates inside the blocked inversion.
template <int TM, int TN, int kRole>
__global__ void __launch_bounds__((TileShape<TM, TN>::kThreads))
zgemm_grouped_kernel(const ZGroup g) {
  using S = TileShape<TM, TN>;
  __shared__ double2 As[2][TK][TM + PAD];
  __shared__ double2 Bs[2][TK][TN + PAD];

  const int total = g.wg_base[g.nprob];
  const int wid = xcd_remap(blockIdx.x, total);
  int p = 0;
  while (p + 1 < g.nprob && wid >= g.wg_base[p + 1]) ++p;
  const ZProblem& P = g.p[p];
  const int local = wid - g.wg_base[p];
  const int e = local / g.tiles_mn[p];
  const int t = local % g.tiles_mn[p];
  const int m0 = (t / g.tiles_n[p]) * TM;
  const int n0 = (t % g.tiles_n[p]) * TN;
  const int m = P.m, n = P.n;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / S::kWavesN;
  const int wn = wave % S::kWavesN;

  double4_t accr[2][2], acci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accr[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
      acci[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
    }

  for (int ti = 0; ti < P.nterms; ++ti) {
    const ZTerm& T = P.t[ti];
    const int nk = (T.k + TK - 1) / TK;
    double2 ra[S::kAPer], rb[S::kBPer];
    load_a<TM, TN>(T, e, m, m0, 0, tid, ra);
    load_b<TM, TN>(T, e, n, n0, 0, tid, rb);
    if (ti > 0) __syncthreads();  // previous term's last LDS reads are done
    store_a<TM, TN>(T, tid, ra, As[0]);
    store_b<TM, TN>(T, tid, rb, Bs[0]);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      const bool more = (kt + 1 < nk);
      if (more) {
        load_a<TM, TN>(T, e, m, m0, (kt + 1) * TK, tid, ra);
        load_b<TM, TN>(T, e, n, n0, (kt + 1) * TK, tid, rb);
      }
#pragma unroll
      for (int ks = 0; ks < TK / 4; ++ks) {
        const int kk = ks * 4 + (lane >> 4);
        double2 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = As[cur][kk][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = Bs[cur][kk][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, accr[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].y, acci[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[i].y, b[j].y, accr[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].x, acci[i][j], 0, 0, 0);
      }
      if (more) {
        store_a<TM, TN>(T, tid, ra, As[cur ^ 1]);
        store_b<TM, TN>(T, tid, rb, Bs[cur ^ 1]);
      }
      __syncthreads();
    }
  }

  // Epilogue: for v_mfma_f64_16x16x4f64 register r of a lane holds
  // D[4*r + lane/16][lane%16] of each 16x16 sub-tile (probed on gfx950).
  double2* D = P.D + (int64_t)e * P.sD;
  double2* D2 = P.D2 ? P.D2 + (int64_t)e * P.sD2 : nullptr;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + r * 4 + (lane >> 4);
        const int col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row >= m || col >= n) continue;
        double2 v = cmk(accr[i][j][r], acci[i][j][r]);
        for (int a = 0; a < P.naddends; ++a) {
          if (a == P.npre && P.d2mode == kD2Partial) D2[(int64_t)row * P.ldd2 + col] = v;
          const ZAddend& A = P.ad[a];
          const double2* F = A.F + (int64_t)e * A.sF;
          if (cnz(A.c1)) v = cadd(v, cmul(A.c1, F[(int64_t)row * A.ldf + col]));
          if (cnz(A.c2)) v = cadd(v, cmul(A.c2, cconj(F[(int64_t)col * A.ldf + row])));
        }
        if (P.npre >= P.naddends && P.d2mode == kD2Partial) D2[(int64_t)row * P.ldd2 + col] = v;
        D[(int64_t)row * P.ldd + col] = v;
        if (P.d2mode == kD2Copy) D2[(int64_t)row * P.ldd2 + col] = cmul(P.s2, v);
        else if (P.d2mode == kD2ConjT) D2[(int64_t)col * P.ldd2 + row] = cmul(P.s2, cconj(v));
      }
}

template <int TM, int TN, int kRole>
hipError_t launch_cfg(const ZGroup& g0, hipStream_t stream) {
  ZGroup g = g0;
  int base = 0;
  for (int p = 0; p < g.nprob; ++p) {
    const int tm = (g.p[p].m + TM - 1) / TM, tn = (g.p[p].n + TN - 1) / TN;
    g.tiles_n[p] = tn;
    g.tiles_mn[p] = tm * tn;
    g.wg_base[p] = base;
    base += tm * tn * g.nE;
  }
  g.wg_base[g.nprob] = base;
  if (base == 0) return hipSuccess;
  hipLaunchKernelGGL((zgemm_grouped_kernel<TM, TN, kRole>), dim3(base), dim3(TileShape<TM, TN>::kThreads),
                     0, stream, g);
  return hipGetLastError();
}

}  // namespace

static_assert(sizeof(ZGroup) <= 4000, "kernel argument block too large");

int tile_m(int cfg) {
  switch (cfg) {
    case kTile32x32: return 32;
    case kTile64x32: return 64;
    case kTile128x64: return 128;
    default: return 64;
  }
}
int tile_n(int cfg) {
  switch (cfg) {
    case kTile32x32: return 32;
    case kTile64x32: return 32;
    case kTile128x64: return 64;
    default: return 64;
  }
}

hipError_t launch_zgemm(const ZGroup& g, int cfg, hipStream_t stream) {
  switch (cfg) {
    case kTile32x32: return launch_cfg<32, 32, 0>(g, stream);
    case kTile64x32: return launch_cfg<64, 32, 0>(g, stream);
    case kTile128x64: return launch_cfg<128, 64, 0>(g, stream);
    default: return launch_cfg<64, 64, 0>(g, stream);
  }
}

hipError_t launch_zgemm_update(const ZGroup& g, int cfg, hipStream_t stream) {
  switch (cfg) {
    case kTile32x32: return launch_cfg<32, 32, 1>(g, stream);
    default: return launch_cfg<64, 64, 1>(g, stream);
  }
}

}  // namespace qx
