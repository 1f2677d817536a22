// Grouped, energy-batched complex GEMM on CDNA4 FP64 matrix cores.
//
// Each launch carries up to kMaxProblems independent problems
//   D = sum_t alpha_t op(A_t) op(B_t) + sum_a (c1_a F_a + c2_a F_a^H)
// over nE energies.  Complex products are four real MFMAs
// (v_mfma_f64_16x16x4f64): re += ar*br - ai*bi, im += ar*bi + ai*br; or,
// in the 3M configurations, three: P1 = ar*br, P2 = ai*bi and
// P3 = (ar + ai)(br + bi) accumulate separately and re = P1 - P2,
// im = P3 - P1 - P2 in the epilogue (25% fewer matrix-core operations, one
// more accumulator set, two FP64 adds per fragment pair).
// Tiles of op(A) (TM x TK) and op(B) (TK x TN) are staged through LDS,
// k-major so that every MFMA fragment read is 16 contiguous complex values;
// conjugation and alpha are applied while staging, so the inner loop is the
// same for every op combination.  Global loads are issued for the next k-tile
// before the MFMAs of the current one (register double buffering, two LDS
// stages, one barrier per k-tile).
#include <type_traits>

#include "qx_common.h"

namespace qx {

namespace {

constexpr int PAD = 1;

// kTK: k-depth of one LDS stage.  kSlab: epilogue rows per LDS pass (the rows
// of one wave row).  The one-wave 32x32 tile (blocks below 64, where the
// products are bound by HBM) halves both for FP64, so that its LDS (8.4 KB
// instead of 16.9 KB) does not cap the number of workgroups per CU below what
// its registers allow (measured: equal at b = 32, +13-34% at b = 48); FP32
// products, whose MFMAs take half as long, lose more to the extra barriers
// than they gain and keep the deeper stage.  kTKo overrides the depth: the
// 3M 128x64 tile (one workgroup per CU by its registers either way) stages
// k = 16, so that each stage's MFMAs cover the next stage's global-load
// latency and the barrier count halves (measured 2-4% faster than k = 8 at
// b = 128-512; the same depth made the 4M 128x64 and the 3M 64x64 tiles
// slower: 99 KB and 66 KB of LDS cost them workgroups per CU).
template <int TM, int TN, bool kF32 = false, int kTKo = 0>
struct TileShape {
  static constexpr bool kSmall = TM == 32 && TN == 32 && !kF32;
  static constexpr int kTK = kTKo ? kTKo : (kSmall ? 4 : 8);
  static constexpr int kSlab = kSmall ? 16 : 32;
  static constexpr int kWavesM = TM / 32;
  static constexpr int kWavesN = TN / 32;
  static constexpr int kThreads = 64 * kWavesM * kWavesN;
  static constexpr int kAPer = TM * kTK / kThreads;
  static constexpr int kBPer = kTK * TN / kThreads;
};

template <int TM, int TN, bool F32, int TKo>
__device__ inline void load_a(const ZTerm& t, int e, int m, int m0, int k0, int tid,
                              double2 (&ra)[TileShape<TM, TN, F32, TKo>::kAPer]) {
  using S = TileShape<TM, TN, F32, TKo>;
  const double2* A = t.A + (int64_t)e * t.sA;
  const double2 zero = cmk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < S::kAPer; ++i) {
    const int idx = tid + i * S::kThreads;
    int mm, kk;
    if (t.opA == kOpN) { mm = idx / S::kTK; kk = idx % S::kTK; }
    else { mm = idx % TM; kk = idx / TM; }
    const int gm = m0 + mm, gk = k0 + kk;
    double2 v = zero;
    if (gm < m && gk < t.k) {
      v = (t.opA == kOpN) ? A[(int64_t)gm * t.lda + gk] : A[(int64_t)gk * t.lda + gm];
    }
    ra[i] = v;
  }
}

__device__ inline void put(double2* p, double2 v) { *p = v; }
__device__ inline void put(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }

template <int TM, int TN, bool F32, int TKo, typename LT>
__device__ inline void store_a(const ZTerm& t, int tid,
                               const double2 (&ra)[TileShape<TM, TN, F32, TKo>::kAPer],
                               LT (*As)[TM + PAD]) {
  using S = TileShape<TM, TN, F32, TKo>;
  const bool cj = (t.opA == kOpC);
#pragma unroll
  for (int i = 0; i < S::kAPer; ++i) {
    const int idx = tid + i * S::kThreads;
    int mm, kk;
    if (!cj) { mm = idx / S::kTK; kk = idx % S::kTK; }
    else { mm = idx % TM; kk = idx / TM; }
    double2 v = ra[i];
    if (cj) v.y = -v.y;
    put(&As[kk][mm], cmul(t.alpha, v));
  }
}

template <int TM, int TN, bool F32, int TKo>
__device__ inline void load_b(const ZTerm& t, int e, int n, int n0, int k0, int tid,
                              double2 (&rb)[TileShape<TM, TN, F32, TKo>::kBPer]) {
  using S = TileShape<TM, TN, F32, TKo>;
  const double2* B = t.B + (int64_t)e * t.sB;
  const double2 zero = cmk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < S::kBPer; ++i) {
    const int idx = tid + i * S::kThreads;
    int kk, nn;
    if (t.opB == kOpN) { kk = idx / TN; nn = idx % TN; }
    else { kk = idx % S::kTK; nn = idx / S::kTK; }
    const int gk = k0 + kk, gn = n0 + nn;
    double2 v = zero;
    if (gn < n && gk < t.k) {
      v = (t.opB == kOpN) ? B[(int64_t)gk * t.ldb + gn] : B[(int64_t)gn * t.ldb + gk];
    }
    rb[i] = v;
  }
}

template <int TM, int TN, bool F32, int TKo, typename LT>
__device__ inline void store_b(const ZTerm& t, int tid,
                               const double2 (&rb)[TileShape<TM, TN, F32, TKo>::kBPer],
                               LT (*Bs)[TN + PAD]) {
  using S = TileShape<TM, TN, F32, TKo>;
  const bool cj = (t.opB == kOpC);
#pragma unroll
  for (int i = 0; i < S::kBPer; ++i) {
    const int idx = tid + i * S::kThreads;
    int kk, nn;
    if (!cj) { kk = idx / TN; nn = idx % TN; }
    else { kk = idx % S::kTK; nn = idx / S::kTK; }
    double2 v = rb[i];
    if (cj) v.y = -v.y;
    put(&Bs[kk][nn], v);
  }
}

template <int TM, int TN, typename LT, int TKo>
struct Smem {
  using S = TileShape<TM, TN, std::is_same<LT, float2>::value, TKo>;
  static constexpr int kA = S::kTK * (TM + PAD);  // one LDS stage of op(A), k-major (LT elements)
  static constexpr int kB = S::kTK * (TN + PAD);  // one LDS stage of op(B), k-major
  static constexpr int kMainBytes = (2 * kA + 2 * kB) * (int)sizeof(LT);
  // the epilogue stages one kSlab-row slab of the output tile at a time
  static constexpr int kTileBytes = S::kSlab * (TN + 1) * (int)sizeof(double2);
  static constexpr int kBytes = kMainBytes > kTileBytes ? kMainBytes : kTileBytes;
};

// f64 MFMA with the BLGP field used as the negate-A modifier.
__device__ inline double4_t mfma(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline double4_t mfma_nega(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
}

// kRole only separates kernel names in traces: 0 = sweep products,
// 1 = trailing updates inside the blocked inversion.
// kF32: operands are rounded to complex64 while staged into LDS and the
// products run on v_mfma_f32_16x16x4f32 with FP32 accumulation; loads,
// addends and stores stay complex128 (the mixed-precision sweeps).
// Occupancy target: four 64x64 workgroups per CU (33 KB LDS each, <= 128
// VGPRs), measured faster than three.
// k3M: FP64 products by three real MFMAs per complex product (see the file
// header); the extra accumulator set lowers the 64x64 target to three
// workgroups per CU.
template <int TM, int TN, int kRole, bool kF32, bool k3M = false, int kTKo = 0>
__global__ void __launch_bounds__((TileShape<TM, TN, kF32, kTKo>::kThreads),
                                  (TM * TN > 4096 ? 1 : (TM * TN == 4096 ? (k3M ? 3 : 4) : 2)))
zgemm_grouped_kernel(const ZGroup g) {
  static_assert(!(k3M && kF32), "3M products are built for FP64 only");
  using S = TileShape<TM, TN, kF32, kTKo>;
  using LT = typename std::conditional<kF32, float2, double2>::type;
  using AccT = typename std::conditional<kF32, float4_t, double4_t>::type;
  using L = Smem<TM, TN, LT, kTKo>;
  constexpr int TK = S::kTK;
  constexpr int kSlab = S::kSlab;
  __shared__ __align__(16) char smem_raw[L::kBytes];
  LT* smem = reinterpret_cast<LT*>(smem_raw);
  auto As = [&](int buf) { return reinterpret_cast<LT(*)[TM + PAD]>(smem + buf * L::kA); };
  auto Bs = [&](int buf) {
    return reinterpret_cast<LT(*)[TN + PAD]>(smem + 2 * L::kA + buf * L::kB);
  };

  // Energy-major order: logical id = e * per + (tile offset of the problem
  // within the energy) + tile.  The XCD remap gives every XCD a contiguous
  // range of energies with all their problems, so XCDs receive equal work even
  // when the problems of a launch differ in cost (number of terms, addends).
  const int per = g.wg_base[g.nprob];
  const int wid = xcd_remap(blockIdx.x, per * g.nE);
  const int e = wid / per;
  const int r = wid - e * per;
  int p = 0;
  while (p + 1 < g.nprob && r >= g.wg_base[p + 1]) ++p;
  const ZProblem& P = g.p[p];
  const int t = r - g.wg_base[p];
  const int m0 = (t / g.tiles_n[p]) * TM;
  const int n0 = (t % g.tiles_n[p]) * TN;
  const int m = P.m, n = P.n;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / S::kWavesN;
  const int wn = wave % S::kWavesN;

  // 4M: accr = re, acci = im.  3M: accr = P1, acci = P2, accs = P3.
  AccT accr[2][2], acci[2][2], accs[k3M ? 2 : 1][k3M ? 2 : 1];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accr[i][j] = AccT{0, 0, 0, 0};
      acci[i][j] = AccT{0, 0, 0, 0};
      if constexpr (k3M) accs[i][j] = AccT{0, 0, 0, 0};
    }

  for (int ti = 0; ti < P.nterms; ++ti) {
    // by value: the term's fields are loaded once into scalar registers
    // instead of being re-read from the kernel-argument segment in the loop
    const ZTerm T = P.t[ti];
    const int nk = (T.k + TK - 1) / TK;
    double2 ra[S::kAPer], rb[S::kBPer];
    load_a<TM, TN, kF32, kTKo>(T, e, m, m0, 0, tid, ra);
    load_b<TM, TN, kF32, kTKo>(T, e, n, n0, 0, tid, rb);
    store_a<TM, TN, kF32, kTKo, LT>(T, tid, ra, As(0));
    store_b<TM, TN, kF32, kTKo, LT>(T, tid, rb, Bs(0));
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      const bool more = (kt + 1 < nk);
      if (more) {
        load_a<TM, TN, kF32, kTKo>(T, e, m, m0, (kt + 1) * TK, tid, ra);
        load_b<TM, TN, kF32, kTKo>(T, e, n, n0, (kt + 1) * TK, tid, rb);
      }
      auto as = As(cur);
      auto bs = Bs(cur);
      // fragments double-buffered across the k-steps of this LDS stage
      LT fa[2][2], fb[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i) fa[0][i] = as[lane >> 4][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) fb[0][j] = bs[lane >> 4][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int ks = 0; ks < TK / 4; ++ks) {
        // 3M: no fragment prefetch (its registers hold the third accumulator set)
        const int c = k3M ? 0 : (ks & 1);
        if constexpr (k3M) {
          if (ks > 0) {
            const int kk = ks * 4 + (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; ++i) fa[0][i] = as[kk][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) fb[0][j] = bs[kk][wn * 32 + j * 16 + (lane & 15)];
          }
        } else if (ks + 1 < TK / 4) {
          const int kk = (ks + 1) * 4 + (lane >> 4);
#pragma unroll
          for (int i = 0; i < 2; ++i) fa[c ^ 1][i] = as[kk][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
          for (int j = 0; j < 2; ++j) fb[c ^ 1][j] = bs[kk][wn * 32 + j * 16 + (lane & 15)];
        }
        if constexpr (kF32) {
          float nai[2];
#pragma unroll
          for (int i = 0; i < 2; ++i) nai[i] = -fa[c][i].y;
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
              accr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[c][i].x, fb[c][j].x, accr[i][j], 0, 0, 0);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
              acci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[c][i].x, fb[c][j].y, acci[i][j], 0, 0, 0);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
              accr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(nai[i], fb[c][j].y, accr[i][j], 0, 0, 0);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
              acci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[c][i].y, fb[c][j].x, acci[i][j], 0, 0, 0);
        } else if constexpr (k3M) {
          double sa[2], sb[2];
#pragma unroll
          for (int i = 0; i < 2; ++i) sa[i] = fa[c][i].x + fa[c][i].y;
#pragma unroll
          for (int j = 0; j < 2; ++j) sb[j] = fb[c][j].x + fb[c][j].y;
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) accr[i][j] = mfma(fa[c][i].x, fb[c][j].x, accr[i][j]);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acci[i][j] = mfma(fa[c][i].y, fb[c][j].y, acci[i][j]);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) accs[i][j] = mfma(sa[i], sb[j], accs[i][j]);
        } else {
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) accr[i][j] = mfma(fa[c][i].x, fb[c][j].x, accr[i][j]);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acci[i][j] = mfma(fa[c][i].x, fb[c][j].y, acci[i][j]);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) accr[i][j] = mfma_nega(fa[c][i].y, fb[c][j].y, accr[i][j]);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acci[i][j] = mfma(fa[c][i].y, fb[c][j].x, acci[i][j]);
        }
      }
      if (more) {
        store_a<TM, TN, kF32, kTKo, LT>(T, tid, ra, As(cur ^ 1));
        store_b<TM, TN, kF32, kTKo, LT>(T, tid, rb, Bs(cur ^ 1));
      }
      __syncthreads();
    }
  }

  // ---- epilogue: one 32-row slab (the rows of one wave row) at a time,
  // staged through LDS so that global accesses are row-coalesced ----
  // For v_mfma_f64_16x16x4f64 register r of a lane holds D[4*r + lane/16][lane%16]
  // of each 16x16 sub-tile (probed on gfx950).
  // (v_mfma_f32_16x16x4f32: register r of lane l holds D[4*(l/16) + r][l%16].)
  constexpr int kPer = kSlab * TN / S::kThreads;
  static_assert(kPer % 2 == 0, "addends are applied in pairs");
  auto Cs = reinterpret_cast<double2(*)[TN + 1]>(smem_raw);
  const int64_t eo = (int64_t)e;
  double2* D = P.D + eo * P.sD;
  double2* D2 = P.D2 ? P.D2 + eo * P.sD2 : nullptr;
  constexpr int kSlabsPerWaveRow = 32 / kSlab;  // 1, or 2 for the 32x32 tile
#pragma unroll 1
  for (int h = 0; h < TM / kSlab; ++h) {
    const int r0 = m0 + h * kSlab;  // first global row of the slab
    if (wm == h / kSlabsPerWaveRow) {
      const int isub = h % kSlabsPerWaveRow;  // sub-tile row held by a half-row slab
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (kSlabsPerWaveRow == 2 && i != isub) continue;
        const int ri = kSlabsPerWaveRow == 1 ? i * 16 : 0;
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rr = kF32 ? ((lane >> 4) * 4 + r) : (r * 4 + (lane >> 4));
            if constexpr (k3M) {
              const double p1 = accr[i][j][r], p2 = acci[i][j][r];
              Cs[ri + rr][wn * 32 + j * 16 + (lane & 15)] = cmk(p1 - p2, accs[i][j][r] - p1 - p2);
            } else {
              Cs[ri + rr][wn * 32 + j * 16 + (lane & 15)] =
                  cmk((double)accr[i][j][r], (double)acci[i][j][r]);
            }
          }
      }
    }
    __syncthreads();
    // Each thread owns elements idx = tid + q*kThreads of the slab (row-major,
    // so global accesses are coalesced along rows); values live in Cs and are
    // updated in pairs: loads of a pair are issued before its stores (pairs,
    // not larger groups, keep the 64x64 kernel within 128 VGPRs unspilled).
    for (int a = 0; a <= P.naddends; ++a) {
      if (a == P.npre && P.d2mode == kD2Partial) {
#pragma unroll 4
        for (int q = 0; q < kPer; ++q) {
          const int idx = tid + q * S::kThreads;
          const int row = r0 + idx / TN, col = n0 + idx % TN;
          if (row < m && col < n) D2[(int64_t)row * P.ldd2 + col] = Cs[idx / TN][idx % TN];
        }
      }
      if (a == P.naddends) break;
      const ZAddend& A = P.ad[a];
      const double2* F = A.F + eo * A.sF;
      const bool use1 = cnz(A.c1), use2 = cnz(A.c2);
#pragma unroll 1
      for (int q0 = 0; q0 < kPer; q0 += 2) {
        double2 f1[2], f2[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int idx = tid + (q0 + u) * S::kThreads;
          const int row = r0 + idx / TN, col = n0 + idx % TN;
          const bool in = row < m && col < n;
          f1[u] = (use1 && in) ? F[(int64_t)row * A.ldf + col] : cmk(0.0, 0.0);
          f2[u] = (use2 && in) ? F[(int64_t)col * A.ldf + row] : cmk(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int idx = tid + (q0 + u) * S::kThreads;
          double2 v = Cs[idx / TN][idx % TN];
          v = cadd(v, cadd(cmul(A.c1, f1[u]), cmul(A.c2, cconj(f2[u]))));
          Cs[idx / TN][idx % TN] = v;
        }
      }
    }
#pragma unroll 4
    for (int q = 0; q < kPer; ++q) {
      const int idx = tid + q * S::kThreads;
      const int row = r0 + idx / TN, col = n0 + idx % TN;
      if (row < m && col < n) {
        const double2 v = Cs[idx / TN][idx % TN];
        D[(int64_t)row * P.ldd + col] = v;
        if (P.d2mode == kD2Copy) D2[(int64_t)row * P.ldd2 + col] = cmul(P.s2, v);
      }
    }
    if (P.d2mode == kD2ConjT) {
      // D2[c][r] = s2 conj(D[r][c]): read the slab transposed from LDS, store
      // runs of kSlab consecutive elements of rows of D2.  The transposed read
      // takes elements that other waves finished in the addend pass: wait for
      // them (the condition is uniform over the workgroup).
      __syncthreads();
#pragma unroll 4
      for (int q = 0; q < kPer; ++q) {
        const int idx = tid + q * S::kThreads;
        const int tr = idx / kSlab, tc = idx % kSlab;
        const int row2 = n0 + tr, col2 = r0 + tc;
        if (row2 < n && col2 < m) D2[(int64_t)row2 * P.ldd2 + col2] = cmul(P.s2, cconj(Cs[tc][tr]));
      }
    }
    __syncthreads();  // the next slab overwrites Cs
  }
}

template <int TM, int TN, int kRole, bool kF32 = false, bool k3M = false, int kTKo = 0>
hipError_t launch_cfg(const ZGroup& g0, hipStream_t stream) {
  ZGroup g = g0;
  int per = 0;  // workgroups per energy (all problems)
  for (int p = 0; p < g.nprob; ++p) {
    const int tm = (g.p[p].m + TM - 1) / TM, tn = (g.p[p].n + TN - 1) / TN;
    g.tiles_n[p] = tn;
    g.tiles_mn[p] = tm * tn;
    g.wg_base[p] = per;
    per += tm * tn;
  }
  g.wg_base[g.nprob] = per;
  const int64_t total = (int64_t)per * g.nE;
  if (total == 0) return hipSuccess;
  if (total > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL((zgemm_grouped_kernel<TM, TN, kRole, kF32, k3M, kTKo>), dim3((unsigned)total),
                     dim3(TileShape<TM, TN, kF32, kTKo>::kThreads), 0, stream, g);
  return hipGetLastError();
}

}  // namespace

static_assert(sizeof(ZGroup) <= 4000, "kernel argument block too large");

int tile_m(int cfg) {
  switch (cfg) {
    case kTile32x32: case kTile32x32_3M: return 32;
    case kTile64x32: return 64;
    case kTile128x64: case kTile128x64_3M: return 128;
    default: return 64;
  }
}
int tile_n(int cfg) {
  switch (cfg) {
    case kTile32x32: case kTile32x32_3M: return 32;
    case kTile64x32: return 32;
    case kTile128x64: case kTile128x64_3M: return 64;
    default: return 64;
  }
}

hipError_t launch_zgemm(const ZGroup& g, int cfg, hipStream_t stream) {
  switch (cfg) {
    case kTile32x32: return launch_cfg<32, 32, 0>(g, stream);
    case kTile64x32: return launch_cfg<64, 32, 0>(g, stream);
    case kTile128x64: return launch_cfg<128, 64, 0>(g, stream);
    case kTile64x64_3M: return launch_cfg<64, 64, 0, false, true>(g, stream);
    case kTile32x32_3M: return launch_cfg<32, 32, 0, false, true>(g, stream);
    case kTile128x64_3M: return launch_cfg<128, 64, 0, false, true, 16>(g, stream);
    default: return launch_cfg<64, 64, 0>(g, stream);
  }
}

hipError_t launch_zgemm_f32(const ZGroup& g, int cfg, hipStream_t stream) {
  switch (cfg) {
    case kTile32x32: return launch_cfg<32, 32, 0, true>(g, stream);
    case kTile128x64: return launch_cfg<128, 64, 0, true>(g, stream);
    default: return launch_cfg<64, 64, 0, true>(g, stream);
  }
}

hipError_t launch_zgemm_update(const ZGroup& g, int cfg, hipStream_t stream) {
  switch (cfg) {
    case kTile32x32: return launch_cfg<32, 32, 1>(g, stream);
    case kTile32x32_3M: return launch_cfg<32, 32, 1, false, true>(g, stream);
    case kTile64x64_3M: case kTile128x64_3M: return launch_cfg<64, 64, 1, false, true>(g, stream);
    default: return launch_cfg<64, 64, 1>(g, stream);
  }
}

}  // namespace qx
