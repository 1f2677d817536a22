// Shared device helpers and launch descriptors for the RGF kernels (gfx950).
//
// Complex numbers are interleaved (re, im) doubles, i.e. the memory layout of
// numpy/torch complex128, handled as HIP double2.  Matrices are row-major with
// a leading dimension and a per-energy stride (in complex elements).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qx {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__host__ __device__ inline double2 cmk(double re, double im) { return make_double2(re, im); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return cmk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return cmk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cconj(double2 a) { return cmk(a.x, -a.y); }
__host__ __device__ inline bool cnz(double2 a) { return a.x != 0.0 || a.y != 0.0; }
// 1/d via conj(d)/|d|^2 with a power-of-two-free scaling by max(|re|,|im|).
__host__ __device__ inline double2 crecip(double2 d) {
  double s = fmax(fabs(d.x), fabs(d.y));
  double xr = d.x / s, xi = d.y / s;
  double den = s * (xr * xr + xi * xi);
  return cmk(xr / den, -xi / den);
}

// 1/d for a pivot: conj(d) / |d|^2 with the reciprocal of |d|^2 from
// v_rcp_f64 and two Newton steps; outside the range where |d|^2 neither
// overflows nor loses precision the scaled division of crecip.  Callers pass
// the same d in every thread, so the branch is uniform.
static __device__ __noinline__ double2 crecip_slow(double2 d) { return crecip(d); }
static __device__ inline double2 pivot_recip(double2 d) {
  const double m = d.x * d.x + d.y * d.y;
  if (!(m > 1e-280 && m < 1e280)) return crecip_slow(d);
  double r = __builtin_amdgcn_rcp(m);
  r = fma(r, fma(-m, r, 1.0), r);
  r = fma(r, fma(-m, r, 1.0), r);
  return cmk(d.x * r, -d.y * r);
}
// The same with the out-of-range case inline (no call: a call makes the
// caller save its live registers around it).
static __device__ inline double2 pivot_recip_inline(double2 d) {
  const double m = d.x * d.x + d.y * d.y;
  if (!(m > 1e-280 && m < 1e280)) return crecip(d);
  double r = __builtin_amdgcn_rcp(m);
  r = fma(r, fma(-m, r, 1.0), r);
  r = fma(r, fma(-m, r, 1.0), r);
  return cmk(d.x * r, -d.y * r);
}

// Value of lane `src` (uniform) of this wave.
static __device__ inline double readlane_d(double x, int src) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

enum Op : int32_t { kOpN = 0, kOpC = 1 };
enum D2Mode : int32_t { kD2None = 0, kD2Partial = 1, kD2Copy = 2, kD2ConjT = 3 };

constexpr int kMaxProblems = 6;
constexpr int kMaxAddends = 3;

// One product  alpha * op(A) * op(B)  of a problem (op = N or conjugate transpose).
struct ZTerm {
  const double2* A;
  const double2* B;
  int64_t sA, sB;  // per-energy strides
  int32_t lda, ldb, opA, opB, k, pad_;
  double2 alpha;
};

// Elementwise addend  c1 * F[r][c] + c2 * conj(F[c][r]).
struct ZAddend {
  const double2* F;
  int64_t sF;
  int32_t ldf, pad_;
  double2 c1, c2;
};

// D = sum_t term_t + sum_a addend_a   (m x n), with an optional second output:
//   kD2Partial: D2 = value after the first `npre` addends,
//   kD2Copy:    D2 = s2 * D,
//   kD2ConjT:   D2^T = s2 * conj(D)  (i.e. D2[c][r] = s2 * conj(D[r][c])).
struct ZProblem {
  int32_t m, n, nterms, naddends, npre, d2mode;
  ZTerm t[2];
  ZAddend ad[kMaxAddends];
  double2* D;
  int64_t sD;
  int32_t ldd, ldd2;
  double2* D2;
  int64_t sD2;
  double2 s2;
};

// Up to kMaxProblems independent problems, each over nE energies, in one launch.
struct ZGroup {
  int32_t nprob, nE;
  int32_t wg_base[kMaxProblems + 1];  // cumulative tile counts within one energy
  int32_t tiles_n[kMaxProblems];
  int32_t tiles_mn[kMaxProblems];
  ZProblem p[kMaxProblems];
};

// Tile configurations selectable per launch (keyed by block size by the autotuner).
enum TileCfg : int32_t {
  kTile64x64 = 0,  // 256 threads, 4 waves of 32x32
  kTile32x32 = 1,  // 64 threads, 1 wave of 32x32
  kTile64x32 = 2,  // 128 threads, 2 waves of 32x32
  kTile128x64 = 3, // 512 threads, 8 waves of 32x32
  // the same tiles with three real MFMAs per complex product (3M)
  kTile64x64_3M = 4,
  kTile32x32_3M = 5,
  kTile128x64_3M = 6,
  kNumTileCfgs = 7
};
int tile_m(int cfg);
int tile_n(int cfg);
// Configurations the FP32-compute products are built for (the others fall
// back to 64x64 there).
inline bool f32_tile_ok(int cfg) {
  return cfg == kTile64x64 || cfg == kTile32x32 || cfg == kTile128x64;
}

// XCD-aware remap: hardware dispatches consecutive workgroups round-robin over
// the 8 XCDs; give each XCD a contiguous range of logical ids so tiles of one
// energy's matrices share that XCD's L2.
__device__ inline int xcd_remap(int w, int total) {
  const int q = total >> 3, r = total & 7;
  const int x = w & 7, s = w >> 3;
  return x * q + (x < r ? x : r) + s;
}

// Host-side launchers (defined in the .hip translation units).
hipError_t launch_zgemm(const ZGroup& g, int cfg, hipStream_t stream);
hipError_t launch_zgemm_update(const ZGroup& g, int cfg, hipStream_t stream);
// Same problems, products computed in complex64 on FP32 matrix cores.
hipError_t launch_zgemm_f32(const ZGroup& g, int cfg, hipStream_t stream);

// Batched Gauss-Jordan inversion with partial pivoting, in place in S
// (n x n, ld n, energy stride sS); result (column-permuted back) written to
// Out (ld ldo, stride sOut).  aux must hold zinv_aux_bytes(n) per energy.
size_t zinv_aux_bytes(int n);
hipError_t launch_zinv(double2* S, int64_t sS, int n, int nE, double2* Out, int64_t sOut,
                       int ldo, void* aux, int* info, int info_tag, int64_t e0, int tile_cfg,
                       hipStream_t stream);

// In-place anti-Hermitian projection X <- (X - X^H)/2 of nblk diagonal blocks.
hipError_t launch_project(double2* base, int64_t sE, const int64_t* offs_dev, const int* sizes_dev,
                          int nblk, int bmax, int nE, hipStream_t stream);

// Meir-Wingreen current traces.
struct TraceJob {
  const double2* X1; const double2* Y1;  // + tr(X1 Y1)
  const double2* X2; const double2* Y2;  // - tr(X2 Y2)
  int64_t sX1, sY1, sX2, sY2;
  int32_t b, slot;
  double sign;
};
hipError_t launch_traces(const TraceJob* jobs_dev, int njobs, int nE, double2* out, int64_t sOut,
                         hipStream_t stream);

// Frobenius norms of a list of matrices, per energy.
struct NormJob {
  const double2* X;
  int64_t sX;
  int32_t rows, cols, ld, pad_;
};
hipError_t launch_norms(const NormJob* jobs_dev, int njobs, int nE, double* out, int64_t sOut,
                        hipStream_t stream);
// Host tables (geometry offsets, trace / norm job lists) reach the device as
// kernel arguments: launch_store_bytes copies `bytes` from host memory to
// dst (8-byte aligned; the last word is written whole, so reserve bytes
// rounded up to 8) through by-value payloads of up to 2 KiB per launch.  No
// host buffer has to outlive the call and no host synchronisation is needed,
// so the copy also works inside a captured graph.
constexpr int kPayloadWords = 256;
struct WordPayload {
  uint64_t w[kPayloadWords];
};
hipError_t launch_store_bytes(void* dst, const void* src, size_t bytes, hipStream_t stream);
// n x n identity (one matrix; use energy stride 0 to share it).
hipError_t launch_identity(double2* X, int n, hipStream_t stream);

// System-matrix assembly A = z_e I - H - S (per energy e; H and S shared).
hipError_t launch_assemble(const double2* h_diag, const double2* h_upper, const double2* h_lower,
                           int64_t Dsz, int64_t Usz, int nE, const double2* z, const int64_t* offs,
                           const int* sizes, const double2* const* sig, int nblk, int bmax,
                           double2* a_diag, double2* a_upper, double2* a_lower, hipStream_t stream);

// Precision conversion helpers for the mixed path.
hipError_t launch_z2c(const double2* src, int64_t sSrc, float2* dst, int64_t sDst, int64_t count,
                      int nE, hipStream_t stream);
hipError_t launch_c2z(const float2* src, int64_t sSrc, double2* dst, int64_t sDst, int64_t count,
                      int nE, hipStream_t stream);

// Persistent on-chip selected solve for geometries whose blocks are all at
// most kSmallMaxBlock (sweep_small.hip).  One problem = one geometry with
// its energy batch and packed buffers (energy strides Dsz / Usz; block
// tables on the device); jobs job0 .. job0 + nE - 1 of the work queue are its
// energies.  Per-workgroup scratch holds the left-connected g_i when the
// retarded output is absent and the Sigma~ blocks of the current.
constexpr int kSmallMaxBlock = 32;
struct SmallProblem {
  int32_t N, nE, job0, pad_;
  int64_t e_base, Dsz, Usz;
  const int* bs;
  const int64_t* od;
  const int64_t* ou;
  const double2 *Ad, *Au, *Al, *SLd, *SLu, *SGd, *SGu;
  double2 *GRd, *GRu, *GRl, *GLd, *GLu, *GLl, *GGd, *GGu, *GGl;
  const double2* const* obr;
  const double2* const* obl;
  const double2* const* obg;
  double2* cur;  // [nE][N+1] or null
};
struct SmallLaunch {
  int32_t nprob, total_jobs;
  int32_t bmax, pad_;  // largest block over the problems (<= 16: half-depth products)
  int* counter;  // work-queue head, reset by the launcher
  int* info;     // singular-pivot status word
  char* scratch;
  int64_t scratch_per_wg;
  const SmallProblem* probs;  // device array of nprob problems
};
size_t sweep_small_scratch_per_wg(int n_blocks, bool need_g, bool current);
hipError_t launch_sweep_small(const SmallLaunch& L, int grid, hipStream_t stream);

// Block-sparse <-> packed tridiagonal (blocks.hip).  Entry k of a map moves
// data[e][src[k]] to out[e][dst[k]] (densify) or back (sparsify, stored
// entries only); op flags in blocks.hip.
hipError_t launch_fill_zero(double2* p, int64_t count, int64_t stride, int nE, hipStream_t s);
hipError_t launch_densify(const double2* data, int64_t sData, int64_t n, const int64_t* src,
                          const int64_t* dst, const uint8_t* op, double2* out, int64_t sOut,
                          int nE, hipStream_t s);
hipError_t launch_sparsify(const double2* in, int64_t sIn, int64_t n, const int64_t* src,
                           const int64_t* dst, const uint8_t* op, double2* data, int64_t sData,
                           int nE, hipStream_t s);

}  // namespace qx
