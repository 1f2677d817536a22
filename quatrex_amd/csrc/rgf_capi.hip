// C ABI and host-side sweep driver for the RGF selected inversion / solve.
//
// The driver turns one call into a sequence of energy-batched launches on the
// context stream: grouped complex GEMMs (zgemm.hip), blocked Gauss-Jordan
// inversions (zinv.hip) and the projection / current kernels (aux.hip).  All
// energies of a chunk advance through the same block step together, so every
// launch covers (problems x energies x tiles) workgroups.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <shared_mutex>
#include <string>
#include <vector>

#include "../../include/rgf.h"
#include "qx_common.h"
#include "rgf_driver.h"

using namespace qx;

namespace qx {

const char* status_name(int s) {
  switch (s) {
    case RGF_OK: return "ok";
    case RGF_EINVAL: return "invalid argument";
    case RGF_ESINGULAR: return "singular matrix";
    case RGF_EHIP: return "HIP error";
    case RGF_ENOMEM: return "out of device memory";
    default: return "unsupported";
  }
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

Ctx::~Ctx() {
  if (side) { (void)hipStreamSynchronize(side); (void)hipStreamDestroy(side); }
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  for (auto& t : timed) { event_pool.push_back(t.a); event_pool.push_back(t.b); }
  for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
  if (ws_owned && ws) (void)hipFree(ws);
  if (meta) (void)hipFree(meta);
  if (info_dev) (void)hipFree(info_dev);
}

int Ctx::fail(int code, const std::string& msg) {
  last_error = msg;
  return code;
}

int Ctx::hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return RGF_OK;
  last_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? RGF_ENOMEM : RGF_EHIP;
}

int Ctx::ensure_workspace(size_t bytes, char** out) {
  if (bytes <= ws_bytes && ws) {
    *out = static_cast<char*>(ws);
    return RGF_OK;
  }
  if (!ws_owned && ws) {
    return fail(RGF_ENOMEM, "caller workspace too small: need " + std::to_string(bytes) +
                                " bytes, have " + std::to_string(ws_bytes));
  }
  if (ws) {
    // the old buffer may still be in use by queued work
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip(e, "hipStreamSynchronize");
    (void)hipFree(ws);
    ws = nullptr;
    ws_bytes = 0;
  }
  hipError_t e = hipMalloc(&ws, bytes);
  if (e != hipSuccess) {
    ws = nullptr;
    return hip(e, "hipMalloc(workspace)");
  }
  ws_bytes = bytes;
  ws_owned = true;
  *out = static_cast<char*>(ws);
  return RGF_OK;
}

int Ctx::ensure_meta(size_t bytes) {
  if (bytes <= meta_bytes && meta) return RGF_OK;
  if (meta) {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip(e, "hipStreamSynchronize");
    (void)hipFree(meta);
    meta = nullptr;
  }
  size_t want = std::max(bytes, (size_t)1 << 16);
  hipError_t e = hipMalloc(&meta, want);
  if (e != hipSuccess) {
    meta = nullptr;
    return hip(e, "hipMalloc(meta)");
  }
  meta_bytes = want;
  return RGF_OK;
}

// Untuned tile: the one the autotuner picks at production batch sizes on
// MI355X (3M FP64 products; the FP32 products have no 3M variant).
static int default_tile(int b, bool f32) {
  if (f32) return b < 64 ? kTile32x32 : kTile64x64;
  return b < 64 ? kTile32x32_3M : kTile64x64_3M;
}

int Ctx::tile_for(int b) const {
  if (tile_override >= 0) return tile_override;
  const std::vector<int>& t = prec32 ? tile_table32 : tile_table;
  if (b >= 0 && b < (int)t.size() && t[b] >= 0) return t[b];
  return default_tile(b, prec32);
}

// ---------------------------------------------------------------------------
// Geometry
// ---------------------------------------------------------------------------

int Geo::init(const rgf_geometry* g, std::string* err) {
  if (!g || g->n_blocks < 1 || !g->block_sizes) {
    *err = "geometry: n_blocks >= 1 and block_sizes required";
    return RGF_EINVAL;
  }
  N = g->n_blocks;
  bs.assign(g->block_sizes, g->block_sizes + N);
  od.assign(N + 1, 0);
  ou.assign(N, 0);
  bmax = 0;
  for (int i = 0; i < N; ++i) {
    if (bs[i] < 1 || bs[i] > 1024) {
      *err = "geometry: block sizes must be in [1, 1024]";
      return RGF_EUNSUPPORTED;
    }
    od[i + 1] = od[i] + bs[i] * bs[i];
    if (i + 1 < N) ou[i + 1] = ou[i] + bs[i] * bs[i + 1];
    bmax = std::max<int64_t>(bmax, bs[i]);
  }
  Dsz = od[N];
  Usz = N > 1 ? ou[N - 1] : 0;
  return RGF_OK;
}

// ---------------------------------------------------------------------------
// Grouped GEMM helper
// ---------------------------------------------------------------------------

PB::PB(int m, int n) {
  std::memset(&P, 0, sizeof(P));
  P.m = m;
  P.n = n;
  P.s2 = cmk(1.0, 0.0);
  P.npre = -1;
}
PB& PB::term(Mat A, int opA, Mat B, int opB, int k, double2 alpha) {
  ZTerm& t = P.t[P.nterms++];
  t.A = A.p; t.sA = A.s; t.lda = A.ld; t.opA = opA;
  t.B = B.p; t.sB = B.s; t.ldb = B.ld; t.opB = opB;
  t.k = k; t.alpha = alpha;
  return *this;
}
PB& PB::add(Mat F, double2 c1, double2 c2) {
  if (!F.p) return *this;
  ZAddend& a = P.ad[P.naddends++];
  a.F = F.p; a.sF = F.s; a.ldf = F.ld; a.c1 = c1; a.c2 = c2;
  return *this;
}
PB& PB::out(Mat D) {
  P.D = D.p; P.sD = D.s; P.ldd = D.ld;
  return *this;
}
PB& PB::out2(int mode, Mat D2, double2 s2) {
  if (!D2.p) return *this;
  P.d2mode = mode; P.D2 = D2.p; P.sD2 = D2.s; P.ldd2 = D2.ld; P.s2 = s2;
  return *this;
}
PB& PB::partial_here() {
  P.npre = P.naddends;
  return *this;
}
ZProblem PB::done() const {
  ZProblem q = P;
  if (q.npre < 0) q.npre = q.naddends;
  if (q.d2mode != kD2Partial && q.npre > q.naddends) q.npre = q.naddends;
  return q;
}

hipEvent_t Ctx::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int Ctx::prof_begin(hipEvent_t* a, hipStream_t s) {
  *a = nullptr;
  if (!profile) return RGF_OK;
  *a = take_event();
  if (!*a) return fail(RGF_EHIP, "hipEventCreate");
  return hip(hipEventRecord(*a, s ? s : stream), "hipEventRecord");
}

int Ctx::prof_end(hipEvent_t a, double flops, int kind, hipStream_t s, double bytes) {
  if (!profile || !a) return RGF_OK;
  hipEvent_t b = take_event();
  if (!b) return fail(RGF_EHIP, "hipEventCreate");
  int rc = hip(hipEventRecord(b, s ? s : stream), "hipEventRecord");
  if (rc) return rc;
  timed.push_back(Timed{a, b, flops, bytes, kind});
  if (timed.size() >= 4096) return prof_collect();
  return RGF_OK;
}

int Ctx::prof_collect() {
  for (auto& t : timed) {
    hipError_t e = hipEventSynchronize(t.b);
    if (e != hipSuccess) return hip(e, "hipEventSynchronize");
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, t.a, t.b);
    if (e != hipSuccess) return hip(e, "hipEventElapsedTime");
    if (t.kind == 0) {
      prof_gemm_launches++;
      prof_gemm_ms += ms;
      prof_gemm_flops += t.flops;
      prof_gemm_bytes += t.bytes;
    }
    else { prof_inv_calls++; prof_inv_ms += ms; prof_inv_flops += t.flops; }
    event_pool.push_back(t.a);
    event_pool.push_back(t.b);
  }
  timed.clear();
  return RGF_OK;
}

static double group_flops(const ZGroup& g) {
  double f = 0;
  for (int p = 0; p < g.nprob; ++p)
    for (int t = 0; t < g.p[p].nterms; ++t)
      f += 8.0 * g.p[p].m * (double)g.p[p].n * g.p[p].t[t].k;
  return f * g.nE;
}

// Algorithmic HBM bytes of one launch: every problem reads op(A) and op(B)
// of each term and each addend once and writes D (and D2) once.
static double group_bytes(const ZGroup& g) {
  double el = 0;
  for (int p = 0; p < g.nprob; ++p) {
    const ZProblem& P = g.p[p];
    const double mn = (double)P.m * P.n;
    for (int t = 0; t < P.nterms; ++t) el += (double)P.t[t].k * (P.m + P.n);
    el += mn * (P.naddends + 1 + (P.d2mode != kD2None ? 1 : 0));
  }
  return el * sizeof(double2) * g.nE;
}

int Ctx::gemm(const std::vector<ZProblem>& probs, int nE, int b) {
  size_t i = 0;
  while (i < probs.size()) {
    ZGroup g;
    std::memset(&g, 0, sizeof(g));
    g.nE = nE;
    int np = 0;
    const int cap = unfused ? 1 : kMaxProblems;
    while (i < probs.size() && np < cap) {
      if (probs[i].D) g.p[np++] = probs[i];
      ++i;
    }
    g.nprob = np;
    if (np == 0) continue;
    hipEvent_t ev;
    int rc = prof_begin(&ev);
    if (rc) return rc;
    rc = hip(prec32 ? launch_zgemm_f32(g, tile_for(b), stream) : launch_zgemm(g, tile_for(b), stream),
             "zgemm launch");
    if (rc) return rc;
    rc = prof_end(ev, group_flops(g), 0, nullptr, group_bytes(g));
    if (rc) return rc;
  }
  return RGF_OK;
}

int Ctx::inv(double2* S, int64_t sS, int n, int nE, Mat out, char* aux, int tag, int64_t e0,
             hipStream_t on) {
  hipStream_t s = on ? on : stream;
  hipEvent_t ev;
  int rc = prof_begin(&ev, s);
  if (rc) return rc;
  rc = hip(launch_zinv(S, sS, n, nE, out.p, out.s, out.ld, aux, info_dev, tag, e0, tile_for(n), s),
           "zinv launch");
  if (rc) return rc;
  return prof_end(ev, 8.0 * n * (double)n * n * nE, 1, s);
}

int Ctx::fork() {
  if (!side) {
    int rc = hip(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "side stream");
    if (!rc) rc = hip(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "event");
    if (!rc) rc = hip(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "event");
    if (rc) return rc;
  }
  int rc = hip(hipEventRecord(ev_fork, stream), "fork record");
  if (rc) return rc;
  return hip(hipStreamWaitEvent(side, ev_fork, 0), "fork wait");
}

int Ctx::join() {
  int rc = hip(hipEventRecord(ev_join, side), "join record");
  if (rc) return rc;
  return hip(hipStreamWaitEvent(stream, ev_join, 0), "join wait");
}

}  // namespace qx

// ---------------------------------------------------------------------------
// Sweeps (FP64)
// ---------------------------------------------------------------------------

namespace {

constexpr double2 ONE = {1.0, 0.0};
constexpr double2 MONE = {-1.0, 0.0};
constexpr double2 ZERO = {0.0, 0.0};

struct Work {
  double2* slots = nullptr;  // [E][kSlots][bmax^2]
  int64_t slot_elems = 0;
  int64_t sE = 0;
  double2* gr = nullptr;     // [E][Dsz] retarded diagonal storage when not an output
  double2* stl = nullptr;    // [E][Dsz] Sigma~^< per block (current)
  double2* stg = nullptr;
  char* aux = nullptr;       // [E][aux_bytes]
  // mixed precision only
  double2* ident = nullptr;  // bmax x bmax identity, shared by all energies (stride 0)
  double2* gru = nullptr;    // [E][Usz] off-diagonal G^R when the caller does not want it
  double2* grl = nullptr;
  double* norms = nullptr;   // [E][kMaxNormJobs]
  Mat slot(int q) const { return Mat{slots + q * slot_elems, sE, 0}; }
};

constexpr int kMaxNormJobs = 64;

constexpr int kSlots = 11;

// Table bytes of one problem of the persistent sweep: its descriptor, block
// sizes and offsets, and the three OBC pointer lists.
size_t small_table_bytes(const Geo& g) {
  auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return r(sizeof(SmallProblem)) + r(g.N * sizeof(int)) + 2 * r(g.N * sizeof(int64_t)) +
         r(3 * g.N * sizeof(void*));
}

size_t work_bytes(const Geo& g, int64_t E, bool solve, bool current, bool have_gr,
                  bool mixed = false, bool have_gr_off = true) {
  const size_t bb = (size_t)g.bmax * g.bmax * sizeof(double2);
  size_t per = (solve ? kSlots : 4) * bb + zinv_aux_bytes((int)g.bmax);
  if (!have_gr) per += g.Dsz * sizeof(double2);
  if (current) per += 2 * g.Dsz * sizeof(double2);
  size_t fixed = 256;
  if (mixed) {
    fixed += bb + 256;
    per += kMaxNormJobs * sizeof(double);
    if (!have_gr_off) per += 2 * g.Usz * sizeof(double2);
  }
  return per * (size_t)E + fixed;
}

void carve(const Geo& g, int64_t E, bool solve, bool current, bool have_gr, char* base, Work* w,
           bool mixed = false, bool have_gr_off = true) {
  const int64_t bb = g.bmax * g.bmax;
  const int nsl = solve ? kSlots : 4;
  char* p = base;
  w->slots = reinterpret_cast<double2*>(p);
  w->slot_elems = bb;
  w->sE = nsl * bb;
  p += (size_t)E * nsl * bb * sizeof(double2);
  if (!have_gr) {
    w->gr = reinterpret_cast<double2*>(p);
    p += (size_t)E * g.Dsz * sizeof(double2);
  }
  if (current) {
    w->stl = reinterpret_cast<double2*>(p);
    p += (size_t)E * g.Dsz * sizeof(double2);
    w->stg = reinterpret_cast<double2*>(p);
    p += (size_t)E * g.Dsz * sizeof(double2);
  }
  w->aux = p;
  p += (size_t)E * zinv_aux_bytes((int)g.bmax);
  if (mixed) {
    p = base + (((size_t)(p - base) + 255) & ~(size_t)255);
    w->ident = reinterpret_cast<double2*>(p);
    p += (size_t)g.bmax * g.bmax * sizeof(double2);
    p = base + (((size_t)(p - base) + 255) & ~(size_t)255);
    w->norms = reinterpret_cast<double*>(p);
    p += (size_t)E * kMaxNormJobs * sizeof(double);
    if (!have_gr_off) {
      w->gru = reinterpret_cast<double2*>(p);
      p += (size_t)E * g.Usz * sizeof(double2);
      w->grl = reinterpret_cast<double2*>(p);
      p += (size_t)E * g.Usz * sizeof(double2);
    }
  }
}

// Energy-offset views of packed operands (each part has its own energy base:
// caller arrays are indexed by global energy, workspace arrays by chunk).
struct BTV {
  const Geo* g;
  double2* d;
  double2* u;
  double2* l;
  int64_t e0;
  int64_t eu = -1, el = -1;  // -1: same as e0
  int64_t bu() const { return eu < 0 ? e0 : eu; }
  int64_t bl() const { return el < 0 ? e0 : el; }
  Mat D(int i) const {
    return d ? Mat{d + e0 * g->Dsz + g->od[i], g->Dsz, (int)g->bs[i]} : Mat{};
  }
  Mat U(int i) const {
    return u ? Mat{u + bu() * g->Usz + g->ou[i], g->Usz, (int)g->bs[i + 1]} : Mat{};
  }
  Mat L(int i) const {
    return l ? Mat{l + bl() * g->Usz + g->ou[i], g->Usz, (int)g->bs[i]} : Mat{};
  }
};

BTV view(const Geo& g, const void* d, const void* u, const void* l, int64_t e0) {
  return BTV{&g, (double2*)d, (double2*)u, (double2*)l, e0, -1, -1};
}

Mat obc_mat(const void* const* arr, const Geo& g, int i, int64_t e0) {
  if (!arr || !arr[i]) return Mat{};
  const int64_t b = g.bs[i];
  return Mat{(double2*)arr[i] + e0 * b * b, b * b, (int)b};
}

Mat with_ld(Mat m, int ld) {
  m.ld = ld;
  return m;
}

// Forward elimination shared by selected_inv and selected_solve: for block j
// writes g_j = (A~_jj - A_ji g_i A_ij)^{-1} into GR(j).
int sweep_chunk_inv(Ctx* c, const Geo& g, int64_t e0, int E, const BTV& A,
                    const void* const* obcr, const BTV& X, const Work& w) {
  const int N = g.N;
  auto S = [&](int b) { return with_ld(w.slot(0), b); };
  auto P = [&](int b) { return with_ld(w.slot(1), b); };
  int rc;
  // block 0
  {
    const int b0 = (int)g.bs[0];
    rc = c->gemm({PB(b0, b0).add(A.D(0), ONE).add(obc_mat(obcr, g, 0, e0), MONE).out(S(b0)).done()},
                 E, b0);
    if (rc) return rc;
    rc = c->inv(S(b0).p, S(b0).s, b0, E, X.D(0), w.aux, 1, e0);
    if (rc) return rc;
  }
  for (int j = 1; j < N; ++j) {
    const int i = j - 1;
    const int bi = (int)g.bs[i], bj = (int)g.bs[j];
    rc = c->gemm({PB(bj, bi).term(A.L(i), kOpN, X.D(i), kOpN, bi).out(P(bi)).done()}, E, bj);
    if (rc) return rc;
    rc = c->gemm({PB(bj, bj)
                      .term(P(bi), kOpN, A.U(i), kOpN, bi, MONE)
                      .add(A.D(j), ONE)
                      .add(obc_mat(obcr, g, j, e0), MONE)
                      .out(S(bj))
                      .done()},
                 E, bj);
    if (rc) return rc;
    rc = c->inv(S(bj).p, S(bj).s, bj, E, X.D(j), w.aux, j + 1, e0);
    if (rc) return rc;
  }
  // backward (rgf.py:98-110)
  for (int i = N - 2; i >= 0; --i) {
    const int j = i + 1;
    const int bi = (int)g.bs[i], bj = (int)g.bs[j];
    Mat U = with_ld(w.slot(2), bj);  // g_i A_ij        (bi x bj)
    Mat R = with_ld(w.slot(3), bi);  // X_jj A_ji       (bj x bi)
    Mat Xu = X.U(i), Xl = X.L(i);
    Mat Xl_tmp = with_ld(w.slot(1), bi);
    if (!Xl.p) Xl = Xl_tmp;
    rc = c->gemm({PB(bi, bj).term(X.D(i), kOpN, A.U(i), kOpN, bi).out(U).done(),
                  PB(bj, bi).term(X.D(j), kOpN, A.L(i), kOpN, bj).out(R).done()},
                 E, bi);
    if (rc) return rc;
    std::vector<ZProblem> l2;
    l2.push_back(PB(bj, bi).term(R, kOpN, X.D(i), kOpN, bi, MONE).out(Xl).done());
    if (Xu.p) l2.push_back(PB(bi, bj).term(U, kOpN, X.D(j), kOpN, bj, MONE).out(Xu).done());
    rc = c->gemm(l2, E, bi);
    if (rc) return rc;
    rc = c->gemm({PB(bi, bi).term(U, kOpN, Xl, kOpN, bj, MONE).add(X.D(i), ONE).out(X.D(i)).done()},
                 E, bi);
    if (rc) return rc;
  }
  return RGF_OK;
}

struct SolveIO {
  BTV A, SL, SG;
  const rgf_obc* obc;
  BTV GR, GL, GG;
  double2* current;  // [E_total][N+1] or null
};

int sweep_chunk_solve(Ctx* c, const Geo& g, int64_t e0, int E, const SolveIO& io, const Work& w) {
  const int N = g.N;
  const BTV& A = io.A;
  const BTV& SL = io.SL;
  const BTV& SG = io.SG;
  const BTV& GL = io.GL;
  const BTV& GG = io.GG;
  BTV GR = io.GR;
  if (!GR.d) GR = BTV{&g, w.gr, GR.u, GR.l, 0, GR.u ? GR.bu() : -1, GR.l ? GR.bl() : -1};
  const void* const* obr = io.obc ? io.obc->retarded : nullptr;
  const void* const* obl = io.obc ? io.obc->lesser : nullptr;
  const void* const* obg = io.obc ? io.obc->greater : nullptr;
  auto OR = [&](int i) { return obc_mat(obr, g, i, e0); };
  auto OL = [&](int i) { return obc_mat(obl, g, i, e0); };
  auto OG = [&](int i) { return obc_mat(obg, g, i, e0); };
  auto ST = [&](double2* base, int i) {
    return base ? Mat{base + g.od[i], g.Dsz, (int)g.bs[i]} : Mat{};
  };
  auto sl = [&](int q, int ld) { return with_ld(w.slot(q), ld); };
  int rc;

  // ---- block 0 (rgf.py:216-239) ----
  {
    const int b = (int)g.bs[0];
    Mat S = sl(0, b), Vl = sl(1, b), Vg = sl(2, b);
    rc = c->gemm({PB(b, b).add(A.D(0), ONE).add(OR(0), MONE).out(S).done()}, E, b);
    if (rc) return rc;
    rc = c->inv(S.p, S.s, b, E, GR.D(0), w.aux, 1, e0);
    if (rc) return rc;
    PB vl(b, b), vg(b, b);
    vl.term(GR.D(0), kOpN, SL.D(0), kOpN, b);
    vg.term(GR.D(0), kOpN, SG.D(0), kOpN, b);
    if (OL(0).p) vl.term(GR.D(0), kOpN, OL(0), kOpN, b);
    if (OG(0).p) vg.term(GR.D(0), kOpN, OG(0), kOpN, b);
    rc = c->gemm({vl.out(Vl).done(), vg.out(Vg).done()}, E, b);
    if (rc) return rc;
    rc = c->gemm({PB(b, b).term(Vl, kOpN, GR.D(0), kOpC, b).out(GL.D(0)).done(),
                  PB(b, b).term(Vg, kOpN, GR.D(0), kOpC, b).out(GG.D(0)).done()},
                 E, b);
    if (rc) return rc;
  }

  // ---- forward sweep (rgf.py:241-302) ----
  for (int j = 1; j < N; ++j) {
    const int i = j - 1;
    const int bi = (int)g.bs[i], bj = (int)g.bs[j];
    Mat S = sl(0, bj), P = sl(1, bi), Ql = sl(2, bj), Qg = sl(3, bj), Yl = sl(4, bi),
        Yg = sl(5, bi), Wl = sl(6, bj), Wg = sl(7, bj);
    const ZProblem pP = PB(bj, bi).term(A.L(i), kOpN, GR.D(i), kOpN, bi).out(P).done();
    const ZProblem pYl = PB(bj, bi).term(A.L(i), kOpN, GL.D(i), kOpN, bi).out(Yl).done();
    const ZProblem pYg = PB(bj, bi).term(A.L(i), kOpN, GG.D(i), kOpN, bi).out(Yg).done();
    const ZProblem pS = PB(bj, bj)
                            .term(P, kOpN, A.U(i), kOpN, bi, MONE)
                            .add(A.D(j), ONE)
                            .add(OR(j), MONE)
                            .out(S)
                            .done();
    const ZProblem pQl = PB(bj, bj).term(P, kOpN, SL.U(i), kOpN, bi).out(Ql).done();
    const ZProblem pQg = PB(bj, bj).term(P, kOpN, SG.U(i), kOpN, bi).out(Qg).done();
    if (c->prec32 && c->refine < 0) {
      // Mixed precision with FP64 Schur complements: P and S in FP64, so the
      // inversion gets the exact FP64 S (the Newton steps' fixed point);
      // Y and Q stay FP32 products, in one launch after P.
      c->prec32 = false;
      rc = c->gemm({pP}, E, bj);
      if (!rc) rc = c->gemm({pS}, E, bj);
      c->prec32 = true;
      if (rc) return rc;
      rc = c->gemm({pYl, pYg, pQl, pQg}, E, bj);
      if (rc) return rc;
    } else {
      // P = A_ji g_i ; Y = A_ji g^<>_i
      rc = c->gemm({pP, pYl, pYg}, E, bj);
      if (rc) return rc;
      // S = A_jj - B^R_j - P A_ij ; Q = P Sigma_ij
      rc = c->gemm({pS, pQl, pQg}, E, bj);
      if (rc) return rc;
    }
    // g_j = S^{-1} is independent of W: with overlap it runs on the side stream
    const bool ovl = c->overlap;
    if (ovl) {
      rc = c->fork();
      if (rc) return rc;
      rc = c->inv(S.p, S.s, bj, E, GR.D(j), w.aux, j + 1, e0, c->side);
      if (rc) return rc;
    }
    // W = Y A_ji^H + Q^H - Q  [-> Sigma~_j]  + Sigma_jj + B_j
    rc = c->gemm({PB(bj, bj)
                      .term(Yl, kOpN, A.L(i), kOpC, bi)
                      .add(Ql, MONE, ONE)
                      .partial_here()
                      .add(SL.D(j), ONE)
                      .add(OL(j), ONE)
                      .out(Wl)
                      .out2(kD2Partial, ST(w.stl, j))
                      .done(),
                  PB(bj, bj)
                      .term(Yg, kOpN, A.L(i), kOpC, bi)
                      .add(Qg, MONE, ONE)
                      .partial_here()
                      .add(SG.D(j), ONE)
                      .add(OG(j), ONE)
                      .out(Wg)
                      .out2(kD2Partial, ST(w.stg, j))
                      .done()},
                 E, bj);
    if (rc) return rc;
    rc = ovl ? c->join() : c->inv(S.p, S.s, bj, E, GR.D(j), w.aux, j + 1, e0);
    if (rc) return rc;
    // Mixed precision: Newton steps g_j <- g_j + g_j (I - S64 g_j) with the Schur
    // complement recomputed in FP64 from the stored g_i (products in FP64,
    // the small correction product in FP32).
    for (int it = 0; it < c->refine && c->prec32; ++it) {
      Mat I = Mat{w.ident, 0, (int)g.bmax};
      Mat P64 = sl(1, bi), S64 = sl(0, bj), Rj = sl(8, bj), Gt = sl(9, bj);
      c->prec32 = false;
      rc = c->gemm({PB(bj, bi).term(A.L(i), kOpN, GR.D(i), kOpN, bi).out(P64).done()}, E, bj);
      if (!rc)
        rc = c->gemm({PB(bj, bj)
                          .term(P64, kOpN, A.U(i), kOpN, bi, MONE)
                          .add(A.D(j), ONE)
                          .add(OR(j), MONE)
                          .out(S64)
                          .done()},
                     E, bj);
      if (!rc)
        rc = c->gemm({PB(bj, bj).term(S64, kOpN, GR.D(j), kOpN, bj, MONE).add(I, ONE).out(Rj).done()},
                     E, bj);
      c->prec32 = true;
      if (rc) return rc;
      rc = c->gemm({PB(bj, bj).term(GR.D(j), kOpN, Rj, kOpN, bj).add(GR.D(j), ONE).out(Gt).done()}, E,
                   bj);
      if (rc) return rc;
      rc = c->gemm({PB(bj, bj).add(Gt, ONE).out(GR.D(j)).done()}, E, bj);
      if (rc) return rc;
    }
    // V = g_j W ; g^<>_j = V g_j^H   (V reuses the Y slots)
    Mat Vl = sl(4, bj), Vg = sl(5, bj);
    rc = c->gemm({PB(bj, bj).term(GR.D(j), kOpN, Wl, kOpN, bj).out(Vl).done(),
                  PB(bj, bj).term(GR.D(j), kOpN, Wg, kOpN, bj).out(Vg).done()},
                 E, bj);
    if (rc) return rc;
    rc = c->gemm({PB(bj, bj).term(Vl, kOpN, GR.D(j), kOpC, bj).out(GL.D(j)).done(),
                  PB(bj, bj).term(Vg, kOpN, GR.D(j), kOpC, bj).out(GG.D(j)).done()},
                 E, bj);
    if (rc) return rc;
  }

  // ---- backward sweep (rgf.py:315-412) ----
  for (int i = N - 2; i >= 0; --i) {
    const int j = i + 1;
    const int bi = (int)g.bs[i], bj = (int)g.bs[j];
    Mat U = sl(0, bj), R = sl(1, bi), M = sl(2, bj), K = sl(3, bi), Zl = sl(4, bj),
        Zg = sl(5, bj), T2l = sl(6, bj), T2g = sl(7, bj), T1l = sl(8, bi), T1g = sl(9, bi),
        Gn = sl(10, bi);
    // U = g_i A_ij ; R = G_jj A_ji
    rc = c->gemm({PB(bi, bj).term(GR.D(i), kOpN, A.U(i), kOpN, bi).out(U).done(),
                  PB(bj, bi).term(GR.D(j), kOpN, A.L(i), kOpN, bj).out(R).done()},
                 E, bi);
    if (rc) return rc;
    // M = U G_jj (G^R_ij = -M) ; K = U R ; Z = g_i Sigma_ij ; T2 = U G^<>_jj
    rc = c->gemm({PB(bi, bj).term(U, kOpN, GR.D(j), kOpN, bj).out(M).out2(kD2Copy, GR.U(i), MONE).done(),
                  PB(bi, bi).term(U, kOpN, R, kOpN, bj).out(K).done(),
                  PB(bi, bj).term(GR.D(i), kOpN, SL.U(i), kOpN, bi).out(Zl).done(),
                  PB(bi, bj).term(GR.D(i), kOpN, SG.U(i), kOpN, bi).out(Zg).done(),
                  PB(bi, bj).term(U, kOpN, GL.D(j), kOpN, bj).out(T2l).done(),
                  PB(bi, bj).term(U, kOpN, GG.D(j), kOpN, bj).out(T2g).done()},
                 E, bi);
    if (rc) return rc;
    // T1 = K g^<>_i - Z M^H ; G^<>_ij = -g^<>_i R^H + Z G_jj^H - T2 ; G_new = g_i + K g_i ;
    // G^R_ji = -R g_i
    std::vector<ZProblem> l3;
    l3.push_back(PB(bi, bi)
                     .term(K, kOpN, GL.D(i), kOpN, bi)
                     .term(Zl, kOpN, M, kOpC, bj, MONE)
                     .out(T1l)
                     .done());
    l3.push_back(PB(bi, bi)
                     .term(K, kOpN, GG.D(i), kOpN, bi)
                     .term(Zg, kOpN, M, kOpC, bj, MONE)
                     .out(T1g)
                     .done());
    l3.push_back(PB(bi, bj)
                     .term(GL.D(i), kOpN, R, kOpC, bi, MONE)
                     .term(Zl, kOpN, GR.D(j), kOpC, bj)
                     .add(T2l, MONE)
                     .out(GL.U(i))
                     .out2(kD2ConjT, GL.L(i), MONE)
                     .done());
    l3.push_back(PB(bi, bj)
                     .term(GG.D(i), kOpN, R, kOpC, bi, MONE)
                     .term(Zg, kOpN, GR.D(j), kOpC, bj)
                     .add(T2g, MONE)
                     .out(GG.U(i))
                     .out2(kD2ConjT, GG.L(i), MONE)
                     .done());
    l3.push_back(PB(bi, bi).term(K, kOpN, GR.D(i), kOpN, bi).add(GR.D(i), ONE).out(Gn).done());
    if (GR.L(i).p)
      l3.push_back(PB(bj, bi).term(R, kOpN, GR.D(i), kOpN, bi, MONE).out(GR.L(i)).done());
    rc = c->gemm(l3, E, bi);
    if (rc) return rc;
    // G^<>_ii = g^<>_i + T2 U^H + (T1 - T1^H) ; G_ii = G_new
    rc = c->gemm({PB(bi, bi)
                      .term(T2l, kOpN, U, kOpC, bj)
                      .add(GL.D(i), ONE)
                      .add(T1l, ONE, MONE)
                      .out(GL.D(i))
                      .done(),
                  PB(bi, bi)
                      .term(T2g, kOpN, U, kOpC, bj)
                      .add(GG.D(i), ONE)
                      .add(T1g, ONE, MONE)
                      .out(GG.D(i))
                      .done(),
                  PB(bi, bi).add(Gn, ONE).out(GR.D(i)).done()},
                 E, bi);
    if (rc) return rc;
  }

  // ---- current (rgf.py:388-428), from carried unprojected G^<> ----
  if (io.current) {
    std::vector<TraceJob> jobs;
    auto job = [&](Mat X1, Mat Y1, Mat X2, Mat Y2, int b, int slot, double sign) {
      TraceJob t{};
      t.X1 = X1.p; t.Y1 = Y1.p; t.X2 = X2.p; t.Y2 = Y2.p;
      t.sX1 = X1.s; t.sY1 = Y1.s; t.sX2 = X2.s; t.sY2 = Y2.s;
      t.b = b; t.slot = slot; t.sign = sign;
      jobs.push_back(t);
    };
    for (int j = 1; j < N; ++j)
      job(ST(w.stg, j), GL.D(j), GG.D(j), ST(w.stl, j), (int)g.bs[j], j, 1.0);
    job(OG(0), GL.D(0), GG.D(0), OL(0), (int)g.bs[0], 0, 1.0);
    job(OG(N - 1), GL.D(N - 1), GG.D(N - 1), OL(N - 1), (int)g.bs[N - 1], N, -1.0);
    // the job list is written by a kernel with the list as its argument: no
    // host wait, and a captured graph rewrites it on every replay; the area
    // is reused by the next chunk (stream order)
    const size_t mark = c->tables_used;
    void* dev = nullptr;
    rc = c->put_table(jobs.data(), jobs.size() * sizeof(TraceJob), &dev);
    if (rc) return rc;
    rc = c->hip(launch_traces(static_cast<TraceJob*>(dev), (int)jobs.size(), E,
                              io.current + e0 * (N + 1), N + 1, c->stream),
                "trace launch");
    c->tables_used = mark;
    if (rc) return rc;
  }

  // ---- stored diagonal G^<> = (X - X^H)/2 (rgf.py:305-310, 362-364, 384-386) ----
  rc = c->hip(launch_project(GL.d + e0 * g.Dsz, g.Dsz, c->offs_dev, c->sizes_dev, N, (int)g.bmax, E,
                             c->stream),
              "project launch");
  if (rc) return rc;
  rc = c->hip(launch_project(GG.d + e0 * g.Dsz, g.Dsz, c->offs_dev, c->sizes_dev, N, (int)g.bmax, E,
                             c->stream),
              "project launch");
  return rc;
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// Persistent on-chip sweep (sweep_small.hip) for geometries whose blocks are
// all at most kSmallMaxBlock
// ---------------------------------------------------------------------------

// Largest persistent grid assumed when sizing workspaces without a context.
constexpr int kMaxPersistentGrid = 1024;

// One problem of a persistent launch: a geometry, its energy count and the
// energy-0 views of its buffers.
struct SmallSpec {
  const Geo* g;
  int64_t nE;
  SolveIO io;
};

bool small_geometry(const Geo& g) { return g.bmax <= kSmallMaxBlock; }

size_t small_per_wg(const Geo& g, bool have_gr, bool current) {
  return sweep_small_scratch_per_wg(g.N, !have_gr, current);
}

// Workspace of a persistent launch: per-workgroup scratch for at most
// min(energies, grid) workgroups.
size_t small_work_bytes(const Geo& g, int64_t nE, bool have_gr, bool current, int grid) {
  const size_t per = small_per_wg(g, have_gr, current);
  if (per == 0) return 0;
  return per * (size_t)std::min<int64_t>(nE, grid) + 256;
}

// Relative cost of one energy of a geometry (block steps x block volume).
double small_cost(const Geo& g) {
  double c = 0;
  for (int i = 0; i < g.N; ++i) c += (double)g.bs[i] * g.bs[i] * g.bs[i];
  return c;
}

// Tables of every problem go to the table area (no host wait), then one
// launch: the work queue holds the energies of all problems, the most
// expensive problems first.
int run_small(Ctx* c, std::vector<SmallSpec> specs, char* scratch) {
  std::stable_sort(specs.begin(), specs.end(), [](const SmallSpec& a, const SmallSpec& b) {
    return small_cost(*a.g) > small_cost(*b.g);
  });
  size_t need = 256;
  int64_t total = 0;
  size_t per_wg = 0;
  for (const SmallSpec& sp : specs) {
    need += small_table_bytes(*sp.g);
    total += sp.nE;
    per_wg = std::max(per_wg, small_per_wg(*sp.g, sp.io.GR.d != nullptr, sp.io.current != nullptr));
  }
  if (total == 0) return RGF_OK;
  if (total > 0x7fffffff) return c->fail(RGF_EUNSUPPORTED, "persistent sweep: too many energies");
  if (!c->tables_fixed) {
    int rc = c->ensure_meta(need);
    if (rc) return rc;
    c->tables = static_cast<char*>(c->meta);
    c->tables_cap = c->meta_bytes;
    c->tables_used = 0;
  }
  const size_t mark = c->tables_used;
  std::vector<SmallProblem> probs;
  int rc = RGF_OK;
  int job0 = 0;
  for (const SmallSpec& sp : specs) {
    const Geo& g = *sp.g;
    const SolveIO& io = sp.io;
    std::vector<int> bs(g.N);
    for (int i = 0; i < g.N; ++i) bs[i] = (int)g.bs[i];
    void *d_bs = nullptr, *d_od = nullptr, *d_ou = nullptr;
    rc = c->put_table(bs.data(), g.N * sizeof(int), &d_bs);
    if (!rc) rc = c->put_table(g.od.data(), g.N * sizeof(int64_t), &d_od);
    if (!rc) rc = c->put_table(g.ou.data(), g.N * sizeof(int64_t), &d_ou);
    void* d_ob[3] = {nullptr, nullptr, nullptr};
    if (!rc && io.obc) {
      const void* const* lists[3] = {io.obc->retarded, io.obc->lesser, io.obc->greater};
      std::vector<const void*> all(3 * g.N, nullptr);
      for (int k = 0; k < 3; ++k)
        for (int i = 0; lists[k] && i < g.N; ++i) all[k * g.N + i] = lists[k][i];
      void* d_all = nullptr;
      rc = c->put_table(all.data(), all.size() * sizeof(void*), &d_all);
      for (int k = 0; !rc && k < 3; ++k)
        if (lists[k]) d_ob[k] = static_cast<char*>(d_all) + k * g.N * sizeof(void*);
    }
    if (rc) {
      c->tables_used = mark;
      return rc;
    }
    SmallProblem q{};
    q.N = g.N;
    q.nE = (int32_t)sp.nE;
    q.job0 = job0;
    q.e_base = 0;
    q.Dsz = g.Dsz;
    q.Usz = g.Usz;
    q.bs = static_cast<const int*>(d_bs);
    q.od = static_cast<const int64_t*>(d_od);
    q.ou = static_cast<const int64_t*>(d_ou);
    q.Ad = io.A.d; q.Au = io.A.u; q.Al = io.A.l;
    q.SLd = io.SL.d; q.SLu = io.SL.u;
    q.SGd = io.SG.d; q.SGu = io.SG.u;
    q.GRd = io.GR.d; q.GRu = io.GR.u; q.GRl = io.GR.l;
    q.GLd = io.GL.d; q.GLu = io.GL.u; q.GLl = io.GL.l;
    q.GGd = io.GG.d; q.GGu = io.GG.u; q.GGl = io.GG.l;
    q.obr = static_cast<const double2* const*>(d_ob[0]);
    q.obl = static_cast<const double2* const*>(d_ob[1]);
    q.obg = static_cast<const double2* const*>(d_ob[2]);
    q.cur = io.current;
    probs.push_back(q);
    job0 += (int)sp.nE;
  }
  void *d_probs = nullptr, *d_counter = nullptr;
  const int zero = 0;
  rc = c->put_table(probs.data(), probs.size() * sizeof(SmallProblem), &d_probs);
  if (!rc) rc = c->put_table(&zero, sizeof(int), &d_counter);
  if (rc) {
    c->tables_used = mark;
    return rc;
  }
  SmallLaunch L{};
  L.nprob = (int)probs.size();
  L.total_jobs = (int)total;
  L.bmax = 0;
  for (const SmallSpec& sp : specs) L.bmax = std::max(L.bmax, (int32_t)sp.g->bmax);
  L.counter = static_cast<int*>(d_counter);
  L.info = c->info_dev;
  L.scratch = per_wg ? scratch : nullptr;
  L.scratch_per_wg = (int64_t)per_wg;
  L.probs = static_cast<const SmallProblem*>(d_probs);
  const int grid = (int)std::min<int64_t>({total, (int64_t)c->n_cu, (int64_t)kMaxPersistentGrid});
  rc = c->hip(launch_sweep_small(L, grid, c->stream), "persistent sweep launch");
  c->tables_used = mark;
  return rc;
}

}  // namespace

// FP64 a-posteriori check of the mixed-precision result on a few diagonal
// blocks i, through exact block-column identities that involve selected
// blocks only (G^A_{k,i} = G^R_{i,k}^H, G^<_{i+1,i} = -G^<_{i,i+1}^H):
//   (A~ G^R)_ii = I
//   (A~ G^<)_ii = (S^< G^A)_ii,  S^< = Sigma^< + B^<   (and the same for >)
// with (A~ X)_ii = A_{i,i-1} X_{i-1,i} + A~_ii X_ii + A_{i,i+1} X_{i+1,i}.
// Relative residuals: ||R^R|| / sqrt(b_i) and
// ||R^<|| / (||A_ii|| ||G^<_ii|| + ||S^<_ii|| ||G^R_ii||); the energy's value is
// the maximum over the sampled blocks and the three functions.
int check_chunk(Ctx* c, const Geo& g, int64_t e0, int E, const SolveIO& io, const BTV& GR,
                const Work& w, std::vector<double>* resid) {
  const int N = g.N;
  const BTV& A = io.A;
  const void* const* obr = io.obc ? io.obc->retarded : nullptr;
  const void* const* obl = io.obc ? io.obc->lesser : nullptr;
  const void* const* obg = io.obc ? io.obc->greater : nullptr;
  std::vector<int> blocks = {0};
  if (N > 2) blocks.push_back(N / 2);
  if (N > 1) blocks.push_back(N - 1);
  auto sl = [&](int q, int ld) { return with_ld(w.slot(q), ld); };
  Mat I = Mat{w.ident, 0, (int)g.bmax};
  std::vector<NormJob> jobs;
  resid->assign(E, 0.0);
  const int per_block = 9;
  std::vector<double> h((size_t)E * kMaxNormJobs);
  int rc;
  for (size_t bi_ = 0; bi_ < blocks.size(); ++bi_) {
    const int i = blocks[bi_];
    const int b = (int)g.bs[i];
    Mat OR = obc_mat(obr, g, i, e0);
    // (A~ G^R)_ii - I
    {
      PB pa(b, b);
      pa.term(A.D(i), kOpN, GR.D(i), kOpN, b);
      if (i > 0) pa.term(A.L(i - 1), kOpN, GR.U(i - 1), kOpN, (int)g.bs[i - 1]);
      PB pb(b, b);
      if (i + 1 < N) pb.term(A.U(i), kOpN, GR.L(i), kOpN, (int)g.bs[i + 1]);
      if (OR.p) pb.term(OR, kOpN, GR.D(i), kOpN, b, MONE);
      pb.add(sl(0, b), ONE).add(I, MONE);
      std::vector<ZProblem> l1 = {pa.out(sl(0, b)).done()}, l2 = {pb.out(sl(1, b)).done()},
                            l3, l4;
      // lesser and greater
      const BTV* Xs[2] = {&io.GL, &io.GG};
      const BTV* Ss[2] = {&io.SL, &io.SG};
      const void* const* Bs[2] = {obl, obg};
      for (int k = 0; k < 2; ++k) {
        const BTV& X = *Xs[k];
        const BTV& Sg = *Ss[k];
        Mat OB = obc_mat(Bs[k], g, i, e0);
        const int q = 2 + 4 * k;
        PB pc(b, b);
        pc.term(A.D(i), kOpN, X.D(i), kOpN, b);
        if (i > 0) pc.term(A.L(i - 1), kOpN, X.U(i - 1), kOpN, (int)g.bs[i - 1]);
        l1.push_back(pc.out(sl(q, b)).done());
        PB pd(b, b);
        if (i + 1 < N) pd.term(A.U(i), kOpN, X.U(i), kOpC, (int)g.bs[i + 1], MONE);
        if (OR.p) pd.term(OR, kOpN, X.D(i), kOpN, b, MONE);
        l2.push_back(pd.add(sl(q, b), ONE).out(sl(q + 1, b)).done());
        PB pe(b, b);
        pe.term(Sg.D(i), kOpN, GR.D(i), kOpC, b, MONE);
        if (i > 0) pe.term(Sg.U(i - 1), kOpC, GR.L(i - 1), kOpC, (int)g.bs[i - 1]);
        l3.push_back(pe.add(sl(q + 1, b), ONE).out(sl(q + 2, b)).done());
        PB pf(b, b);
        if (i + 1 < N) pf.term(Sg.U(i), kOpN, GR.U(i), kOpC, (int)g.bs[i + 1], MONE);
        if (OB.p) pf.term(OB, kOpN, GR.D(i), kOpC, b, MONE);
        l4.push_back(pf.add(sl(q + 2, b), ONE).out(sl(q + 3, b)).done());
      }
      for (auto* L : {&l1, &l2, &l3, &l4}) {
        rc = c->gemm(*L, E, b);
        if (rc) return rc;
      }
    }
    // norms: residuals, then the scales
    jobs.clear();
    auto nj = [&](Mat m, int rows, int cols) {
      NormJob j{};
      j.X = m.p; j.sX = m.s; j.rows = rows; j.cols = cols; j.ld = m.ld;
      jobs.push_back(j);
    };
    nj(sl(1, b), b, b);                 // R^R
    nj(sl(5, b), b, b);                 // R^<
    nj(sl(9, b), b, b);                 // R^>
    nj(A.D(i), b, b);
    nj(GR.D(i), b, b);
    nj(io.GL.D(i), b, b);
    nj(io.GG.D(i), b, b);
    nj(io.SL.D(i), b, b);
    nj(io.SG.D(i), b, b);
    const size_t mark = c->tables_used;
    void* dev = nullptr;
    rc = c->put_table(jobs.data(), jobs.size() * sizeof(NormJob), &dev);
    if (rc) return rc;
    rc = c->hip(launch_norms(static_cast<NormJob*>(dev), (int)jobs.size(), E,
                             w.norms + bi_ * per_block, kMaxNormJobs, c->stream),
                "norm launch");
    c->tables_used = mark;
    if (rc) return rc;
  }
  rc = c->hip(hipMemcpyAsync(h.data(), w.norms, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream),
              "read norms");
  if (rc) return rc;
  rc = c->hip(hipStreamSynchronize(c->stream), "read norms sync");
  if (rc) return rc;
  for (int e = 0; e < E; ++e) {
    double worst = 0.0;
    for (size_t bi_ = 0; bi_ < blocks.size(); ++bi_) {
      const double* n = h.data() + (size_t)e * kMaxNormJobs + bi_ * per_block;
      const double b = (double)g.bs[blocks[bi_]];
      const double rr = n[0] / std::sqrt(b);
      const double rl = n[1] / std::max(n[3] * n[5] + n[7] * n[4], 1e-300);
      const double rg = n[2] / std::max(n[3] * n[6] + n[8] * n[4], 1e-300);
      worst = std::max({worst, rr, rl, rg});
    }
    (*resid)[e] = worst;
  }
  return RGF_OK;
}

namespace qx {

__global__ void zero_ints_kernel(int* p, int n) {
  int i = threadIdx.x;
  if (i < n) p[i] = 0;
}

int Ctx::reset_info() {
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, stream, info_dev, 4);
  return hip(hipGetLastError(), "zero info");
}

// Bytes of table memory one call may use: the geometry tables (block
// offsets and sizes), then room for one chunk's job list (trace jobs, or the
// norm jobs of the mixed-precision check, or the boundary table of the
// assembly), each rounded up to 256 bytes.
size_t table_bytes(const Geo& g) {
  auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t geo = r(g.N * (sizeof(int64_t) + sizeof(int)));
  const size_t jobs = std::max({(g.N + 2) * sizeof(TraceJob), 16 * sizeof(NormJob),
                                g.N * sizeof(double2*), small_table_bytes(g) + 256});
  return geo + r(jobs);
}

int Ctx::put_table(const void* host, size_t bytes, void** dev) {
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (!tables || tables_used + need > tables_cap)
    return fail(RGF_EHIP, "table area too small (" + std::to_string(tables_used + need) + " > " +
                              std::to_string(tables_cap) + " bytes)");
  char* d = tables + tables_used;
  int rc = hip(launch_store_bytes(d, host, bytes, stream), "table upload");
  if (rc) return rc;
  tables_used += need;
  *dev = d;
  return RGF_OK;
}

// Block offsets/sizes used by the projection and assembly kernels, at the
// start of the table area.
int Ctx::upload_geometry(const Geo& g) {
  const size_t need = table_bytes(g);
  if (!tables_fixed) {
    int rc = ensure_meta(need);
    if (rc) return rc;
    tables = static_cast<char*>(meta);
    tables_cap = meta_bytes;
  }
  tables_used = 0;
  const size_t bytes = g.N * (sizeof(int64_t) + sizeof(int));
  std::vector<char> host(bytes);
  std::memcpy(host.data(), g.od.data(), g.N * sizeof(int64_t));
  std::vector<int> sz(g.N);
  for (int i = 0; i < g.N; ++i) sz[i] = (int)g.bs[i];
  std::memcpy(host.data() + g.N * sizeof(int64_t), sz.data(), g.N * sizeof(int));
  void* dev = nullptr;
  int rc = put_table(host.data(), bytes, &dev);
  if (rc) return rc;
  offs_dev = static_cast<int64_t*>(dev);
  sizes_dev = reinterpret_cast<int*>(static_cast<char*>(dev) + g.N * sizeof(int64_t));
  return RGF_OK;
}

int64_t Ctx::chunk_for(const Geo& g, int64_t nE, bool solve, bool current, bool have_gr) {
  int64_t chunk = nE;
  if (max_chunk > 0) chunk = std::min<int64_t>(chunk, max_chunk);
  const size_t per = work_bytes(g, 1, solve, current, have_gr);
  size_t budget = ws_budget;
  if (!ws_owned && ws) budget = ws_bytes;
  int64_t fit = (int64_t)(budget / std::max<size_t>(per, 1));
  if (fit < 1) fit = 1;
  chunk = std::min(chunk, fit);
  // keep gridDim.y/z and workgroup counts in range
  chunk = std::min<int64_t>(chunk, 65535);
  return std::max<int64_t>(chunk, 1);
}

}  // namespace qx

namespace {

// The table-area fields of a context, saved while a graph capture or the
// allocation-free entry point points them elsewhere.
struct TableState {
  char* tables;
  size_t cap, used;
  bool fixed;
  int64_t* offs;
  int* sizes;
  static TableState save(const qx::Ctx& c) {
    return TableState{c.tables, c.tables_cap, c.tables_used, c.tables_fixed, c.offs_dev, c.sizes_dev};
  }
  void restore(qx::Ctx* c) const {
    c->tables = tables;
    c->tables_cap = cap;
    c->tables_used = used;
    c->tables_fixed = fixed;
    c->offs_dev = offs;
    c->sizes_dev = sizes;
  }
};

// The current needs the lesser and greater boundary blocks of both ends
// (rgf.py:414-428; the reference raises TypeError without them).
bool current_obc_ok(const rgf_obc* obc, const qx::Geo& g) {
  const int last = g.N - 1;
  return obc && obc->lesser && obc->greater && obc->lesser[0] && obc->greater[0] &&
         obc->lesser[last] && obc->greater[last];
}

// Argument checks shared by the selected-solve entry points.
int check_solve_args(qx::Ctx* ctx, const qx::Geo& g, const rgf_bt_const* a,
                     const rgf_bt_const* sigma_lesser, const rgf_bt_const* sigma_greater,
                     const rgf_obc* obc, const rgf_bt* g_lesser, const rgf_bt* g_greater,
                     const void* current) {
  const bool multi = g.N > 1;
  if (!a || !a->diag || (multi && (!a->upper || !a->lower)))
    return ctx->fail(RGF_EINVAL, "selected_solve: A needs diag, upper and lower");
  if (!sigma_lesser || !sigma_greater || !sigma_lesser->diag || !sigma_greater->diag ||
      (multi && (!sigma_lesser->upper || !sigma_greater->upper)))
    return ctx->fail(RGF_EINVAL, "selected_solve: Sigma needs diag and upper");
  if (!g_lesser || !g_greater || !g_lesser->diag || !g_greater->diag ||
      (multi && (!g_lesser->upper || !g_greater->upper)))
    return ctx->fail(RGF_EINVAL, "selected_solve: G^< and G^> need diag and upper outputs");
  if (current && !current_obc_ok(obc, g))
    return ctx->fail(RGF_EINVAL,
                     "current needs lesser/greater OBC blocks on the first and last block");
  return RGF_OK;
}

// One chunk's worth of solve I/O views, shared by the FP64 entry points.
SolveIO solve_io(const qx::Geo& g, int64_t e0, const rgf_bt_const* a,
                 const rgf_bt_const* sigma_lesser, const rgf_bt_const* sigma_greater,
                 const rgf_obc* obc, rgf_bt* g_retarded, rgf_bt* g_lesser, rgf_bt* g_greater,
                 void* current) {
  const bool have_gr = g_retarded && g_retarded->diag;
  SolveIO io;
  io.A = view(g, a->diag, a->upper, a->lower, e0);
  io.SL = view(g, sigma_lesser->diag, sigma_lesser->upper, nullptr, e0);
  io.SG = view(g, sigma_greater->diag, sigma_greater->upper, nullptr, e0);
  io.obc = obc;
  io.GR = have_gr ? view(g, g_retarded->diag, g_retarded->upper, g_retarded->lower, e0)
                  : BTV{&g, nullptr, nullptr, nullptr, 0, -1, -1};
  io.GL = view(g, g_lesser->diag, g_lesser->upper, g_lesser->lower, e0);
  io.GG = view(g, g_greater->diag, g_greater->upper, g_greater->lower, e0);
  io.current = static_cast<double2*>(current);
  return io;
}

}  // namespace

namespace {

// FP64 selected solve by the launch-per-product-group sweep, all energies in
// chunks that fit the workspace budget (the status word is not reset here).
int solve_chunked(qx::Ctx* ctx, const qx::Geo& g, int64_t n_energies, const rgf_bt_const* a,
                  const rgf_bt_const* sigma_lesser, const rgf_bt_const* sigma_greater,
                  const rgf_obc* obc, rgf_bt* g_retarded, rgf_bt* g_lesser, rgf_bt* g_greater,
                  void* current) {
  const bool have_gr = g_retarded && g_retarded->diag;
  const int64_t chunk = ctx->chunk_for(g, n_energies, true, current != nullptr, have_gr);
  int rc = ctx->upload_geometry(g);
  if (rc) return rc;
  char* ws = nullptr;
  rc = ctx->ensure_workspace(work_bytes(g, chunk, true, current != nullptr, have_gr), &ws);
  if (rc) return rc;
  Work w;
  carve(g, chunk, true, current != nullptr, have_gr, ws, &w);
  for (int64_t e0 = 0; e0 < n_energies; e0 += chunk) {
    const int E = (int)std::min<int64_t>(chunk, n_energies - e0);
    const SolveIO io = solve_io(g, e0, a, sigma_lesser, sigma_greater, obc, g_retarded, g_lesser,
                                g_greater, current);
    // workspace-resident retarded storage is indexed from the chunk start
    rc = sweep_chunk_solve(ctx, g, e0, E, io, w);
    if (rc) return rc;
  }
  return RGF_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// Graph capture and graph launches never overlap across host threads: a
// launch in one thread while another thread captured has ended the process
// with SIGSEGV inside hipGraphLaunch.  Captures hold this lock exclusively,
// launches shared (launches from several threads still run concurrently).
static std::shared_mutex g_graph_mutex;

// Workspace and tables of rgf_selected_solve_noalloc_z.  In host code the
// name `scratch` is a host-side stand-in, never a device address: host code
// uses only the per-device address that hipGetSymbolAddress returned
// (Ctx::noalloc_scratch).  Each GPU holds its own copy.
static __device__ __align__(256) char scratch[RGF_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mutex;
std::vector<char*> g_scratch_ptr;  // by device ordinal

// Device address of `scratch` on the current device (looked up once per
// device; the caller has made `device` current).
int scratch_address(int device, char** out, std::string* err) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (device < 0) {
    *err = "negative device ordinal";
    return RGF_EINVAL;
  }
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return RGF_EHIP;
    }
    g_scratch_ptr[device] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[device];
  return RGF_OK;
}

}  // namespace

struct rgf_graph {
  rgf_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* ws = nullptr;
  void* tables = nullptr;
  // recorded after every launch: destruction waits on it rather than on the
  // context's stream, so a graph may outlive its context
  hipEvent_t done = nullptr;
  ~rgf_graph() {
    if (done) {
      (void)hipEventSynchronize(done);
      (void)hipEventDestroy(done);
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ws) (void)hipFree(ws);
    if (tables) (void)hipFree(tables);
  }
};

extern "C" {

const char* rgf_version(void) { return "quatrex_amd-rgf 0.1 (gfx950)"; }

int rgf_create(int device, void* stream, rgf_ctx** out) {
  if (!out) return RGF_EINVAL;
  *out = nullptr;
  rgf_ctx* ctx = nullptr;
  try {
    ctx = new rgf_ctx();
  } catch (...) {
    return RGF_ENOMEM;
  }
  ctx->device = device;
  ctx->stream = static_cast<hipStream_t>(stream);
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return RGF_EHIP;
  }
  if (hipMalloc(&ctx->info_dev, 4 * sizeof(int)) != hipSuccess) {
    delete ctx;
    return RGF_ENOMEM;
  }
  if (hipMemsetAsync(ctx->info_dev, 0, 4 * sizeof(int), ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    delete ctx;
    return RGF_EHIP;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    ctx->n_cu = cus;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0)
    ctx->ws_budget = std::min<size_t>((size_t)16 << 30, free_b / 4);
  // the allocation-free entry point's array on this device (a failure is
  // reported by that entry point only)
  if (scratch_address(device, &ctx->noalloc_scratch, &ctx->noalloc_error) != RGF_OK)
    ctx->noalloc_scratch = nullptr;
  *out = ctx;
  return RGF_OK;
}

int rgf_destroy(rgf_ctx* ctx) {
  if (!ctx) return RGF_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  delete ctx;
  return RGF_OK;
}

int rgf_set_stream(rgf_ctx* ctx, void* stream) {
  if (!ctx) return RGF_EINVAL;
  ctx->stream = static_cast<hipStream_t>(stream);
  return RGF_OK;
}

int rgf_set_option(rgf_ctx* ctx, int32_t option, int64_t value) {
  if (!ctx) return RGF_EINVAL;
  switch (option) {
    case RGF_OPT_MAX_CHUNK: ctx->max_chunk = value; return RGF_OK;
    case RGF_OPT_WORKSPACE_BUDGET:
      if (value <= 0) return ctx->fail(RGF_EINVAL, "workspace budget must be positive");
      ctx->ws_budget = (size_t)value;
      return RGF_OK;
    case RGF_OPT_PROFILE: ctx->profile = value != 0; return RGF_OK;
    case RGF_OPT_UNFUSED: ctx->unfused = value != 0; return RGF_OK;
    case RGF_OPT_OVERLAP: ctx->overlap = value != 0; return RGF_OK;
    case RGF_OPT_PERSISTENT: ctx->persistent = value != 0; return RGF_OK;
    case RGF_OPT_TILE_CFG:
      if (value >= qx::kNumTileCfgs) return ctx->fail(RGF_EINVAL, "unknown tile config");
      ctx->tile_override = (int)value;
      return RGF_OK;
    default:
      // 1000 + b (FP64) / 3000 + b (FP32 compute): tile config for block size b
      if ((option >= 1000 && option < 1000 + 1025) || (option >= 3000 && option < 3000 + 1025)) {
        const bool f32 = option >= 3000;
        const int b = option - (f32 ? 3000 : 1000);
        if (value >= qx::kNumTileCfgs) return ctx->fail(RGF_EINVAL, "unknown tile config");
        if (f32 && value >= 0 && !qx::f32_tile_ok((int)value))
          return ctx->fail(RGF_EINVAL, "FP32-compute products support tile configs 0, 1 and 3");
        std::vector<int>& t = f32 ? ctx->tile_table32 : ctx->tile_table;
        if ((int)t.size() <= b) t.resize(b + 1, -1);
        t[b] = (int)value;
        return RGF_OK;
      }
      return ctx->fail(RGF_EINVAL, "unknown option");
  }
}

int rgf_last_error(const rgf_ctx* ctx, char* buf, size_t len) {
  if (!ctx || !buf || len == 0) return RGF_EINVAL;
  std::snprintf(buf, len, "%s", ctx->last_error.c_str());
  return RGF_OK;
}

int rgf_status(rgf_ctx* ctx, int64_t* info) {
  if (!ctx) return RGF_EINVAL;
  int h[4] = {0, 0, 0, 0};
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return ctx->hip(e, "hipStreamSynchronize");
  // stream-ordered copies only: a legacy-stream call would wait on every
  // blocking stream, including one another thread is capturing
  e = hipMemcpyAsync(h, ctx->info_dev, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return ctx->hip(e, "read status");
  if (int rc = ctx->reset_info()) return rc;
  if (info) {
    info[0] = h[0] ? h[1] : -1;
    info[1] = h[0] ? h[0] - 1 : -1;
    info[2] = h[0] ? h[2] : -1;
  }
  if (h[0]) {
    ctx->last_error = "Singular matrix (energy " + std::to_string(h[1]) + ", block " +
                      std::to_string(h[0] - 1) + ", column " + std::to_string(h[2]) + ")";
    return RGF_ESINGULAR;
  }
  return RGF_OK;
}

int rgf_workspace_size(const rgf_geometry* geom, int64_t n_energies, int32_t solve,
                       int32_t with_current, int32_t with_retarded, size_t* bytes) {
  if (!bytes || n_energies < 0) return RGF_EINVAL;
  qx::Geo g;
  std::string err;
  int rc = g.init(geom, &err);
  if (rc) return rc;
  *bytes = work_bytes(g, n_energies, solve != 0, with_current != 0, with_retarded != 0);
  if (solve && small_geometry(g))
    *bytes = std::max(*bytes, small_work_bytes(g, n_energies, with_retarded != 0, with_current != 0,
                                               kMaxPersistentGrid));
  return RGF_OK;
}

int rgf_set_workspace(rgf_ctx* ctx, void* ptr, size_t bytes) {
  if (!ctx) return RGF_EINVAL;
  if (ctx->ws_owned && ctx->ws) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->ws);
  }
  ctx->ws = ptr;
  ctx->ws_bytes = ptr ? bytes : 0;
  ctx->ws_owned = false;
  return RGF_OK;
}

int rgf_inv_z(rgf_ctx* ctx, int64_t n, int64_t n_mats, const void* a, void* out) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  if (n < 1 || n > 1024) return ctx->fail(RGF_EUNSUPPORTED, "rgf_inv_z: n must be in [1, 1024]");
  if (n_mats < 0 || (n_mats > 0 && (!a || !out))) return ctx->fail(RGF_EINVAL, "rgf_inv_z: null buffer");
  if (n_mats == 0) return RGF_OK;
  int rc = hipSetDevice(ctx->device) == hipSuccess ? RGF_OK : RGF_EHIP;
  if (rc) return ctx->fail(rc, "hipSetDevice");
  const size_t per = (size_t)n * n * sizeof(double2) + qx::zinv_aux_bytes((int)n);
  // the workspace holds chunk matrices plus 256 bytes of alignment slack
  const size_t budget = (!ctx->ws_owned && ctx->ws) ? ctx->ws_bytes : ctx->ws_budget;
  int64_t chunk = std::max<int64_t>(1, (int64_t)((budget > 256 ? budget - 256 : 0) / per));
  chunk = std::min<int64_t>({chunk, n_mats, 65535});
  char* ws = nullptr;
  rc = ctx->ensure_workspace(per * chunk + 256, &ws);
  if (rc) return rc;
  rc = ctx->reset_info();
  if (rc) return rc;
  const double2* A = static_cast<const double2*>(a);
  double2* O = static_cast<double2*>(out);
  double2* S = reinterpret_cast<double2*>(ws);
  char* aux = ws + (size_t)chunk * n * n * sizeof(double2);
  for (int64_t e0 = 0; e0 < n_mats; e0 += chunk) {
    const int E = (int)std::min<int64_t>(chunk, n_mats - e0);
    rc = ctx->hip(hipMemcpyAsync(S, A + e0 * n * n, (size_t)E * n * n * sizeof(double2),
                                 hipMemcpyDeviceToDevice, ctx->stream),
                  "copy to workspace");
    if (rc) return rc;
    rc = ctx->inv(S, n * n, (int)n, E, qx::Mat{O + e0 * n * n, n * n, (int)n}, aux, 1, e0);
    if (rc) return rc;
  }
  return RGF_OK;
  GUARD_END
}

int rgf_assemble_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies, const void* z,
                   const rgf_bt_const* h, const void* const* sigma_b, rgf_bt* a) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  qx::Geo g;
  int rc = g.init(geom, &ctx->last_error);
  if (rc) return rc;
  if (n_energies < 0) return ctx->fail(RGF_EINVAL, "n_energies < 0");
  if (n_energies == 0) return RGF_OK;
  if (!z || !h || !h->diag || !a || !a->diag || (g.N > 1 && (!h->upper || !h->lower || !a->upper || !a->lower)))
    return ctx->fail(RGF_EINVAL, "assemble: z, H and A buffers required");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  rc = ctx->upload_geometry(g);
  if (rc) return rc;
  // per-block boundary pointers go to the table area after the geometry tables
  std::vector<const double2*> sig(g.N, nullptr);
  if (sigma_b)
    for (int i = 0; i < g.N; ++i) sig[i] = static_cast<const double2*>(sigma_b[i]);
  void* dev = nullptr;
  rc = ctx->put_table(sig.data(), g.N * sizeof(double2*), &dev);
  if (rc) return rc;
  const double2** dsig = static_cast<const double2**>(dev);
  return ctx->hip(launch_assemble((const double2*)h->diag, (const double2*)h->upper,
                                  (const double2*)h->lower, g.Dsz, g.Usz, (int)n_energies,
                                  (const double2*)z, ctx->offs_dev, ctx->sizes_dev, dsig, g.N,
                                  (int)g.bmax, (double2*)a->diag, (double2*)a->upper,
                                  (double2*)a->lower, ctx->stream),
                  "assemble");
  GUARD_END
}

int rgf_selected_inv_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                       const rgf_bt_const* a, const void* const* obc_retarded, rgf_bt* x) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  qx::Geo g;
  int rc = g.init(geom, &ctx->last_error);
  if (rc) return rc;
  if (n_energies < 0) return ctx->fail(RGF_EINVAL, "n_energies < 0");
  if (n_energies == 0) return RGF_OK;
  if (!a || !a->diag || (g.N > 1 && (!a->upper || !a->lower)))
    return ctx->fail(RGF_EINVAL, "selected_inv: A needs diag, upper and lower");
  if (!x || !x->diag) return ctx->fail(RGF_EINVAL, "selected_inv: X diag output required");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  const int64_t chunk = ctx->chunk_for(g, n_energies, false, false, true);
  char* ws = nullptr;
  rc = ctx->ensure_workspace(work_bytes(g, chunk, false, false, true), &ws);
  if (rc) return rc;
  rc = ctx->reset_info();
  if (rc) return rc;
  Work w;
  carve(g, chunk, false, false, true, ws, &w);
  for (int64_t e0 = 0; e0 < n_energies; e0 += chunk) {
    const int E = (int)std::min<int64_t>(chunk, n_energies - e0);
    BTV A = view(g, a->diag, a->upper, a->lower, e0);
    BTV X = view(g, x->diag, x->upper, x->lower, e0);
    rc = sweep_chunk_inv(ctx, g, e0, E, A, obc_retarded, X, w);
    if (rc) return rc;
  }
  return RGF_OK;
  GUARD_END
}

int rgf_selected_solve_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                         const rgf_bt_const* a, const rgf_bt_const* sigma_lesser,
                         const rgf_bt_const* sigma_greater, const rgf_obc* obc,
                         rgf_bt* g_retarded, rgf_bt* g_lesser, rgf_bt* g_greater, void* current) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  qx::Geo g;
  int rc = g.init(geom, &ctx->last_error);
  if (rc) return rc;
  if (n_energies < 0) return ctx->fail(RGF_EINVAL, "n_energies < 0");
  if (n_energies == 0) return RGF_OK;
  rc = check_solve_args(ctx, g, a, sigma_lesser, sigma_greater, obc, g_lesser, g_greater, current);
  if (rc) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  const bool have_gr = g_retarded && g_retarded->diag;
  rc = ctx->reset_info();
  if (rc) return rc;
  if (ctx->persistent && small_geometry(g)) {
    // every block <= 32: one persistent launch, intermediates in LDS
    const size_t bytes = small_work_bytes(g, n_energies, have_gr, current != nullptr, ctx->n_cu);
    char* ws = nullptr;
    if (bytes) {
      rc = ctx->ensure_workspace(bytes, &ws);
      if (rc) return rc;
    }
    return run_small(ctx, {SmallSpec{&g, n_energies, solve_io(g, 0, a, sigma_lesser, sigma_greater, obc,
                                                               g_retarded, g_lesser, g_greater, current)}},
                     ws);
  }
  return solve_chunked(ctx, g, n_energies, a, sigma_lesser, sigma_greater, obc, g_retarded, g_lesser,
                       g_greater, current);
  GUARD_END
}

int rgf_selected_solve_noalloc_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                                 const rgf_bt_const* a, const rgf_bt_const* sigma_lesser,
                                 const rgf_bt_const* sigma_greater, const rgf_obc* obc,
                                 rgf_bt* g_retarded, rgf_bt* g_lesser, rgf_bt* g_greater,
                                 void* current) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  if (!ctx->noalloc_scratch)
    return ctx->fail(RGF_EHIP, "allocation-free scratch array unavailable: " + ctx->noalloc_error);
  qx::Geo g;
  int rc = g.init(geom, &ctx->last_error);
  if (rc) return rc;
  if (n_energies < 0) return ctx->fail(RGF_EINVAL, "n_energies < 0");
  if (n_energies == 0) return RGF_OK;
  rc = check_solve_args(ctx, g, a, sigma_lesser, sigma_greater, obc, g_lesser, g_greater, current);
  if (rc) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  const bool have_gr = g_retarded && g_retarded->diag;
  const bool cur = current != nullptr;
  const bool small = ctx->persistent && small_geometry(g);
  // tables first, then the workspace; the batched sweep runs in chunks that
  // fit, the persistent sweep needs its whole per-workgroup scratch
  const size_t total = RGF_NOALLOC_SCRATCH_BYTES;
  const size_t tbytes = (table_bytes(g) + 255) & ~(size_t)255;
  const size_t need = small ? small_work_bytes(g, n_energies, have_gr, cur, ctx->n_cu)
                            : work_bytes(g, 1, true, cur, have_gr);
  if (tbytes + need > total)
    return ctx->fail(RGF_EUNSUPPORTED,
                     "allocation-free solve: this geometry needs " + std::to_string(tbytes + need) +
                         " bytes of scratch (" + (small ? "persistent sweep" : "one energy point") +
                         "), the fixed array holds " + std::to_string(total));
  // point the context at the array for this call, then restore it
  void* const saved_ws = ctx->ws;
  const size_t saved_ws_bytes = ctx->ws_bytes;
  const bool saved_ws_owned = ctx->ws_owned;
  const TableState saved_tables = TableState::save(*ctx);
  const bool saved_overlap = ctx->overlap, saved_profile = ctx->profile;
  ctx->ws = ctx->noalloc_scratch + tbytes;
  ctx->ws_bytes = total - tbytes;
  ctx->ws_owned = false;
  ctx->tables = ctx->noalloc_scratch;
  ctx->tables_cap = tbytes;
  ctx->tables_used = 0;
  ctx->tables_fixed = true;
  ctx->overlap = false;  // the side stream and its events are created lazily
  ctx->profile = false;  // profiling creates events
  rc = ctx->reset_info();
  if (!rc) {
    if (small) {
      rc = run_small(ctx, {SmallSpec{&g, n_energies, solve_io(g, 0, a, sigma_lesser, sigma_greater, obc,
                                                               g_retarded, g_lesser, g_greater, current)}},
                     static_cast<char*>(ctx->ws));
    } else {
      rc = solve_chunked(ctx, g, n_energies, a, sigma_lesser, sigma_greater, obc, g_retarded, g_lesser,
                         g_greater, current);
    }
  }
  ctx->ws = saved_ws;
  ctx->ws_bytes = saved_ws_bytes;
  ctx->ws_owned = saved_ws_owned;
  saved_tables.restore(ctx);
  ctx->overlap = saved_overlap;
  ctx->profile = saved_profile;
  return rc;
  GUARD_END
}

int rgf_selected_solve_batch_z(rgf_ctx* ctx, int32_t n_problems, const rgf_solve_problem* problems) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  if (n_problems < 0 || (n_problems > 0 && !problems)) return ctx->fail(RGF_EINVAL, "batch: problems required");
  std::vector<qx::Geo> geos(n_problems);
  for (int p = 0; p < n_problems; ++p) {
    const rgf_solve_problem& q = problems[p];
    int rc = geos[p].init(&q.geom, &ctx->last_error);
    if (rc) return rc;
    if (q.n_energies < 0) return ctx->fail(RGF_EINVAL, "batch: n_energies < 0");
    rc = check_solve_args(ctx, geos[p], &q.a, &q.sigma_lesser, &q.sigma_greater, q.obc, &q.g_lesser,
                          &q.g_greater, q.current);
    if (rc) return rc;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  int rc = ctx->reset_info();
  if (rc) return rc;
  // geometries with every block <= 32 share one persistent launch
  std::vector<SmallSpec> small;
  size_t per_wg = 0;
  int64_t small_energies = 0;
  for (int p = 0; p < n_problems; ++p) {
    const rgf_solve_problem& q = problems[p];
    if (!ctx->persistent || !small_geometry(geos[p]) || q.n_energies == 0) continue;
    rgf_bt gr = q.g_retarded, gl = q.g_lesser, gg = q.g_greater;
    small.push_back(SmallSpec{&geos[p], q.n_energies,
                              solve_io(geos[p], 0, &q.a, &q.sigma_lesser, &q.sigma_greater, q.obc,
                                       gr.diag ? &gr : nullptr, &gl, &gg, q.current)});
    per_wg = std::max(per_wg, small_per_wg(geos[p], gr.diag != nullptr, q.current != nullptr));
    small_energies += q.n_energies;
  }
  if (!small.empty()) {
    char* ws = nullptr;
    if (per_wg) {
      const size_t grid = (size_t)std::min<int64_t>(small_energies, ctx->n_cu);
      rc = ctx->ensure_workspace(per_wg * grid + 256, &ws);
      if (rc) return rc;
    }
    rc = run_small(ctx, small, ws);
    if (rc) return rc;
  }
  for (int p = 0; p < n_problems; ++p) {
    const rgf_solve_problem& q = problems[p];
    if ((ctx->persistent && small_geometry(geos[p])) || q.n_energies == 0) continue;
    rgf_bt gr = q.g_retarded, gl = q.g_lesser, gg = q.g_greater;
    rc = solve_chunked(ctx, geos[p], q.n_energies, &q.a, &q.sigma_lesser, &q.sigma_greater, q.obc,
                       gr.diag ? &gr : nullptr, &gl, &gg, q.current);
    if (rc) return rc;
  }
  return RGF_OK;
  GUARD_END
}

int rgf_autotune(rgf_ctx* ctx, int64_t b, int64_t n_energies, int32_t f32, int32_t* best_cfg,
                 float* ms_per_cfg) {
  if (!ctx || !best_cfg) return RGF_EINVAL;
  GUARD_BEGIN
  if (b < 1 || b > 1024 || n_energies < 1) return ctx->fail(RGF_EINVAL, "autotune: bad sizes");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  const size_t el = (size_t)b * b * n_energies;
  double2 *A = nullptr, *B = nullptr, *D = nullptr;
  int rc = ctx->hip(hipMalloc(&A, el * sizeof(double2)), "autotune alloc");
  if (!rc) rc = ctx->hip(hipMalloc(&B, el * sizeof(double2)), "autotune alloc");
  if (!rc) rc = ctx->hip(hipMalloc(&D, el * sizeof(double2)), "autotune alloc");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc = ctx->hip(hipEventCreate(&e0), "event");
  if (!rc) rc = ctx->hip(hipEventCreate(&e1), "event");
  if (!rc) rc = ctx->hip(hipMemsetAsync(A, 0, el * sizeof(double2), ctx->stream), "memset");
  if (!rc) rc = ctx->hip(hipMemsetAsync(B, 0, el * sizeof(double2), ctx->stream), "memset");
  float best = 1e30f;
  int best_c = -1;
  const int64_t s = b * b;
  // a representative step: three products (N.N, N.C with an addend, two-term)
  std::vector<ZProblem> probs = {
      PB((int)b, (int)b).term(Mat{A, s, (int)b}, kOpN, Mat{B, s, (int)b}, kOpN, (int)b).out(Mat{D, s, (int)b}).done(),
      PB((int)b, (int)b).term(Mat{A, s, (int)b}, kOpN, Mat{B, s, (int)b}, kOpC, (int)b)
          .add(Mat{A, s, (int)b}, cmk(1.0, 0.0)).out(Mat{D, s, (int)b}).done(),
      PB((int)b, (int)b).term(Mat{B, s, (int)b}, kOpN, Mat{A, s, (int)b}, kOpN, (int)b)
          .term(Mat{A, s, (int)b}, kOpC, Mat{B, s, (int)b}, kOpN, (int)b, cmk(-1.0, 0.0))
          .out(Mat{D, s, (int)b}).done()};
  const bool saved32 = ctx->prec32;
  const int saved_over = ctx->tile_override;
  const bool saved_prof = ctx->profile;
  ctx->prec32 = f32 != 0;
  ctx->profile = false;
  // Configurations are timed round-robin over several rounds (minimum per
  // configuration), so clock drift cannot favour one; another configuration
  // replaces the untuned default only when it is at least 2% faster.
  float tmin[qx::kNumTileCfgs];
  for (int cfg = 0; cfg < qx::kNumTileCfgs; ++cfg) tmin[cfg] = -1.f;
  for (int round = 0; round < 5 && !rc; ++round) {
    for (int cfg = 0; cfg < qx::kNumTileCfgs && !rc; ++cfg) {
      const bool ok = f32 ? qx::f32_tile_ok(cfg) : true;
      if (!ok) continue;
      ctx->tile_override = cfg;
      if (round == 0) rc = ctx->gemm(probs, (int)n_energies, (int)b);  // warm-up
      if (!rc) rc = ctx->hip(hipEventRecord(e0, ctx->stream), "event");
      for (int r = 0; r < 2 && !rc; ++r) rc = ctx->gemm(probs, (int)n_energies, (int)b);
      if (!rc) rc = ctx->hip(hipEventRecord(e1, ctx->stream), "event");
      if (!rc) rc = ctx->hip(hipEventSynchronize(e1), "event sync");
      float ms = 0.f;
      if (!rc) rc = ctx->hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      ms *= 0.5f;
      if (!rc && (tmin[cfg] < 0.f || ms < tmin[cfg])) tmin[cfg] = ms;
    }
  }
  for (int cfg = 0; cfg < qx::kNumTileCfgs; ++cfg) {
    if (ms_per_cfg) ms_per_cfg[cfg] = tmin[cfg];
    if (tmin[cfg] >= 0.f && tmin[cfg] < best) { best = tmin[cfg]; best_c = cfg; }
  }
  const int def_c = qx::default_tile((int)b, f32 != 0);
  if (best_c != def_c && tmin[def_c] >= 0.f && best > 0.98f * tmin[def_c]) best_c = def_c;
  ctx->prec32 = saved32;
  ctx->tile_override = saved_over;
  ctx->profile = saved_prof;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(ctx->stream);
  if (A) (void)hipFree(A);
  if (B) (void)hipFree(B);
  if (D) (void)hipFree(D);
  if (rc) return rc;
  *best_cfg = best_c;
  std::vector<int>& t = f32 ? ctx->tile_table32 : ctx->tile_table;
  if ((int)t.size() <= b) t.resize(b + 1, -1);
  t[b] = best_c;
  return RGF_OK;
  GUARD_END
}

int rgf_graph_capture_solve(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                            const rgf_bt_const* a, const rgf_bt_const* sigma_lesser,
                            const rgf_bt_const* sigma_greater, const rgf_obc* obc,
                            int32_t precision, int32_t refine, rgf_bt* g_retarded,
                            rgf_bt* g_lesser, rgf_bt* g_greater, void* current,
                            rgf_graph** out) {
  if (!ctx || !out) return RGF_EINVAL;
  *out = nullptr;
  GUARD_BEGIN
  qx::Geo g;
  int rc = g.init(geom, &ctx->last_error);
  if (rc) return rc;
  if (n_energies < 1 || n_energies > 65535) return ctx->fail(RGF_EINVAL, "graph: 1 <= n_energies <= 65535");
  if (precision != 0 && precision != 1) return ctx->fail(RGF_EINVAL, "graph: precision 0 or 1");
  if (refine < -1) return ctx->fail(RGF_EINVAL, "graph: refine must be >= -1");
  const bool multi = g.N > 1;
  if (!a || !a->diag || (multi && (!a->upper || !a->lower)) || !sigma_lesser || !sigma_greater ||
      !sigma_lesser->diag || !sigma_greater->diag ||
      (multi && (!sigma_lesser->upper || !sigma_greater->upper)) || !g_lesser || !g_greater ||
      !g_lesser->diag || !g_greater->diag || (multi && (!g_lesser->upper || !g_greater->upper)))
    return ctx->fail(RGF_EINVAL, "graph: missing buffers");
  if (current && !current_obc_ok(obc, g))
    return ctx->fail(RGF_EINVAL, "current needs lesser/greater OBC blocks on the first and last block");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  std::unique_lock<std::shared_mutex> lock(g_graph_mutex);
  rgf_graph* gr = new rgf_graph();
  gr->ctx = ctx;
  if (hipEventCreateWithFlags(&gr->done, hipEventDisableTiming) != hipSuccess) {
    delete gr;
    return ctx->fail(RGF_EHIP, "graph event");
  }
  const bool have_gr = g_retarded && g_retarded->diag;
  const bool mixed = precision == 1;
  const bool cur = current != nullptr;
  const bool small = !mixed && ctx->persistent && small_geometry(g);
  const size_t wbytes = small ? std::max<size_t>(256, small_work_bytes(g, n_energies, have_gr, cur, ctx->n_cu))
                              : work_bytes(g, n_energies, true, cur, have_gr, mixed, true);
  const size_t tbytes = table_bytes(g);
  // An FP64 graph shares a caller-provided context workspace that is large
  // enough (it must then outlive the graph; graphs and solves on the context
  // stay serial on its stream).  Otherwise - and always for the mixed sweep,
  // whose identity block is written once before capture - the graph owns one.
  char* wsp = nullptr;
  if (!mixed && !ctx->ws_owned && ctx->ws && ctx->ws_bytes >= wbytes) {
    wsp = static_cast<char*>(ctx->ws);
  } else {
    rc = ctx->hip(hipMalloc(&gr->ws, wbytes), "graph workspace");
    if (rc) { delete gr; return rc; }
    wsp = static_cast<char*>(gr->ws);
  }
  rc = ctx->hip(hipMalloc(&gr->tables, tbytes), "graph tables");
  if (rc) { delete gr; return rc; }
  // The graph records with its own table memory: geometry tables written
  // now, the job list of the current by a kernel node of the graph itself.
  const TableState saved_tables = TableState::save(*ctx);
  ctx->tables = static_cast<char*>(gr->tables);
  ctx->tables_cap = tbytes;
  ctx->tables_fixed = true;
  rc = ctx->upload_geometry(g);
  if (!rc) rc = ctx->hip(hipStreamSynchronize(ctx->stream), "graph tables sync");
  if (rc) { saved_tables.restore(ctx); delete gr; return rc; }
  Work w;
  carve(g, n_energies, true, cur, have_gr, wsp, &w, mixed, true);
  if (mixed) {
    rc = ctx->hip(launch_identity(w.ident, (int)g.bmax, ctx->stream), "identity");
    if (!rc) rc = ctx->hip(hipStreamSynchronize(ctx->stream), "identity sync");
    if (rc) { saved_tables.restore(ctx); delete gr; return rc; }
  }
  SolveIO io;
  io.A = view(g, a->diag, a->upper, a->lower, 0);
  io.SL = view(g, sigma_lesser->diag, sigma_lesser->upper, nullptr, 0);
  io.SG = view(g, sigma_greater->diag, sigma_greater->upper, nullptr, 0);
  io.obc = obc;
  io.GR = have_gr ? view(g, g_retarded->diag, g_retarded->upper, g_retarded->lower, 0)
                  : BTV{&g, nullptr, nullptr, nullptr, 0, -1, -1};
  io.GL = view(g, g_lesser->diag, g_lesser->upper, g_lesser->lower, 0);
  io.GG = view(g, g_greater->diag, g_greater->upper, g_greater->lower, 0);
  io.current = static_cast<double2*>(current);
  // record with the graph's own tables and flags; restore the context afterwards
  const bool saved_prof = ctx->profile, saved32 = ctx->prec32;
  const int saved_ref = ctx->refine;
  ctx->profile = false;
  ctx->prec32 = mixed;
  ctx->refine = mixed ? refine : 0;
  // Capture is recorded on a private stream (the caller's stream may be the
  // legacy NULL stream, which cannot be captured); the graph is launched on
  // the context's stream.
  hipGraph_t graph = nullptr;
  hipStream_t saved_stream = ctx->stream;
  hipStream_t cap = nullptr;
  rc = ctx->hip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "capture stream");
  if (!rc) {
    ctx->stream = cap;
    rc = ctx->hip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture");
  }
  int rc_sweep = RGF_OK;
  if (!rc) {
    rc_sweep = ctx->reset_info();
    if (!rc_sweep)
      rc_sweep = small ? run_small(ctx, {SmallSpec{&g, n_energies, io}}, wsp)
                       : sweep_chunk_solve(ctx, g, 0, (int)n_energies, io, w);
    rc = ctx->hip(hipStreamEndCapture(cap, &graph), "end capture");
  }
  ctx->stream = saved_stream;
  if (cap) hipStreamDestroy(cap);
  saved_tables.restore(ctx);
  ctx->profile = saved_prof;
  ctx->prec32 = saved32;
  ctx->refine = saved_ref;
  if (rc_sweep) rc = rc_sweep;
  if (!rc) rc = ctx->hip(hipGraphInstantiate(&gr->exec, graph, nullptr, nullptr, 0), "instantiate");
  gr->graph = graph;
  if (rc) { delete gr; return rc; }
  *out = gr;
  return RGF_OK;
  GUARD_END
}

int rgf_graph_launch(rgf_graph* graph) {
  if (!graph || !graph->exec) return RGF_EINVAL;
  rgf_ctx* ctx = graph->ctx;
  std::shared_lock<std::shared_mutex> lock(g_graph_mutex);
  int rc = ctx->hip(hipGraphLaunch(graph->exec, ctx->stream), "graph launch");
  if (!rc) rc = ctx->hip(hipEventRecord(graph->done, ctx->stream), "graph done record");
  return rc;
}

int rgf_graph_destroy(rgf_graph* graph) {
  if (!graph) return RGF_OK;
  delete graph;
  return RGF_OK;
}

int rgf_densify_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies, const void* data,
                  int64_t data_stride, const rgf_block_map* map, rgf_bt* out) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  qx::Geo g;
  int rc = g.init(geom, &ctx->last_error);
  if (rc) return rc;
  if (n_energies < 0 || n_energies > 65535) return ctx->fail(RGF_EINVAL, "densify: 0 <= n_energies <= 65535");
  if (n_energies == 0) return RGF_OK;
  if (!data || !map || !out || data_stride < 0) return ctx->fail(RGF_EINVAL, "densify: data, map and out required");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  double2* parts[3] = {static_cast<double2*>(out->diag), static_cast<double2*>(out->upper),
                       static_cast<double2*>(out->lower)};
  const int64_t size[3] = {g.Dsz, g.Usz, g.Usz};
  for (int p = 0; p < 3; ++p) {
    const rgf_block_map_part& m = map->part[p];
    if (m.n > 0 && (!parts[p] || !m.src || !m.dst || !m.op))
      return ctx->fail(RGF_EINVAL, "densify: map part without output or index arrays");
  }
  for (int p = 0; p < 3; ++p) {
    if (!parts[p]) continue;
    const rgf_block_map_part& m = map->part[p];
    if (!(m.covers_part && m.n > 0)) {
      rc = ctx->hip(launch_fill_zero(parts[p], size[p], size[p], (int)n_energies, ctx->stream), "densify zero");
      if (rc) return rc;
    }
    rc = ctx->hip(launch_densify(static_cast<const double2*>(data), data_stride, m.n, m.src, m.dst, m.op,
                                 parts[p], size[p], (int)n_energies, ctx->stream),
                  "densify");
    if (rc) return rc;
  }
  return RGF_OK;
  GUARD_END
}

int rgf_sparsify_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                   const rgf_bt_const* in, const rgf_block_map* map, void* data,
                   int64_t data_stride) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  qx::Geo g;
  int rc = g.init(geom, &ctx->last_error);
  if (rc) return rc;
  if (n_energies < 0 || n_energies > 65535) return ctx->fail(RGF_EINVAL, "sparsify: 0 <= n_energies <= 65535");
  if (n_energies == 0) return RGF_OK;
  if (!data || !map || !in || data_stride < 0) return ctx->fail(RGF_EINVAL, "sparsify: data, map and in required");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  const double2* parts[3] = {static_cast<const double2*>(in->diag), static_cast<const double2*>(in->upper),
                             static_cast<const double2*>(in->lower)};
  const int64_t size[3] = {g.Dsz, g.Usz, g.Usz};
  for (int p = 0; p < 3; ++p) {
    const rgf_block_map_part& m = map->part[p];
    if (m.n <= 0 || !parts[p]) continue;  // parts not given are left untouched in data
    if (!m.src || !m.dst || !m.op) return ctx->fail(RGF_EINVAL, "sparsify: map part without index arrays");
    rc = ctx->hip(launch_sparsify(parts[p], size[p], m.n, m.src, m.dst, m.op,
                                  static_cast<double2*>(data), data_stride, (int)n_energies, ctx->stream),
                  "sparsify");
    if (rc) return rc;
  }
  return RGF_OK;
  GUARD_END
}

int rgf_profile_read(rgf_ctx* ctx, rgf_profile* out, int32_t reset) {
  if (!ctx || !out) return RGF_EINVAL;
  int rc = ctx->prof_collect();
  if (rc) return rc;
  out->gemm_launches = ctx->prof_gemm_launches;
  out->gemm_ms = ctx->prof_gemm_ms;
  out->gemm_flops = ctx->prof_gemm_flops;
  out->gemm_bytes = ctx->prof_gemm_bytes;
  out->inv_calls = ctx->prof_inv_calls;
  out->inv_ms = ctx->prof_inv_ms;
  out->inv_flops = ctx->prof_inv_flops;
  if (reset) {
    ctx->prof_gemm_launches = ctx->prof_inv_calls = 0;
    ctx->prof_gemm_ms = ctx->prof_gemm_flops = ctx->prof_gemm_bytes = 0;
    ctx->prof_inv_ms = ctx->prof_inv_flops = 0;
  }
  return RGF_OK;
}

int rgf_selected_solve_mixed(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                             const rgf_bt_const* a, const rgf_bt_const* sigma_lesser,
                             const rgf_bt_const* sigma_greater, const rgf_obc* obc,
                             const rgf_mixed_opts* opts, rgf_bt* g_retarded, rgf_bt* g_lesser,
                             rgf_bt* g_greater, void* current) {
  if (!ctx) return RGF_EINVAL;
  GUARD_BEGIN
  qx::Geo g;
  int rc = g.init(geom, &ctx->last_error);
  if (rc) return rc;
  if (!opts) return ctx->fail(RGF_EINVAL, "mixed: options required");
  if (opts->refine < -1) return ctx->fail(RGF_EINVAL, "mixed: refine must be >= -1");
  if (n_energies < 0) return ctx->fail(RGF_EINVAL, "n_energies < 0");
  if (n_energies == 0) return RGF_OK;
  const bool multi = g.N > 1;
  rc = check_solve_args(ctx, g, a, sigma_lesser, sigma_greater, obc, g_lesser, g_greater, current);
  if (rc) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(RGF_EHIP, "hipSetDevice");
  const bool have_gr = g_retarded && g_retarded->diag;
  const bool have_off = have_gr && (!multi || (g_retarded->upper && g_retarded->lower));
  const bool cur = current != nullptr;
  const size_t per = work_bytes(g, 1, true, cur, have_gr, true, have_off);
  int64_t chunk = ctx->chunk_for(g, n_energies, true, cur, have_gr);
  chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, (int64_t)(ctx->ws_budget / per)));
  rc = ctx->upload_geometry(g);
  if (rc) return rc;
  char* ws = nullptr;
  rc = ctx->ensure_workspace(work_bytes(g, chunk, true, cur, have_gr, true, have_off), &ws);
  if (rc) return rc;
  rc = ctx->reset_info();
  if (rc) return rc;
  Work w;
  carve(g, chunk, true, cur, have_gr, ws, &w, true, have_off);
  rc = ctx->hip(launch_identity(w.ident, (int)g.bmax, ctx->stream), "identity");
  if (rc) return rc;
  std::vector<double> resid;
  std::vector<int32_t> flags;
  for (int64_t e0 = 0; e0 < n_energies; e0 += chunk) {
    const int E = (int)std::min<int64_t>(chunk, n_energies - e0);
    SolveIO io;
    io.A = view(g, a->diag, a->upper, a->lower, e0);
    io.SL = view(g, sigma_lesser->diag, sigma_lesser->upper, nullptr, e0);
    io.SG = view(g, sigma_greater->diag, sigma_greater->upper, nullptr, e0);
    io.obc = obc;
    if (have_gr) {
      io.GR = view(g, g_retarded->diag, have_off ? g_retarded->upper : w.gru,
                   have_off ? g_retarded->lower : w.grl, e0);
      if (!have_off) { io.GR.eu = 0; io.GR.el = 0; }
    } else {
      io.GR = BTV{&g, nullptr, w.gru, w.grl, 0, 0, 0};
    }
    io.GL = view(g, g_lesser->diag, g_lesser->upper, g_lesser->lower, e0);
    io.GG = view(g, g_greater->diag, g_greater->upper, g_greater->lower, e0);
    io.current = static_cast<double2*>(current);
    ctx->prec32 = true;
    ctx->refine = opts->refine;
    rc = sweep_chunk_solve(ctx, g, e0, E, io, w);
    ctx->prec32 = false;
    ctx->refine = 0;
    if (rc) return rc;
    const bool want_check = opts->fallback_tol > 0 || opts->residuals || opts->fallback_flags;
    if (!want_check) continue;
    BTV GRc = io.GR;
    if (!GRc.d) GRc = BTV{&g, w.gr, io.GR.u, io.GR.l, 0, io.GR.bu(), io.GR.bl()};
    rc = check_chunk(ctx, g, e0, E, io, GRc, w, &resid);
    if (rc) return rc;
    flags.assign(E, 0);
    std::vector<int> redo;
    if (opts->fallback_tol > 0)
      for (int e = 0; e < E; ++e)
        if (!(resid[e] <= opts->fallback_tol)) { flags[e] = 1; redo.push_back(e); }
    if (!redo.empty()) {
      if ((int)redo.size() * 4 > E) {
        rc = sweep_chunk_solve(ctx, g, e0, E, io, w);
        if (rc) return rc;
      } else {
        for (int e : redo) {
          SolveIO one = io;
          const int64_t ge = e0 + e;
          one.A = view(g, a->diag, a->upper, a->lower, ge);
          one.SL = view(g, sigma_lesser->diag, sigma_lesser->upper, nullptr, ge);
          one.SG = view(g, sigma_greater->diag, sigma_greater->upper, nullptr, ge);
          one.GL = view(g, g_lesser->diag, g_lesser->upper, g_lesser->lower, ge);
          one.GG = view(g, g_greater->diag, g_greater->upper, g_greater->lower, ge);
          one.GR.e0 = have_gr ? ge : 0;
          if (!have_off) { one.GR.eu = 0; one.GR.el = 0; }
          else if (have_gr) { one.GR.eu = -1; one.GR.el = -1; }
          rc = sweep_chunk_solve(ctx, g, ge, 1, one, w);
          if (rc) return rc;
        }
      }
    }
    if (opts->residuals) {
      rc = ctx->hip(hipMemcpyAsync(opts->residuals + e0, resid.data(), E * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->stream),
                    "write residuals");
      if (rc) return rc;
    }
    if (opts->fallback_flags) {
      rc = ctx->hip(hipMemcpyAsync(opts->fallback_flags + e0, flags.data(), E * sizeof(int32_t),
                                   hipMemcpyHostToDevice, ctx->stream),
                    "write flags");
      if (rc) return rc;
    }
    rc = ctx->hip(hipStreamSynchronize(ctx->stream), "mixed sync");  // host vectors reused
    if (rc) return rc;
  }
  return RGF_OK;
  GUARD_END
}

}  // extern "C"
