// Host-side driver types shared by the FP64 and mixed-precision sweeps.
#pragma once

#include <new>
#include <string>
#include <vector>

#include "../../include/rgf.h"
#include "qx_common.h"

namespace qx {

// Energy-strided matrix view: element (r, c) of energy e is p[e*s + r*ld + c].
struct Mat {
  double2* p = nullptr;
  int64_t s = 0;
  int ld = 0;
};

struct Geo {
  int N = 0;
  std::vector<int64_t> bs, od, ou;
  int64_t Dsz = 0, Usz = 0, bmax = 0;
  int init(const rgf_geometry* g, std::string* err);
};

// Builder for one grouped-GEMM problem.
struct PB {
  ZProblem P;
  PB(int m, int n);
  PB& term(Mat A, int opA, Mat B, int opB, int k, double2 alpha = double2{1.0, 0.0});
  PB& add(Mat F, double2 c1, double2 c2 = double2{0.0, 0.0});
  PB& out(Mat D);
  PB& out2(int mode, Mat D2, double2 s2 = double2{1.0, 0.0});
  PB& partial_here();
  ZProblem done() const;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  bool ws_owned = true;
  size_t ws_budget = (size_t)4 << 30;
  int64_t max_chunk = 0;
  int tile_override = -1;
  std::vector<int> tile_table;    // FP64 products, by block size
  std::vector<int> tile_table32;  // FP32-compute products, by block size
  int* info_dev = nullptr;
  // context-owned table memory, grown on demand
  void* meta = nullptr;
  size_t meta_bytes = 0;
  // Where the current call's geometry tables and job lists (trace / norm
  // jobs, boundary tables) go: normally `meta`; a graph's own table memory
  // while the graph is captured; this device's fixed array in the
  // allocation-free entry point.  `tables_fixed`: never grown or replaced.
  char* tables = nullptr;
  size_t tables_cap = 0;
  size_t tables_used = 0;
  bool tables_fixed = false;
  int64_t* offs_dev = nullptr;
  int* sizes_dev = nullptr;
  std::string last_error;
  // device address of this GPU's fixed scratch array (the allocation-free
  // entry point), looked up when the context is created; null with the
  // reason in `noalloc_error` when the lookup failed
  char* noalloc_scratch = nullptr;
  std::string noalloc_error;

  // mixed precision: sweep products on FP32 matrix cores; `refine` FP64
  // Newton steps per left-connected inverse
  bool prec32 = false;
  int refine = 0;
  // one product group per launch (reference op order, no grouping)
  bool unfused = false;
  // geometries with every block <= kSmallMaxBlock run the persistent on-chip
  // sweep (one workgroup per energy, intermediates in LDS)
  bool persistent = true;
  int n_cu = 256;  // compute units of the device (persistent grid size)
  // run each block inversion on a second stream, overlapped with the
  // independent lesser/greater products of the same forward step
  bool overlap = true;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int fork();   // side stream waits for the work queued so far on `stream`
  int join();   // `stream` waits for the work queued so far on `side`

  // launch profiling (RGF_OPT_PROFILE)
  bool profile = false;
  struct Timed { hipEvent_t a, b; double flops, bytes; int kind; };
  std::vector<Timed> timed;
  std::vector<hipEvent_t> event_pool;
  int64_t prof_gemm_launches = 0, prof_inv_calls = 0;
  double prof_gemm_ms = 0, prof_gemm_flops = 0, prof_gemm_bytes = 0, prof_inv_ms = 0,
         prof_inv_flops = 0;
  hipEvent_t take_event();
  int prof_begin(hipEvent_t* a, hipStream_t s = nullptr);
  int prof_end(hipEvent_t a, double flops, int kind, hipStream_t s = nullptr, double bytes = 0);
  int prof_collect();

  virtual ~Ctx();
  int fail(int code, const std::string& msg);
  int hip(hipError_t e, const char* what);
  int ensure_workspace(size_t bytes, char** out);
  int ensure_meta(size_t bytes);
  int tile_for(int b) const;
  int gemm(const std::vector<ZProblem>& probs, int nE, int b);
  int inv(double2* S, int64_t sS, int n, int nE, Mat out, char* aux, int tag, int64_t e0,
          hipStream_t on = nullptr);
  int reset_info();
  int upload_geometry(const Geo& g);
  // appends a host table to `tables` (stream-ordered, no host wait)
  int put_table(const void* host, size_t bytes, void** dev);
  int64_t chunk_for(const Geo& g, int64_t nE, bool solve, bool current, bool have_gr);
};

}  // namespace qx

// The opaque handle of the C ABI.
struct rgf_ctx : public qx::Ctx {};

// C ABI bodies: no C++ exception crosses the interface.
#define GUARD_BEGIN try {
#define GUARD_END                                              \
  }                                                            \
  catch (const std::bad_alloc&) {                              \
    return ctx ? ctx->fail(RGF_ENOMEM, "host allocation failed") : RGF_ENOMEM; \
  }                                                            \
  catch (...) {                                                \
    return ctx ? ctx->fail(RGF_EHIP, "internal error") : RGF_EHIP; \
  }
