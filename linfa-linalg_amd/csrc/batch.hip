// Batched solves of many SPD systems of uneven sizes (include/llh.h:
// llh_batch_*): the planner, the captured hipGraph and the autotuner.
#include <algorithm>
#include <climits>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "llh.h"
#include "llh_context.h"
#include "llh_kernels.h"

using namespace llh;

namespace llh {
namespace {

// The init/finish kernels read the status pointer from the device base so a
// replayed graph follows the buffers of the current call.
template <typename T>
__global__ void batch_init_from_base(const BatchBase<T>* base, int64_t count, int ncounters) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const BatchBase<T> bb = *base;
  if (i < count) bb.status[i] = kNoFail;
  if (i < ncounters) bb.counters[i] = 0;
}
template <typename T>
__global__ void batch_finish_from_base(const BatchBase<T>* base, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const BatchBase<T> bb = *base;
  if (i < count && bb.status[i] == kNoFail) bb.status[i] = -1;
}

}  // namespace
}  // namespace llh

struct llh_batch_plan {
  static constexpr uint32_t kMagic = 0x4c4c4842;  // "LLHB"
  uint32_t magic = kMagic;
  llh_precision prec = LLH_PREC_F64;
  int device = 0;
  size_t count = 0;
  std::vector<int32_t> n, m, lda;
  std::vector<uint64_t> a_off, b_off;
  size_t a_elems = 0, b_elems = 0;
  llh_batch_config cfg{0, 0, 0, 0};
  int cus = 256;

  // schedule (rebuilt when the configuration changes)
  struct SmallLaunch {
    int first, count, nmax, grid, counter;
  };
  // The grouped blocked factor as a sequence of launches over work lists:
  // diagonal blocks (StepWork, also the forward substitution's blocks), panel
  // tiles and trailing tiles (TileWork).
  enum PhaseKind { kDiagPhase, kPanelPhase, kSyrkPhase };
  struct Phase {
    PhaseKind kind;
    int off, count;
  };
  std::vector<SmallLaunch> small;
  std::vector<Phase> phases;
  std::vector<std::pair<int, int>> bwd;  // StepWork ranges of the back-solve steps
  int res_off = 0, nres = 0;             // StepWork range of the residual blocks (refined mode)
  int nclasses = 0;                      // small-system launches per pass
  int med_first = 0, nmed = 0;           // items of the blocked (medium) bucket
  bool lds_solve = false;                // every medium n <= kSolveLdsMaxN: one substitution launch
  bool med_fused = false;                // the medium bucket's factor runs as one fused per-system launch
  bool med_fused_solve = false;          // ... and its substitutions in the same launch
  bool fused_fwd = false;                // grouped medium bucket: the factor's steps carry the forward sweep
  void* d_y = nullptr;                   // ... into this workspace (B's layout)
  int med_counter = 0, med_grid = 0;     // its work-queue counter and resident grid
  int ncounters = 0;
  size_t linv_elems = 0;

  // device memory
  void* d_items = nullptr;
  void* d_step = nullptr;
  void* d_tile = nullptr;
  void* d_base = nullptr;
  void* d_mbase = nullptr;  // MixedBase (refined mode)
  void* d_linv = nullptr;
  // refined mode: f32 copy of A, f32 right-hand side / residual / correction,
  // f64 copy of B
  float* a32 = nullptr;
  float* x32 = nullptr;
  double* bk64 = nullptr;
  int32_t* d_counters = nullptr;
  size_t linv_bytes = 0;

  hipStream_t cap = nullptr;   // capture stream (a caller's stream may be the null stream)
  hipStream_t side = nullptr;  // second branch of the captured graph
  hipEvent_t fork = nullptr, join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BatchBase<float> h_base32{};
  MixedBase h_mbase{};
};

namespace {

#define CHECK_HIP(expr, where)                             \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail(err, e_, where); \
  } while (0)

bool bad_ctx(llh_context* ctx) { return ctx == nullptr || ctx->magic != llh_context::kMagic; }
bool bad_plan(const llh_batch_plan* p) { return p == nullptr || p->magic != llh_batch_plan::kMagic; }

// Element size of the caller's buffers, and of the arithmetic of the factor.
size_t esize(const llh_batch_plan* p) { return p->prec == LLH_PREC_F32 ? 4 : 8; }
size_t compute_esize(const llh_batch_plan* p) { return p->prec == LLH_PREC_F64 ? 8 : 4; }
bool refined(const llh_batch_plan* p) { return p->prec == LLH_PREC_F32_REFINED; }
int ir_steps(const llh_batch_plan* p) { return p->cfg.ir_iterations > 0 ? p->cfg.ir_iterations : 2; }

void free_schedule(llh_batch_plan* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
  for (void** q : {&p->d_items, &p->d_step, &p->d_tile}) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
  if (p->d_counters) (void)hipFree(p->d_counters);
  p->d_counters = nullptr;
  if (p->d_y) (void)hipFree(p->d_y);
  p->d_y = nullptr;
}

// One pass over the batch: pass 0 factors and solves; a later pass (the
// refinement of the mixed mode) solves again against the stored factors.  The
// small systems run on the side stream, the blocked steps on the main stream;
// each pass has its own work-queue counters.
template <typename T>
hipError_t record_pass(llh_batch_plan* p, hipStream_t s, int pass) {
  const auto* base = static_cast<const BatchBase<T>*>(p->d_base);
  const auto* items = static_cast<const BatchItem*>(p->d_items);
  const auto* sw = static_cast<const StepWork*>(p->d_step);
  const auto* tw = static_cast<const TileWork*>(p->d_tile);
  hipError_t e;
  if ((e = hipEventRecord(p->fork, s)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(p->side, p->fork, 0)) != hipSuccess) return e;
  for (const auto& sl : p->small)
    if ((e = launch_small_batch<T>(base, items + sl.first, sl.count, sl.nmax, sl.grid,
                                   sl.counter + pass * p->nclasses, p->side, pass == 0 ? 1 : 2)) != hipSuccess)
      return e;
  if (pass == 0 && p->med_fused) {
    // every medium system's factor (and substitutions), one workgroup each
    if ((e = launch_med<T>(base, items + p->med_first, p->nmed, p->med_counter, p->med_grid,
                           p->med_fused_solve, s)) != hipSuccess)
      return e;
    if (!p->med_fused_solve)
      if ((e = launch_grp_solve<T>(base, items + p->med_first, p->nmed, s)) != hipSuccess) return e;
    if ((e = hipEventRecord(p->join, p->side)) != hipSuccess) return e;
    return hipStreamWaitEvent(s, p->join, 0);
  }
  // a diagonal-block phase is followed by that block's panel phase; the
  // forward substitution of a block (multi-launch solves) comes after both
  const StepWork* fwd_pending = nullptr;
  int fwd_count = 0;
  for (const auto& ph : p->phases) {
    switch (ph.kind) {
      case llh_batch_plan::kDiagPhase:
        if (pass == 0 && (e = launch_grp_diag_fast<T>(base, items, sw + ph.off, ph.count, s)) != hipSuccess)
          return e;
        fwd_pending = sw + ph.off;
        fwd_count = ph.count;
        break;
      case llh_batch_plan::kPanelPhase:
        if (pass == 0 && (e = launch_grp_panel<T>(base, items, tw + ph.off, ph.count, s)) != hipSuccess) return e;
        if (!p->lds_solve && fwd_pending != nullptr)
          if ((e = launch_grp_fwd<T>(base, items, fwd_pending, fwd_count, s)) != hipSuccess) return e;
        fwd_pending = nullptr;
        break;
      case llh_batch_plan::kSyrkPhase:
        if (pass == 0 && (e = launch_grp_syrk<T>(base, items, tw + ph.off, ph.count, s)) != hipSuccess) return e;
        break;
    }
  }
  if (p->lds_solve) {
    // the whole forward and back substitution of every medium system, one
    // workgroup each, after its factor
    if ((e = launch_grp_solve<T>(base, items + p->med_first, p->nmed, s)) != hipSuccess) return e;
  } else {
    for (const auto& bw : p->bwd)
      if ((e = launch_grp_bwd<T>(base, items, sw + bw.first, bw.second, s)) != hipSuccess) return e;
  }
  if ((e = hipEventRecord(p->join, p->side)) != hipSuccess) return e;
  return hipStreamWaitEvent(s, p->join, 0);
}

template <typename T>
hipError_t record(llh_batch_plan* p, hipStream_t s) {
  const auto* base = static_cast<const BatchBase<T>*>(p->d_base);
  const int64_t cnt = (int64_t)p->count;
  const int64_t initn = std::max<int64_t>(cnt, p->ncounters);
  hipLaunchKernelGGL((batch_init_from_base<T>), dim3((initn + 255) / 256), dim3(256), 0, s, base, cnt,
                     p->ncounters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (!refined(p)) {
    if ((e = record_pass<T>(p, s, 0)) != hipSuccess) return e;
  } else {
    // f32 factor and solve from copies, then refinement in f64:
    // x = x32; repeat { x32 = r = b - A x (f64 sums); x32 = solve32(x32); x += x32 },
    // the updates of x skipping systems whose factor failed (their B stays the caller's)
    const auto* mb = static_cast<const MixedBase*>(p->d_mbase);
    const auto* items = static_cast<const BatchItem*>(p->d_items);
    const auto* sw = static_cast<const StepWork*>(p->d_step);
    if ((e = launch_mixed_flat(mb, 0, (int64_t)p->a_elems, s)) != hipSuccess) return e;
    if ((e = launch_mixed_flat(mb, 1, (int64_t)p->b_elems, s)) != hipSuccess) return e;
    if ((e = record_pass<T>(p, s, 0)) != hipSuccess) return e;
    if ((e = launch_grp_update(mb, items, sw + p->res_off, p->nres, false, s)) != hipSuccess) return e;
    for (int it = 1; it <= ir_steps(p); ++it) {
      if ((e = launch_grp_residual(mb, items, sw + p->res_off, p->nres, s)) != hipSuccess) return e;
      if ((e = record_pass<T>(p, s, it)) != hipSuccess) return e;
      if ((e = launch_grp_update(mb, items, sw + p->res_off, p->nres, true, s)) != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL((batch_finish_from_base<T>), dim3((cnt + 255) / 256), dim3(256), 0, s, base, cnt);
  return hipGetLastError();
}

// Builds the bucketed schedule for the plan's configuration, uploads the work
// lists and captures the graph on the context's stream.
llh_status build_schedule(llh_context* ctx, llh_batch_plan* p, llh_error* err) {
  free_schedule(p);
  const int smax = p->prec == LLH_PREC_F64 ? kSmallMaxD : kSmallMaxS;
  p->cfg.small_limit = std::max(0, std::min(p->cfg.small_limit, smax));
  p->cfg.small_waves = std::max(1, std::min(p->cfg.small_waves, 16));
  p->cfg.medium_mode = std::max(0, std::min(p->cfg.medium_mode, 2));
  std::vector<int> order(p->count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return p->n[x] > p->n[y]; });
  std::vector<BatchItem> items;
  std::vector<int> small_idx, med_idx;
  for (int i : order) {
    if (p->n[i] <= 0) continue;
    (p->n[i] <= p->cfg.small_limit ? small_idx : med_idx).push_back(i);
  }
  auto mk = [&](int i, int64_t linv_off) {
    return BatchItem{(int64_t)p->a_off[i], (int64_t)p->b_off[i], linv_off, p->n[i], p->m[i], p->lda[i], i};
  };
  // small systems: one persistent launch per LDS class, largest first
  p->small.clear();
  int counter = 0;
  const int classes[3] = {32, 64, smax};
  size_t pos = 0;
  for (int ci = 0; ci < 3; ++ci) {
    const int lo = ci == 0 ? 0 : classes[ci - 1];
    const int first = (int)items.size();
    // small_idx is sorted by decreasing n: take the systems of this class
    std::vector<int> cls;
    for (size_t q = 0; q < small_idx.size(); ++q)
      if (p->n[small_idx[q]] > lo && p->n[small_idx[q]] <= classes[ci]) cls.push_back(small_idx[q]);
    if (cls.empty()) continue;
    for (int i : cls) items.push_back(mk(i, 0));
    const int per_cu = classes[ci] <= 32 ? 8 : (classes[ci] <= 64 ? 4 : 1);
    const int grid = std::max(1, p->cus * std::min(per_cu, p->cfg.small_waves));
    p->small.push_back({first, (int)cls.size(), classes[ci], grid, counter++});
    pos += cls.size();
  }
  (void)pos;
  p->nclasses = counter;
  p->ncounters = std::max(counter * (refined(p) ? 1 + ir_steps(p) : 1), 1);
  // larger systems: grouped blocked steps
  const int med_first = (int)items.size();
  int64_t linv = 0;
  int maxn = 0;
  for (int i : med_idx) {
    items.push_back(mk(i, linv));
    linv += (int64_t)((p->n[i] + kNB - 1) / kNB) * kNB * kNB;
    maxn = std::max(maxn, p->n[i]);
  }
  p->linv_elems = (size_t)linv;
  p->med_first = med_first;
  p->nmed = (int)items.size() - med_first;
  p->lds_solve = maxn <= kSolveLdsMaxN;
  p->med_fused = p->lds_solve && p->nmed > 0 && p->cfg.medium_mode != 2;
  p->med_fused_solve = p->med_fused && p->cfg.medium_mode == 0;
  // Grouped steps with one substitution launch: each diagonal step also forms
  // Y_k = inv(L_kk) Y_k and each panel tile Y_i -= L_ik Y_k from the tile it
  // has just written, so the substitution launch reads L once (back sweep
  // only).  Same products in the same order as its forward sweep.  Not in the
  // refined mode, whose solve-only passes have no factor steps.
  p->fused_fwd = p->lds_solve && p->nmed > 0 && p->cfg.medium_mode == 2 && !refined(p);
  p->med_counter = p->ncounters;
  if (p->med_fused) {
    p->ncounters += 1;
    p->med_grid = p->cus * (p->prec == LLH_PREC_F64 ? med_resident_per_cu<double>(p->med_fused_solve)
                                                    : med_resident_per_cu<float>(p->med_fused_solve));
  }
  std::vector<StepWork> sw;
  std::vector<TileWork> tw;
  p->phases.clear();
  p->bwd.clear();
  // Panels of kBatchPanel columns: 64-column blocks (diagonal block and panel
  // tiles), each followed by the update of the next blocks' columns with the
  // recursive-halving rank, and at the panel's end one update of the rest of
  // that rank -- each trailing tile is read and written once per panel (with
  // one rank-64 update per block, the C traffic of the trailing tiles was three
  // quarters of theirs).
  auto diag_phase = [&](int k0) {
    const int off = (int)sw.size();
    for (int q = med_first; q < (int)items.size(); ++q)
      if (items[q].n > k0) sw.push_back({q, k0});
    p->phases.push_back({llh_batch_plan::kDiagPhase, off, (int)sw.size() - off});
  };
  auto panel_phase = [&](int k0) {
    const int off = (int)tw.size();
    for (int q = med_first; q < (int)items.size(); ++q) {
      const int below = items[q].n - k0 - kNB;
      for (int m0 = 0; m0 < below; m0 += 64) tw.push_back({q, k0, m0, 0, kNB});
    }
    p->phases.push_back({llh_batch_plan::kPanelPhase, off, (int)tw.size() - off});
  };
  auto syrk_phase = [&](int k0, int kb, int max_cols) {
    const int off = (int)tw.size();
    for (int q = med_first; q < (int)items.size(); ++q) {
      const int m2 = items[q].n - k0 - kb;
      for (int m0 = 0; m0 < m2; m0 += 64)
        for (int n0 = 0; n0 <= m0 && n0 < max_cols; n0 += 64) tw.push_back({q, k0, m0, n0, kb});
    }
    p->phases.push_back({llh_batch_plan::kSyrkPhase, off, (int)tw.size() - off});
  };
  constexpr int nbp = kBatchPanel / kNB;
  for (int k0 = 0; k0 < maxn; k0 += kBatchPanel)
    for (int j = 0; j < nbp; ++j) {
      const int kj = k0 + kNB * j;
      if (kj >= maxn) break;
      diag_phase(kj);
      panel_phase(kj);
      if (kj + kNB >= maxn) break;
      if (j + 1 == nbp) {
        syrk_phase(k0, kBatchPanel, INT_MAX);
      } else {
        const int r = lowbit(j + 1);
        syrk_phase(k0 + kNB * (j + 1 - r), kNB * r, kNB * std::min(r, nbp - j - 1));
      }
    }
  for (int k0 = ((maxn - 1) / kNB) * kNB; maxn > 0 && k0 >= 0; k0 -= kNB) {
    const int off = (int)sw.size();
    for (int q = med_first; q < (int)items.size(); ++q)
      if (items[q].n > k0) sw.push_back({q, k0});
    p->bwd.push_back({off, (int)sw.size() - off});
  }
  // refined mode: the residual of every system, 64 rows per workgroup
  p->res_off = (int)sw.size();
  if (refined(p))
    for (int q = 0; q < (int)items.size(); ++q)
      for (int r0 = 0; r0 < items[q].n; r0 += 64) sw.push_back({q, r0});
  p->nres = (int)sw.size() - p->res_off;
  auto upload = [&](void** dst, const void* src, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 64));
    if (e != hipSuccess) return e;
    if (bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    return e;
  };
  CHECK_HIP(upload(&p->d_items, items.data(), items.size() * sizeof(BatchItem)), "batch items");
  CHECK_HIP(upload(&p->d_step, sw.data(), sw.size() * sizeof(StepWork)), "batch work");
  CHECK_HIP(upload(&p->d_tile, tw.data(), tw.size() * sizeof(TileWork)), "batch tiles");
  CHECK_HIP(hipMalloc(&p->d_counters, std::max(64, p->ncounters) * sizeof(int32_t)), "batch counters");
  if (p->fused_fwd) CHECK_HIP(hipMalloc(&p->d_y, std::max<size_t>(p->b_elems * esize(p), 64)), "batch forward");
  const size_t lb = std::max<size_t>(p->linv_elems * compute_esize(p), 64);
  if (p->linv_bytes < lb) {
    if (p->d_linv) (void)hipFree(p->d_linv);
    p->d_linv = nullptr;
    p->linv_bytes = 0;
    CHECK_HIP(hipMalloc(&p->d_linv, lb), "batch inverses");
    p->linv_bytes = lb;
  }
  // capture on the plan's own stream; the instantiated graph is launched on
  // the context's stream
  hipStream_t s = p->cap;
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
  hipError_t e = p->prec == LLH_PREC_F64 ? record<double>(p, s) : record<float>(p, s);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(s, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(err, e, "record batch graph");
  }
  CHECK_HIP(e2, "end capture");
  p->graph = g;
  CHECK_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0), "instantiate batch graph");
  return ok(err);
}

llh_status run(llh_context* ctx, llh_batch_plan* p, void* a, void* b, int32_t* status, bool sync,
               llh_error* err) {
  // base pointers (outside the graph), then one graph launch
  if (refined(p)) {
    // host staging of the two bases lives in the plan: the copies are async
    p->h_base32 = BatchBase<float>{p->a32, p->x32, static_cast<float*>(p->d_linv), status, p->d_counters, nullptr};
    p->h_mbase = MixedBase{static_cast<const double*>(a), static_cast<double*>(b), p->a32, p->x32, p->bk64,
                           status};
    CHECK_HIP(hipMemcpyAsync(p->d_base, &p->h_base32, sizeof(p->h_base32), hipMemcpyHostToDevice, ctx->stream),
              "batch base");
    CHECK_HIP(hipMemcpyAsync(p->d_mbase, &p->h_mbase, sizeof(p->h_mbase), hipMemcpyHostToDevice, ctx->stream),
              "batch base");
  } else if (p->prec == LLH_PREC_F32) {
    BatchBase<float> bb{static_cast<float*>(a), static_cast<float*>(b), static_cast<float*>(p->d_linv), status,
                        p->d_counters, static_cast<float*>(p->d_y)};
    CHECK_HIP(hipMemcpyAsync(p->d_base, &bb, sizeof(bb), hipMemcpyHostToDevice, ctx->stream), "batch base");
  } else {
    BatchBase<double> bb{static_cast<double*>(a), static_cast<double*>(b), static_cast<double*>(p->d_linv),
                         status, p->d_counters, static_cast<double*>(p->d_y)};
    CHECK_HIP(hipMemcpyAsync(p->d_base, &bb, sizeof(bb), hipMemcpyHostToDevice, ctx->stream), "batch base");
  }
  CHECK_HIP(hipGraphLaunch(p->exec, ctx->stream), "graph launch");
  if (sync) CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  return ok(err);
}

}  // namespace

extern "C" {

llh_status llh_batch_plan_create(llh_context* ctx, llh_precision prec, size_t count, const int32_t* n,
                                 const int32_t* m, const llh_batch_config* config, llh_batch_plan** out,
                                 llh_error* err) {
  if (out == nullptr) return invalid(err, "null output pointer");
  *out = nullptr;
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (prec != LLH_PREC_F64 && prec != LLH_PREC_F32 && prec != LLH_PREC_F32_REFINED)
    return invalid(err, "unknown precision");
  if (count > 0 && (n == nullptr || m == nullptr)) return invalid(err, "null size arrays");
  if (count > (1u << 28)) return invalid(err, "too many systems");
  auto* p = new llh_batch_plan();
  p->prec = prec;
  p->device = ctx->device;
  p->count = count;
  p->n.assign(n, n + count);
  p->m.assign(m, m + count);
  for (size_t i = 0; i < count; ++i) {
    if (p->n[i] < 0 || p->m[i] < 0 || p->n[i] > (1 << 20)) {
      delete p;
      return invalid(err, "system sizes must be in [0, 2^20] with m >= 0");
    }
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) p->cus = prop.multiProcessorCount;
  p->cfg = config ? *config : llh_batch_config{prec == LLH_PREC_F64 ? kSmallMaxD : kSmallMaxS, 4, 0, 0};
  // packed layout: 256-byte aligned starts, lda rounded up to 4
  const size_t es = esize(p), align = 256 / es;
  size_t ao = 0, bo = 0;
  p->lda.resize(count);
  p->a_off.resize(count);
  p->b_off.resize(count);
  for (size_t i = 0; i < count; ++i) {
    p->lda[i] = (p->n[i] + 3) & ~3;
    p->a_off[i] = ao;
    ao += ((size_t)p->n[i] * p->lda[i] + align - 1) / align * align;
    p->b_off[i] = bo;
    bo += ((size_t)p->n[i] * p->m[i] + 3) / 4 * 4;
  }
  p->a_elems = ao;
  p->b_elems = bo;
  hipError_t e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&p->d_base, sizeof(BatchBase<double>));
  if (e == hipSuccess && refined(p)) {
    e = hipMalloc(&p->d_mbase, sizeof(MixedBase));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&p->a32), std::max<size_t>(ao, 16) * 4);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&p->x32), std::max<size_t>(bo, 16) * 4);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&p->bk64), std::max<size_t>(bo, 16) * 8);
  }
  if (e != hipSuccess) {
    llh_batch_plan_destroy(p);
    return hip_fail(err, e, "batch plan");
  }
  llh_status st = build_schedule(ctx, p, err);
  if (st != LLH_OK) {
    llh_batch_plan_destroy(p);
    return st;
  }
  *out = p;
  return ok(err);
}

size_t llh_batch_a_elems(const llh_batch_plan* p) { return bad_plan(p) ? 0 : p->a_elems; }
size_t llh_batch_b_elems(const llh_batch_plan* p) { return bad_plan(p) ? 0 : p->b_elems; }

llh_status llh_batch_layout(const llh_batch_plan* p, uint64_t* a_off, int32_t* lda, uint64_t* b_off,
                            llh_error* err) {
  if (bad_plan(p)) return invalid(err, "null or destroyed plan");
  if (a_off) std::memcpy(a_off, p->a_off.data(), p->count * sizeof(uint64_t));
  if (lda) std::memcpy(lda, p->lda.data(), p->count * sizeof(int32_t));
  if (b_off) std::memcpy(b_off, p->b_off.data(), p->count * sizeof(uint64_t));
  return ok(err);
}

llh_batch_config llh_batch_get_config(const llh_batch_plan* p) {
  return bad_plan(p) ? llh_batch_config{0, 0, 0, 0} : p->cfg;
}

llh_status llh_batch_solve(llh_context* ctx, llh_batch_plan* p, void* a, void* b, int32_t* status,
                           llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (bad_plan(p)) return invalid(err, "null or destroyed plan");
  if (p->count == 0) return ok(err);
  if (a == nullptr || status == nullptr || (p->b_elems > 0 && b == nullptr)) return invalid(err, "null buffer");
  return run(ctx, p, a, b, status, true, err);
}

llh_status llh_batch_autotune(llh_context* ctx, llh_batch_plan* p, const void* a, const void* b,
                              int32_t* status, double* best_ms, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (bad_plan(p)) return invalid(err, "null or destroyed plan");
  if (p->count == 0) return ok(err);
  const size_t es = esize(p), ab = p->a_elems * es, bb = p->b_elems * es;
  void *wa = nullptr, *wb = nullptr;
  CHECK_HIP(hipMalloc(&wa, std::max<size_t>(ab, 64)), "autotune copy");
  hipError_t e = hipMalloc(&wb, std::max<size_t>(bb, 64));
  if (e != hipSuccess) {
    (void)hipFree(wa);
    return hip_fail(err, e, "autotune copy");
  }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  const int smax = p->prec == LLH_PREC_F64 ? kSmallMaxD : kSmallMaxS;
  const llh_batch_config keep = p->cfg;
  llh_batch_config best = keep;
  double best_t = 1e300;
  llh_status st = LLH_OK;
  const int limits[] = {32, 64, smax};
  const int waves[] = {1, 2, 4, 8};
  // candidates: bucket bound x resident small workgroups x the medium bucket's
  // form (fused with its substitutions, fused factor, grouped steps)
  std::vector<llh_batch_config> cands;
  for (int lim : limits)
    for (int wv : waves) {
      if (lim > 64 && wv > 1) continue;  // one resident 128-class workgroup per CU fills its LDS
      for (int mode = 0; mode < 3; ++mode) cands.push_back({lim, wv, keep.ir_iterations, mode});
    }
  for (const llh_batch_config& cand : cands) {
    p->cfg = cand;
    st = build_schedule(ctx, p, err);
    if (st != LLH_OK) break;
    // without a fusable medium bucket the three forms are the same schedule
    if (cand.medium_mode != 2 && !(p->lds_solve && p->nmed > 0)) continue;
    float ms_best = 1e30f;
    for (int rep = 0; rep < 2 && st == LLH_OK; ++rep) {
      e = hipMemcpyAsync(wa, a, ab, hipMemcpyDeviceToDevice, ctx->stream);
      if (e == hipSuccess && bb) e = hipMemcpyAsync(wb, b, bb, hipMemcpyDeviceToDevice, ctx->stream);
      if (e == hipSuccess) e = hipEventRecord(t0, ctx->stream);
      if (e != hipSuccess) {
        st = hip_fail(err, e, "autotune");
        break;
      }
      st = run(ctx, p, wa, wb, status, false, err);
      if (st != LLH_OK) break;
      e = hipEventRecord(t1, ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(t1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
      if (e != hipSuccess) {
        st = hip_fail(err, e, "autotune timing");
        break;
      }
      ms_best = std::min(ms_best, ms);
    }
    if (st != LLH_OK) break;
    if (ms_best < best_t) {
      best_t = ms_best;
      best = p->cfg;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  (void)hipFree(wa);
  (void)hipFree(wb);
  p->cfg = st == LLH_OK ? best : keep;
  llh_status st2 = build_schedule(ctx, p, err);
  if (st != LLH_OK) return st;
  if (st2 != LLH_OK) return st2;
  if (best_ms) *best_ms = best_t;
  return ok(err);
}

void llh_batch_plan_destroy(llh_batch_plan* p) {
  if (bad_plan(p)) return;
  free_schedule(p);
  if (p->d_linv) (void)hipFree(p->d_linv);
  if (p->d_base) (void)hipFree(p->d_base);
  if (p->d_mbase) (void)hipFree(p->d_mbase);
  for (void* q : {(void*)p->a32, (void*)p->x32, (void*)p->bk64})
    if (q) (void)hipFree(q);
  if (p->fork) (void)hipEventDestroy(p->fork);
  if (p->join) (void)hipEventDestroy(p->join);
  if (p->side) (void)hipStreamDestroy(p->side);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  p->magic = 0;
  delete p;
}

}  // extern "C"
