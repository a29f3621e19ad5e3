// Solve server (include/llh.h: llh_server_*): continuous batching of SPD solve
// requests with HBM admission and a resident factor cache.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "llh.h"
#include "llh_context.h"
#include "llh_driver.h"
#include "llh_kernels.h"

using namespace llh;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// A resident factor: L (f64 or f32), its diagonal-block inverses, and (f32
// entries) the server's own f64 copy of A that every refinement of a later
// request reads: a caller's buffer only has to outlive its own request.
struct Entry {
  uint64_t id = 0;
  int n = 0;
  llh_precision prec = LLH_PREC_F64;
  void* l = nullptr;
  void* dinv = nullptr;
  double* a_own = nullptr;    // f32 entries: f64 copy of A
  int32_t* d_info = nullptr;  // factor status word; ||A||_inf (f64) at byte 8
  bool orphan = false;        // replaced in the cache; freed when its last job ends
  size_t bytes = 0;
  hipEvent_t ready = nullptr;
  int refs = 0;
  uint64_t last_use = 0;
  bool failed = false;
  int64_t failed_row = -1;
  hipError_t error = hipSuccess;  // an enqueue of its factor failed
  bool dinv_ready = false;
  bool in_step = false;  // factored by the batched step that has not been launched yet
};

// Pinned host staging, recycled by power-of-two size class.
struct PinnedPool {
  std::map<size_t, std::vector<void*>> free_;
  std::vector<void*> all;
  void* get(size_t bytes, size_t* cls) {
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    *cls = c;
    auto& v = free_[c];
    if (!v.empty()) {
      void* p = v.back();
      v.pop_back();
      return p;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) return nullptr;
    all.push_back(p);
    return p;
  }
  void put(void* p, size_t cls) { free_[cls].push_back(p); }
  ~PinnedPool() {
    for (void* p : all) (void)hipHostFree(p);
  }
};

// Host staging of one batched step's descriptors: it must outlive the step's
// asynchronous upload, so it is released with the step's event.
struct StepStage {
  hipEvent_t done = nullptr;
  void* host = nullptr;
  size_t cls = 0;
  int remaining = 0;
};

struct Job {
  uint64_t ticket = 0;
  llh_request req{};
  double t_submit = 0, t_admit = 0;
  int stream = 0;
  Entry* entry = nullptr;
  bool hit = false, batched = false;
  llh_precision prec_used = LLH_PREC_F64;
  void* dev = nullptr;  // b | x | r | d | norms
  size_t dev_bytes = 0;
  void* host = nullptr;  // pinned: x | norms(2m) | info
  size_t host_cls = 0;
  hipEvent_t done = nullptr;
  size_t reserved = 0;
  StepStage* step = nullptr;  // batched jobs share their step's staging and event
  hipError_t error = hipSuccess;  // a failed enqueue: reported instead of reading the staging
  bool settled = false;           // no event: its stream was synchronized, its work is complete
};

// Host staging of a job (pinned): x (n m doubles) | residual and solution
// norms (2 m doubles) | factor status (int32) and solve error word (int32) |
// ||A||_inf (double).
inline int32_t* stage_info(double* hx, size_t nm, int m) { return reinterpret_cast<int32_t*>(hx + nm + 2 * m); }

struct StreamState {
  hipStream_t s = nullptr;
  SolveSync sync;
  int32_t epoch = 1;  // the flags' epoch counter of this stream's solves
  void* sync_buf = nullptr;
  size_t sync_bytes = 0;
};

}  // namespace

struct llh_server {
  static constexpr uint32_t kMagic = 0x4c4c4853;  // "LLHS"
  uint32_t magic = kMagic;
  int device = 0;
  llh_server_config cfg{};
  std::vector<StreamState> streams;
  int next_stream = 0;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  bool stop = false;
  std::deque<Job> incoming;  // submitted, not yet seen by the worker
  std::unordered_map<uint64_t, llh_result> results;
  uint64_t next_ticket = 1;
  // worker-owned state
  std::deque<Job> pending;
  std::list<Job> jobs;
  std::unordered_map<uint64_t, Entry*> cache;
  uint64_t clock = 0;
  size_t used = 0, peak = 0;
  PinnedPool pinned;
  std::vector<hipEvent_t> event_pool;
  llh_server_stats stats{};
  llh_error last_error{};
  bool evict_requested = false;
  // test hooks (llh_server_inject): failures to force from the host
  std::atomic<int> inject_staging{0}, inject_event{0};
  std::atomic<int64_t> spin_limit{kSpinLimit};
};

namespace {

// An event from the pool or a new one; *out is null when creation failed.
hipError_t get_event(llh_server* s, hipEvent_t* out) {
  *out = nullptr;
  const int countdown = s->inject_event.load();
  if (countdown > 0) {  // test hook: the countdown-th request from now fails
    s->inject_event.store(countdown - 1);
    if (countdown == 1) return hipErrorOutOfMemory;
  }
  if (!s->event_pool.empty()) {
    *out = s->event_pool.back();
    s->event_pool.pop_back();
    return hipSuccess;
  }
  hipEvent_t e = nullptr;
  hipError_t x = hipEventCreateWithFlags(&e, hipEventDisableTiming);
  if (x != hipSuccess) return x;
  *out = e;
  return hipSuccess;
}

size_t es_of(llh_precision p) { return p == LLH_PREC_F32 ? 4 : 8; }

size_t entry_bytes(const llh_request& r, llh_precision prec) {
  const size_t n = (size_t)r.n;
  size_t b = n * n * es_of(prec) + (size_t)dinv_elems(r.n) * es_of(prec) + 256;
  if (prec == LLH_PREC_F32) b += n * n * 8;  // the entry's own f64 A
  return b;
}

size_t job_dev_bytes(const llh_request& r) {
  const size_t nm = (size_t)r.n * r.m;
  return round_up(nm * 8, 256) * 3 + round_up(nm * 4, 256) + round_up((2 * r.m + 2) * 8, 256);
}

void dev_free(hipStream_t st, void* p) {
  if (p) (void)hipFreeAsync(p, st);
}

void free_entry(llh_server* s, Entry* e, hipStream_t st) {
  dev_free(st, e->l);
  dev_free(st, e->dinv);
  dev_free(st, e->a_own);
  dev_free(st, e->d_info);
  if (e->ready) s->event_pool.push_back(e->ready);
  s->used -= e->bytes;
  delete e;
}

// Frees least-recently-used idle factors until `need` more bytes fit.
bool make_room(llh_server* s, size_t need) {
  while (s->used + need > s->cfg.hbm_budget_bytes) {
    Entry* victim = nullptr;
    for (auto& kv : s->cache)
      if (kv.second->refs == 0 && (!victim || kv.second->last_use < victim->last_use)) victim = kv.second;
    if (!victim) return false;
    s->cache.erase(victim->id);
    free_entry(s, victim, s->streams[0].s);
  }
  return true;
}

hipError_t ensure_sync(llh_server* s, StreamState& st, int n, int m) {
  const size_t nb = (n + kNB - 1) / kNB, nq = (m + kSolveChunkHost - 1) / kSolveChunkHost;
  const size_t want = (2 + nb * nq) * sizeof(int32_t);
  st.sync.epoch = &st.epoch;
  st.sync.spin_limit = s->spin_limit.load();
  if (st.sync_bytes >= want) return hipSuccess;
  hipError_t e = hipStreamSynchronize(st.s);
  if (e != hipSuccess) return e;
  if (st.sync_buf) (void)hipFree(st.sync_buf);
  st.sync_buf = nullptr;
  const size_t bytes = std::max<size_t>(want, 4096);
  if ((e = hipMalloc(&st.sync_buf, bytes)) != hipSuccess) return e;
  st.sync_bytes = bytes;
  if ((e = launch_fill_i32(static_cast<int32_t*>(st.sync_buf), 0, (int64_t)(bytes / 4), st.s)) != hipSuccess)
    return e;
  st.sync.words = static_cast<int32_t*>(st.sync_buf);
  st.sync.nflags = (int64_t)(bytes / 4) - 2;
  st.epoch = 1;
  return hipSuccess;
}

template <typename T>
hipError_t solve_with(StreamState& st, const Entry* e, int m, T* d) {
  T* l = static_cast<T*>(e->l);
  T* dinv = static_cast<T*>(e->dinv);
  return solve_pair_enqueue<T>(e->n, l, e->n, m, d, m, dinv, &st.sync, st.s);
}

// Enqueues the factor of a new entry on stream st (scratch: n doubles).  The
// request's A (host or device) is read only here, on st, before the request
// completes.
hipError_t enqueue_factor(StreamState& st, Entry* e, const llh_request& r, double* scratch) {
  const int n = e->n;
  const size_t bytes = (size_t)n * n * 8;
  const hipMemcpyKind kind = r.a_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  hipError_t x;
  if (e->prec == LLH_PREC_F32) {
    if ((x = hipMemcpyAsync(e->a_own, r.a, bytes, kind, st.s)) != hipSuccess) return x;
    float* l = static_cast<float*>(e->l);
    double* anorm = reinterpret_cast<double*>(reinterpret_cast<char*>(e->d_info) + 8);
    if ((x = launch_symm_abs_rowsum(n, e->a_own, n, scratch, st.s)) != hipSuccess) return x;
    if ((x = launch_colmax(n, 1, scratch, 1, anorm, st.s)) != hipSuccess) return x;
    if ((x = launch_convert_d2s(n, n, e->a_own, n, l, n, true, st.s)) != hipSuccess) return x;
    x = potrf_enqueue<float>(l, n, n, e->d_info, false, static_cast<float*>(e->dinv), st.s, nullptr, true);
  } else {
    double* l = static_cast<double*>(e->l);
    if ((x = hipMemcpyAsync(l, r.a, bytes, kind, st.s)) != hipSuccess) return x;
    x = potrf_enqueue<double>(l, n, n, e->d_info, false, static_cast<double*>(e->dinv), st.s);
  }
  if (x != hipSuccess) return x;
  e->dinv_ready = true;
  return hipEventRecord(e->ready, st.s);
}

// Enqueues the solve of one job (its entry is factored or being factored on
// the same or an earlier-ordered stream).
hipError_t enqueue_solve(llh_server* s, Job& j) {
  StreamState& st = s->streams[j.stream];
  Entry* e = j.entry;
  const int n = e->n, m = j.req.m;
  const size_t nm = (size_t)n * m;
  char* base = static_cast<char*>(j.dev);
  double* db = reinterpret_cast<double*>(base);
  double* dx = reinterpret_cast<double*>(base + round_up(nm * 8, 256));
  double* dr = reinterpret_cast<double*>(base + 2 * round_up(nm * 8, 256));
  float* dd = reinterpret_cast<float*>(base + 3 * round_up(nm * 8, 256));
  double* dn = reinterpret_cast<double*>(base + 3 * round_up(nm * 8, 256) + round_up(nm * 4, 256));
  double* hx = static_cast<double*>(j.host);
  hipError_t x;
  if ((x = ensure_sync(s, st, n, m)) != hipSuccess) return x;
  // this job's error word: cleared before its first sweep, read back after its last
  if ((x = solve_sync_begin(st.sync, st.s)) != hipSuccess) return x;
  std::memcpy(hx, j.req.b, nm * 8);
  if ((x = hipMemcpyAsync(db, hx, nm * 8, hipMemcpyHostToDevice, st.s)) != hipSuccess) return x;
  if (e->prec == LLH_PREC_F64) {
    if ((x = hipMemcpyAsync(dx, db, nm * 8, hipMemcpyDeviceToDevice, st.s)) != hipSuccess) return x;
    if ((x = solve_with<double>(st, e, m, dx)) != hipSuccess) return x;
  } else {
    if ((x = launch_convert_d2s(n, m, db, m, dd, m, false, st.s)) != hipSuccess) return x;
    if ((x = solve_with<float>(st, e, m, dd)) != hipSuccess) return x;
    if ((x = launch_fill_i32(reinterpret_cast<int32_t*>(dx), 0, 2 * (int64_t)nm, st.s)) != hipSuccess) return x;
    if ((x = launch_axpy_s2d(n, m, dd, m, dx, m, st.s)) != hipSuccess) return x;
    for (int it = 0; it < s->cfg.ir_iterations; ++it) {
      if ((x = launch_symm_residual(n, m, e->a_own, n, dx, m, db, m, dr, m, st.s)) != hipSuccess) return x;
      if ((x = launch_convert_d2s(n, m, dr, m, dd, m, false, st.s)) != hipSuccess) return x;
      if ((x = solve_with<float>(st, e, m, dd)) != hipSuccess) return x;
      if ((x = launch_axpy_s2d(n, m, dd, m, dx, m, st.s)) != hipSuccess) return x;
    }
    if ((x = launch_symm_residual(n, m, e->a_own, n, dx, m, db, m, dr, m, st.s)) != hipSuccess) return x;
    if ((x = launch_colmax(n, m, dr, m, dn, st.s)) != hipSuccess) return x;
    if ((x = launch_colmax(n, m, dx, m, dn + m, st.s)) != hipSuccess) return x;
    if ((x = hipMemcpyAsync(hx + nm, dn, 2 * m * 8, hipMemcpyDeviceToHost, st.s)) != hipSuccess) return x;
    if ((x = hipMemcpyAsync(hx + nm + 2 * m + 1, reinterpret_cast<char*>(e->d_info) + 8, 8, hipMemcpyDeviceToHost,
                            st.s)) != hipSuccess)
      return x;
  }
  if ((x = hipMemcpyAsync(hx, dx, nm * 8, hipMemcpyDeviceToHost, st.s)) != hipSuccess) return x;
  int32_t* info = stage_info(hx, nm, m);
  if ((x = hipMemcpyAsync(info, e->d_info, 4, hipMemcpyDeviceToHost, st.s)) != hipSuccess) return x;
  if ((x = hipMemcpyAsync(info + 1, st.sync.words + 1, 4, hipMemcpyDeviceToHost, st.s)) != hipSuccess) return x;
  return hipEventRecord(j.done, st.s);
}

llh_status fail_job(llh_server* s, Job& j, llh_status code, const char* msg, hipError_t hip = hipSuccess) {
  llh_result r{};
  r.status = code;
  r.failed_row = code == LLH_ERR_HIP ? (int64_t)hip : -1;
  r.submit_s = j.t_submit;
  r.admit_s = j.t_admit;
  r.done_s = now_s();
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->results[j.ticket] = r;
    s->stats.completed++;
    s->stats.rejected += code == LLH_ERR_REJECTED;
  }
  s->cv_done.notify_all();
  set_error(&s->last_error, code, 0, 0, -1, msg, 0);
  return code;
}

void finish(llh_server* s, Job& j) {
  Entry* e = j.entry;
  const size_t nm = (size_t)e->n * j.req.m;
  double* hx = static_cast<double*>(j.host);
  // a job whose enqueue failed never wrote its staging block
  const int32_t info = j.error == hipSuccess ? stage_info(hx, nm, j.req.m)[0] : kNoFail;
  const int32_t solve_err = j.error == hipSuccess ? stage_info(hx, nm, j.req.m)[1] : 0;
  llh_result r{};
  r.cache_hit = j.hit;
  r.batched = j.batched;
  r.precision_used = j.prec_used;
  r.submit_s = j.t_submit;
  r.admit_s = j.t_admit;
  r.failed_row = -1;
  if (j.error != hipSuccess) {
    r.status = LLH_ERR_HIP;
    r.failed_row = (int64_t)j.error;  // the HIP error code
    e->failed = true;                 // its factor may be incomplete: drop the entry when idle
    set_error(&s->last_error, LLH_ERR_HIP, 0, 0, -1, hipGetErrorString(j.error), (int)j.error);
  } else if (info != kNoFail) {
    r.status = LLH_ERR_NOT_POSITIVE_DEFINITE;
    r.failed_row = info;
    e->failed = true;
    e->failed_row = info;
  } else if (solve_err != 0) {
    // a sweep's bounded wait gave up: x is unfinished, the factor is sound
    r.status = LLH_ERR_HIP;
    r.failed_row = 0;
    set_error(&s->last_error, LLH_ERR_HIP, 0, 0, -1, "triangular solve: a block wait exceeded its bound", 0);
  } else {
    r.status = LLH_OK;
    std::memcpy(j.req.x, hx, nm * 8);
    if (e->prec == LLH_PREC_F32 && !j.batched) {
      const double anrm = hx[nm + 2 * j.req.m + 1];
      double worst = 0.0;
      for (int c = 0; c < j.req.m; ++c) {
        const double rn = hx[nm + c], xn = hx[nm + j.req.m + c];
        const double rel = (anrm > 0 && xn > 0) ? rn / (anrm * xn) : (rn == 0 ? 0.0 : 1e300);
        worst = std::max(worst, rel != rel ? 1e300 : rel);
      }
      r.residual = worst;
    }
  }
  StreamState& st = s->streams[j.stream];
  dev_free(st.s, j.dev);
  s->used -= j.reserved;
  s->pinned.put(j.host, j.host_cls);
  if (j.done && !j.batched) s->event_pool.push_back(j.done);
  e->refs -= 1;
  e->last_use = ++s->clock;
  if ((e->failed || e->orphan) && e->refs == 0) {
    auto it = s->cache.find(e->id);
    if (it != s->cache.end() && it->second == e) s->cache.erase(it);
    free_entry(s, e, st.s);
  }
  r.done_s = now_s();
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->results[j.ticket] = r;
    s->stats.completed++;
  }
  s->cv_done.notify_all();
}

// Admits one pending job: reserves memory, attaches (or creates) its entry.
// Returns false when it does not fit yet.
bool admit(llh_server* s, Job& j) {
  const llh_request& r = j.req;
  const llh_precision want = (llh_precision)r.precision;
  auto it = s->cache.find(r.matrix_id);
  Entry* e = nullptr;
  if (it != s->cache.end() && it->second->n == r.n && !it->second->failed &&
      (it->second->prec == want || r.n <= s->cfg.small_limit))
    e = it->second;
  const bool small_new = e == nullptr && r.n <= s->cfg.small_limit;
  const llh_precision prec = small_new ? LLH_PREC_F64 : (e ? e->prec : want);
  const size_t jb = job_dev_bytes(r);
  const size_t eb = e ? 0 : entry_bytes(r, prec);
  if (!make_room(s, jb + eb)) return false;
  if (e == nullptr && r.a == nullptr) {
    fail_job(s, j, LLH_ERR_INVALID_ARGUMENT, "matrix not resident and no matrix given");
    return true;
  }
  // Small new systems run in the step's batched launch on stream 0, so their
  // stream-ordered allocations are made there; everything else round-robins.
  if (small_new || s->streams.size() == 1) {
    j.stream = 0;
  } else {
    j.stream = 1 + s->next_stream % ((int)s->streams.size() - 1);
    s->next_stream = (s->next_stream + 1) % ((int)s->streams.size() - 1);
  }
  hipStream_t st = s->streams[j.stream].s;
  size_t hb = (size_t)r.n * r.m * 8 + (2 * r.m + 2) * 8;
  j.host = s->pinned.get(hb, &j.host_cls);
  if (j.host == nullptr || hipMallocAsync(&j.dev, jb, st) != hipSuccess) {
    if (j.host) s->pinned.put(j.host, j.host_cls);
    fail_job(s, j, LLH_ERR_OUT_OF_MEMORY, "request buffers");
    return true;
  }
  j.dev_bytes = jb;
  j.reserved = jb;
  s->used += jb;
  if (e == nullptr) {
    e = new Entry();
    e->id = r.matrix_id;
    e->n = r.n;
    e->prec = prec;
    const size_t n2 = (size_t)r.n * r.n;
    bool okm = hipMallocAsync(&e->l, n2 * es_of(prec), st) == hipSuccess &&
               hipMallocAsync(&e->dinv, (size_t)dinv_elems(r.n) * es_of(prec), st) == hipSuccess &&
               hipMallocAsync(reinterpret_cast<void**>(&e->d_info), 256, st) == hipSuccess;
    if (okm && prec == LLH_PREC_F32)
      okm = hipMallocAsync(reinterpret_cast<void**>(&e->a_own), n2 * 8, st) == hipSuccess;
    const hipError_t xev = okm ? get_event(s, &e->ready) : hipSuccess;
    e->bytes = eb;
    s->used += eb;
    if (!okm || xev != hipSuccess) {
      (void)hipGetLastError();
      free_entry(s, e, st);
      dev_free(st, j.dev);
      s->used -= jb;
      s->pinned.put(j.host, j.host_cls);
      if (!okm)
        fail_job(s, j, LLH_ERR_OUT_OF_MEMORY, "factor buffers");
      else
        fail_job(s, j, LLH_ERR_HIP, "factor event", xev);
      return true;
    }
    auto old = s->cache.find(r.matrix_id);
    if (old != s->cache.end()) {
      if (old->second->refs == 0) free_entry(s, old->second, st);
      else old->second->orphan = true;  // freed when its last job ends
    }
    s->cache[r.matrix_id] = e;
    j.hit = false;
  } else {
    j.hit = true;
    std::lock_guard<std::mutex> g(s->mu);
    s->stats.cache_hits++;
  }
  e->refs += 1;
  e->last_use = ++s->clock;
  j.entry = e;
  j.prec_used = e->prec;
  j.t_admit = now_s();
  s->peak = std::max(s->peak, s->used);
  return true;
}

void launch_small_step(llh_server* s, std::vector<Job*>& batch) {
  // One launch of the exact-order kernel over every new small system, on the
  // stream their buffers were allocated on.  Any failed enqueue fails every job
  // of the step (and marks its entries failed): none of them reads a status
  // word or a solution that nothing wrote.
  const int si = 0;
  StreamState& st = s->streams[si];
  std::vector<SmallDesc<double>> descs;
  int nmax = 0;
  hipError_t x = hipSuccess;
  for (Job* j : batch) {
    j->stream = si;
    Entry* e = j->entry;
    const int n = e->n, m = j->req.m;
    const size_t nm = (size_t)n * m;
    double* hx = static_cast<double*>(j->host);
    double* dx = reinterpret_cast<double*>(static_cast<char*>(j->dev) + round_up(nm * 8, 256));
    const hipMemcpyKind kind = j->req.a_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (x == hipSuccess) x = hipMemcpyAsync(e->l, j->req.a, (size_t)n * n * 8, kind, st.s);
    std::memcpy(hx, j->req.b, nm * 8);
    stage_info(hx, nm, m)[1] = 0;  // the batched kernel has no flag waits
    if (x == hipSuccess) x = hipMemcpyAsync(dx, hx, nm * 8, hipMemcpyHostToDevice, st.s);
    descs.push_back({static_cast<double*>(e->l), dx, n, m, n, m, e->d_info});
    nmax = std::max(nmax, n);
  }
  // descriptors through a pinned staging block
  const size_t dbytes = descs.size() * sizeof(SmallDesc<double>);
  size_t cls = 0;
  void* hd = nullptr;
  void* dd = nullptr;
  if (x == hipSuccess) {
    if (s->inject_staging.load() > 0) {
      s->inject_staging.fetch_sub(1);
      x = hipErrorOutOfMemory;
    } else if ((hd = s->pinned.get(dbytes, &cls)) == nullptr) {
      x = hipErrorOutOfMemory;
    } else {
      x = hipMallocAsync(&dd, dbytes, st.s);
      if (x != hipSuccess) dd = nullptr;
    }
  }
  if (x == hipSuccess) {
    std::memcpy(hd, descs.data(), dbytes);
    x = hipMemcpyAsync(dd, hd, dbytes, hipMemcpyHostToDevice, st.s);
  }
  if (x == hipSuccess)
    x = launch_potrf_potrs_exact_batched<double>(static_cast<SmallDesc<double>*>(dd), (int)descs.size(), nmax, 1,
                                                 st.s);
  if (dd) dev_free(st.s, dd);
  if (x != hipSuccess) (void)hipGetLastError();
  for (Job* j : batch) {
    Entry* e = j->entry;
    const size_t nm = (size_t)e->n * j->req.m;
    double* hx = static_cast<double*>(j->host);
    double* dx = reinterpret_cast<double*>(static_cast<char*>(j->dev) + round_up(nm * 8, 256));
    if (x == hipSuccess) x = hipMemcpyAsync(hx, dx, nm * 8, hipMemcpyDeviceToHost, st.s);
    if (x == hipSuccess) x = hipMemcpyAsync(stage_info(hx, nm, j->req.m), e->d_info, 4, hipMemcpyDeviceToHost, st.s);
  }
  if (x != hipSuccess) (void)hipGetLastError();
  hipEvent_t ev = nullptr;
  const hipError_t xev = get_event(s, &ev);
  auto* stage = new StepStage{ev, hd, cls, (int)batch.size()};
  for (Job* j : batch) {
    Entry* e = j->entry;
    if (x != hipSuccess) {
      j->error = x;
      e->failed = true;  // no later request may use its factor
      e->error = x;
    }
    (void)hipEventRecord(e->ready, st.s);
    e->in_step = false;
    j->batched = true;
    j->done = ev;
    j->step = stage;
  }
  if (xev != hipSuccess || hipEventRecord(ev, st.s) != hipSuccess) {
    // no event to poll: wait for the step here; the jobs are then complete
    (void)hipGetLastError();
    const hipError_t xs = hipStreamSynchronize(st.s);
    for (Job* j : batch) {
      j->settled = true;
      if (j->error == hipSuccess && xs != hipSuccess) j->error = xs;
    }
  }
  // the staging block and the shared event are recycled after the step's last
  // job finishes
  std::lock_guard<std::mutex> g(s->mu);
  s->stats.batched += batch.size();
  s->stats.steps++;
}

void worker_loop(llh_server* s) {
  (void)hipSetDevice(s->device);
  std::vector<hipEvent_t> shared_events;
  for (;;) {
    {
      std::unique_lock<std::mutex> g(s->mu);
      while (!s->incoming.empty()) {
        s->pending.push_back(std::move(s->incoming.front()));
        s->incoming.pop_front();
      }
      if (s->stop && s->pending.empty() && s->jobs.empty()) break;
      if (s->evict_requested) {
        s->evict_requested = false;
        for (auto it = s->cache.begin(); it != s->cache.end();) {
          if (it->second->refs == 0) {
            free_entry(s, it->second, s->streams[0].s);
            it = s->cache.erase(it);
          } else {
            ++it;
          }
        }
      }
    }
    bool progressed = false;
    // completions
    for (auto it = s->jobs.begin(); it != s->jobs.end();) {
      if (it->settled || (it->done != nullptr && hipEventQuery(it->done) == hipSuccess)) {
        StepStage* stage = it->step;
        finish(s, *it);
        it = s->jobs.erase(it);
        if (stage && --stage->remaining == 0) {
          if (stage->host) s->pinned.put(stage->host, stage->cls);
          if (stage->done) s->event_pool.push_back(stage->done);
          delete stage;
        }
        progressed = true;
      } else {
        ++it;
      }
    }
    // admissions, in arrival order
    std::vector<Job*> small;
    while (!s->pending.empty() && (int)s->jobs.size() < s->cfg.max_inflight) {
      Job j = std::move(s->pending.front());
      const size_t before = s->results.size();
      (void)before;
      if (!admit(s, j)) {
        s->pending.push_front(std::move(j));
        break;
      }
      s->pending.pop_front();
      progressed = true;
      if (j.entry == nullptr) continue;  // failed during admission
      j.done = nullptr;
      s->jobs.push_back(std::move(j));
      Job& q = s->jobs.back();
      if (!q.hit && q.entry->n <= s->cfg.small_limit) {
        q.entry->in_step = true;
        small.push_back(&q);
        continue;
      }
      // a hit on a system the pending step factors: launch that step first, so
      // the entry's ready event is recorded before this job waits on it
      if (q.entry->in_step) {
        launch_small_step(s, small);
        small.clear();
      }
      StreamState& st = s->streams[q.stream];
      hipError_t x = hipSuccess;
      if (!q.hit) {
        const size_t nm = (size_t)q.req.n * q.req.m;
        double* scratch = reinterpret_cast<double*>(static_cast<char*>(q.dev) + 2 * round_up(nm * 8, 256));
        x = enqueue_factor(st, q.entry, q.req, scratch);
      } else {
        x = hipStreamWaitEvent(st.s, q.entry->ready, 0);
        if (x == hipSuccess && !q.entry->dinv_ready) {  // factored by the small batch path
          x = launch_trtri_blocks<double>(q.entry->n, static_cast<double*>(q.entry->l), q.entry->n, false, false,
                                          static_cast<double*>(q.entry->dinv), nullptr, st.s);
          q.entry->dinv_ready = true;
          if (x == hipSuccess) x = hipEventRecord(q.entry->ready, st.s);  // later users wait for the inverses too
        }
      }
      hipEvent_t done = nullptr;
      const hipError_t xev = get_event(s, &done);
      q.done = done;
      // a hit on an entry whose batched step failed to enqueue
      if (x == hipSuccess && q.entry->failed && q.entry->error != hipSuccess) x = q.entry->error;
      if (x == hipSuccess) x = xev;
      if (x == hipSuccess) x = enqueue_solve(s, q);
      if (x != hipSuccess) {
        q.error = x;
        (void)hipGetLastError();  // clear the sticky launch error before the next enqueue
        if (!q.hit) {
          // the factor may be incomplete: no later request may use the entry
          q.entry->failed = true;
          q.entry->error = x;
          (void)hipEventRecord(q.entry->ready, st.s);
        }
        if (q.done == nullptr || hipEventRecord(q.done, st.s) != hipSuccess) {
          // no event to poll: wait for what was enqueued; the job is then complete
          (void)hipGetLastError();
          (void)hipStreamSynchronize(st.s);
          q.settled = true;
        }
      }
    }
    if (!small.empty()) launch_small_step(s, small);
    if (progressed) {
      std::lock_guard<std::mutex> g(s->mu);
      s->stats.bytes_in_use = s->used;
      s->stats.bytes_peak = s->peak;
      s->stats.resident_factors = s->cache.size();
    }
    if (!progressed) {
      std::unique_lock<std::mutex> g(s->mu);
      s->cv_submit.wait_for(g, std::chrono::microseconds(s->jobs.empty() ? 2000 : 20),
                            [&] { return !s->incoming.empty() || s->stop; });
    }
  }
}

}  // namespace

extern "C" {

llh_status llh_server_create(llh_context* ctx, const llh_server_config* config, llh_server** out,
                             llh_error* err) {
  if (out == nullptr) return invalid(err, "null output pointer");
  *out = nullptr;
  if (ctx == nullptr || ctx->magic != llh_context::kMagic) return invalid(err, "null or destroyed context");
  auto* s = new llh_server();
  s->device = ctx->device;
  s->cfg = config ? *config : llh_server_config{0, 4, 64, 128, 3};
  if (s->cfg.streams <= 0) s->cfg.streams = 4;
  if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 64;
  if (s->cfg.small_limit < 0 || s->cfg.small_limit > kSmallMaxD) s->cfg.small_limit = kSmallMaxD;
  if (s->cfg.ir_iterations < 0) s->cfg.ir_iterations = 3;
  (void)hipSetDevice(s->device);
  if (s->cfg.hbm_budget_bytes == 0) {
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    s->cfg.hbm_budget_bytes = (uint64_t)(fr * 0.8);
  }
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, s->device) == hipSuccess) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  s->streams.resize(s->cfg.streams);
  for (auto& st : s->streams) {
    hipError_t e = hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking);
    if (e != hipSuccess) {
      llh_server_destroy(s);
      return hip_fail(err, e, "server streams");
    }
  }
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return ok(err);
}

llh_status llh_server_submit(llh_server* s, const llh_request* req, uint64_t* ticket, llh_error* err) {
  if (s == nullptr || s->magic != llh_server::kMagic) return invalid(err, "null or destroyed server");
  if (req == nullptr || ticket == nullptr) return invalid(err, "null request");
  if (req->n <= 0 || req->n > 65536 || req->m <= 0 || req->m > 64 || req->b == nullptr || req->x == nullptr ||
      (req->precision != LLH_PREC_F64 && req->precision != LLH_PREC_F32))
    return set_error(err, LLH_ERR_REJECTED, 0, 0, -1, "request sizes or precision out of range", 0);
  const size_t need = job_dev_bytes(*req) + entry_bytes(*req, (llh_precision)req->precision);
  if (need > s->cfg.hbm_budget_bytes) return set_error(err, LLH_ERR_REJECTED, 0, 0, -1, "request exceeds budget", 0);
  Job j;
  j.req = *req;
  j.t_submit = now_s();
  {
    std::lock_guard<std::mutex> g(s->mu);
    j.ticket = s->next_ticket++;
    *ticket = j.ticket;
    s->incoming.push_back(std::move(j));
    s->stats.submitted++;
  }
  s->cv_submit.notify_one();
  return ok(err);
}

int llh_server_poll(llh_server* s, uint64_t ticket, llh_result* out) {
  if (s == nullptr || s->magic != llh_server::kMagic) return 0;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->results.find(ticket);
  if (it == s->results.end()) return 0;
  if (out) *out = it->second;
  s->results.erase(it);
  return 1;
}

llh_status llh_server_wait(llh_server* s, uint64_t ticket, llh_result* out, llh_error* err) {
  if (s == nullptr || s->magic != llh_server::kMagic) return invalid(err, "null or destroyed server");
  std::unique_lock<std::mutex> g(s->mu);
  s->cv_done.wait(g, [&] { return s->results.count(ticket) > 0; });
  llh_result r = s->results[ticket];
  s->results.erase(ticket);
  if (out) *out = r;
  if (r.status == LLH_ERR_NOT_POSITIVE_DEFINITE) return not_pd(err, r.failed_row);
  if (r.status != LLH_OK) return set_error(err, (llh_status)r.status, 0, 0, -1, "request failed", 0);
  return ok(err);
}

void llh_server_get_stats(llh_server* s, llh_server_stats* out) {
  if (s == nullptr || s->magic != llh_server::kMagic || out == nullptr) return;
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->stats;  // the worker publishes the memory figures into stats under mu
}

llh_status llh_server_inject(llh_server* s, int kind, int64_t value, llh_error* err) {
  if (s == nullptr || s->magic != llh_server::kMagic) return invalid(err, "null or destroyed server");
  switch (kind) {
    case LLH_SERVER_INJECT_STAGING: s->inject_staging.store((int)value); break;
    case LLH_SERVER_INJECT_EVENT: s->inject_event.store((int)value); break;
    case LLH_SERVER_INJECT_SOLVE_SPIN: s->spin_limit.store(value == 0 ? kSpinLimit : value); break;
    default: return invalid(err, "unknown injection");
  }
  return ok(err);
}

void llh_server_evict_all(llh_server* s) {
  if (s == nullptr || s->magic != llh_server::kMagic) return;
  {  // the worker owns the cache: it evicts on its next pass
    std::lock_guard<std::mutex> g(s->mu);
    s->evict_requested = true;
  }
  s->cv_submit.notify_all();
}

void llh_server_destroy(llh_server* s) {
  if (s == nullptr || s->magic != llh_server::kMagic) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_submit.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto& st : s->streams)
    if (st.s) (void)hipStreamSynchronize(st.s);
  for (auto& kv : s->cache) free_entry(s, kv.second, s->streams.empty() ? nullptr : s->streams[0].s);
  s->cache.clear();
  for (auto& st : s->streams) {
    if (st.s) (void)hipStreamSynchronize(st.s);
    if (st.sync_buf) (void)hipFree(st.sync_buf);
    if (st.s) (void)hipStreamDestroy(st.s);
  }
  for (auto e : s->event_pool) (void)hipEventDestroy(e);
  s->magic = 0;
  delete s;
}

}  // extern "C"
