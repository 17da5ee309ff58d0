// Serving loop of libstencils_hip (include/stencils_hip.h, sth_server_*):
// continuous batching of iterated stencil solves with admission by an HBM
// budget.  Part of api.hip's translation unit (it drives the same plans,
// tuner and batched launcher).
//
// Scheduler thread, one round at a time:
//   1. retire  -- requests whose result copy has completed on the device get
//                 a completion record (device timestamps) and free their
//                 buffers;
//   2. admit   -- pending requests, in arrival order, whose arrival time has
//                 passed and whose two buffers fit in the arena (strict FIFO:
//                 a request that does not fit holds back the ones behind it,
//                 so large requests are not starved);
//   3. advance -- K = min(round_steps, fewest remaining steps) steps; each
//                 step is one batched launch over the active requests
//                 (requests join and leave only between rounds), buffers
//                 ping-pong per request.  Requests of 2^21 cells or more
//                 that have a two-step kernel run their own launches
//                 instead, two steps each, up to 4 x round_steps steps per
//                 round.  Requests that reach their step count get
//                 their state copied to `result` and an event recorded.
// Launches are stream-ordered, so a freed buffer is reused only by work
// enqueued after the copy that last read it.  The host runs at most two
// rounds ahead of the device (events per round and stream), which bounds the
// admission delay.
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>

struct sth_server {
  struct Req {
    sth_request r{};
    std::vector<int32_t> offs;
    std::vector<unsigned char> weights;
    uint64_t padval = 0;
    int64_t cells = 0;
    size_t lead = 0;        // bytes before the parent's first element (Halo interior on 16 bytes)
    size_t pitch = 0;       // bytes per parent row in the arena
    size_t row_bytes = 0;   // bytes per parent row in the caller's layout
    int64_t rows = 0;       // parent rows (all dims but the first)
    size_t bytes_one = 0;   // one buffer
    size_t off[2] = {0, 0};
    sth::Plan plan[2];      // plan[c]: buffer c -> buffer 1 - c
    int64_t done = 0;
    int cur = 0;
    double admitted = 0;
    hipEvent_t fin = nullptr;
    int lane = -1;              // solo requests: index of their stream
    hipEvent_t join = nullptr;  // orders a solo stream after the main one and back
  };

  // Solo requests run on a few streams of their own (default priority), so
  // requests too small to fill the device alone overlap each other and the
  // shared batch.  Each is pinned to one stream; events order its work after
  // its copy-in and its copy-out after its work.
  static constexpr int kSoloStreams = 4;
  hipStream_t solo_streams[kSoloStreams] = {};
  int next_lane = 0;

  // The scheduler enqueues at most kLookahead rounds ahead of the device:
  // a round starts once the round kLookahead before it has finished on
  // every stream, so a newly admitted request waits behind little queued work.
  static constexpr int kLookahead = 2;
  hipEvent_t marks[kLookahead][1 + kSoloStreams] = {};
  int64_t round_no = 0;
  // launch counts of the solo requests (scheduler thread only; published in
  // `stats` under `mu` after each round)
  int64_t solo_launches = 0, fused_launches = 0;

  sth_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  sth_server_config cfg{};
  char* arena = nullptr;
  std::map<size_t, size_t> free_blocks;  // offset -> bytes
  size_t in_use = 0, peak = 0;

  std::mutex mu;
  std::condition_variable cv;       // wakes the scheduler (submit / stop)
  std::condition_variable done_cv;  // wakes pollers
  std::deque<std::unique_ptr<Req>> pending;
  std::vector<std::unique_ptr<Req>> active;
  std::vector<std::unique_ptr<Req>> retiring;
  std::deque<sth_completion> completed;
  sth_server_stats stats{};
  int32_t error = STH_OK;
  std::string error_msg;
  bool started = false, stop = false;
  std::thread thr;
  std::chrono::steady_clock::time_point t0;
  hipEvent_t ev0 = nullptr;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  bool alloc(size_t bytes, size_t* off) {
    for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t at = it->first + bytes;
      free_blocks.erase(it);
      if (rest) free_blocks[at] = rest;
      in_use += bytes;
      peak = std::max(peak, in_use);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    in_use -= bytes;
    auto it = free_blocks.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_blocks.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_blocks.erase(nx);
    }
    if (it != free_blocks.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_blocks.erase(it);
      }
    }
  }

  sth_array buffer_desc(const Req& q, int c) const {
    sth_array a = q.r.state;
    a.data = arena + q.off[c] + q.lead;
    const int64_t es = sth::dtype_size(a.dtype);
    const int64_t h = a.halo;
    a.stride[0] = 1;
    a.stride[1] = a.ndim > 1 ? static_cast<int64_t>(q.pitch) / es : 0;
    a.stride[2] = a.ndim > 2 ? a.stride[1] * (a.size[1] + 2 * h) : 0;
    return a;
  }

  int fail_all(int rc) {
    error = rc;
    error_msg = sth::g_err;
    return rc;
  }

  // --- 2. admission (called with `mu` held)
  int admit(double t) {
    int rc = STH_OK;
    while (!pending.empty() && pending.front()->r.arrival <= t &&
           (cfg.max_active <= 0 || static_cast<int32_t>(active.size()) < cfg.max_active)) {
      Req& q = *pending.front();
      size_t o0 = 0, o1 = 0;
      if (!alloc(q.bytes_one, &o0)) break;
      if (!alloc(q.bytes_one, &o1)) {
        release(o0, q.bytes_one);
        break;
      }
      q.off[0] = o0;
      q.off[1] = o1;
      sth_stencil st = q.r.stencil;
      st.offsets = q.offs.data();
      st.weights = q.weights.empty() ? nullptr : q.weights.data();
      const sth_array a = buffer_desc(q, 0), b = buffer_desc(q, 1);
      int prc = STH_OK;
      for (int c = 0; c < 2 && prc == STH_OK; ++c) {
        std::vector<sth::Variant> cands;
        std::string group;
        prc = sth::make_plan(c ? &b : &a, c ? &a : &b, &st, q.r.boundary, &q.padval, q.r.padding, &q.r.op, 0,
                             &q.plan[c], &cands, &group);
        int idx = 0;
        if (prc == STH_OK) prc = tune_pick(ctx, q.plan[c], cands, group, 0, stream, false, true, &idx);
        if (prc == STH_OK) sth::apply_variant(q.plan[c], cands[idx]);
      }
      if (prc == STH_EINVAL || prc == STH_ENOTSUP) {
        // this request cannot run (submit validated it against the same
        // layout, so this is a safeguard): it completes with that status and
        // the others go on
        release(o0, q.bytes_one);
        release(o1, q.bytes_one);
        sth_completion c{};
        c.tag = q.r.tag;
        c.status = prc;
        c.arrival = q.r.arrival;
        c.admitted = t;
        c.finished = t;
        c.steps = q.r.steps;
        c.cells = q.cells;
        completed.push_back(c);
        stats.completed += 1;
        pending.pop_front();
        done_cv.notify_all();
        continue;
      }
      if (prc) return prc;
      // both buffers start as the state, as a SwitchingStencilArray's source
      // and dest do (src/array.jl:593-598)
      for (const size_t o : {o0, o1}) {
        const hipError_t ce = hipMemcpy2DAsync(arena + o + q.lead, q.pitch, q.r.state.data, q.row_bytes, q.row_bytes,
                                               static_cast<size_t>(q.rows), hipMemcpyDefault, stream);
        if (ce != hipSuccess) return sth::fail(STH_EDEVICE, "copy-in: %s", hipGetErrorString(ce));
      }
      hipError_t e;
      e = hipEventCreate(&q.fin);
      if (e != hipSuccess) return sth::fail(STH_EDEVICE, "hipEventCreate: %s", hipGetErrorString(e));
      if (solo(q)) {
        q.lane = next_lane++ % kSoloStreams;
        e = hipEventCreateWithFlags(&q.join, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(q.join, stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(solo_streams[q.lane], q.join, 0);
        if (e != hipSuccess) return sth::fail(STH_EDEVICE, "solo stream join: %s", hipGetErrorString(e));
      }
      q.admitted = t;
      active.push_back(std::move(pending.front()));
      pending.pop_front();
    }
    return rc;
  }

  // --- 1. retirement (called with `mu` held)
  int retire() {
    for (size_t i = 0; i < retiring.size();) {
      Req& q = *retiring[i];
      const hipError_t e = hipEventQuery(q.fin);
      if (e == hipErrorNotReady) {
        ++i;
        continue;
      }
      if (e != hipSuccess) return sth::fail(STH_EDEVICE, "request event: %s", hipGetErrorString(e));
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev0, q.fin) != hipSuccess) ms = static_cast<float>(now() * 1e3);
      sth_completion c{};
      c.tag = q.r.tag;
      c.status = STH_OK;
      c.arrival = q.r.arrival;
      c.admitted = q.admitted;
      c.finished = ms * 1e-3;
      c.steps = q.r.steps;
      c.cells = q.cells;
      completed.push_back(c);
      stats.completed += 1;
      stats.cell_steps += q.cells * q.r.steps;
      release(q.off[0], q.bytes_one);
      release(q.off[1], q.bytes_one);
      (void)hipEventDestroy(q.fin);
      if (q.join) (void)hipEventDestroy(q.join);
      retiring.erase(retiring.begin() + static_cast<std::ptrdiff_t>(i));
      done_cv.notify_all();
    }
    return STH_OK;
  }

  int finish(Req& q) {
    const int c = q.cur;
    if (q.lane >= 0) {  // the copy-out waits for the request's own stream
      hipError_t je = hipEventRecord(q.join, solo_streams[q.lane]);
      if (je == hipSuccess) je = hipStreamWaitEvent(stream, q.join, 0);
      if (je != hipSuccess) return sth::fail(STH_EDEVICE, "solo stream join: %s", hipGetErrorString(je));
    }
    hipError_t e = hipMemcpy2DAsync(q.r.result, q.row_bytes, arena + q.off[c] + q.lead, q.pitch, q.row_bytes,
                                    static_cast<size_t>(q.rows), hipMemcpyDefault, stream);
    if (e != hipSuccess) return sth::fail(STH_EDEVICE, "copy-out: %s", hipGetErrorString(e));
    e = hipEventRecord(q.fin, stream);
    if (e != hipSuccess) return sth::fail(STH_EDEVICE, "hipEventRecord: %s", hipGetErrorString(e));
    return STH_OK;
  }

  // --- 3. one round (called without `mu`: only the scheduler touches `active`)
  // A request with a two-step kernel (3-D and 2-D radius-1 stencils, Life)
  // and at least 2^21 cells runs as
  // launches of its own: two steps per launch, and more steps per round than
  // the shared batch.  Below that size batching the steps pays more.
  // (STH_SERVER_SOLO_CELLS overrides the size)
  static int64_t solo_cells() {
    static const int64_t v = [] {
      const char* e = std::getenv("STH_SERVER_SOLO_CELLS");
      return e ? std::atoll(e) : (int64_t(1) << 21);
    }();
    return v;
  }
  bool solo(const Req& q) const {
    return (q.plan[0].fn_fused || q.plan[0].jit_fused) && (q.plan[1].fn_fused || q.plan[1].jit_fused) &&
           q.cells >= solo_cells();
  }

  // `n` steps of a solo request.  Fused launches leave the buffers holding
  // generations one apart from the step-by-step order, which later steps do
  // not see (each reads only its source's interior, refreshing halos first);
  // a request's last two steps are always single steps, so its final state,
  // halo included, is the step-by-step one.
  int advance_solo(Req& q, int64_t n, int64_t* launches) {
    const hipStream_t st = solo_streams[q.lane];
    if (q.done == 0 && n > 0) {
      // the first step is a single step, as in sth_iterate: the two-step
      // kernels assume data this library wrote (Life: bytes already 0 or 1)
      if (int rc = sth::run_plan(q.plan[q.cur], st)) return rc;
      q.cur ^= 1;
      q.done += 1;
      *launches += 1;
      solo_launches += 1;
      n -= 1;
    }
    const int64_t rem = q.r.steps - q.done;
    const int64_t G = std::max<int32_t>(1, q.plan[0].fused_gens);  // steps per fused launch
    const int64_t fused = n == rem ? std::max<int64_t>(0, (n - 2) / G) : n / G;
    for (int64_t i = 0; i < fused; ++i) {
      if (int rc = sth::run_fused(q.plan[q.cur], st)) return rc;
      q.cur ^= 1;
      q.done += G;
      *launches += 1;
      solo_launches += 1;
      fused_launches += 1;
    }
    for (int64_t i = G * fused; i < n; ++i) {
      if (int rc = sth::run_plan(q.plan[q.cur], st)) return rc;
      q.cur ^= 1;
      q.done += 1;
      *launches += 1;
      solo_launches += 1;
    }
    return STH_OK;
  }

  int advance(int64_t* launches) {
    hipEvent_t* mark = marks[round_no % kLookahead];
    if (round_no >= kLookahead)
      for (int i = 0; i <= kSoloStreams; ++i)
        if (hipEventSynchronize(mark[i]) != hipSuccess) return sth::fail(STH_EDEVICE, "round mark wait failed");
    const int rc = advance_round(launches);
    if (rc) return rc;
    for (int i = 0; i <= kSoloStreams; ++i)
      if (hipEventRecord(mark[i], i ? solo_streams[i - 1] : stream) != hipSuccess)
        return sth::fail(STH_EDEVICE, "round mark record failed");
    ++round_no;
    return STH_OK;
  }

  int advance_round(int64_t* launches) {
    const int64_t R = cfg.round_steps > 0 ? cfg.round_steps : 8;
    const bool serial = (cfg.flags & STH_SERVER_SERIAL) != 0;
    int rc = STH_OK;
    if (serial) {  // the plain per-request path: one request, one launch per step
      if (active.empty()) return STH_OK;
      Req& q = *active.front();
      for (; q.done < q.r.steps; ++q.done, q.cur ^= 1, *launches += 1)
        if ((rc = sth::run_plan(q.plan[q.cur], stream))) return rc;
      return STH_OK;
    }
    std::vector<Req*> shared, alone;
    for (auto& q : active) (q->lane >= 0 ? alone : shared).push_back(q.get());
    int64_t K = R;
    for (Req* q : shared) K = std::min(K, q->r.steps - q->done);
    if (shared.empty())
      for (Req* q : alone) K = std::min(K, q->r.steps - q->done);
    // the shared batch: one launch per step over every shared request
    std::vector<const sth::Plan*> plans(shared.size());
    for (int64_t k = 0; k < K && !shared.empty(); ++k) {
      for (size_t i = 0; i < shared.size(); ++i) plans[i] = &shared[i]->plan[shared[i]->cur];
      if ((rc = launch_batched(ctx, plans.data(), static_cast<int32_t>(plans.size()), false, stream))) return rc;
      for (Req* q : shared) {
        q->cur ^= 1;
        q->done += 1;
      }
      *launches += 1;
    }
    // solo requests: up to 4 rounds' worth of steps, never leaving a single
    // step for the next round (the last two steps are single steps)
    for (Req* q : alone) {
      const int64_t rem = q->r.steps - q->done;
      int64_t n = std::min(rem, std::max(K, 4 * R));
      if (rem - n == 1) n = n > 1 ? n - 1 : rem;
      if ((rc = advance_solo(*q, n, launches))) return rc;
    }
    return STH_OK;
  }

  void loop() {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      int rc = retire();
      if (rc == STH_OK) rc = admit(now());
      if (rc) {
        fail_all(rc);
        done_cv.notify_all();
        return;
      }
      // requests with nothing left to do (zero steps) finish at once
      for (size_t i = 0; i < active.size();) {
        if (active[i]->done >= active[i]->r.steps) {
          if ((rc = finish(*active[i]))) {
            fail_all(rc);
            done_cv.notify_all();
            return;
          }
          retiring.push_back(std::move(active[i]));
          active.erase(active.begin() + static_cast<std::ptrdiff_t>(i));
        } else {
          ++i;
        }
      }
      stats.active = static_cast<int64_t>(active.size());
      stats.pending = static_cast<int64_t>(pending.size());
      stats.hbm_in_use = in_use;
      stats.hbm_peak = peak;
      if (active.empty()) {
        if (stop && pending.empty() && retiring.empty()) return;
        // wait for the next arrival, a submission, or a retiring copy
        double wait = 1e-3;
        if (!pending.empty()) wait = std::min(wait, std::max(0.0, pending.front()->r.arrival - now()));
        if (!retiring.empty()) wait = std::min(wait, 50e-6);
        cv.wait_for(lk, std::chrono::duration<double>(wait));
        continue;
      }
      lk.unlock();
      int64_t launches = 0;
      rc = advance(&launches);
      lk.lock();
      stats.launches += launches;
      stats.rounds += 1;
      stats.solo_launches = solo_launches;
      stats.fused_launches = fused_launches;
      if (rc) {
        fail_all(rc);
        done_cv.notify_all();
        return;
      }
    }
  }
};

namespace {

int server_validate(const sth_request* q) {
  if (!q) return sth::fail(STH_EINVAL, "request is NULL");
  if (!q->state.data || !q->result) return sth::fail(STH_EINVAL, "request state or result is NULL");
  if (q->steps < 0) return sth::fail(STH_EINVAL, "negative step count");
  if (q->stencil.length < 1 || !q->stencil.offsets || q->stencil.ndim < 1)
    return sth::fail(STH_EINVAL, "request stencil has no offsets");
  if (q->state.ndim < 1 || q->state.ndim > 3 || sth::dtype_size(q->state.dtype) == 0)
    return sth::fail(STH_EINVAL, "request state descriptor is invalid");
  if (q->op.kind == STH_OP_MEAN && !sth::is_float(q->state.dtype))
    return sth::fail(STH_EINVAL, "mapstencil! cannot store a Float64 mean into an integer buffer");
  return STH_OK;
}

}  // namespace

extern "C" {

int sth_server_create(sth_ctx* ctx, const sth_server_config* cfg, void* stream, sth_server** out) {
  STH_GUARD_BEGIN
  int rc = set_device(ctx);
  if (rc) return rc;
  if (!cfg || !out) return fail(STH_EINVAL, "config or out is NULL");
  if (cfg->hbm_budget < 4096) return fail(STH_EINVAL, "hbm_budget of %zu bytes is too small", cfg->hbm_budget);
  auto srv = std::make_unique<sth_server>();
  srv->ctx = ctx;
  srv->stream = static_cast<hipStream_t>(stream);
  srv->cfg = *cfg;
  if (cfg->flags & STH_SERVER_SERIAL) srv->cfg.max_active = 1;  // one request at a time
  const size_t bytes = cfg->hbm_budget & ~static_cast<size_t>(255);
  hipError_t e = hipMalloc(&srv->arena, bytes);
  if (e != hipSuccess) return fail(STH_ENOMEM, "arena of %zu bytes: %s", bytes, hipGetErrorString(e));
  srv->free_blocks[0] = bytes;
  e = hipEventCreate(&srv->ev0);
  for (int i = 0; i < sth_server::kSoloStreams && e == hipSuccess; ++i)
    e = hipStreamCreateWithFlags(&srv->solo_streams[i], hipStreamNonBlocking);
  for (auto& row : srv->marks)
    for (hipEvent_t& m : row)
      if (e == hipSuccess) e = hipEventCreateWithFlags(&m, hipEventDisableTiming);
  if (e != hipSuccess) {
    for (hipStream_t x : srv->solo_streams)
      if (x) (void)hipStreamDestroy(x);
    for (auto& row : srv->marks)
      for (hipEvent_t m : row)
        if (m) (void)hipEventDestroy(m);
    if (srv->ev0) (void)hipEventDestroy(srv->ev0);
    (void)hipFree(srv->arena);
    return fail(STH_EDEVICE, "server events and streams: %s", hipGetErrorString(e));
  }
  *out = srv.release();
  return STH_OK;
  STH_GUARD_END
}

int sth_server_submit(sth_server* srv, const sth_request* q) {
  STH_GUARD_BEGIN
  if (!srv) return fail(STH_EINVAL, "server is NULL");
  int rc = server_validate(q);
  if (rc) return rc;
  auto r = std::make_unique<sth_server::Req>();
  r->r = *q;
  const int nd = q->stencil.ndim;
  r->offs.assign(q->stencil.offsets, q->stencil.offsets + static_cast<size_t>(q->stencil.length) * nd);
  if (q->stencil.weights) {
    const int ws = dtype_size(q->stencil.weights_dtype);
    if (ws == 0) return fail(STH_EINVAL, "unknown weights dtype %d", q->stencil.weights_dtype);
    const auto* w = static_cast<const unsigned char*>(q->stencil.weights);
    r->weights.assign(w, w + static_cast<size_t>(q->stencil.length) * ws);
  }
  const int es = dtype_size(q->state.dtype);
  if (q->padval) std::memcpy(&r->padval, q->padval, static_cast<size_t>(es));
  const int64_t h = q->state.halo;
  const int ndim = q->state.ndim;
  int64_t cells = 1, rows = 1;
  for (int d = 0; d < ndim; ++d) {
    if (q->state.size[d] < 1) return fail(STH_EINVAL, "request state has an empty dim");
    cells *= q->state.size[d];
    if (d > 0) rows *= q->state.size[d] + 2 * h;
  }
  r->cells = cells;
  r->rows = rows;
  r->row_bytes = static_cast<size_t>((q->state.size[0] + 2 * h) * es);
  if (ndim > 1 && q->state.stride[1] != q->state.size[0] + 2 * h)
    return fail(STH_EINVAL, "request state must be dense (stride[1] = size[0] + 2 halo)");
  r->pitch = (r->row_bytes + 63) & ~static_cast<size_t>(63);
  // buffers start on 256 bytes and rows and planes are 64-byte multiples, so
  // a lead of this many bytes puts the first interior cell on 16 bytes (the
  // vector kernels' alignment), as a Halo StencilArray's own allocation does
  r->lead = (16 - static_cast<size_t>(h * es) % 16) % 16;
  r->bytes_one = (r->lead + r->pitch * static_cast<size_t>(rows) + 255) & ~static_cast<size_t>(255);
  // validate the call once, against stand-in buffers of the arena layout
  {
    sth_array a = srv->buffer_desc(*r, 0), b = a;
    a.data = reinterpret_cast<void*>((uintptr_t(1) << 20) + r->lead);
    b.data = reinterpret_cast<void*>((uintptr_t(1) << 20) + 2 * r->bytes_one + 4096 + r->lead);
    sth_stencil st = q->stencil;
    st.offsets = r->offs.data();
    st.weights = r->weights.empty() ? nullptr : r->weights.data();
    Plan tmp;
    if ((rc = make_plan(&a, &b, &st, q->boundary, &r->padval, q->padding, &q->op, 0, &tmp))) return rc;
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  if (2 * r->bytes_one > (srv->cfg.hbm_budget & ~static_cast<size_t>(255)))
    return fail(STH_ENOMEM, "request needs %zu bytes; the server budget is %zu", 2 * r->bytes_one,
                srv->cfg.hbm_budget);
  // keep `pending` in arrival order (stable for equal arrivals)
  auto it = srv->pending.end();
  while (it != srv->pending.begin() && (*std::prev(it))->r.arrival > r->r.arrival) --it;
  srv->pending.insert(it, std::move(r));
  srv->stats.submitted += 1;
  srv->cv.notify_all();
  return STH_OK;
  STH_GUARD_END
}

int sth_server_start(sth_server* srv) {
  STH_GUARD_BEGIN
  if (!srv) return fail(STH_EINVAL, "server is NULL");
  int rc = set_device(srv->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->started) return fail(STH_EINVAL, "server already started");
  STH_HIP(hipStreamSynchronize(srv->stream));
  STH_HIP(hipEventRecord(srv->ev0, srv->stream));
  STH_HIP(hipEventSynchronize(srv->ev0));
  srv->t0 = std::chrono::steady_clock::now();
  srv->started = true;
  srv->thr = std::thread([srv] { srv->loop(); });
  return STH_OK;
  STH_GUARD_END
}

int sth_server_poll(sth_server* srv, sth_completion* out, int32_t max, int32_t* n, int32_t wait_ms) {
  STH_GUARD_BEGIN
  if (!srv || !n || (max > 0 && !out)) return fail(STH_EINVAL, "bad poll arguments");
  std::unique_lock<std::mutex> lk(srv->mu);
  if (srv->completed.empty() && wait_ms > 0 && srv->error == STH_OK)
    srv->done_cv.wait_for(lk, std::chrono::milliseconds(wait_ms),
                          [srv] { return !srv->completed.empty() || srv->error != STH_OK; });
  int32_t k = 0;
  while (k < max && !srv->completed.empty()) {
    out[k++] = srv->completed.front();
    srv->completed.pop_front();
  }
  *n = k;
  if (srv->error) return fail(srv->error, "server stopped: %s", srv->error_msg.c_str());
  return STH_OK;
  STH_GUARD_END
}

int sth_server_drain(sth_server* srv, double timeout_s) {
  STH_GUARD_BEGIN
  if (!srv) return fail(STH_EINVAL, "server is NULL");
  std::unique_lock<std::mutex> lk(srv->mu);
  if (!srv->started) return fail(STH_EINVAL, "server not started");
  const bool ok = srv->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), [srv] {
    return srv->error != STH_OK || srv->stats.completed == srv->stats.submitted;
  });
  if (srv->error) return fail(srv->error, "server stopped: %s", srv->error_msg.c_str());
  if (!ok) return fail(STH_EDEVICE, "drain timed out with %lld of %lld requests done",
                       static_cast<long long>(srv->stats.completed), static_cast<long long>(srv->stats.submitted));
  return STH_OK;
  STH_GUARD_END
}

int sth_server_stats_get(sth_server* srv, sth_server_stats* out) {
  STH_GUARD_BEGIN
  if (!srv || !out) return fail(STH_EINVAL, "bad arguments");
  std::lock_guard<std::mutex> lk(srv->mu);
  *out = srv->stats;
  out->hbm_in_use = srv->in_use;
  out->hbm_peak = srv->peak;
  out->started = srv->started ? srv->now() : 0.0;
  return STH_OK;
  STH_GUARD_END
}

int sth_server_destroy(sth_server* srv) {
  STH_GUARD_BEGIN
  if (!srv) return STH_OK;
  (void)hipSetDevice(srv->ctx->device);
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
    srv->cv.notify_all();
  }
  if (srv->thr.joinable()) srv->thr.join();
  (void)hipStreamSynchronize(srv->stream);
  for (hipStream_t x : srv->solo_streams)
    if (x) {
      (void)hipStreamSynchronize(x);
      (void)hipStreamDestroy(x);
    }
  for (auto* v : {&srv->active, &srv->retiring})
    for (auto& q : *v) {
      if (q->fin) (void)hipEventDestroy(q->fin);
      if (q->join) (void)hipEventDestroy(q->join);
    }
  for (auto& row : srv->marks)
    for (hipEvent_t m : row)
      if (m) (void)hipEventDestroy(m);
  if (srv->ev0) (void)hipEventDestroy(srv->ev0);
  if (srv->arena) (void)hipFree(srv->arena);
  delete srv;
  return STH_OK;
  STH_GUARD_END
}

}  // extern "C"
