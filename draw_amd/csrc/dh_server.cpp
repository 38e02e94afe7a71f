// dh_server.cpp -- continuous-batching render server over libdrawhip.
//
// Jobs (uniforms + meshes of one frame) queue in arrival order.  Each worker
// thread owns a context that shares the server context's resources (own HIP
// stream and scratch), admits the longest prefix of the queue that fits its
// launch limits and the HBM budget, renders it with one dh_render_batch launch
// and publishes the results.  While one worker's launch runs on the device the
// others admit and launch what arrived meanwhile, so launches overlap.
//
// Memory accounting (bytes held on the device on the server's behalf):
//   canvases  every canvas the server allocated, in use by a job or pooled
//             for reuse (pooled canvases of other sizes are freed when a job
//             needs room);
//   scratch   per worker, the larger of its context's scratch buffers and the
//             estimate for the launch it is admitting.
// A job is admitted only if the total stays within the budget.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "dh_internal.hpp"

namespace {

struct Job {
  uint64_t ticket = 0;
  dh_frame_params fr;
  std::vector<dh_handle> meshes;
  uint32_t flags = 0;
  uint64_t n_tri = 0;
  uint64_t canvas_bytes = 0;
  uint64_t need = 0;  // scratch a launch of this job alone asked for (0 = not known)
  int32_t status = DH_JOB_QUEUED;
  uint32_t batch_views = 0;
  dh_handle canvas = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
};

struct Worker {
  dh_ctx ctx = nullptr;
  std::thread th;
  uint64_t scratch_reserved = 0;  // bytes counted for this worker's scratch
};

}  // namespace

struct dh_server_s {
  dh_ctx ctx = nullptr;  // resources; also frees pooled canvases
  dh_server_config cfg;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work;  // queue or memory changed
  std::condition_variable cv_done;  // a job finished
  std::deque<Job*> queue;
  std::unordered_map<uint64_t, Job*> jobs;
  std::multimap<std::pair<int32_t, int32_t>, dh_handle> free_canvases;  // (W, H) -> canvas
  uint64_t canvas_bytes = 0;   // every canvas allocated by the server
  uint64_t scratch_bytes = 0;  // sum of workers' reservations
  uint64_t next_ticket = 1;
  uint64_t epoch = 0;          // bumped whenever admission could change
  uint32_t running = 0;        // launches on the device now
  bool stop = false;
  dh_server_stats st = {};
  std::vector<std::unique_ptr<Worker>> workers;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  uint64_t in_use() const { return canvas_bytes + scratch_bytes; }
  bool fits(uint64_t extra) const { return cfg.budget_bytes == 0 || in_use() + extra <= cfg.budget_bytes; }
  void note_peak() { st.bytes_peak = std::max<uint64_t>(st.bytes_peak, in_use()); }
};

namespace {

uint64_t canvas_bytes_of(int32_t W, int32_t H) { return 12ull * (uint64_t)W * (uint64_t)H; }

uint64_t tiles_of(int32_t W, int32_t H, int TS) {
  return (uint64_t)((W + TS - 1) / TS) * (uint64_t)((H + TS - 1) / TS);
}

// Free pooled canvases (not the given size) until `extra` more bytes fit.
void evict_for(dh_server s, uint64_t extra, int32_t W, int32_t H) {
  for (auto it = s->free_canvases.begin(); it != s->free_canvases.end() && !s->fits(extra);) {
    if (it->first == std::make_pair(W, H)) {
      ++it;
      continue;
    }
    dh_release(s->ctx, it->second);
    s->canvas_bytes -= canvas_bytes_of(it->first.first, it->first.second);
    it = s->free_canvases.erase(it);
  }
}

void worker_loop(dh_server s, Worker* w) {
  const int TS = dh::ctx_tile_size(w->ctx);
  std::unique_lock<std::mutex> lk(s->mu);
  uint64_t seen = ~0ull;
  for (;;) {
    s->cv_work.wait(lk, [&] { return (!s->queue.empty() && s->epoch != seen) || (s->stop && s->queue.empty()); });
    if (s->queue.empty()) break;  // stopping, nothing left
    seen = s->epoch;
    // Admit the longest prefix of the queue that fits.
    std::vector<Job*> batch;
    uint64_t tris = 0, tiles = 0, px = 0, draws = 0;
    while (!s->queue.empty() && batch.size() < s->cfg.max_views) {
      Job* j = s->queue.front();
      if (!batch.empty() && tris + j->n_tri > s->cfg.max_tris) break;
      const uint64_t est = dh::scratch_estimate(tris + j->n_tri, tiles + tiles_of(j->fr.W, j->fr.H, TS),
                                                px + (uint64_t)j->fr.W * j->fr.H, draws + j->meshes.size(),
                                                batch.size() + 1);
      const uint64_t want_scratch =
          std::max<uint64_t>(std::max<uint64_t>(dh::ctx_scratch_bytes(w->ctx), est), j->need);
      const uint64_t scratch_extra = want_scratch > w->scratch_reserved ? want_scratch - w->scratch_reserved : 0;
      auto fc = s->free_canvases.find(std::make_pair(j->fr.W, j->fr.H));
      const uint64_t canvas_extra = (fc == s->free_canvases.end()) ? j->canvas_bytes : 0;
      if (!s->fits(scratch_extra + canvas_extra)) {
        evict_for(s, scratch_extra + canvas_extra, j->fr.W, j->fr.H);
        if (!s->fits(scratch_extra + canvas_extra)) break;  // waits for memory
      }
      if (fc != s->free_canvases.end()) {
        j->canvas = fc->second;
        s->free_canvases.erase(fc);
      } else {
        const int rc = dh_canvas_create(s->ctx, j->fr.W, j->fr.H, &j->canvas);
        if (rc != DH_OK) {  // device allocation failed: fail this job, keep going
          s->queue.pop_front();
          j->status = rc;
          j->t_done = s->now_ms();
          s->st.failed++;
          s->cv_done.notify_all();
          continue;
        }
        s->canvas_bytes += j->canvas_bytes;
      }
      s->scratch_bytes += scratch_extra;
      w->scratch_reserved += scratch_extra;
      s->note_peak();
      s->queue.pop_front();
      j->status = DH_JOB_RUNNING;
      j->t_admit = s->now_ms();
      tris += j->n_tri;
      tiles += tiles_of(j->fr.W, j->fr.H, TS);
      px += (uint64_t)j->fr.W * j->fr.H;
      draws += j->meshes.size();
      batch.push_back(j);
    }
    if (batch.empty()) {
      if (w->scratch_reserved > 0) {
        // An idle worker's scratch may be what blocks the queue: free it so
        // that this or another worker can admit.
        dh::ctx_release_scratch(w->ctx);
        s->scratch_bytes -= w->scratch_reserved;
        w->scratch_reserved = 0;
        s->epoch++;
        s->cv_work.notify_all();
      } else if (s->running == 0 && s->cfg.budget_bytes && !s->stop) {
        // Nothing runs and this worker holds no scratch.  If no unreleased
        // result holds a canvas and no other worker holds scratch, nothing
        // will ever free memory: the front job cannot run.
        bool held = s->scratch_bytes != 0;
        for (auto& kv : s->jobs) held |= kv.second->canvas != 0;
        if (!held) {
          Job* j = s->queue.front();
          s->queue.pop_front();
          j->status = DH_E_BUDGET;
          j->t_done = s->now_ms();
          s->st.failed++;
          s->epoch++;
          s->cv_done.notify_all();
          s->cv_work.notify_all();
        }
      } else if (s->stop && s->running == 0) {
        // Shutting down and nothing running will free memory: the rest
        // cannot be admitted.
        const double t = s->now_ms();
        for (Job* j : s->queue) {
          j->status = DH_E_BUDGET;
          j->t_done = t;
          s->st.failed++;
        }
        s->queue.clear();
        s->cv_done.notify_all();
        s->cv_work.notify_all();
      }
      continue;  // otherwise wait for a release or a completion
    }
    // Render outside the lock; other workers admit meanwhile.
    std::vector<dh_view> views(batch.size());
    for (size_t k = 0; k < batch.size(); k++) {
      memset(&views[k], 0, sizeof(dh_view));
      views[k].fr = batch[k]->fr;
      views[k].canvas = batch[k]->canvas;
      views[k].meshes = batch[k]->meshes.data();
      views[k].n_meshes = (uint32_t)batch[k]->meshes.size();
    }
    s->running++;
    // Launch with the scratch capped at this worker's reservation.  A launch
    // whose bin lists outgrow the estimate asks for more: grow the
    // reservation if the budget allows, else hand back all but the first job
    // and relaunch; a single job that cannot grow goes back to the queue
    // front and waits for memory.
    int rc;
    for (;;) {
      dh::ctx_set_scratch_cap(w->ctx, s->cfg.budget_bytes ? w->scratch_reserved : 0);
      lk.unlock();
      rc = dh_render_batch(w->ctx, views.data(), (uint32_t)views.size(), 0);
      lk.lock();
      if (rc != DH_E_BUDGET) break;
      const uint64_t need = dh::ctx_scratch_need(w->ctx);
      const uint64_t extra = need > w->scratch_reserved ? need - w->scratch_reserved : 0;
      if (!s->fits(extra)) evict_for(s, extra, 0, 0);
      if (s->fits(extra)) {
        s->scratch_bytes += extra;
        w->scratch_reserved += extra;
        s->note_peak();
        continue;
      }
      const size_t keep = batch.size() > 1 ? 1 : 0;
      for (size_t k = batch.size(); k-- > keep;) {  // back to the queue front, in order
        Job* j = batch[k];
        s->free_canvases.emplace(std::make_pair(j->fr.W, j->fr.H), j->canvas);
        j->canvas = 0;
        j->status = DH_JOB_QUEUED;
        s->queue.push_front(j);
      }
      if (keep == 0) batch[0]->need = need;  // admit it again only with room for that
      batch.resize(keep);
      views.resize(keep);
      if (keep == 0) break;
    }
    s->running--;
    const uint64_t scratch_now = dh::ctx_scratch_bytes(w->ctx);
    if (batch.empty()) {
      // the front job did not fit: wait for completions or releases
      s->scratch_bytes = s->scratch_bytes - w->scratch_reserved + scratch_now;
      w->scratch_reserved = scratch_now;
      s->epoch++;
      s->cv_work.notify_all();
      continue;
    }
    // the reservation follows the buffers actually held
    s->scratch_bytes = s->scratch_bytes - w->scratch_reserved + scratch_now;
    w->scratch_reserved = scratch_now;
    s->note_peak();
    const double t = s->now_ms();
    s->st.launches++;
    for (Job* j : batch) {
      j->status = rc;
      j->t_done = t;
      j->batch_views = (uint32_t)batch.size();
      if (rc == DH_OK) s->st.completed++;
      else s->st.failed++;
      if ((j->flags & DH_JOB_AUTORELEASE) && rc == DH_OK) {
        s->free_canvases.emplace(std::make_pair(j->fr.W, j->fr.H), j->canvas);
        j->canvas = 0;
      }
    }
    s->epoch++;
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

}  // namespace

extern "C" int dh_srv_create(dh_ctx ctx, const dh_server_config* cfg, dh_server* out) {
  if (!ctx || !out) return DH_E_ARG;
  *out = nullptr;
  dh_server s = new dh_server_s();
  s->ctx = ctx;
  s->cfg = cfg ? *cfg : dh_server_config{};
  if (s->cfg.workers == 0) s->cfg.workers = 2;
  if (s->cfg.max_views == 0) s->cfg.max_views = 64;
  if (s->cfg.max_tris == 0) s->cfg.max_tris = 16ull << 20;
  s->t0 = std::chrono::steady_clock::now();
  for (uint32_t i = 0; i < s->cfg.workers; i++) {
    auto w = std::make_unique<Worker>();
    const int rc = dh_ctx_create_shared(ctx, &w->ctx);
    if (rc != DH_OK) {
      for (auto& x : s->workers) dh_ctx_destroy(x->ctx);
      delete s;
      return rc;
    }
    s->workers.push_back(std::move(w));
  }
  for (auto& w : s->workers) w->th = std::thread(worker_loop, s, w.get());
  *out = s;
  return DH_OK;
}

extern "C" int dh_srv_submit(dh_server s, const dh_job* job, uint64_t* ticket) {
  if (!s || !job || !ticket || (job->n_meshes && !job->meshes) || job->fr.W <= 0 || job->fr.H <= 0)
    return DH_E_ARG;
  auto j = std::make_unique<Job>();
  j->fr = job->fr;
  j->meshes.assign(job->meshes, job->meshes + job->n_meshes);
  j->flags = job->flags;
  for (dh_handle m : j->meshes) {
    uint32_t n = 0;
    const int rc = dh::mesh_triangles(s->ctx, m, &n);
    if (rc != DH_OK) return rc;
    j->n_tri += n;
  }
  j->canvas_bytes = canvas_bytes_of(j->fr.W, j->fr.H);
  const int TS = dh::ctx_tile_size(s->ctx);
  const uint64_t alone = j->canvas_bytes + dh::scratch_estimate(j->n_tri, tiles_of(j->fr.W, j->fr.H, TS),
                                                                 (uint64_t)j->fr.W * j->fr.H, j->meshes.size(), 1);
  if (s->cfg.budget_bytes && alone > s->cfg.budget_bytes) return DH_E_BUDGET;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return DH_E_ARG;
  j->ticket = s->next_ticket++;
  j->t_submit = s->now_ms();
  *ticket = j->ticket;
  Job* p = j.release();
  s->jobs[p->ticket] = p;
  s->queue.push_back(p);
  s->st.submitted++;
  s->epoch++;
  s->cv_work.notify_all();
  return DH_OK;
}

extern "C" int dh_srv_poll(dh_server s, uint64_t ticket, int32_t wait, dh_job_info* out) {
  if (!s || !out) return DH_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end()) return DH_E_HANDLE;
  Job* j = it->second;
  if (wait)
    s->cv_done.wait(lk, [&] { return j->status != DH_JOB_QUEUED && j->status != DH_JOB_RUNNING; });
  out->status = j->status;
  out->batch_views = j->batch_views;
  out->canvas = j->canvas;
  out->t_submit_ms = j->t_submit;
  out->t_admit_ms = j->t_admit;
  out->t_done_ms = j->t_done;
  return DH_OK;
}

extern "C" int dh_srv_release(dh_server s, uint64_t ticket) {
  if (!s) return DH_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end()) return DH_E_HANDLE;
  Job* j = it->second;
  if (j->status == DH_JOB_QUEUED || j->status == DH_JOB_RUNNING) return DH_E_ARG;  // not finished
  if (j->canvas) s->free_canvases.emplace(std::make_pair(j->fr.W, j->fr.H), j->canvas);
  s->jobs.erase(it);
  delete j;
  s->epoch++;
  s->cv_work.notify_all();
  return DH_OK;
}

extern "C" int dh_srv_stats(dh_server s, dh_server_stats* out) {
  if (!s || !out) return DH_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->bytes_in_use = s->in_use();
  out->queued = s->queue.size();
  return DH_OK;
}

extern "C" int dh_srv_destroy(dh_server s) {
  if (!s) return DH_E_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;  // workers drain the queue, then exit
    s->epoch++;
    s->cv_work.notify_all();
  }
  for (auto& w : s->workers) w->th.join();
  for (auto& w : s->workers) dh_ctx_destroy(w->ctx);
  for (auto& kv : s->free_canvases) dh_release(s->ctx, kv.second);
  for (auto& kv : s->jobs) {
    if (kv.second->canvas) dh_release(s->ctx, kv.second->canvas);
    delete kv.second;
  }
  delete s;
  return DH_OK;
}
