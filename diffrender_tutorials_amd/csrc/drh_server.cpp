// drh_server.cpp -- continuous-batching server for mesh-fitting jobs (one MI355X).
//
// A job = a scene, its cameras, target images and an iteration count.  Jobs are queued by
// drh_server_submit (inputs copied).  Threads serve them: admitters set queued jobs up while
// a slot and the HBM budget allow (handle, commit with its LBVH build, uploads); the reaper
// tears finished ones down (downloads, frees), which admits the next.
// A finished job's handle is kept idle, with its device buffers, streams and events, for the
// next job (one with the same settings first; another's settings are reset in place), whose
// set-up then re-commits and re-captures instead of creating and destroying HIP objects.
// The stepper steps the admitted jobs round-robin: every job keeps up to two launches of
// four fused fwd+bwd+Adam steps each in flight on its own HIP stream (one graph launch per
// four steps after its first, eager, step), so the device overlaps the jobs' kernels while
// set-up and teardown of others go on beside them.  Per-iteration latency = HIP event time
// of a launch on its stream over its steps, contention included.
//
// The reference has no serving layer (its optimiser runs one scene at a time,
// src/optimizer_simple.cpp:272-309); this is the "serving" pattern of the path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/drh.h"
#include "drh_internal.h"

namespace {

struct Job {
  int id = 0;
  drh_settings st{};
  // owned copies of the scene arrays
  int n_meshes = 0;
  std::vector<int> vcount, fcount, tex_w, tex_h, diff;
  std::vector<float> transforms, pos, col, tc, nrm, texels;
  std::vector<uint32_t> idx;
  bool has_col = false, has_tc = false, has_nrm = false;
  std::vector<drh_camera> cams;
  std::vector<float> targets;
  int edge_samples = 0, iterations = 0;
  float lr_pos = 1e-3f, lr_col = 1e-3f;
  int flags = 0;     // DRH_JOB_*
  size_t bytes = 0;  // admission estimate
  size_t hit_bytes = 0;   // DRH_JOB_ALIGNED_INTERIOR: stored hits charged at admission (0: recompute)
  // bytes of the render's stored hits (16 per sample of every view)
  size_t hits_needed() const {
    const int sq = (int)std::sqrt((float)std::max(st.spp, 1));
    size_t px = 0;
    for (auto& c : cams) px += (size_t)c.width * (size_t)c.height;
    return 16 * px * (size_t)(sq * sq);
  }
  // run state
  drh_handle h = nullptr;
  bool reset_handle = false;   // h is an idle handle of other settings, reset at admission
  // pinned host memory the last launch's downloads land in: positions, colours (3V floats
  // each), the loss, the fault word; `exported` once they are queued
  float* host_out = nullptr;
  size_t host_out_cap = 0;
  bool exported = false;
  size_t host_out_bytes() const { return 4 * (6 * (size_t)V() + 2); }
  void* d_targets = nullptr;   // the handle's job buffers (drh_job_upload)
  void* d_loss = nullptr;
  int issued = 0, completed = 0;
  struct Launch { hipEvent_t a, b; int steps; };
  std::deque<Launch> inflight;
  // results
  int status = 0;  // 0 queued / running, 1 done, <0 failed (drh_status)
  std::string error;
  std::vector<float> out_pos, out_col;
  float final_loss = 0.0f;
  std::vector<float> latencies_ms;
  std::chrono::steady_clock::time_point t_submit, t_admit, t_done;

  int V() const { int v = 0; for (int c : vcount) v += c; return v; }
  int F() const { int f = 0; for (int c : fcount) f += c; return f; }
  drh_scene_desc desc() const {
    return drh_scene_desc{n_meshes, vcount.data(), fcount.data(), transforms.data(), pos.data(),
                          has_col ? col.data() : nullptr, has_tc ? tc.data() : nullptr, has_nrm ? nrm.data() : nullptr,
                          idx.data(), tex_w.data(), tex_h.data(), texels.empty() ? nullptr : texels.data(), diff.data(),
                          nullptr, nullptr};   // no RESTRICTED_TRANSFORM meshes in jobs
  }
};

// Launches a job keeps queued on its stream (graph replays back to back, so the device stays
// fed while the stepping thread serves the other jobs), and steps per launch: after its first
// (eager) step a job runs STEPS_PER_LAUNCH fused steps per graph launch (drh_opt_steps), which
// cuts the HIP calls per step -- the calls of the stepping, admitting and teardown threads
// serialise in the runtime.
constexpr size_t MAX_INFLIGHT = 2;
constexpr int STEPS_PER_LAUNCH = 4;
// Threads setting jobs up (each job's set-up is serial host work: handle, uploads, edge set,
// LBVH build); DRH_SERVER_ADMITTERS (1..16) overrides the default when the server starts.
constexpr int ADMITTERS = 1;
int admitter_threads() {
  const char* e = std::getenv("DRH_SERVER_ADMITTERS");
  const int n = e ? std::atoi(e) : ADMITTERS;
  return n >= 1 && n <= 16 ? n : ADMITTERS;
}

bool same_settings(const drh_settings& a, const drh_settings& b) {
  return a.mode == b.mode && a.spp == b.spp && a.rng_mode == b.rng_mode && a.tracer == b.tracer &&
         a.device == b.device && a.light_custom == b.light_custom &&
         std::memcmp(a.light_dir, b.light_dir, sizeof a.light_dir) == 0 &&
         std::memcmp(a.light_color, b.light_color, sizeof a.light_color) == 0 &&
         std::memcmp(a.ambient, b.ambient, sizeof a.ambient) == 0;
}

// A finished job's handle waiting for the next job (settings reset if they differ).
struct Idle {
  drh_handle h;
  drh_settings st;
  size_t bytes;   // charged to the budget while idle: the footprint of its last job
};

void destroy_idle(const Idle& e) {
  drh_destroy(e.h);
}

// HBM a job holds while active (the buffers drh_renderer sizes for it), with 25% headroom.
size_t estimate_bytes(const Job& j) {
  const size_t V = j.V(), F = j.F();
  size_t px = 0;
  for (auto& c : j.cams) px += (size_t)c.width * (size_t)c.height;
  const size_t N = (size_t)j.edge_samples;
  const size_t E = 3 * F / 2 + 16;
  size_t b = 0;
  b += 12 * V * 7 + 8 * V + 16 * V;               // vertex arrays, projected vertices
  b += 12 * F + 4 * F;                            // indices, triangle mesh ids
  b += 16 * F * 6 + 4 * F * 8 + 16 * F * 4 + 64 * F + 16 * F * j.cams.size() + 16 * F;  // LBVH build + suspects
  b += 24 * E + 8 * N + 16 * N;                   // edges, cdf; rng table, side colour differences
  if (N < E) b += 24 * N + px / 16 + 64;          // pixel-bin order of the edge samples
  b += 12 * px * 4;                               // image, adjoint, target (+ loss partials)
  b += 4 * (size_t)(6 * V + 64) * 4;              // gradient + Adam moments
  b += 4 * j.texels.size();
  return b + b / 4 + (1u << 20);
}

}  // namespace

struct drh_server_s {
  int device = 0;
  size_t budget = 0, used = 0;
  int max_active = 4;
  std::mutex mu;
  std::condition_variable cv_work;   // stepper: admitted jobs ready, or stop
  std::condition_variable cv_help;   // admitter: a queued job may be set up now, or stop
  std::condition_variable cv_reap;   // reaper: jobs to finish, or stop
  std::condition_variable cv_done;   // drh_server_wait
  std::deque<Job*> pending;          // submitted, not yet admitted (FIFO)
  std::deque<Job*> ready;            // set up by the admitter, not yet stepped
  struct Done { Job* j; int status; std::string err; };
  std::deque<Done> to_finish;        // stepping over: downloads and teardown for the reaper
  int admitting = 0;                 // jobs the admitter is setting up
  int stepping = 0;                  // jobs the stepper holds (mirror for the slot count)
  std::vector<Job*> active;          // stepper only
  // idle handles (at most max_active), oldest first; their bytes stay in `used` (idle_bytes
  // of it) until a job takes the handle over or admission evicts it to make room
  std::vector<Idle> idle;
  size_t idle_bytes = 0;
  std::vector<hipEvent_t> spare_events;   // stepper only: recycled timing events
  hipEvent_t take_event() {
    if (spare_events.empty()) { hipEvent_t e = nullptr; (void)hipEventCreate(&e); return e; }
    hipEvent_t e = spare_events.back();
    spare_events.pop_back();
    return e;
  }
  // pinned result buffers of finished jobs, by capacity, for the next jobs (freed at destroy:
  // hipHostFree may wait for the device)
  std::multimap<size_t, float*> pinned_free;   // under mu
  std::map<int, std::unique_ptr<Job>> jobs;
  int next_id = 1;
  bool stop = false, stepper_done = false;
  std::thread worker, reaper;
  std::vector<std::thread> admitters;
  // stats
  std::vector<float> lat_ms;
  long long iters_done = 0;
  int jobs_done = 0, jobs_failed = 0, peak_active = 0;
  double active_sum = 0.0;
  long long active_samples = 0;
  double admit_s = 0.0, step_s = 0.0, finish_s = 0.0;   // thread time by activity
  // admission by phase (handle and options, targets, cameras + sample table, commit,
  // optimiser set-up); printed at destroy when DRH_SERVER_TRACE is set
  double phase_s[5] = {0, 0, 0, 0, 0};
  int handles_created = 0, handles_reused = 0, handles_reset = 0, handles_evicted = 0;
  // DRH_SERVER_TIMELINE=<path>: every launch's device interval (HIP events against one
  // reference event recorded at start), written to <path> as CSV at destroy
  struct Span { int job, steps; float t0, t1; };
  std::string timeline_path;
  hipEvent_t t_ref = nullptr;
  std::vector<Span> timeline;   // stepper only
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  void run();         // stepping thread
  void admit_run();   // admission thread
  void reap_run();    // teardown thread
  bool admit_one(Job* j, double phase[5]);
  void finish(Job* j, int status, const std::string& err);
  // with mu held: may the admitter set up the next queued job now?
  bool can_admit() const {
    if (stop || pending.empty()) return false;
    if (pending.front()->bytes > budget) return true;   // refused at once
    return stepping + (int)ready.size() + admitting < max_active &&
           used - idle_bytes + pending.front()->bytes <= budget;   // idle handles can be evicted
  }
};

void drh_server_s::finish(Job* j, int status, const std::string& err) {
  bool kept = false;
  std::string fault_err;
  if (j->h) {
    if (status == 1) {
      const size_t V = (size_t)j->V();
      j->out_pos.resize(3 * V);
      j->out_col.resize(3 * V);
      if (j->exported) {
        // the last launch's event has passed, and with it the downloads queued before it
        std::memcpy(j->out_pos.data(), j->host_out, 12 * V);
        std::memcpy(j->out_col.data(), j->host_out + 3 * V, 12 * V);
        std::memcpy(&j->final_loss, j->host_out + 6 * V, 4);
        uint32_t f;
        std::memcpy(&f, j->host_out + 6 * V + 1, 4);
        if (f) {
          status = DRH_E_DEVICE;
          fault_err = "device check failed (fault word set)";
        }
      } else {
        if (drh_get_params(j->h, j->out_pos.data(), j->out_col.data()) != DRH_OK) { status = DRH_E_HIP; }
        if (drh_memcpy_dtoh(j->h, &j->final_loss, j->d_loss, 4) != DRH_OK) status = DRH_E_HIP;
      }
    }
    for (auto& ev : j->inflight) { (void)hipEventDestroy(ev.a); (void)hipEventDestroy(ev.b); }
    j->inflight.clear();
    // a healthy handle (its last step completed: the downloads above waited for it) goes idle
    // for the next job; the oldest idle one makes room if need be
    const Idle e{j->h, j->st, j->bytes + j->hit_bytes};
    std::vector<Idle> drop;
    if (status == 1 && (j->exported || drh_sync(j->h) == DRH_OK)) {
      std::lock_guard<std::mutex> lk(mu);
      if (!stop) {
        if ((int)idle.size() >= max_active) {
          drop.push_back(idle.front());
          idle_bytes -= idle.front().bytes;
          used -= std::min(used, idle.front().bytes);
          idle.erase(idle.begin());
        }
        idle.push_back(e);
        idle_bytes += e.bytes;
        kept = true;
      }
    }
    if (!kept) drop.push_back(e);
    for (const Idle& d : drop) destroy_idle(d);   // waits for the handle's queued copies
    j->h = nullptr;
    j->d_targets = j->d_loss = nullptr;
  }
  std::lock_guard<std::mutex> lk(mu);
  if (j->host_out) {   // nothing is queued into it any more (kept: drained; else destroyed)
    pinned_free.emplace(j->host_out_cap, j->host_out);
    j->host_out = nullptr;
  }
  if (!kept) used -= std::min(used, j->bytes + j->hit_bytes);   // a kept handle's charge stays
  j->status = status;
  j->error = fault_err.empty() ? err : fault_err;
  j->t_done = std::chrono::steady_clock::now();
  if (status == 1) jobs_done++; else jobs_failed++;
  cv_done.notify_all();
  cv_help.notify_all();   // memory freed: the next queued job may fit
}

bool drh_server_s::admit_one(Job* j, double phase[5]) {
  drh_status s = DRH_OK;
  auto phase_t = std::chrono::steady_clock::now();
  drh_handle h = j->h;   // an idle handle taken over at admission, or none
  if (!h) {
    if ((s = drh_create(&j->st, &h)) != DRH_OK) { finish(j, s, "drh_create failed"); return false; }
    j->h = h;
  } else if (j->reset_handle && (s = drh_reset_settings(h, &j->st)) != DRH_OK) {
    finish(j, s, drh_last_error(h));
    return false;
  }
  // every option the server sets, set for this job (an idle handle keeps its last job's)
  (void)drh_set_device_sharing(h, max_active);
  (void)drh_opt_use_graph(h, 1);
  // a short job refits its commit-time tree throughout: one captured step form instead of
  // two (capture and instantiation cost host time comparable to a short job's steps)
  (void)drh_set_bvh_rebuild(h, j->iterations <= 64 ? j->iterations + 1 : 8);
  (void)drh_set_deterministic(h, (j->flags & DRH_JOB_DETERMINISTIC) ? 1 : 0);
  (void)drh_set_edge_grad_mode(h, (j->flags & DRH_JOB_FULL_JACOBIAN) ? DRH_EDGE_GRAD_FULL_JACOBIAN
                                                                       : DRH_EDGE_GRAD_REFERENCE_DIAGONAL);
  if (j->flags & DRH_JOB_ALIGNED_INTERIOR) {
    // the hit buffer was charged to the job at admission if it fitted the budget; else recompute
    (void)drh_set_interior_samples(h, 1, DRH_HITS_AUTO);
    (void)drh_set_hit_buffer_limit(h, j->hit_bytes);
  } else {
    (void)drh_set_interior_samples(h, 0, DRH_HITS_RECOMPUTE);
    (void)drh_set_hit_buffer_limit(h, ~(size_t)0);
  }
  const drh_scene_desc d = j->desc();
  {
    const size_t need = j->host_out_bytes();
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = pinned_free.lower_bound(need);
      if (it != pinned_free.end() && it->first <= 4 * need + (64 << 10)) {   // no far larger one
        j->host_out_cap = it->first;
        j->host_out = it->second;
        pinned_free.erase(it);
      }
    }
    if (!j->host_out) {
      size_t cap = 64 << 10;
      while (cap < need) cap *= 2;
      if (hipHostMalloc((void**)&j->host_out, cap, hipHostMallocDefault) != hipSuccess) {
        j->host_out = nullptr;
        finish(j, DRH_E_OOM, "pinned host memory for the results");
        return false;
      }
      j->host_out_cap = cap;
    }
  }
  // uploads (targets, cameras, edge-sample table), the commit's LBVH build and the optimiser
  // set-up are all queued on the handle's stream without a synchronisation: in a busy server
  // one would wait for whatever shares the stream's hardware queue
  auto lap = [&](int k) {
    const auto now = std::chrono::steady_clock::now();
    phase[k] += std::chrono::duration<double>(now - phase_t).count();
    phase_t = now;
  };
  lap(0);
  if ((s = drh_job_upload(h, j->targets.data(), 4 * j->targets.size(), &j->d_targets, &j->d_loss)) != DRH_OK) {
    finish(j, s, drh_last_error(h));
    return false;
  }
  lap(1);
  if ((s = drh_prepare_steps(h, j->cams.data(), (int)j->cams.size(), j->edge_samples)) != DRH_OK) { finish(j, s, drh_last_error(h)); return false; }
  lap(2);
  if ((s = drh_commit(h, &d)) != DRH_OK) { finish(j, s, drh_last_error(h)); return false; }
  lap(3);
  if ((s = drh_opt_init(h, DRH_OPT_ADAM, j->lr_pos, j->lr_col)) != DRH_OK) { finish(j, s, drh_last_error(h)); return false; }
  lap(4);
  j->t_admit = std::chrono::steady_clock::now();
  return true;
}

// Admission (handle, commit with its LBVH build, uploads) runs on its own thread beside the
// stepping thread, and so does teardown (downloads, frees): neither holds up the other
// jobs' steps, and a burst of finishing jobs does not delay the next admissions.
void drh_server_s::admit_run() {
  (void)hipSetDevice(device);
  while (true) {
    Job* adm = nullptr;
    std::vector<Idle> evict;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_help.wait(lk, [&] { return can_admit() || stop; });
      if (!can_admit()) break;   // stopped
      Job* j = pending.front();
      pending.pop_front();
      if (j->bytes > budget) {           // can never fit
        j->status = DRH_E_OOM;
        j->error = "job exceeds the server's HBM budget";
        jobs_failed++;
        cv_done.notify_all();
        continue;
      }
      // take over an idle handle whose charge fits, one with the same settings first, else
      // any (its settings are reset: cheaper than a new handle's streams, events and first
      // allocations); then evict idle handles (oldest first) until the job fits the budget
      size_t take = idle.size();
      for (size_t k = 0; k < idle.size(); k++) {
        if (used - idle_bytes + std::max(j->bytes, idle[k].bytes) > budget) continue;
        if (same_settings(idle[k].st, j->st)) { take = k; break; }
        if (take == idle.size()) take = k;
      }
      if (take < idle.size()) {
        const size_t k = take;
        const Idle e = idle[k];
        j->reset_handle = !same_settings(e.st, j->st);
        idle.erase(idle.begin() + k);
        idle_bytes -= e.bytes;
        used -= std::min(used, e.bytes);
        j->h = e.h;
        j->bytes = std::max(j->bytes, e.bytes);   // its buffers are at least the last job's
        handles_reused++;
        if (j->reset_handle) handles_reset++;
      }
      used += j->bytes;
      if (!j->h) handles_created++;
      while (used > budget && !idle.empty()) {
        handles_evicted++;
        evict.push_back(idle.front());
        idle_bytes -= idle.front().bytes;
        used -= std::min(used, idle.front().bytes);
        idle.erase(idle.begin());
      }
      // store the render's hits when they fit what is left of the budget (multi-view
      // adjoint: storing saves the interior pass's rays, recomputing saves the HBM)
      j->hit_bytes = 0;
      if ((j->flags & DRH_JOB_ALIGNED_INTERIOR) && used + j->hits_needed() <= budget) {
        j->hit_bytes = j->hits_needed();
        used += j->hit_bytes;
      }
      admitting++;
      adm = j;
    }
    const auto t = std::chrono::steady_clock::now();
    for (const Idle& e : evict) destroy_idle(e);
    double phase[5] = {0, 0, 0, 0, 0};
    const bool ok = admit_one(adm, phase);
    std::lock_guard<std::mutex> lk(mu);
    for (int k = 0; k < 5; k++) phase_s[k] += phase[k];
    admit_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    admitting--;
    if (ok) ready.push_back(adm);
    cv_work.notify_all();   // the stepper also waits here for admissions to drain at stop
  }
}

void drh_server_s::reap_run() {
  (void)hipSetDevice(device);
  while (true) {
    Done fin{nullptr, 0, std::string()};
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_reap.wait(lk, [&] { return !to_finish.empty() || (stop && stepper_done); });
      if (to_finish.empty()) break;   // stopped, and the stepper has handed over everything
      fin = std::move(to_finish.front());
      to_finish.pop_front();
    }
    const auto t = std::chrono::steady_clock::now();
    finish(fin.j, fin.status, fin.err);
    std::lock_guard<std::mutex> lk(mu);
    finish_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
  }
}

void drh_server_s::run() {
  (void)hipSetDevice(device);
  auto hand_over = [&](size_t k, int status, const std::string& err) {
    Job* j = active[k];
    active.erase(active.begin() + k);
    std::lock_guard<std::mutex> lk(mu);
    to_finish.push_back(Done{j, status, err});
    stepping = (int)active.size();
    cv_reap.notify_all();
    cv_help.notify_all();   // a slot is free
  };
  while (true) {
    {
      std::unique_lock<std::mutex> lk(mu);
      while (!ready.empty()) { active.push_back(ready.front()); ready.pop_front(); }
      stepping = (int)active.size();
      if (stop) break;
      if (active.empty()) {
        cv_work.wait_for(lk, std::chrono::milliseconds(50));
        continue;
      }
      peak_active = std::max(peak_active, (int)active.size());
      active_sum += active.size();
      active_samples++;
    }
    // stepping: keep up to MAX_INFLIGHT launches in flight per job; harvest completed ones
    bool progressed = false;
    for (size_t k = 0; k < active.size();) {
      Job* j = active[k];
      // harvest
      while (!j->inflight.empty() && hipEventQuery(j->inflight.front().b) == hipSuccess) {
        const Job::Launch l = j->inflight.front();
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, l.a, l.b);
        if (t_ref) {
          float t0 = 0.0f;
          if (hipEventElapsedTime(&t0, t_ref, l.a) == hipSuccess) timeline.push_back(Span{j->id, l.steps, t0, t0 + ms});
        }
        spare_events.push_back(l.a);
        spare_events.push_back(l.b);
        j->inflight.pop_front();
        j->completed += l.steps;
        // per-iteration latency: the launch's device time over its steps
        for (int q = 0; q < l.steps; q++) j->latencies_ms.push_back(ms / l.steps);
        {
          std::lock_guard<std::mutex> lk(mu);
          for (int q = 0; q < l.steps; q++) lat_ms.push_back(ms / l.steps);
          iters_done += l.steps;
        }
        progressed = true;
      }
      if (j->completed >= j->iterations) {
        hand_over(k, 1, "");
        progressed = true;
        continue;
      }
      // issue
      if (j->issued < j->iterations && j->inflight.size() < MAX_INFLIGHT) {
        const auto ts = std::chrono::steady_clock::now();
        hipEvent_t a = take_event(), b = take_event();
        hipStream_t st = (hipStream_t)drh_stream(j->h);
        const int steps = j->issued == 0 ? 1 : std::min(STEPS_PER_LAUNCH, j->iterations - j->issued);
        (void)hipEventRecord(a, st);
        drh_status s = drh_opt_steps(j->h, j->cams.data(), (int)j->cams.size(), (const float*)j->d_targets,
                                     j->edge_samples, (float*)j->d_loss, steps);
        if (s == DRH_OK && j->issued + steps == j->iterations) {
          // the last launch: its results go to pinned memory behind it, read at teardown
          const size_t V = (size_t)j->V();
          s = drh_export_params(j->h, j->host_out, j->host_out + 3 * V, j->d_loss, j->host_out + 6 * V,
                                (uint32_t*)(j->host_out + 6 * V + 1));
          j->exported = s == DRH_OK;
        }
        (void)hipEventRecord(b, st);
        if (s != DRH_OK) {
          spare_events.push_back(a);
          spare_events.push_back(b);
          hand_over(k, s, drh_last_error(j->h));
          continue;
        }
        j->inflight.push_back(Job::Launch{a, b, steps});
        j->issued += steps;
        std::lock_guard<std::mutex> lk(mu);
        step_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
        progressed = true;
      }
      k++;
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  // Stopped.  An admitter may still be setting up a job it took before the stop (outside the
  // lock); it would push that job into `ready` after this thread had left, where nothing
  // drains it any more (its handle, pooled HBM and budget share would never be released).  So
  // wait for every admission in progress to end, then hand the held jobs and whatever the
  // admitters left in `ready` to the reaper.
  {
    std::unique_lock<std::mutex> lk(mu);
    cv_work.wait(lk, [&] { return admitting == 0; });
    while (!ready.empty()) { active.push_back(ready.front()); ready.pop_front(); }
  }
  while (!active.empty()) hand_over(active.size() - 1, DRH_E_INVALID, "server stopped");
  for (hipEvent_t e : spare_events) (void)hipEventDestroy(e);
  spare_events.clear();
  std::lock_guard<std::mutex> lk(mu);
  stepper_done = true;
  cv_reap.notify_all();
}

extern "C" {

drh_status drh_server_create(int device, size_t hbm_budget_bytes, int max_active, drh_server* out) {
  if (!out || max_active <= 0) return DRH_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return DRH_E_HIP;
  auto* s = new drh_server_s();
  s->device = device;
  s->budget = hbm_budget_bytes;
  s->max_active = max_active;
  if (const char* tl = std::getenv("DRH_SERVER_TIMELINE")) {
    (void)hipSetDevice(device);
    if (hipEventCreate(&s->t_ref) == hipSuccess && hipEventRecord(s->t_ref, nullptr) == hipSuccess) {
      s->timeline_path = tl;
    } else {
      (void)hipGetLastError();
      if (s->t_ref) (void)hipEventDestroy(s->t_ref);
      s->t_ref = nullptr;
    }
  }
  s->worker = std::thread([s] { s->run(); });
  for (int k = 0, n = admitter_threads(); k < n; k++) s->admitters.emplace_back([s] { s->admit_run(); });
  s->reaper = std::thread([s] { s->reap_run(); });
  *out = s;
  return DRH_OK;
}

drh_status drh_server_submit(drh_server s, const drh_job_desc* d, int* job_id) {
  if (!s || !d || !job_id || !d->cams || d->views <= 0 || !d->targets || d->iterations <= 0 || d->edge_samples < 0)
    return DRH_E_INVALID;
  const drh_scene_desc& sc = d->scene;
  if (sc.n_meshes <= 0 || !sc.vcount || !sc.fcount || !sc.transforms || !sc.pos || !sc.idx) return DRH_E_INVALID;
  // jobs fit vertex positions and colours; RESTRICTED_TRANSFORM meshes go through a handle
  if (sc.restricted)
    for (int i = 0; i < sc.n_meshes; i++)
      if (sc.restricted[i]) return DRH_E_UNSUPPORTED;
  auto j = std::make_unique<Job>();
  j->st = d->settings;
  j->n_meshes = sc.n_meshes;
  j->vcount.assign(sc.vcount, sc.vcount + sc.n_meshes);
  j->fcount.assign(sc.fcount, sc.fcount + sc.n_meshes);
  int V = 0, F = 0, T = 0;
  for (int i = 0; i < sc.n_meshes; i++) {
    V += sc.vcount[i];
    F += sc.fcount[i];
    const int tw = sc.tex_w ? sc.tex_w[i] : 0, th = sc.tex_h ? sc.tex_h[i] : 0;
    j->tex_w.push_back(tw);
    j->tex_h.push_back(th);
    T += 3 * tw * th;
    j->diff.push_back(sc.differentiable ? sc.differentiable[i] : 1);
  }
  j->transforms.assign(sc.transforms, sc.transforms + 16 * sc.n_meshes);
  j->pos.assign(sc.pos, sc.pos + 3 * (size_t)V);
  if ((j->has_col = sc.col != nullptr)) j->col.assign(sc.col, sc.col + 3 * (size_t)V);
  if ((j->has_tc = sc.tc != nullptr)) j->tc.assign(sc.tc, sc.tc + 2 * (size_t)V);
  if ((j->has_nrm = sc.nrm != nullptr)) j->nrm.assign(sc.nrm, sc.nrm + 3 * (size_t)V);
  j->idx.assign(sc.idx, sc.idx + 3 * (size_t)F);
  if (T && sc.texels) j->texels.assign(sc.texels, sc.texels + T);
  j->cams.assign(d->cams, d->cams + d->views);
  size_t px = 0;
  for (auto& c : j->cams) px += (size_t)c.width * (size_t)c.height;
  j->targets.assign(d->targets, d->targets + 3 * px);
  j->edge_samples = d->edge_samples;
  j->iterations = d->iterations;
  j->lr_pos = d->lr_pos;
  j->lr_col = d->lr_col;
  j->flags = d->flags;
  j->bytes = estimate_bytes(*j);
  j->t_submit = std::chrono::steady_clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  j->id = s->next_id++;
  *job_id = j->id;
  s->pending.push_back(j.get());
  s->jobs[j->id] = std::move(j);
  s->cv_help.notify_all();
  return DRH_OK;
}

drh_status drh_server_wait(drh_server s, int job_id, double timeout_s, drh_job_result* r) {
  if (!s) return DRH_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return DRH_E_INVALID;
  Job* j = it->second.get();
  const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return j->status != 0; });
  if (!ok) return DRH_E_INVALID;
  if (r) {
    r->status = j->status == 1 ? DRH_OK : (drh_status)j->status;
    r->iterations = j->completed;
    r->final_loss = j->final_loss;
    r->queue_ms = std::chrono::duration<double, std::milli>(j->t_admit - j->t_submit).count();
    r->run_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_admit).count();
    if (r->pos && !j->out_pos.empty()) std::memcpy(r->pos, j->out_pos.data(), 4 * j->out_pos.size());
    if (r->col && !j->out_col.empty()) std::memcpy(r->col, j->out_col.data(), 4 * j->out_col.size());
    if (r->latencies_ms && r->latencies_cap > 0) {
      const int n = std::min<int>(r->latencies_cap, (int)j->latencies_ms.size());
      std::memcpy(r->latencies_ms, j->latencies_ms.data(), 4 * (size_t)n);
    }
    r->hbm_bytes = j->bytes + j->hit_bytes;
    r->hit_bytes = j->hit_bytes;
  }
  return j->status == 1 ? DRH_OK : (drh_status)j->status;
}

drh_status drh_server_stats(drh_server s, drh_server_stats_t* out) {
  if (!s || !out) return DRH_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<float> l = s->lat_ms;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) -> double {
    if (l.empty()) return 0.0;
    const size_t i = std::min(l.size() - 1, (size_t)(p * (double)(l.size() - 1) + 0.5));
    return l[i];
  };
  out->iterations = s->iters_done;
  out->jobs_done = s->jobs_done;
  out->jobs_failed = s->jobs_failed;
  out->jobs_queued = (int)s->pending.size();
  out->p50_ms = pct(0.50);
  out->p99_ms = pct(0.99);
  out->peak_active = s->peak_active;
  out->mean_active = s->active_samples ? s->active_sum / (double)s->active_samples : 0.0;
  out->hbm_used = s->used;
  out->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
  out->admit_s = s->admit_s;
  out->step_s = s->step_s;
  out->finish_s = s->finish_s;
  return DRH_OK;
}

void drh_server_destroy(drh_server s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (Job* j : s->pending) { j->status = DRH_E_INVALID; j->error = "server stopped"; }
    s->pending.clear();
    s->cv_work.notify_all();
    s->cv_help.notify_all();
  }
  for (auto& t : s->admitters) t.join();   // they set up no new job once stopped
  s->worker.join();     // hands its jobs (and those admitted meanwhile) to the reaper
  s->reaper.join();     // finishes them, then exits
  if (std::getenv("DRH_SERVER_TRACE"))
    std::fprintf(stderr, "drh_server admission s: handle %.4f targets %.4f cameras+table %.4f commit %.4f opt_init %.4f; "
                 "handles created %d reused %d (of which reset %d) evicted %d\n",
                 s->phase_s[0], s->phase_s[1], s->phase_s[2], s->phase_s[3], s->phase_s[4], s->handles_created,
                 s->handles_reused, s->handles_reset, s->handles_evicted);
  (void)hipSetDevice(s->device);
  for (const Idle& e : s->idle) destroy_idle(e);   // no thread is left to take them
  s->idle.clear();
  for (auto& p : s->pinned_free) (void)hipHostFree(p.second);
  s->pinned_free.clear();
  if (s->t_ref) {
    if (FILE* f = std::fopen(s->timeline_path.c_str(), "w")) {
      std::fprintf(f, "job,steps,start_ms,end_ms\n");
      for (const auto& sp : s->timeline) std::fprintf(f, "%d,%d,%.4f,%.4f\n", sp.job, sp.steps, sp.t0, sp.t1);
      std::fclose(f);
    }
    (void)hipEventDestroy(s->t_ref);
  }
  delete s;
}

}  // extern "C"

extern "C" size_t drh_server_job_bytes(const drh_job_desc* d) {
  if (!d || d->scene.n_meshes <= 0 || !d->scene.vcount || !d->scene.fcount || !d->cams || d->views <= 0) return 0;
  Job j;
  j.n_meshes = d->scene.n_meshes;
  j.vcount.assign(d->scene.vcount, d->scene.vcount + j.n_meshes);
  j.fcount.assign(d->scene.fcount, d->scene.fcount + j.n_meshes);
  j.cams.assign(d->cams, d->cams + d->views);
  j.edge_samples = d->edge_samples;
  size_t T = 0;
  for (int i = 0; i < j.n_meshes; i++) T += 3 * (size_t)(d->scene.tex_w ? d->scene.tex_w[i] : 0) * (d->scene.tex_h ? d->scene.tex_h[i] : 0);
  j.texels.resize(T);
  return estimate_bytes(j);
}
