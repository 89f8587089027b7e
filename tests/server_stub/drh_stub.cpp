// tests/server_stub/drh_stub.cpp -- host-only stand-ins for the drh_* calls the server makes
// (test code, built only by tests/test_server_threads.py).  Every call checks that its handle
// is alive and writes to it without locking, so ThreadSanitizer reports two threads using one
// handle at once and AddressSanitizer a handle used after drh_destroy.  Timings: a commit
// takes ~0.3 ms of host time (a wide admission window), a fused step queues ~40-200 us of
// simulated device work on the handle's stream.
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/drh.h"
#include "../../diffrender_tutorials_amd/csrc/drh_internal.h"

struct drh_renderer {
  ihipStream_t stream;
  int magic = 0x0D44;
  int V = 0;
  int steps = 0;
  int calls = 0;           // written by every call (unsynchronised on purpose)
  std::string err;
  std::vector<float> job_tgt;
  float job_loss[4] = {0, 0, 0, 0};
};

std::atomic<int> g_live_handles{0};
std::atomic<int> g_created_handles{0}, g_reset_handles{0};
std::atomic<long long> g_steps{0};

static drh_renderer* live(drh_handle h, const char* what) {
  if (!h || h->magic != 0x0D44) { fprintf(stderr, "stub: %s on a dead handle\n", what); abort(); }
  h->calls++;
  return h;
}
static void host_work(int us) { std::this_thread::sleep_for(std::chrono::microseconds(us)); }

extern "C" {
drh_status drh_create(const drh_settings* s, drh_handle* out) {
  if (!s || !out) return DRH_E_INVALID;
  *out = new drh_renderer();
  g_live_handles++;
  g_created_handles++;
  return DRH_OK;
}
void drh_destroy(drh_handle h) {
  live(h, "drh_destroy");
  while (stub_now_ns() < h->stream.busy_until.load()) host_work(20);   // hipStreamSynchronize
  h->magic = 0;
  delete h;
  g_live_handles--;
}
const char* drh_last_error(drh_handle h) { return live(h, "drh_last_error")->err.c_str(); }
drh_status drh_commit(drh_handle h, const drh_scene_desc* d) {
  live(h, "drh_commit");
  h->V = 0;
  for (int i = 0; i < d->n_meshes; i++) h->V += d->vcount[i];
  host_work(300);
  return DRH_OK;
}
drh_status drh_set_device_sharing(drh_handle h, int) { live(h, "drh_set_device_sharing"); return DRH_OK; }
drh_status drh_set_bvh_rebuild(drh_handle h, int) { live(h, "drh_set_bvh_rebuild"); return DRH_OK; }
drh_status drh_set_deterministic(drh_handle h, int) { live(h, "drh_set_deterministic"); return DRH_OK; }
drh_status drh_set_edge_grad_mode(drh_handle h, int) { live(h, "drh_set_edge_grad_mode"); return DRH_OK; }
drh_status drh_set_interior_samples(drh_handle h, int, int) { live(h, "drh_set_interior_samples"); return DRH_OK; }
drh_status drh_set_hit_buffer_limit(drh_handle h, size_t) { live(h, "drh_set_hit_buffer_limit"); return DRH_OK; }
drh_status drh_malloc(drh_handle h, size_t bytes, void** p) {
  live(h, "drh_malloc");
  *p = calloc(1, bytes ? bytes : 4);
  return *p ? DRH_OK : DRH_E_OOM;
}
drh_status drh_free(drh_handle h, void* p) {
  live(h, "drh_free");
  while (stub_now_ns() < h->stream.busy_until.load()) host_work(20);   // waits for the stream
  free(p);
  return DRH_OK;
}
drh_status drh_memcpy_htod(drh_handle h, void* d, const void* s, size_t n) {
  live(h, "drh_memcpy_htod");
  std::memcpy(d, s, n);
  return DRH_OK;
}
drh_status drh_memcpy_dtoh(drh_handle h, void* d, const void* s, size_t n) {
  live(h, "drh_memcpy_dtoh");
  while (stub_now_ns() < h->stream.busy_until.load()) host_work(20);
  std::memcpy(d, s, n);
  return DRH_OK;
}
drh_status drh_prepare_steps(drh_handle h, const drh_camera*, int, int) {
  live(h, "drh_prepare_steps");
  host_work(50);
  return DRH_OK;
}
drh_status drh_opt_init(drh_handle h, int, float, float) { live(h, "drh_opt_init"); h->steps = 0; return DRH_OK; }
// edge_samples == 13 tags a job whose third step fails (the stepper's failure path)
drh_status drh_opt_step(drh_handle h, const drh_camera*, int, const float* d_targets, int edge_samples, float* d_loss) {
  live(h, "drh_opt_step");
  if (edge_samples == 13 && h->steps == 2) { h->err = "injected failure"; return DRH_E_HIP; }
  h->steps++;
  g_steps++;
  h->stream.enqueue(40000 + 40000 * (h->V % 5));
  if (d_loss) *d_loss = (float)h->steps + 0.0f * d_targets[0];
  return DRH_OK;
}
drh_status drh_opt_steps(drh_handle h, const drh_camera* c, int v, const float* d_targets, int edge_samples, float* d_loss,
                         int steps) {
  for (int k = 0; k < steps; k++) {
    const drh_status s = drh_opt_step(h, c, v, d_targets, edge_samples, d_loss);
    if (s != DRH_OK) return s;
  }
  return DRH_OK;
}
void* drh_stream(drh_handle h) { return &live(h, "drh_stream")->stream; }
drh_status drh_reset_settings(drh_handle h, const drh_settings* s) {
  live(h, "drh_reset_settings");
  if (!s) return DRH_E_INVALID;
  g_reset_handles++;
  return DRH_OK;
}
// the downloads land when the stream passes them: written here at once, with the values the
// queued steps leave (the server reads them only after the last launch's event)
drh_status drh_export_params(drh_handle h, float* pos, float* col, const void* d_loss, float* loss, uint32_t* fault) {
  live(h, "drh_export_params");
  for (int k = 0; k < 3 * h->V; k++) { pos[k] = (float)h->steps; col[k] = 0.0f; }
  std::memcpy(loss, d_loss, 4);
  *fault = 0;
  return DRH_OK;
}
drh_status drh_job_upload(drh_handle h, const void* targets, size_t bytes, void** d_targets, void** d_loss) {
  live(h, "drh_job_upload");
  h->job_tgt.resize(bytes / 4 + 1);
  std::memcpy(h->job_tgt.data(), targets, bytes);
  *d_targets = h->job_tgt.data();
  *d_loss = h->job_loss;
  return DRH_OK;
}
drh_status drh_sync(drh_handle h) {
  live(h, "drh_sync");
  while (stub_now_ns() < h->stream.busy_until.load()) host_work(20);   // hipStreamSynchronize
  return DRH_OK;
}
drh_status drh_opt_use_graph(drh_handle h, int) { live(h, "drh_opt_use_graph"); return DRH_OK; }
drh_status drh_get_params(drh_handle h, float* pos, float* col) {
  live(h, "drh_get_params");
  while (stub_now_ns() < h->stream.busy_until.load()) host_work(20);
  for (int k = 0; k < 3 * h->V; k++) { if (pos) pos[k] = (float)h->steps; if (col) col[k] = 0.0f; }
  return DRH_OK;
}
}  // extern "C"
