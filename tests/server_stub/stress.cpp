// tests/server_stub/stress.cpp -- drives drh_server.cpp (built against the stubs in this
// folder) through its thread hand-offs: jobs of uneven length completing, a job whose step
// fails, admission refused by budget, and servers destroyed while admissions and steps are in
// flight.  Exit status 0 when every check holds; ThreadSanitizer / AddressSanitizer abort the
// run on a data race or a use after free.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/drh.h"

extern std::atomic<int> g_live_handles, g_created_handles, g_reset_handles;

#define CHECK(c) do { if (!(c)) { fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); return 1; } } while (0)

struct TinyJob {
  int vcount = 3, fcount = 1;
  float tr[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float pos[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
  uint32_t idx[3] = {0, 1, 2};
  int diff = 1;
  drh_camera cam{};
  std::vector<float> target = std::vector<float>(4 * 4 * 3, 0.5f);
  drh_job_desc desc(int iters, int tag) {
    cam.width = cam.height = 4;
    drh_job_desc d{};
    d.settings.mode = DRH_VERTEX_COLOR;
    d.settings.spp = 1;
    d.scene = drh_scene_desc{1, &vcount, &fcount, tr, pos, nullptr, nullptr, nullptr, idx, nullptr, nullptr, nullptr,
                             &diff, nullptr, nullptr};
    d.cams = &cam;
    d.views = 1;
    d.targets = target.data();
    d.edge_samples = tag;
    d.iterations = iters;
    d.lr_pos = d.lr_col = 1e-3f;
    return d;
  }
};

int main() {
  TinyJob tj;
  // 1. uneven jobs, all complete; a failing job fails alone
  {
    drh_server s;
    CHECK(drh_server_create(0, 1ull << 30, 4, &s) == DRH_OK);
    std::vector<int> ids;
    for (int k = 0; k < 24; k++) {
      drh_job_desc d = tj.desc(1 + (k * 7) % 19, k == 5 ? 13 : 64);
      int id;
      CHECK(drh_server_submit(s, &d, &id) == DRH_OK);
      ids.push_back(id);
    }
    for (int k = 0; k < 24; k++) {
      float pos[9];
      drh_job_result r{};
      r.pos = pos;
      const drh_status st = drh_server_wait(s, ids[k], 60.0, &r);
      if (k == 5) { CHECK(st == DRH_E_HIP); continue; }
      CHECK(st == DRH_OK);
      CHECK(r.iterations == 1 + (k * 7) % 19);
      CHECK(pos[0] == (float)r.iterations);
    }
    drh_server_stats_t stt;
    CHECK(drh_server_stats(s, &stt) == DRH_OK);
    CHECK(stt.jobs_done == 23 && stt.jobs_failed == 1 && stt.peak_active <= 4);
    drh_server_destroy(s);
    CHECK(g_live_handles.load() == 0);
  }
  // 2. a job above the whole budget is refused; the others run one at a time
  {
    drh_job_desc d = tj.desc(3, 64);
    const size_t need = drh_server_job_bytes(&d);
    drh_server s;
    CHECK(drh_server_create(0, need + need / 2, 4, &s) == DRH_OK);
    int small1, small2, big;
    CHECK(drh_server_submit(s, &d, &small1) == DRH_OK);
    CHECK(drh_server_submit(s, &d, &small2) == DRH_OK);
    drh_server_destroy(s);   // possibly while the first admission is in flight
    CHECK(g_live_handles.load() == 0);
    CHECK(drh_server_create(0, need / 2, 4, &s) == DRH_OK);
    CHECK(drh_server_submit(s, &d, &big) == DRH_OK);
    CHECK(drh_server_wait(s, big, 60.0, nullptr) == DRH_E_OOM);
    drh_server_destroy(s);
    (void)small1; (void)small2;
  }
  // 3. servers destroyed while jobs are being admitted and stepped
  for (int rep = 0; rep < 30; rep++) {
    drh_server s;
    CHECK(drh_server_create(0, 1ull << 30, 1 + rep % 5, &s) == DRH_OK);
    std::vector<int> ids;
    for (int k = 0; k < 12; k++) {
      drh_job_desc d = tj.desc(5 + k % 9, 64);
      int id;
      CHECK(drh_server_submit(s, &d, &id) == DRH_OK);
      ids.push_back(id);
    }
    if (rep % 3 == 1) CHECK(drh_server_wait(s, ids[0], 60.0, nullptr) == DRH_OK);
    drh_server_destroy(s);
    CHECK(g_live_handles.load() == 0);   // every admitted job was torn down
  }
  // 4. idle handles: under a budget of one and a half jobs every job takes the last job's
  // handle over, those of other settings (alternating spp) resetting it: one handle in all
  {
    drh_job_desc d = tj.desc(4, 64);
    const size_t need = drh_server_job_bytes(&d);
    const int created0 = g_created_handles.load(), reset0 = g_reset_handles.load();
    drh_server s;
    CHECK(drh_server_create(0, need + need / 2, 4, &s) == DRH_OK);
    std::vector<int> ids;
    for (int k = 0; k < 16; k++) {
      drh_job_desc dk = tj.desc(2 + k % 3, 64);
      if (k >= 8) dk.settings.spp = (k % 2) ? 1 : 4;
      int id;
      CHECK(drh_server_submit(s, &dk, &id) == DRH_OK);
      ids.push_back(id);
    }
    for (int k = 0; k < 16; k++) {
      float pos[9];
      drh_job_result r{};
      r.pos = pos;
      CHECK(drh_server_wait(s, ids[k], 60.0, &r) == DRH_OK);
      CHECK(r.iterations == 2 + k % 3 && pos[0] == (float)r.iterations);
    }
    drh_server_stats_t stt;
    CHECK(drh_server_stats(s, &stt) == DRH_OK);
    CHECK(stt.hbm_used <= need + need / 2);
    CHECK(g_live_handles.load() == 1);   // the last job's handle, idle
    CHECK(g_created_handles.load() - created0 == 1);
    CHECK(g_reset_handles.load() - reset0 >= 7);   // jobs 8..15 alternate spp 4 / 1
    drh_server_destroy(s);
    CHECK(g_live_handles.load() == 0);
  }
  // 5. two idle handles of different settings: a job takes the one with its own settings
  // (no reset) although the other went idle first
  {
    drh_job_desc d = tj.desc(4, 64);
    const size_t need = drh_server_job_bytes(&d);
    drh_server s;
    CHECK(drh_server_create(0, 3 * need, 2, &s) == DRH_OK);
    drh_job_desc a = tj.desc(3, 64), b = tj.desc(3, 64);
    b.settings.spp = 4;
    int ia, ib;
    CHECK(drh_server_submit(s, &a, &ia) == DRH_OK);
    CHECK(drh_server_submit(s, &b, &ib) == DRH_OK);
    CHECK(drh_server_wait(s, ia, 60.0, nullptr) == DRH_OK);
    CHECK(drh_server_wait(s, ib, 60.0, nullptr) == DRH_OK);
    const int created0 = g_created_handles.load(), reset0 = g_reset_handles.load();
    int ic;
    CHECK(drh_server_submit(s, &b, &ic) == DRH_OK);
    CHECK(drh_server_wait(s, ic, 60.0, nullptr) == DRH_OK);
    CHECK(g_created_handles.load() == created0 && g_reset_handles.load() == reset0);
    drh_server_destroy(s);
    CHECK(g_live_handles.load() == 0);
  }
  printf("server stress ok\n");
  return 0;
}
