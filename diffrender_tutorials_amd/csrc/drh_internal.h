// drh_internal.h -- entry points the library uses between its own translation units; not
// part of the C ABI (hidden from the shared library's exports).
#pragma once
#include "../../include/drh.h"

// Upload the cameras and the edge-sample table of a job's fused steps while the handle's
// stream holds no kernels yet, so the first step finds them in place and does not wait for
// the commit's kernels (the server calls it before drh_commit).
extern "C" __attribute__((visibility("hidden"))) drh_status drh_prepare_steps(drh_handle h, const drh_camera* cams,
                                                                              int views, int edge_samples);

// Replace a handle's settings, keeping its streams, events and buffers (as drh_create for
// those settings would leave it, before a commit); the server reuses idle handles this way.
extern "C" __attribute__((visibility("hidden"))) drh_status drh_reset_settings(drh_handle h, const drh_settings* settings);

// Queue, behind the handle's queued work, the downloads of the parameters (positions and
// colours, 3 floats per vertex each), the loss word at d_loss and the kernels' fault word into
// host memory, without waiting: pinned destinations (hipHostMalloc) are complete once the
// stream has passed them.  The server queues them with a job's last steps, so its teardown
// reads the results without a device wait.
extern "C" __attribute__((visibility("hidden"))) drh_status drh_export_params(drh_handle h, float* pos, float* col,
                                                                              const void* d_loss, float* loss,
                                                                              uint32_t* fault);

// A served job's target images (`bytes` from `targets`) and loss word in buffers the handle
// owns (kept and regrown across jobs, freed with the handle): allocation ordered on the
// handle's stream and the upload staged, without waiting.
extern "C" __attribute__((visibility("hidden"))) drh_status drh_job_upload(drh_handle h, const void* targets,
                                                                           size_t bytes, void** d_targets,
                                                                           void** d_loss);
