// drh_internal.h -- entry points the library uses between its own translation units; not
// part of the C ABI (hidden from the shared library's exports).
#pragma once
#include "../../include/drh.h"

// Upload the cameras and the edge-sample table of a job's fused steps while the handle's
// stream holds no kernels yet, so the first step finds them in place and does not wait for
// the commit's kernels (the server calls it before drh_commit).
extern "C" __attribute__((visibility("hidden"))) drh_status drh_prepare_steps(drh_handle h, const drh_camera* cams,
                                                                              int views, int edge_samples);
