// dress_capi_internal.h -- side-table helpers shared by the C ABI files.
#pragma once

#include <algorithm>
#include <cmath>

#include "dress_hip_internal.h"

namespace dh {

// device graph registered for g by dress_init_graph, or built now from g
DevGraph *device_graph(p_dress_graph_t g);
DevGraph *lookup(const dress_graph_t *g);
int default_precision();
void record_stats(DevGraph *G, const FitResult &r);
void record_lifted(int reruns, double device_ms, int64_t subgraphs, int64_t values);

}  // namespace dh
