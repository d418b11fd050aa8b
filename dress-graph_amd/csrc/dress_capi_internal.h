// dress_capi_internal.h -- side-table helpers shared by the C ABI files.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

#include "dress_hip_internal.h"

namespace dh {

// device graph registered for g by dress_init_graph, or built now from g
DevGraph *device_graph(p_dress_graph_t g);
DevGraph *lookup(const dress_graph_t *g);
int default_precision();
// precision used by this thread's calls that take none (lifted fits);
// -1 restores the process-wide setting
void set_thread_precision(int precision);
// cancellation flag watched by this thread's long host loops (the
// per-subgraph lifted path); nullptr: none
void set_thread_cancel(const std::atomic<bool> *flag);
bool thread_cancelled();
// device bytes per subgraph value while a delta/nabla call holds it: the
// value, the histogram assembly's keys, segments, run flags and positions,
// runs and the sort's temporary storage
constexpr int64_t LIFT_BYTES_PER_VALUE = 112;
// most values one chunk of a delta/nabla call keeps on the device; 0 clears
// the calling thread's limit
void set_thread_lifted_values(int64_t max_values);
int64_t lifted_chunk_values();
void record_stats(DevGraph *G, const FitResult &r);
void record_lifted(int reruns, double device_ms, int64_t subgraphs, int64_t values);
bool lds_eligible_size(int N, int E);
// dress_multi.hip: delta/nabla items fitted many at a time on the parent's device graph
void set_thread_grid_share(int percent);
void set_thread_max_slots(int slots);
int thread_max_slots();
void set_thread_pressure(const std::atomic<int> *p);
void yield_to_pressure(double max_seconds);
int thread_grid_share();
int multi_max_k();
int multi_slots(const DevGraph *G);
size_t multi_pass_bytes(const DevGraph *G);
bool multi_items(DevGraph *G, int kind, int k, const std::vector<int> &verts, const std::vector<int64_t> &live,
                 const std::vector<int64_t> &sel, const double *marks, int max_it, double eps, double *d_vals,
                 double *ms_rows);
// dress_masked.hip: delta/nabla items as masked fits of the parent's device graph
bool masked_items(p_dress_graph_t g, DevGraph *G, int kind, int k, const std::vector<int> &verts,
                  const std::vector<int64_t> &live, const std::vector<int64_t> &sel, const double *marks, int max_it,
                  double eps, int precision, double *d_vals, int64_t cap, int64_t *n_vals, double *ms_rows,
                  const std::vector<int64_t> &patch_off, const std::vector<int> &patch_e,
                  const std::vector<double> &patch_w, const std::vector<double> &patch_parent_w);

}  // namespace dh
