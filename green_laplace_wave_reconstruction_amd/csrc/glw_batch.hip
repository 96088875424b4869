// Batched evaluation of many plans in one HIP graph.
//
// glw_batch_create captures one execute call per item on `branches` forked
// streams (fork/join with events on a private origin stream) and instantiates
// the graph once; glw_batch_launch replays it with one hipGraphLaunch.  The
// plans' kernels clear their counters with fill kernels (no memset nodes) and
// take every table from device memory, so replays are independent.  The batch
// holds every plan it captured (glw_plan_hold) until it is destroyed, so no
// later call on those plans can move the scratch the graph replays into;
// destroy a batch before its plans.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "glw.h"
#include "glw_internal.hpp"

struct glw_batch_s {
  int device = 0;
  hipGraphExec_t exec = nullptr;
  std::vector<glw_plan*> held;  // distinct plans of the graph, each holding one hold count
  ~glw_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    for (glw_plan* p : held) (void)glw_plan_hold(p, 0);
  }
};

namespace {

struct CaptureStreams {
  hipStream_t origin = nullptr;
  std::vector<hipStream_t> br;
  std::vector<hipEvent_t> ev;  // fork, then one join per branch
  ~CaptureStreams() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (hipStream_t s : br) (void)hipStreamDestroy(s);
    if (origin) (void)hipStreamDestroy(origin);
  }
};

int check_items(const glw_batch_item* items, int n) {
  for (int i = 0; i < n; ++i) {
    const glw_batch_item& it = items[i];
    const std::string at = "batch item " + std::to_string(i) + ": ";
    if (!it.plan) return glw_report_error(GLW_EINVAL, at + "null plan");
    if (it.kind < GLW_KIND_LINEAR || it.kind > GLW_KIND_ORDER4) return glw_report_error(GLW_EINVAL, at + "bad kind");
    if (it.nt < 1) return glw_report_error(GLW_EINVAL, at + "nt must be positive");
    if (!it.s0 || !it.times || !it.a || (!it.b && it.kind != GLW_KIND_ORDER4))
      return glw_report_error(GLW_EINVAL, at + "null array");
    if (glw_plan_profiling(it.plan))
      return glw_report_error(GLW_EINVAL, at + "stage profiling is enabled on the plan (disable it before capture)");
  }
  return GLW_OK;
}

// Longest-processing-time assignment; items of one plan stay on one branch
// (they share the plan's workspace).
std::vector<int> assign_branches(const glw_batch_item* items, int n, int nb) {
  std::map<glw_plan*, double> cost;
  std::vector<glw_plan*> order;
  for (int i = 0; i < n; ++i) {
    if (!cost.count(items[i].plan)) order.push_back(items[i].plan);
    cost[items[i].plan] += glw_plan_cost(items[i].plan, items[i].kind, items[i].nt);
  }
  std::stable_sort(order.begin(), order.end(), [&](glw_plan* x, glw_plan* y) { return cost[x] > cost[y]; });
  std::vector<double> load(nb, 0.0);
  std::map<glw_plan*, int> where;
  for (glw_plan* p : order) {
    const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    where[p] = b;
    load[b] += cost[p];
  }
  std::vector<int> out(n);
  for (int i = 0; i < n; ++i) out[i] = where[items[i].plan];
  return out;
}

}  // namespace

extern "C" {

int glw_batch_create(const glw_batch_item* items, int32_t n, int32_t branches, glw_batch** batch) {
  if (!batch) return glw_report_error(GLW_EINVAL, "null batch pointer");
  *batch = nullptr;
  if (!items || n < 1) return glw_report_error(GLW_EINVAL, "empty batch");
  if (branches < 1 || branches > 16) return glw_report_error(GLW_EINVAL, "branches must be 1..16");
  int rc = check_items(items, n);
  if (rc) return rc;
  const int nb = std::min<int>(branches, n);
  const std::vector<int> where = assign_branches(items, n, nb);

  glw_batch_s* B = new glw_batch_s();
  auto bail = [&](int code) {
    delete B;
    return code;
  };
  if (hipGetDevice(&B->device) != hipSuccess) return bail(glw_report_error(GLW_EDEVICE, "no HIP device"));
  CaptureStreams cs;
  auto hip = [&](hipError_t e, const char* what) {
    return e == hipSuccess ? GLW_OK : glw_report_error(GLW_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
  };
  if ((rc = hip(hipStreamCreateWithFlags(&cs.origin, hipStreamNonBlocking), "hipStreamCreateWithFlags")))
    return bail(rc);
  cs.br.resize(nb, nullptr);
  cs.ev.resize(nb + 1, nullptr);
  for (int b = 0; b < nb; ++b)
    if ((rc = hip(hipStreamCreateWithFlags(&cs.br[b], hipStreamNonBlocking), "hipStreamCreateWithFlags")))
      return bail(rc);
  for (int e = 0; e <= nb; ++e)
    if ((rc = hip(hipEventCreateWithFlags(&cs.ev[e], hipEventDisableTiming), "hipEventCreateWithFlags")))
      return bail(rc);

  if ((rc = hip(hipStreamBeginCapture(cs.origin, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")))
    return bail(rc);
  int err = hip(hipEventRecord(cs.ev[0], cs.origin), "hipEventRecord");
  for (int b = 0; b < nb && !err; ++b) err = hip(hipStreamWaitEvent(cs.br[b], cs.ev[0], 0), "hipStreamWaitEvent");
  for (int i = 0; i < n && !err; ++i) {
    const glw_batch_item& it = items[i];
    err = glw_execute_kind(it.plan, it.kind, it.s0, it.times, it.nt, it.a, it.b, cs.br[where[i]]);
  }
  // join every branch even after an error, so the capture can end cleanly
  for (int b = 0; b < nb; ++b) {
    const int e1 = hip(hipEventRecord(cs.ev[b + 1], cs.br[b]), "hipEventRecord");
    const int e2 = hip(hipStreamWaitEvent(cs.origin, cs.ev[b + 1], 0), "hipStreamWaitEvent");
    if (!err) err = e1 ? e1 : e2;
  }
  hipGraph_t graph = nullptr;
  const int ec = hip(hipStreamEndCapture(cs.origin, &graph), "hipStreamEndCapture");
  if (!err) err = ec;
  if (!err) err = hip(hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  if (graph) (void)hipGraphDestroy(graph);
  if (err) return bail(err);
  for (int i = 0; i < n; ++i)
    if (std::find(B->held.begin(), B->held.end(), items[i].plan) == B->held.end()) {
      (void)glw_plan_hold(items[i].plan, 1);
      B->held.push_back(items[i].plan);
    }
  *batch = B;
  return GLW_OK;
}

int glw_batch_launch(glw_batch* batch, void* stream) {
  if (!batch || !batch->exec) return glw_report_error(GLW_EINVAL, "null batch");
  int dev = -1;
  GLW_HIP_CHECK(hipGetDevice(&dev));
  if (dev != batch->device) GLW_HIP_CHECK(hipSetDevice(batch->device));
  GLW_HIP_CHECK(hipGraphLaunch(batch->exec, (hipStream_t)stream));
  return GLW_OK;
}

int glw_batch_destroy(glw_batch* batch) {
  delete batch;
  return GLW_OK;
}

}  // extern "C"
