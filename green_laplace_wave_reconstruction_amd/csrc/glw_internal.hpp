// Internal helpers shared by the glw translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "glw.h"

// Record `msg` as the thread's last error (glw_last_error) and return `code`.
int glw_report_error(int code, const std::string& msg);

// Evaluate `kind` (GLW_KIND_*) of `plan` at `nt` instants on `st`: the body of
// the glw_*_execute entry points.  Launches only (capturable in a graph once
// the plan has run once eagerly with its current schedule).
int glw_execute_kind(glw_plan* plan, int kind, const double* s0, const double* times, int nt, double* a, double* b,
                     hipStream_t st);
// True while stage timing events are enabled on the plan.
bool glw_plan_profiling(const glw_plan* plan);
// Relative work of one evaluation (used to balance graph branches).
double glw_plan_cost(const glw_plan* plan, int kind, int nt);
// Free the split-row scratch (it is re-grown by the next eager execution).
int glw_plan_release_scratch(glw_plan* plan);

#define GLW_HIP_CHECK(call)                                                                          \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      return glw_report_error(GLW_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)
