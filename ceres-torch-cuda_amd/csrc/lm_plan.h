// Host-side driver of a device-resident LM solve, shared by the solver plans.
// A plan supplies how to enqueue one LM iteration, the init and finalize
// launches, and where its LMState and records live; the driver runs the loop
// host-driven (one synchronisation per iteration) or device-driven (graph
// replays of iterations_per_graph iterations; kernels return at once after
// the LM has finished).  The device loop keeps one replay queued ahead of the
// one it waits for: after each replay the state's done flag is copied to
// pinned memory behind an event, and the host queues replay k only after the
// flag of replay k - 2 has arrived (and is unset), so at most one replay runs
// past the end of the solve.
#pragma once

#include <functional>

#include "ba_device.h"

namespace ct {

struct LMDriver {
  hipStream_t stream = nullptr;          // the caller's stream
  hipStream_t capture_stream = nullptr;  // device loop: graph capture / replay stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_iterations = 0;
  int max_iter = 0;
  int loop_mode = CT_LOOP_HOST;
  LMState* host_state = nullptr;  // pinned
  int32_t* host_done = nullptr;   // pinned [2]: done flag after the replays of even / odd index
  hipEvent_t done_event[2] = {nullptr, nullptr};
  LMState* dev_state = nullptr;
  double* dev_records = nullptr;
  std::function<int(hipStream_t, bool)> enqueue_iteration;  // (stream, evaluation only)
  std::function<int(hipStream_t)> launch_init;
  std::function<int(hipStream_t)> launch_finalize;
};

int lm_driver_setup(LMDriver& d, const ct_lm_options& o, hipStream_t stream, LMState* dev_state,
                    double* dev_records);
int lm_driver_read_state(LMDriver& d, hipStream_t s);
int lm_driver_run(LMDriver& d, ct_lm_report* report, double* records);
void lm_driver_release(LMDriver& d);

}  // namespace ct
