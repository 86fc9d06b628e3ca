// Host-side LM loop driver (see lm_plan.h).
#include <map>
#include <mutex>
#include <vector>

#include "lm_plan.h"

namespace ct {

namespace {

// Pinned host buffers of released drivers, kept for the next plan: a server
// creates and destroys a plan per request, and pinned allocation and release
// are slow calls that serialise with other threads' work.
std::mutex g_pinned_lock;
std::map<size_t, std::vector<void*>> g_pinned_free;

int pinned_take(size_t bytes, void** out) {
  {
    std::lock_guard<std::mutex> g(g_pinned_lock);
    std::vector<void*>& v = g_pinned_free[bytes];
    if (!v.empty()) {
      *out = v.back();
      v.pop_back();
      return CT_OK;
    }
  }
  CT_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
  return CT_OK;
}

void pinned_give(void* p, size_t bytes) {
  if (p == nullptr) return;
  std::lock_guard<std::mutex> g(g_pinned_lock);
  g_pinned_free[bytes].push_back(p);
}

}  // namespace

int lm_driver_setup(LMDriver& d, const ct_lm_options& o, hipStream_t stream, LMState* dev_state,
                    double* dev_records) {
  d.stream = stream;
  d.dev_state = dev_state;
  d.dev_records = dev_records;
  d.max_iter = o.max_num_iterations;
  d.loop_mode = o.loop_mode;
  int per = o.iterations_per_graph > 0 ? o.iterations_per_graph : 8;
  if (per > o.max_num_iterations + 2) per = o.max_num_iterations + 2;
  d.graph_iterations = per;
  CT_TRY(pinned_take(sizeof(LMState), reinterpret_cast<void**>(&d.host_state)));
  if (d.loop_mode == CT_LOOP_DEVICE) {
    CT_HIP(hipStreamCreateWithFlags(&d.capture_stream, hipStreamNonBlocking));
    CT_TRY(pinned_take(2 * sizeof(int32_t), reinterpret_cast<void**>(&d.host_done)));
    for (hipEvent_t& e : d.done_event) CT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  return CT_OK;
}

void lm_driver_release(LMDriver& d) {
  if (d.graph_exec) (void)hipGraphExecDestroy(d.graph_exec);
  if (d.graph) (void)hipGraphDestroy(d.graph);
  if (d.capture_stream) (void)hipStreamDestroy(d.capture_stream);
  pinned_give(d.host_state, sizeof(LMState));
  pinned_give(d.host_done, 2 * sizeof(int32_t));
  for (hipEvent_t& e : d.done_event) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  d.host_done = nullptr;
  d.graph_exec = nullptr;
  d.graph = nullptr;
  d.capture_stream = nullptr;
  d.host_state = nullptr;
}

int lm_driver_read_state(LMDriver& d, hipStream_t s) {
  CT_HIP(hipMemcpyAsync(d.host_state, d.dev_state, sizeof(LMState), hipMemcpyDeviceToHost, s));
  CT_HIP(hipStreamSynchronize(s));
  return CT_OK;
}

namespace {

int build_graph(LMDriver& d) {
  if (d.graph_exec) return CT_OK;
  CT_HIP(hipStreamBeginCapture(d.capture_stream, hipStreamCaptureModeThreadLocal));
  int st = CT_OK;
  for (int k = 0; k < d.graph_iterations && st == CT_OK; ++k) st = d.enqueue_iteration(d.capture_stream, false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(d.capture_stream, &g);
  if (st != CT_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) {
    set_error(std::string("hipStreamEndCapture failed: ") + hipGetErrorString(e));
    return CT_ERR_HIP;
  }
  d.graph = g;
  CT_HIP(hipGraphInstantiate(&d.graph_exec, g, nullptr, nullptr, 0));
  return CT_OK;
}

}  // namespace

int lm_driver_run(LMDriver& d, ct_lm_report* report, double* records) {
  const int max_passes = d.max_iter + 2;  // iterations 0..max + deferred checks
  hipStream_t s = d.stream;
  if (d.loop_mode == CT_LOOP_DEVICE) {
    // order the capture stream after the caller's stream and back
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    CT_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    CT_HIP(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    CT_HIP(hipEventRecord(ev_in, s));
    CT_HIP(hipStreamWaitEvent(d.capture_stream, ev_in, 0));
    int st = build_graph(d);
    if (st != CT_OK) {
      (void)hipEventDestroy(ev_in);
      (void)hipEventDestroy(ev_out);
      return st;
    }
    hipStream_t cs = d.capture_stream;
    CT_TRY(d.launch_init(cs));
    const int launches = (max_passes + d.graph_iterations - 1) / d.graph_iterations;
    const int32_t* dev_done = &d.dev_state->done;
    for (int k = 0; k < launches; ++k) {
      if (k >= 2) {
        // replay k - 2's flag (replay k - 1 keeps the device busy meanwhile)
        CT_HIP(hipEventSynchronize(d.done_event[k & 1]));
        if (d.host_done[k & 1]) break;
      }
      CT_HIP(hipGraphLaunch(d.graph_exec, cs));
      CT_HIP(hipMemcpyAsync(&d.host_done[k & 1], dev_done, sizeof(int32_t), hipMemcpyDeviceToHost, cs));
      CT_HIP(hipEventRecord(d.done_event[k & 1], cs));
    }
    CT_TRY(d.launch_finalize(cs));
    CT_HIP(hipEventRecord(ev_out, cs));
    CT_HIP(hipStreamWaitEvent(s, ev_out, 0));
    (void)hipEventDestroy(ev_in);
    (void)hipEventDestroy(ev_out);
    CT_TRY(lm_driver_read_state(d, s));
  } else {
    CT_TRY(d.launch_init(s));
    for (int k = 0; k < max_passes; ++k) {
      // pass 0 only evaluates; the last pass only runs the deferred checks of
      // the last accepted step and ends at the iteration limit
      CT_TRY(d.enqueue_iteration(s, k == 0 || k == max_passes - 1));
      CT_TRY(lm_driver_read_state(d, s));  // the one host sync of the iteration
      if (d.host_state->done) break;
    }
    CT_TRY(d.launch_finalize(s));
    CT_TRY(lm_driver_read_state(d, s));
  }
  const LMState& h = *d.host_state;
  if (report) {
    report->termination = h.termination;
    report->reason = h.reason;
    report->num_records = h.num_records;
    report->num_successful_steps = h.num_successful;
    report->num_unsuccessful_steps = h.num_unsuccessful;
    report->num_residual_evaluations = h.num_residual_evals;
    report->num_jacobian_evaluations = h.num_jacobian_evals;
    report->num_linear_solves = h.num_linear_solves;
    report->initial_cost = h.initial_cost;
    report->final_cost = h.best_cost;
    report->device_seconds = (double)(h.t_end - h.t0) * 1e-8;
  }
  if (records) {
    const size_t nrec = (size_t)(d.max_iter + 2) * CT_REC_FIELDS;
    CT_HIP(hipMemcpyAsync(records, d.dev_records, nrec * sizeof(double), hipMemcpyDeviceToHost, s));
    CT_HIP(hipStreamSynchronize(s));
  }
  if (!h.done) {
    set_error("LM loop did not terminate within the pass budget");
    return CT_ERR_HIP;
  }
  return CT_OK;
}

}  // namespace ct
