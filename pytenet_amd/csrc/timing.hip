// Live launch timing of the complex products (ptn_timing_begin / ptn_timing_end).
//
// While timing is on, every complex-product kernel launch is bracketed by two
// HIP events recorded on the stream it is launched on (the device clock of
// that stream, nothing else), together with the algorithmic flops and bytes
// of the launch. ptn_timing_end waits for the events and sums, per kernel
// instance, the launch count, the device time and the flops/bytes. Event
// records cost a few microseconds of host time per launch and do not
// synchronise anything, so a sweep can be timed with it switched on.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <string.h>
#include <vector>
#include "common.h"

namespace ptn {

namespace {
struct Rec {
  const char* name;
  double flop, bytes;
};
std::mutex g_mu;
bool g_on = false;
int64_t g_dropped = 0;
std::vector<hipEvent_t> g_events;   // two per record, reused across timing windows
std::vector<Rec> g_recs;
constexpr size_t MAX_RECS = 1u << 18;
}  // namespace

int timing_pre(hipStream_t s) {
  if (!g_on || noalloc_mode()) return -1;
  // launches recorded into a graph (stream capture) are not timed per launch
  hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cst) != hipSuccess || cst != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_on) return -1;
  const size_t idx = g_recs.size();
  if (idx >= MAX_RECS) { ++g_dropped; return -1; }
  while (g_events.size() < 2 * (idx + 1)) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { ++g_dropped; return -1; }
    g_events.push_back(e);
  }
  if (hipEventRecord(g_events[2 * idx], s) != hipSuccess) { ++g_dropped; return -1; }
  g_recs.push_back({nullptr, 0.0, 0.0});
  return (int)idx;
}

void timing_post(int slot, hipStream_t s, const char* name, double flop, double bytes) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if ((size_t)slot >= g_recs.size()) return;
  if (hipEventRecord(g_events[2 * (size_t)slot + 1], s) != hipSuccess) {
    g_recs[slot].name = nullptr;   // unusable pair: left out of the sums
    ++g_dropped;
    return;
  }
  g_recs[slot] = {name, flop, bytes};
}

}  // namespace ptn

using namespace ptn;

extern "C" {

int ptn_timing_begin(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_recs.clear();
  g_dropped = 0;
  g_on = true;
  return PTN_OK;
}

int ptn_timing_end(ptn_kernel_timing* out, int max_out, int* n_out) {
  std::vector<Rec> recs;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = false;
    recs.swap(g_recs);
  }
  struct Agg { int64_t n = 0; double ms = 0, flop = 0, bytes = 0; };
  std::map<std::string, Agg> agg;
  int st = PTN_OK;
  for (size_t i = 0; i < recs.size(); ++i) {
    if (!recs[i].name) continue;
    hipEvent_t e0 = g_events[2 * i], e1 = g_events[2 * i + 1];
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) { ptn_set_last_hip_error(e); st = PTN_ERR_HIP; continue; }
    Agg& a = agg[recs[i].name];
    a.n += 1;
    a.ms += ms;
    a.flop += recs[i].flop;
    a.bytes += recs[i].bytes;
  }
  std::vector<std::pair<std::string, Agg>> rows(agg.begin(), agg.end());
  std::sort(rows.begin(), rows.end(),
            [](const std::pair<std::string, Agg>& x, const std::pair<std::string, Agg>& y) {
              return x.second.ms > y.second.ms;
            });
  int n = 0;
  for (const auto& r : rows) {
    if (n >= max_out || !out) break;
    ptn_kernel_timing& o = out[n++];
    memset(o.name, 0, sizeof(o.name));
    strncpy(o.name, r.first.c_str(), sizeof(o.name) - 1);
    o.launches = r.second.n;
    o.total_ms = r.second.ms;
    o.flop = r.second.flop;
    o.bytes = r.second.bytes;
  }
  if (n_out) *n_out = n;
  return st;
}

}  // extern "C"
