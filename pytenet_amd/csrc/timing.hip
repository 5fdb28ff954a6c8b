// Live launch timing of the complex products and the Jacobi SVD kernels
// (ptn_timing_begin / ptn_timing_end).
//
// While timing is on, every such kernel is launched with a start and a stop
// HIP event attached to its own dispatch (hipExtLaunchKernel), so the events
// carry the kernel's begin and end on the device clock of its stream, together
// with the algorithmic flops and bytes of the launch. ptn_timing_end waits for the events and sums, per kernel
// instance, the launch count, the device time and the flops/bytes. Event
// records cost a few microseconds of host time per launch and do not
// synchronise anything, so a sweep can be timed with it switched on (beyond
// 32768 pending launches the oldest are read, long after they completed).
#include <algorithm>
#include <deque>
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
  hipEvent_t e0, e1;
  int seq;
};
struct Agg {
  int64_t n = 0;
  double ms = 0, flop = 0, bytes = 0;
};
std::mutex g_mu;
bool g_on = false;
int g_err = PTN_OK;
int g_seq = 0;
std::vector<hipEvent_t> g_free;      // recycled events
std::deque<Rec> g_pending;           // launches whose events are not read yet
std::map<std::string, Agg> g_agg;    // folded records of the current window
// pending records beyond this many are folded (their events waited for and
// read, oldest first), so a long window keeps a bounded number of events
constexpr size_t MAX_PENDING = 1u << 15;

hipEvent_t take_event() {
  if (!g_free.empty()) {
    hipEvent_t e = g_free.back();
    g_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// read the oldest pending record into the aggregates (caller holds g_mu)
void fold_front() {
  Rec r = g_pending.front();
  g_pending.pop_front();
  if (r.name && r.e1) {
    hipError_t e = hipEventSynchronize(r.e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.e0, r.e1);
    if (e != hipSuccess) {
      ptn_set_last_hip_error(e);
      g_err = PTN_ERR_HIP;
    } else {
      Agg& a = g_agg[r.name];
      a.n += 1;
      a.ms += ms;
      a.flop += r.flop;
      a.bytes += r.bytes;
    }
  }
  if (r.e0) g_free.push_back(r.e0);
  if (r.e1) g_free.push_back(r.e1);
}
}  // namespace

int timing_acquire(hipStream_t s, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  if (!g_on || noalloc_mode()) return -1;
  // launches recorded into a graph (stream capture) are not timed per launch
  hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cst) != hipSuccess || cst != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_on) return -1;
  while (g_pending.size() >= MAX_PENDING) fold_front();
  Rec r{nullptr, 0.0, 0.0, take_event(), take_event(), g_seq};
  if (!r.e0 || !r.e1) {
    if (r.e0) g_free.push_back(r.e0);
    if (r.e1) g_free.push_back(r.e1);
    return -1;
  }
  g_seq = (g_seq + 1) & 0x3fffffff;
  g_pending.push_back(r);
  *e0 = r.e0;
  *e1 = r.e1;
  return r.seq;
}

void timing_commit(int slot, const char* name, double flop, double bytes) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  // the record of this launch: normally the newest (searched from the back)
  for (auto it = g_pending.rbegin(); it != g_pending.rend(); ++it) {
    if (it->seq != slot) continue;
    it->name = name;
    it->flop = flop;
    it->bytes = bytes;
    return;
  }
}

}  // namespace ptn

using namespace ptn;

extern "C" {

int ptn_timing_begin(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  while (!g_pending.empty()) fold_front();
  g_agg.clear();
  g_err = PTN_OK;
  g_on = true;
  return PTN_OK;
}

int ptn_timing_end(ptn_kernel_timing* out, int max_out, int* n_out) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = false;
  while (!g_pending.empty()) fold_front();
  std::vector<std::pair<std::string, Agg>> rows(g_agg.begin(), g_agg.end());
  g_agg.clear();
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
  const int st = g_err;
  g_err = PTN_OK;
  return st;
}

}  // extern "C"
