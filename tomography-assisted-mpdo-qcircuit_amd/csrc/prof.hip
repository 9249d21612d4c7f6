#include "prof.h"

#include <array>
#include <map>
#include <mutex>
#include <vector>

#include "mpdo.h"
#include "mpdo_debug.h"

namespace mpdo {
namespace prof {
namespace {

struct Rec {
  int cls;
  hipEvent_t e0, e1;
  double flops, bytes;
  bool has_shape;
  std::array<int, 5> shape;
};
struct ShapeAcc {
  long long calls = 0;
  double ms = 0, flops = 0;
};
struct Acc {
  long long launches = 0;
  double ms = 0, flops = 0, bytes = 0;
};

std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_free;
Acc g_acc[NCLS];
double g_exec_flops[NCLS], g_exec_bytes[NCLS];
std::vector<int> g_eig;   // (n, batch, sweeps) triples
std::map<std::pair<int, std::array<int, 5>>, ShapeAcc> g_shapes;

// fold finished records into the totals (waits for the last one) and recycle their events
void drain_locked() {
  if (g_recs.empty()) return;
  (void)hipEventSynchronize(g_recs.back().e1);   // profiling only: a failed wait shows as a missing time
  for (const Rec& r : g_recs) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) {
      Acc& a = g_acc[r.cls];
      a.launches += 1;
      a.ms += ms;
      a.flops += r.flops;
      a.bytes += r.bytes;
      if (r.has_shape) {
        ShapeAcc& sa = g_shapes[{r.cls, r.shape}];
        sa.calls += 1;
        sa.ms += ms;
        sa.flops += r.flops;
      }
    }
    g_free.push_back(r.e0);
    g_free.push_back(r.e1);
  }
  g_recs.clear();
}

hipEvent_t take_locked() {
  if (!g_free.empty()) {
    hipEvent_t e = g_free.back();
    g_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

}  // namespace

bool enabled() { return g_on; }

void eig_log(int n, int batch, int sweeps) {
  if (!g_on) return;
  std::lock_guard<std::mutex> lk(g_mu);
  g_eig.push_back(n);
  g_eig.push_back(batch);
  g_eig.push_back(sweeps);
}

void add_executed(int cls, double flops, double bytes) {
  if (!g_on || cls < 0 || cls >= NCLS) return;
  std::lock_guard<std::mutex> lk(g_mu);
  g_exec_flops[cls] += flops;
  g_exec_bytes[cls] += bytes;
}

Scope::Scope(int cls, hipStream_t s, double flops, double bytes, const int* shape)
    : cls_(cls), s_(s), flops_(flops), bytes_(bytes) {
  if (!g_on) return;
  if (shape) {
    has_shape_ = true;
    for (int i = 0; i < 5; ++i) shape_[i] = shape[i];
  }
  std::lock_guard<std::mutex> lk(g_mu);
  e0_ = take_locked();
  e1_ = take_locked();
  (void)hipEventRecord(e0_, s_);
}

Scope::~Scope() {
  if (!e0_) return;
  std::lock_guard<std::mutex> lk(g_mu);
  (void)hipEventRecord(e1_, s_);
  std::array<int, 5> sh{shape_[0], shape_[1], shape_[2], shape_[3], shape_[4]};
  g_recs.push_back(Rec{cls_, e0_, e1_, flops_, bytes_, has_shape_, sh});
  if (g_recs.size() >= 4096) drain_locked();
}

}  // namespace prof
}  // namespace mpdo

extern "C" {

mpdo_status mpdo_debug_prof_enable(int on) {
  using namespace mpdo::prof;
  std::lock_guard<std::mutex> lk(g_mu);
  drain_locked();
  if (on) {
    for (auto& a : g_acc) a = Acc{};
    for (int c = 0; c < NCLS; ++c) g_exec_flops[c] = g_exec_bytes[c] = 0;
    g_eig.clear();
    g_shapes.clear();
  }
  g_on = on != 0;
  return MPDO_OK;
}

mpdo_status mpdo_debug_prof_read(int cls, long long* launches, double* ms, double* flops, double* bytes) {
  using namespace mpdo::prof;
  if (cls < 0 || cls >= NCLS) return MPDO_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  drain_locked();
  const Acc& a = g_acc[cls];
  if (launches) *launches = a.launches;
  if (ms) *ms = a.ms;
  if (flops) *flops = a.flops;
  if (bytes) *bytes = a.bytes;
  return MPDO_OK;
}

mpdo_status mpdo_debug_prof_read_executed(int cls, double* flops, double* bytes) {
  using namespace mpdo::prof;
  if (cls < 0 || cls >= NCLS) return MPDO_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  if (flops) *flops = g_exec_flops[cls];
  if (bytes) *bytes = g_exec_bytes[cls];
  return MPDO_OK;
}

int mpdo_debug_prof_shapes(int cls, int* shapes, double* stats, int cap) {
  using namespace mpdo::prof;
  std::lock_guard<std::mutex> lk(g_mu);
  drain_locked();
  int n = 0;
  for (const auto& kv : g_shapes) {
    if (kv.first.first != cls) continue;
    if (n < cap) {
      for (int i = 0; i < 5; ++i) shapes[5 * n + i] = kv.first.second[i];
      stats[3 * n] = (double)kv.second.calls;
      stats[3 * n + 1] = kv.second.ms;
      stats[3 * n + 2] = kv.second.flops;
    }
    ++n;
  }
  return n;
}

int mpdo_debug_prof_eig_log(int* out, int cap) {
  using namespace mpdo::prof;
  std::lock_guard<std::mutex> lk(g_mu);
  const int n = (int)g_eig.size() / 3;
  for (int i = 0; i < 3 * n && i < cap; ++i) out[i] = g_eig[i];
  return n;
}

}  // extern "C"
