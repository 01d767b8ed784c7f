// Per-kernel-class timing registry behind hps_prof_* (see hps_prof.h).
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hps_prof.h"
#include "../../include/hps_mi355x.h"

namespace hps {
namespace prof {

namespace {
struct Rec {
  hipEvent_t e0, e1;
  double flops, bytes;
};
struct Cls {
  std::vector<Rec> recs;
};
std::mutex mu;
bool on = false;
std::map<std::string, Cls> reg;
std::vector<std::string> order;

hipEvent_t new_event() {
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}
}  // namespace

bool enabled() { return on; }

Scope::Scope(hipStream_t s, const char* name, double flops, double bytes)
    : s_(s), name_(name), flops_(flops), bytes_(bytes) {
  if (!on) return;
  e0_ = new_event();
  hipEventRecord(e0_, s_);
}

Scope::~Scope() {
  if (!e0_) return;
  hipEvent_t e1 = new_event();
  hipEventRecord(e1, s_);
  std::lock_guard<std::mutex> g(mu);
  auto it = reg.find(name_);
  if (it == reg.end()) {
    order.push_back(name_);
    it = reg.emplace(name_, Cls{}).first;
  }
  it->second.recs.push_back(Rec{e0_, e1, flops_, bytes_});
}

}  // namespace prof
}  // namespace hps

using namespace hps::prof;

int hps_prof_enable(int enable) {
  std::lock_guard<std::mutex> g(mu);
  on = enable != 0;
  return HPS_OK;
}

int hps_prof_reset(void) {
  std::lock_guard<std::mutex> g(mu);
  for (auto& kv : reg)
    for (auto& r : kv.second.recs) {
      hipEventDestroy(r.e0);
      hipEventDestroy(r.e1);
    }
  reg.clear();
  order.clear();
  return HPS_OK;
}

int hps_prof_count(void) {
  std::lock_guard<std::mutex> g(mu);
  return (int)order.size();
}

int hps_prof_query(int idx, char* name, int name_len, long* launches, double* ms, double* flops, double* bytes) {
  std::lock_guard<std::mutex> g(mu);
  if (idx < 0 || idx >= (int)order.size()) return HPS_ERR_ARG;
  const std::string& nm = order[idx];
  Cls& c = reg[nm];
  double tot = 0, fl = 0, by = 0;
  for (auto& r : c.recs) {
    hipEventSynchronize(r.e1);
    float t = 0;
    hipEventElapsedTime(&t, r.e0, r.e1);
    tot += t;
    fl += r.flops;
    by += r.bytes;
  }
  if (name && name_len > 0) {
    std::strncpy(name, nm.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (launches) *launches = (long)c.recs.size();
  if (ms) *ms = tot;
  if (flops) *flops = fl;
  if (bytes) *bytes = by;
  return HPS_OK;
}
