// timing.h -- per-kernel HIP-event timing on the launch stream.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "common.h"

namespace cube {

class KTimer {
 public:
  struct Rec {
    std::string name;
    hipEvent_t a, b;
    double bytes, flops;
  };
  struct Acc {
    int64_t calls = 0;
    double ms = 0, bytes = 0, flops = 0;
  };
  bool enabled = false;
  hipStream_t stream = nullptr;

  void begin(const char* name, double bytes, double flops) {
    if (!enabled) return;
    Rec r{name, get(), get(), bytes, flops};
    CUBE_HIP(hipEventRecord(r.a, stream));
    open_.push_back(r);
  }
  void end() {
    if (!enabled || open_.empty()) return;
    Rec r = open_.back();
    open_.pop_back();
    CUBE_HIP(hipEventRecord(r.b, stream));
    done_.push_back(r);
  }
  void collect() {
    if (done_.empty()) return;
    CUBE_HIP(hipStreamSynchronize(stream));
    for (auto& r : done_) {
      float ms = 0;
      CUBE_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      Acc& a = acc_[r.name];
      a.calls++;
      a.ms += ms;
      a.bytes += r.bytes;
      a.flops += r.flops;
      pool_.push_back(r.a);
      pool_.push_back(r.b);
    }
    done_.clear();
  }
  void reset() {
    collect();
    acc_.clear();
  }
  std::string report() {
    collect();
    std::string s;
    char line[512];
    for (auto& kv : acc_) {
      snprintf(line, sizeof(line), "%s %lld %.6f %.6e %.6e\n", kv.first.c_str(), (long long)kv.second.calls,
               kv.second.ms, kv.second.bytes, kv.second.flops);
      s += line;
    }
    return s;
  }
  ~KTimer() {
    for (auto e : pool_) (void)hipEventDestroy(e);
    for (auto& r : done_) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  }

 private:
  hipEvent_t get() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e;
    CUBE_HIP(hipEventCreate(&e));
    return e;
  }
  std::vector<Rec> open_, done_;
  std::vector<hipEvent_t> pool_;
  std::map<std::string, Acc> acc_;
};

struct KScope {
  KTimer* t;
  KScope(KTimer* t_, const char* name, double bytes, double flops = 0) : t(t_) { t->begin(name, bytes, flops); }
  ~KScope() { t->end(); }
};

}  // namespace cube
