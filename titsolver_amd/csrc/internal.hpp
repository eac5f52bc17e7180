// Internal interfaces shared by the translation units of libtit_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "device_common.hpp"

namespace tit_amd {

// Kernel timer: an event pair around each launch while enabled, aggregated by
// kernel name (the TIT_PROFILE_SECTION idea, per kernel instead of per scope).
class Profiler {
 public:
  bool enabled = false;
  void begin(hipStream_t s, const char* name);
  void end(hipStream_t s);
  // Synchronises, folds the recorded pairs into the totals.
  void collect();
  void clear();
  ~Profiler();

  struct Total {
    std::string name;
    uint64_t launches = 0;
    double total_ms = 0.0;
  };
  std::vector<Total> totals;

 private:
  struct Rec {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending_;
  std::vector<hipEvent_t> pool_;
  hipEvent_t get_event();
};

struct ProfScope {
  Profiler& p;
  hipStream_t s;
  ProfScope(Profiler& p_, hipStream_t s_, const char* name) : p(p_), s(s_) {
    if (p.enabled) p.begin(s, name);
  }
  ~ProfScope() {
    if (p.enabled) p.end(s);
  }
};

// Device state arrays (SoA, cell-sorted order).
struct DevState {
  double* x[3];
  double* v[3];
  double* rho;
  float2* aux;  // per particle (p / rho^2, cs) of rho, FP32
  float4* pv;   // per particle (v (FP32, z = 0 in 2D), rho - rho0): what the
                // pair terms read of a neighbour besides its position
  float* vol;   // 3D: per particle m / rho (FP32), written with per-particle
                // masses only (in 2D pv.z holds it); null in 2D
};

// ---- radix sort (radix_sort.hip) -------------------------------------------
struct SortScratch {
  uint32_t* hist = nullptr;  // 2048 * max_tiles (digits of up to 11 bits)
  uint32_t* tot = nullptr;   // 2048
  uint32_t max_tiles = 0;
};
uint32_t radix_sort_tiles(uint32_t n);
// Stable LSD sort of (keys, values) by the low `key_bits` bits of the keys.
// The first pass takes values = index (vals is not read).  On return the
// sorted pairs are in keys / vals (the alt arrays are scratch).  Nothing runs
// on the device while ctl[0] is 0.
void radix_sort_pairs(hipStream_t stream, Profiler& prof, uint32_t n,
                      int key_bits, uint32_t* keys, uint32_t* keys_alt,
                      uint32_t* vals, uint32_t* vals_alt,
                      const SortScratch& scratch, const uint32_t* ctl);

}  // namespace tit_amd
