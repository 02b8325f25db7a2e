// pdem_host.hpp -- host-side helpers shared by the engine and the closure
// network: HIP error handling and the network object.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <atomic>
#include <string>

struct HipError {
  std::string msg;
};

#define HIPCHK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw HipError{std::string(#call) + ": " + hipGetErrorString(e_)};              \
  } while (0)

constexpr int TPB = 256;

inline unsigned blocks(int64_t n, int tpb = TPB) { return (unsigned)((n + tpb - 1) / tpb); }

// The closure network (pdem_mlp_* entry points): FP32 master parameters and
// Adam state, BF16 weight copies, normalisation constants, all on the
// device that was current at pdem_mlp_create.
struct pdem_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float* p = nullptr;        // master parameters
  float* m = nullptr;        // Adam first moment
  float* v = nullptr;        // Adam second moment
  short* wb = nullptr;       // BF16 W copies
  short* wtb = nullptr;      // BF16 Wt copies
  float* norm = nullptr;     // in_mean[16] in_scale[16] out_scale[4] out_weight[4]
  float* partial = nullptr;  // per-workgroup gradient partials
  int64_t partial_wg = 0;
  int64_t adam_step = 0;
  // engine contexts whose closure reads this network (their captured steps
  // hold its buffers): pdem_mlp_destroy refuses while it is nonzero
  mutable std::atomic<int> users{0};
  std::string err;
};
