// Internal helpers shared by the HIP translation units of libfp3d_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/fp3d.h"

namespace fp3d {

void set_error(const std::string &msg);

// Launch geometry: 256 work-items = 4 wave64 wavefronts per workgroup.
constexpr int kBlock = 256;
// rows per workgroup of the LDS-staged node-gather matrix assembly
constexpr int kGatherRows = 128;

inline unsigned grid_for(int64_t n, int per_block = kBlock) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

int check_launch(const char *what);

// Workgroups are dealt to the 8 XCDs round-robin; xcd_block() maps them so
// that each XCD walks one contiguous range of row blocks, keeping the
// elements shared by neighbouring rows in that XCD's L2.  The grid is padded
// to a multiple of 8 (xcd_grid) so the mapping is a bijection.
__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nb_pad) {
  int64_t per = nb_pad >> 3;
  return (bid & 7) * per + (bid >> 3);
}
inline unsigned xcd_grid(int64_t n_blocks) {
  return (unsigned)(((n_blocks + 7) / 8) * 8);
}

// Blocks used by the deterministic two-pass reductions; the partial-sum
// layout depends only on n, never on timing, so results are bitwise
// repeatable.
constexpr int kReduceBlocks = 1024;

}  // namespace fp3d

#define FP3D_REQUIRE(cond, msg)        \
  do {                                 \
    if (!(cond)) {                     \
      fp3d::set_error(msg);            \
      return FP3D_EINVAL;              \
    }                                  \
  } while (0)

#define FP3D_HIP(call)                                                   \
  do {                                                                   \
    hipError_t _e = (call);                                              \
    if (_e != hipSuccess) {                                              \
      fp3d::set_error(std::string(#call) + ": " + hipGetErrorString(_e)); \
      return FP3D_EHIP;                                                  \
    }                                                                    \
  } while (0)

#define FP3D_STREAM(s) ((hipStream_t)(s))

// element connectivity rows are loaded as one int4 (16 bytes)
#define FP3D_ALIGNED16(p) ((((uintptr_t)(p)) & 15u) == 0)
