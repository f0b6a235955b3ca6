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
// rows per workgroup of the LDS-staged node-gather matrix assembly, and its
// lanes (four per row)
constexpr int kGatherRows = 64;
constexpr int kGatherLanes = 4 * kGatherRows;

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

// One solve's buffers in a batched residual evaluation (layout shared with
// the host: 24 words of 8 bytes; see fp3d_batch_residual_eval).
struct EvalJob {
  const double *phi;
  double *grad, *q2n, *q2l, *rho;
  uint8_t *lim;
  double *gdotb;
  int32_t *up;
  double *nu, *rt, *se, *su, *terms, *R, *sR;
  const uint8_t *fmask;
  double *work;     // 3 * kReduceBlocks doubles
  double *out;      // [0] limited, [1] nu max, [2] nu > 0, [3] floored, [4] l2, [5] linf
  double m_inf, q2cap;
  int64_t reserved[4];
};
static_assert(sizeof(EvalJob) == 192, "EvalJob layout");

// the masked copy R -> sR and the norms of sR of every job (fp3d_linalg.hip)
int batch_masked_norms(int nb, const EvalJob *jobs, int64_t n, hipStream_t s);

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
