// The fused single-step and pair kernels for bf16_t field storage: one
// translation unit per storage type, so that the three compile in parallel
// (the entry points and their checks are in fused.hip).

#include "fused_kernels.h"

namespace bzh {

void launch_fused_bf16(hipStream_t s, const FusedParams& p, int nblocks) {
  launch_field<bf16_t>(s, p, nblocks);
}

void launch_pair_bf16(hipStream_t s, const PairParams& p, int nblocks) {
  launch_pair_field<bf16_t>(s, p, nblocks);
}

}  // namespace bzh
