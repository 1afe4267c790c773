// The fused single-step and pair kernels for float field storage: one
// translation unit per storage type, so that the three compile in parallel
// (the entry points and their checks are in fused.hip).

#include "fused_kernels.h"

namespace bzh {

void launch_fused_f32(hipStream_t s, const FusedParams& p, int nblocks) {
  launch_field<float>(s, p, nblocks);
}

void launch_pair_f32(hipStream_t s, const PairParams& p, int nblocks) {
  launch_pair_field<float>(s, p, nblocks);
}

}  // namespace bzh
