// The fused single-step and pair kernels for __half field storage: one
// translation unit per storage type, so that the three compile in parallel
// (the entry points and their checks are in fused.hip).

#include "fused_kernels.h"

namespace bzh {

void launch_fused_f16(hipStream_t s, const FusedParams& p, int nblocks) {
  launch_field<__half>(s, p, nblocks);
}

void launch_pair_f16(hipStream_t s, const PairParams& p, int nblocks) {
  launch_pair_field<__half>(s, p, nblocks);
}

}  // namespace bzh
