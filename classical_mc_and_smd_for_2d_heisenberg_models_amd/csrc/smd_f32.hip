// FP32 instantiations of the SMD time-step kernel (batched form), compiled
// apart from the host code and the FP64 ones so the build runs them in
// parallel.
#include "smd_kernel.h"

namespace hsmd {

void smd_launch_f32(const SmdLaunch &c, StepArgs a) {
  if (c.wide) {
    smd_launch_wide<float>(c, a);
    return;
  }
  switch (c.spt) {
    case 1: smd_launch_zc<float, 1>(c, a); break;
    case 2: smd_launch_zc<float, 2>(c, a); break;
    case 4: smd_launch_zc<float, 4>(c, a); break;
    default: smd_launch_zc<float, 8>(c, a); break;
  }
}

}  // namespace hsmd
