// FP64 instantiations of the SMD time-step kernel (batched and plain
// one-instance-per-launch forms), compiled apart from the FP32 ones.
#include "smd_kernel.h"

namespace hsmd {

void smd_launch_f64(const SmdLaunch &c, StepArgs a) {
  if (c.wide) {
    smd_launch_wide<double>(c, a);
    return;
  }
  switch (c.spt) {
    case 1: smd_launch_zc<double, 1>(c, a); break;
    case 2: smd_launch_zc<double, 2>(c, a); break;
    default: smd_launch_zc<double, 4>(c, a); break;
  }
}

}  // namespace hsmd
