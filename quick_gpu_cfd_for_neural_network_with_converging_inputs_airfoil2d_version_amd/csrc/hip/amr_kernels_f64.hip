// amr_kernels_f64.hip -- FP64 instantiation of the time-step kernels (IEEE division) and the
// precision-independent target launch.
#include "amr_kernels_impl.h"

namespace hipamr {

void launch_target(Control* ctrl, int n, double target, long long maxsteps, const double* targets,
                   const long long* maxsteps_each, hipStream_t s) {
  hipLaunchKernelGGL(k_target, dim3(blocks_of(n, 256)), dim3(256), 0, s, ctrl, n, target, maxsteps, targets,
                     maxsteps_each);
}

template int launch_trial<double>(const ArenaPtrs<double>&, hipStream_t, LaunchHook*);
template int launch_residual<double>(const ArenaPtrs<double>&, hipStream_t, LaunchHook*);
template int launch_prims<double>(const ArenaPtrs<double>&, hipStream_t, LaunchHook*);
template void launch_monitor<double>(const ArenaPtrs<double>&, hipStream_t);

}  // namespace hipamr
