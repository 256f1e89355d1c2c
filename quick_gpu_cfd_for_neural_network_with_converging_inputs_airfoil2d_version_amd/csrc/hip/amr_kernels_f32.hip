// amr_kernels_f32.hip -- FP32 instantiation of the time-step kernels.  This unit is compiled
// with -fno-hip-fp32-correctly-rounded-divide-sqrt: FP32 quotients and square roots are about
// 1 ulp (reciprocal and multiply) instead of the scaled IEEE sequences, which were a large part
// of the face loop's ALU time.  The FP64 state, the RK update and the dt control stay FP64.
#include "amr_kernels_impl.h"

namespace hipamr {

template int launch_trial<float>(const ArenaPtrs<float>&, hipStream_t, LaunchHook*);
template int launch_residual<float>(const ArenaPtrs<float>&, hipStream_t, LaunchHook*);
template int launch_prims<float>(const ArenaPtrs<float>&, hipStream_t, LaunchHook*);
template void launch_monitor<float>(const ArenaPtrs<float>&, hipStream_t);

}  // namespace hipamr
