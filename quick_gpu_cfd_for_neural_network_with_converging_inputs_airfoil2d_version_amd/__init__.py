"""MI355X-native (gfx950) explicit cut-cell finite-volume airfoil time step.

Host mesher + HIP time step behind the reference's static_amr2d_gpu.h backend API, with the C
ABI of include/amr_hip.h underneath.  The kernels are in csrc/hip (amr_kernels_impl.h), the
arena layout in csrc/hip/amr_arena.h and amr_device.hip.
"""
__version__ = "0.1.0"
