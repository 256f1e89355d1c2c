// static_amr2d_gpu.h -- the reference's GPU backend API, implemented for MI355X (gfx950).
//
// Same declarations as the reference's src/static_amr2d_gpu.h:6-25, so a caller compiled
// against either header links libamr_hip.so unchanged:
//   static_amr_gpu_create    replaces src/static_amr2d_gpu.h:17-18 (src/static_amr2d_cuda.cu:274-325)
//   static_amr_gpu_advance   replaces src/static_amr2d_gpu.h:19-21 (src/static_amr2d_cuda.cu:327-353)
//   static_amr_gpu_download  replaces src/static_amr2d_gpu.h:22    (src/static_amr2d_cuda.cu:355-363)
//   static_amr_gpu_stats     replaces src/static_amr2d_gpu.h:23    (src/static_amr2d_cuda.cu:364)
//   static_amr_gpu_destroy   replaces src/static_amr2d_gpu.h:24    (src/static_amr2d_cuda.cu:365)
// Return codes: 0 ok, 1 error (message in *error; advance keeps the last accepted state),
// 2 wall budget expired with a resumable state.  advance updates time/steps/rejected only;
// download fills u.  The HIP kernels are always the fused layout, so AmrGpuOptions::fused and
// ::sa_fused are accepted and ignored; the state is FP64 (FP32 is available through
// amr_hip_create's precision option).
#ifndef HIPAMR_STATIC_AMR2D_GPU_H
#define HIPAMR_STATIC_AMR2D_GPU_H

#include <cstdint>
#include <string>

#include "amr2d_types.h"

namespace cfd {
struct StaticAmrGpu;
struct AmrGpuOptions {
  int batch_trials = 32;
  bool graph = true;
  bool fused = false;
  bool event_timing = false;
  bool sa_fused = false;
};
struct AmrGpuStats {
  uint64_t allocated_bytes = 0, pinned_bytes = 0, geometry_h2d_bytes = 0, state_h2d_bytes = 0;
  uint64_t control_h2d_bytes = 0;
  uint64_t control_d2h_bytes = 0, state_d2h_bytes = 0, status_polls = 0, full_downloads = 0;
  double setup_seconds = 0, advance_seconds = 0, download_seconds = 0;
  double advance_event_seconds = 0;
  uint64_t event_batches = 0, graph_nodes_per_trial = 0;
};
int static_amr_gpu_create(const StaticAmrMesh2D&, const Params2D&, const StaticAmrState2D&, const AmrGpuOptions&,
                          StaticAmrGpu**, std::string*);
int static_amr_gpu_advance(StaticAmrGpu*, double tend, long maxsteps, StaticAmrState2D*, std::string*,
                           double wall_budget_seconds = 0);
int static_amr_gpu_download(StaticAmrGpu*, StaticAmrState2D*, std::string*);
AmrGpuStats static_amr_gpu_stats(const StaticAmrGpu*);
void static_amr_gpu_destroy(StaticAmrGpu*);
}  // namespace cfd

#endif
