// amr_device.h -- host-side owner of a device arena: layout, upload, launches, readback.
#ifndef HIPAMR_AMR_DEVICE_H
#define HIPAMR_AMR_DEVICE_H

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "amr_arena.h"
#include "amr_hip.h"

namespace hipamr {

// One mesh to place in an arena: its view (host memory of the caller), parameters and
// initial conserved state (nv * n_cells doubles) with time and counters.
struct MeshInput {
  const amr_mesh_view* view;
  amr_params params;
  const double* u;
  amr_state_meta meta;
};

struct MeshSlot {
  int cell_begin, n_cells, face_begin, n_faces, wall_begin, n_walls, nv;
};

// Validation of one input, mirroring the reference create checks
// (src/static_amr2d_cuda.cu:277-281).  Empty string when valid.
std::string validate_input(const MeshInput& in);

class Arena {
 public:
  virtual ~Arena() = default;
  virtual int precision() const = 0;
  virtual const std::vector<MeshSlot>& slots() const = 0;
  virtual uint64_t device_bytes() const = 0;
  virtual uint64_t geometry_bytes() const = 0;
  virtual uint64_t state_bytes() const = 0;
  virtual int launch_trial(hipStream_t s) = 0;
  virtual int launch_residual(hipStream_t s) = 0;
  virtual Control* control() = 0;
  virtual int n_meshes() const = 0;
  // Copy the state buffer `buffer` (0/1) of mesh m, or R0 when buffer == 2, into host doubles.
  virtual hipError_t read_state(int m, int buffer, double* out, hipStream_t s) = 0;
  // Write host doubles into state buffer `buffer` of mesh m.
  virtual hipError_t write_state(int m, int buffer, const double* in, hipStream_t s) = 0;
  virtual int n_cell_blocks() const = 0;
  virtual int n_face_blocks() const = 0;
  virtual int n_wall_blocks() const = 0;
  virtual int tile() const = 0;
};

// Build and upload an arena (precision 64 or 32, tile 64/128/256).  nullptr + error on failure.
std::unique_ptr<Arena> make_arena(const std::vector<MeshInput>& inputs, int precision, int tile, hipStream_t stream,
                                  std::string* error);

}  // namespace hipamr

#endif
