// amr_device.h -- host-side owner of a device arena: slot layout, upload, launches, readback.
//
// An arena is a fixed set of slots.  Each slot owns tile-aligned ranges of the global cell,
// face and wall index spaces and ranges of the ragged tables (cell->face CSR, gradient terms,
// SRD groups), all sized by the slot's capacity.  A mesh no larger than the capacity can be
// admitted into a slot at any time (its ranges are rewritten and uploaded, nothing else moves),
// which is what lets a serving loop refill slots while other meshes keep running.
#ifndef HIPAMR_AMR_DEVICE_H
#define HIPAMR_AMR_DEVICE_H

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "amr_arena.h"
#include "amr_hip.h"

namespace hipamr {

// One mesh to place in a slot: its view (host memory of the caller), parameters and initial
// conserved state (nv * n_cells doubles) with time and counters.
struct MeshInput {
  const amr_mesh_view* view;
  amr_params params;
  const double* u;
  amr_state_meta meta;
};

// Capacity of a slot in every index space (cells include at least one padding cell).
struct SlotCapacity {
  int cells, faces, walls, csr, terms, groups, members;
};

struct MeshSlot {
  int cell_begin, face_begin, wall_begin, csr_begin, term_begin, grp_begin, mem_begin;
  SlotCapacity cap;
  int n_cells, n_faces, n_walls, nv;  // the admitted mesh (n_cells == 0: empty slot)
};

// Kernels an arena's trial graph must contain (fixed when the graph is captured).
struct ArenaFeatures {
  int laminar, sa, small, visc_group;
};

// Validation of one input, mirroring the reference create checks
// (src/static_amr2d_cuda.cu:277-281).  Empty string when valid.
std::string validate_input(const MeshInput& in);
// Smallest slot capacity that holds the mesh at this tile.
SlotCapacity capacity_of(const amr_mesh_view& v, int tile);
// Kernels the mesh needs.
ArenaFeatures features_of(const amr_mesh_view& v);

class Arena {
 public:
  virtual ~Arena() = default;
  virtual int precision() const = 0;
  virtual const std::vector<MeshSlot>& slots() const = 0;
  virtual const ArenaFeatures& features() const = 0;
  virtual uint64_t device_bytes() const = 0;
  virtual uint64_t geometry_bytes() const = 0;
  virtual uint64_t state_bytes() const = 0;
  virtual int launch_trial(hipStream_t s, LaunchHook* hook = nullptr) = 0;
  virtual int launch_residual(hipStream_t s, LaunchHook* hook = nullptr) = 0;
  // Base primitives of every running slot whose host-written state has none yet; follows every
  // launch_target before trials are launched.
  virtual int launch_prims(hipStream_t s) = 0;
  // Threads per cell of the flux launches (1 or 2, at most 256 threads per block); launches
  // issued after the call use it.  False: not a valid spread at this tile.
  virtual bool set_flux_spread(int spread) = 0;
  virtual int flux_spread() const = 0;
  // Independent branches of the trial (slots m with m % n_branches() == j) and the trial of one
  // branch: its slots' kernels and commit only, for its own stream.
  virtual int n_branches() const = 0;
  virtual int launch_trial_branch(int j, hipStream_t s) = 0;
  virtual Control* control() = 0;
  // Device arrays of per-slot targets / step caps used by launch_target.
  virtual double* target_array() = 0;
  virtual long long* maxsteps_array() = 0;
  // Steady-monitor partials (after a residual pass): per cell block 6 doubles, per wall block 3.
  // frames: n_meshes x 6 host doubles (region xmin, xmax, ymin, ymax; moment point x, y).
  virtual hipError_t launch_monitor(hipStream_t s, const double* frames) = 0;
  virtual hipError_t read_monitor(std::vector<double>* cells, std::vector<double>* walls, hipStream_t s) = 0;
  virtual int n_meshes() const = 0;
  // Copy the state buffer `buffer` (0/1) of slot m, or R0 when buffer == 2, into host doubles.
  virtual hipError_t read_state(int m, int buffer, double* out, hipStream_t s) = 0;
  // Write host doubles into state buffer `buffer` of slot m.
  virtual hipError_t write_state(int m, int buffer, const double* in, hipStream_t s) = 0;
  // Place a mesh into slot m (must fit its capacity and the arena's kernels); its control
  // starts finished at the mesh's time until a target is set.  Empty string on success; an
  // invalid or oversized mesh leaves the slot as it was.
  virtual std::string admit(int m, const MeshInput& in, hipStream_t s) = 0;
  // Empty slot m (its blocks become no-ops).
  virtual std::string release(int m, hipStream_t s) = 0;
  // nutilde_inf of the mesh in slot m (scales the SA residual norm).
  virtual double sa_nu_inf(int m) const = 0;
  virtual int n_cell_blocks() const = 0;
  virtual int n_face_blocks() const = 0;
  virtual int n_wall_blocks() const = 0;
  virtual int tile() const = 0;
};

// Sum the monitor partials of slot m (fixed block order, long double): out = sum V|r_v| for
// v < 5, region volume, fx, fy, mz.
void sum_monitor(const Arena& arena, int m, const std::vector<double>& cells, const std::vector<double>& walls,
                 double* out9);

// Arena with one slot per input, each sized to its mesh, all admitted.  With `slab` the arena
// is laid out in that device memory (slab_bytes long; an error if it needs more) and allocates
// no device or pinned memory, staging through pageable host memory instead.
std::unique_ptr<Arena> make_arena(const std::vector<MeshInput>& inputs, int precision, int tile, hipStream_t stream,
                                  std::string* error, void* slab = nullptr, size_t slab_bytes = 0);
// Arena of empty slots with the given capacities and kernel set.
std::unique_ptr<Arena> make_slot_arena(const std::vector<SlotCapacity>& caps, const ArenaFeatures& features,
                                       int precision, int tile, hipStream_t stream, std::string* error);
// Device bytes of an arena with these slots (the single allocation make_slot_arena makes).
uint64_t arena_bytes(const std::vector<SlotCapacity>& caps, const ArenaFeatures& features, int precision, int tile);
// Host-only layout check of meshes placed one per slot (no device calls).
std::string check_inputs_layout(const std::vector<MeshInput>& inputs, int tile);

}  // namespace hipamr

#endif
