// mesher.h -- host-side preprocessing of the cut-cell time step: static 2:1 quadtree cut-cell
// mesh, least-squares/wall-donor viscous tables, SA wall distance, uniform initial state.
#ifndef HIPAMR_MESHER_H
#define HIPAMR_MESHER_H

#include <array>
#include <string>
#include <vector>

#include "amr2d_types.h"

namespace cfd {
namespace hipamr {

// Reference static_amr2d_build (src/static_amr2d.cpp:229-358).  0 ok, 1 error.
int build_mesh(const StaticAmrOptions2D& options, const std::vector<double>& px,
               const std::vector<double>& py, const std::vector<int>& patches,
               StaticAmrMesh2D* mesh, std::string* error);
// Reference static_amr2d_prepare_viscous + static_amr2d_prepare_sa
// (src/static_amr2d_viscous.cpp:39-76, src/static_amr2d_sa.cpp:5-15).
int prepare_viscous(StaticAmrMesh2D* mesh, const Params2D& params, std::string* error);
// Reference static_amr2d_viscous_ready (src/static_amr2d_viscous.cpp:34-38).
bool viscous_ready(const StaticAmrMesh2D& mesh);
bool transport_valid(const AmrTransport2D& t);
bool boundary_valid(const AmrBoundary2D& b, int legacy, int side);
// Reference static_amr2d_initialize (src/static_amr2d.cpp:360-367).
int initialize_state(const StaticAmrMesh2D& mesh, const Params2D& params,
                     const std::array<double, 4>& primitive, StaticAmrState2D* state);

}  // namespace hipamr
}  // namespace cfd

#endif
