// Host-side (CPU) setup routines of the library: geometry, inflow states,
// input-file parsers and the time-loop driver.
#pragma once

namespace ostc_host {

// Mesh::init_ghost_grid face part (src/Mesh.cpp:789-857) on (ngg+1)-padded arrays
void ghost_faces(int mx, int my, int mz, int ngg, double *x, double *y, double *z);
// Mesh::init_ghost_grid edge/corner part (src/Mesh.cpp:864-937)
void ghost_edges_corners(int mx, int my, int mz, int ngg, double *x, double *y, double *z);
// Block::compute_jac_metric (src/Mesh.cpp:259-521); -1 on a bad interior point
int jac_metric(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z, double *jac,
               double *metric, int *n_neg);

}  // namespace ostc_host
