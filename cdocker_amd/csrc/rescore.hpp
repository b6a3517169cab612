// Grid-energy rescoring and rigid-body refinement of explicit poses.
#pragma once
#include <hip/hip_runtime.h>

namespace cdk {

// coords [npose][natom][3] (A); pch[a] = protein probe channel of atom a
// (-1: none), ech = elec channel; wv = sqrt(eps), wq = charge; gmin = grid
// origin.  energy [npose]; grad [npose][natom][3] (optional, kcal/mol/A);
// inside [npose] = atoms on the grid (optional).
hipError_t launch_rescore(hipStream_t st, int npose, int natom, const double* coords, const int* pch,
                          const float* wv, const float* wq, const float* grid, int N, int ech, const double* gmin,
                          double h, double* energy, double* grad, int* inside);
// Rigid-body descent in place on coords, `steps` trial moves per pose.
hipError_t launch_refine(hipStream_t st, int npose, int natom, double* coords, const int* pch, const float* wv,
                         const float* wq, const float* grid, int N, int ech, const double* gmin, double h, int steps,
                         double penalty, double* energy);

// out[(k*Y + j)*X + i] = in[(i*Y + j)*Z + k] as double.
hipError_t launch_grid_xfast(hipStream_t st, const float* in, int X, int Y, int Z, double* out);

}  // namespace cdk
