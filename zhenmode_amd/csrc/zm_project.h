// Column-divergence projection of the predictor velocity
// (dynamics/projection.py:64-127): conjugate gradients on the 2-D
// area-weighted Poisson problem whose operator is the column divergence of the
// layer-gated adjoint gradient, followed by the 3-D gradient correction and
// the bounded transport-checked refinements.  All solver arithmetic is float64.
#pragma once

#include <hip/hip_runtime.h>

#include "zm_common.h"

namespace zm {

// Taps of the 9-point column operator, stored [tap][j*nx+i] in this order:
// centre, east, west, east+2, west+2, north, south, north+2, south+2.
constexpr int kProjTaps = 9;
constexpr int kProjBlock = 256;
constexpr int kProjMaxBlocks = 1024;  // columns beyond 256 K are grid-strided

// Solver state kept in device memory between the launches of one projection.
// Scalars that a launch derives from the previous launch's block partials are
// written by block 0 only and read by later launches only.
struct ProjScalars {
  double gam[2];     // r . M r, double-buffered by iteration parity
  double atol2;      // a solve stops once r . r <= atol2
  double floor2;     // (tol |b_0|)^2, the absolute floor of the refinement solves
  double threshold;  // refinements run while the surface transport norm^2 exceeds it
  double tol2, ttol2;
  int maxiter;
  int done;          // the running solve has stopped
  int active;        // the running solve's correction is applied
  int its[4];        // iterations of the first solve and of each refinement
};

struct ProjGeo {
  int nx, ny, n2, nblocks;
  const double* coef;  // [kProjTaps][n2]
  const double* inv;   // [n2] inverse Jacobi diagonal, or ones
  const double* wet;   // [n2] surface wet mask
  const double* idx;   // 1/dx per row
  const double* idyc;  // (1/dy)/cos per row
  const double* cfn;   // 0.5 (cos_j + cos_{j+1 mod ny})
  const double* cfs;   // 0.5 (cos_j + cos_{j-1 mod ny})
  const double* area;  // dx dy per row
  const double* dzn;   // node thickness [nz]
};

struct ProjWork {
  double *x, *r, *p0, *p1, *Ap;
  double* part_rhs;  // [nblocks][4]: b.b, b.M b, transport norm^2 of the right-hand side launch
  double* part_a;    // [nblocks]: p.Ap
  double* part_b;    // [nblocks][2]: r.M r, r.r
  ProjScalars* sc;
};

// b = -(column divergence(u, v) / dtg) area, x = 0, r = b and the solve's
// scalars.  mode 0 starts a projection (threshold and floor from this right
// hand side); mode 1 starts a refinement of the corrected velocity, active
// only while its surface transport norm^2 exceeds the threshold.
template <typename real>
void launch_proj_rhs(const Geo<real>& g, const ProjGeo& pg, const ProjWork& w, const real* u, const real* v,
                     double dtg, int mode, hipStream_t s);

// maxiter iterations of preconditioned CG, each launch a no-op once the
// solve has stopped; mode as for launch_proj_rhs.
void launch_proj_cg(const ProjGeo& pg, const ProjWork& w, int maxiter, int mode, int slot, hipStream_t s);

// ud = us - dtg grad_c3d(x), vd likewise, where the solve is active (in place allowed).
template <typename real>
void launch_proj_correct(const Geo<real>& g, const ProjGeo& pg, const ProjWork& w, double dtg, const real* us,
                         const real* vs, real* ud, real* vd, hipStream_t s);

// Host: the nine taps [kProjTaps][n2] of the column operator and the
// diagonal of _column_projection_diagonal (projection.py:10-50).  nwet and
// wet2 are (j*nx+i); dx_row, cos_lat per row; dzn per level.
void proj_stencil_host(int nx, int ny, int nz, const int* nwet, const double* wet2, const double* dx_row, double dy,
                       const double* cos_lat, const double* dzn, double* coef, double* diag);

}  // namespace zm
