// Column-divergence projection (dynamics/projection.py:64-127) on the device.
//
// The operator A(psi) = -column_divergence(grad_c3d(psi)) * area couples each
// column with its neighbours at distance one and two along x and along y, so
// it is held as nine float64 taps per column, built once on the host from the
// wet levels and the metric.  One CG iteration is two launches: k_proj_A
// (stopping test, beta, p = M r + beta p, Ap and the p.Ap partials) and
// k_proj_B (alpha, x, r and the r.M r, r.r partials).  Each launch folds the
// previous launch's per-block partials itself, every block in the same fixed
// order, so every block holds bit-identical scalars, results do not depend on
// block scheduling, and no fence or atomic is needed inside a launch.  Once a
// solve has stopped the remaining launches return at once, which keeps the
// launch sequence fixed for hipGraph capture.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "zm_project.h"

namespace zm {

// block-wide sums of NV values (result valid in every thread)
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
  for (int q = 0; q < NV; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
    for (int q = 0; q < NV; ++q) red[w * NV + q] = v[q];
  __syncthreads();
  for (int q = 0; q < NV; ++q) v[q] = (red[q] + red[NV + q]) + (red[2 * NV + q] + red[3 * NV + q]);
}

// sums of the previous launch's block partials [nb][stride], fixed order
template <int NV>
__device__ __forceinline__ void fold_partials(const double* __restrict__ part, int stride, int nb, double (&t)[NV],
                                              double* red) {
  for (int q = 0; q < NV; ++q) t[q] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kProjBlock)
    for (int q = 0; q < NV; ++q) t[q] += part[b * stride + q];
  block_sum<NV>(t, red);
}

__device__ __forceinline__ int wrap_x(int i, int nx) {
  i %= nx;
  return i < 0 ? i + nx : i;
}

__device__ __forceinline__ int first_column(int nblocks) {
  return (int)(xcd_tile(blockIdx.x, nblocks) * kProjBlock + threadIdx.x);
}

// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(kProjBlock) void k_proj_rhs(Geo<real> g, ProjGeo pg, ProjWork w,
                                                         const real* __restrict__ u, const real* __restrict__ v,
                                                         double dtg, int mode) {
  __shared__ double red[4 * 3];
  double val[3] = {0.0, 0.0, 0.0};
  for (int c = first_column(pg.nblocks); c < pg.n2; c += pg.nblocks * kProjBlock) {
    const int j = c / pg.nx, i = c - j * pg.nx, row = j * pg.nx;
    const int e = row + wrap_x(i + 1, pg.nx), wc = row + wrap_x(i - 1, pg.nx);
    const bool hasN = j + 1 < pg.ny, hasS = j > 0;
    const int n = hasN ? c + pg.nx : c, s = hasS ? c - pg.nx : c;
    const int nwC = g.nwet[c], nwE = g.nwet[e], nwW = g.nwet[wc], nwN = g.nwet[n], nwS = g.nwet[s];
    const double idx = pg.idx[j], idyc = pg.idyc[j], cn = pg.cfn[j], cs = hasS ? pg.cfn[j - 1] : 0.0;
    double cd = 0.0;
    for (int k = 0; k < nwC; ++k) {
      const long long pl = (long long)k * pg.n2;
      const double uc = (double)u[pl + c], vc = (double)v[pl + c];
      const double fxe = k < nwE ? 0.5 * (uc + (double)u[pl + e]) : 0.0;
      const double fxw = k < nwW ? 0.5 * ((double)u[pl + wc] + uc) : 0.0;
      const double fyn = (hasN && k < nwN) ? 0.5 * (vc + (double)v[pl + n]) * cn : 0.0;
      const double fys = (hasS && k < nwS) ? 0.5 * ((double)v[pl + s] + vc) * cs : 0.0;
      cd += ((fxe - fxw) * idx + (fyn - fys) * idyc) * pg.dzn[k];
    }
    const double wet = pg.wet[c];
    cd *= wet;
    const double a = pg.area[j] * wet;
    const double b = -(cd / dtg) * a;
    w.x[c] = 0.0;
    w.r[c] = b;
    val[0] += b * b;
    val[1] += b * (b * pg.inv[c]);
    val[2] += cd * cd * a;
  }
  block_sum<3>(val, red);
  if (threadIdx.x == 0) {
    for (int q = 0; q < 3; ++q) w.part_rhs[blockIdx.x * 4 + q] = val[q];
    if (blockIdx.x == 0) {
      w.sc->done = 0;
      if (mode == 0)
        for (int q = 0; q < 4; ++q) w.sc->its[q] = 0;
    }
  }
}

// One column's operands of k_proj_A: the nine taps and M r, p_old at the
// nine stencil points.  Loaded before the launch folds the previous partials,
// so the loads overlap that fold.
struct ColA {
  double C[kProjTaps], z[kProjTaps], p[kProjTaps];
};

__device__ __forceinline__ void load_col_a(const ProjGeo& pg, const double* __restrict__ r,
                                           const double* __restrict__ po, bool first, int c, ColA& a) {
  const int j = c / pg.nx, i = c - j * pg.nx, row = j * pg.nx;
  const int q[kProjTaps] = {c,
                            row + wrap_x(i + 1, pg.nx),
                            row + wrap_x(i - 1, pg.nx),
                            row + wrap_x(i + 2, pg.nx),
                            row + wrap_x(i - 2, pg.nx),
                            j + 1 < pg.ny ? c + pg.nx : c,
                            j > 0 ? c - pg.nx : c,
                            j + 2 < pg.ny ? c + 2 * pg.nx : c,
                            j > 1 ? c - 2 * pg.nx : c};
  const long long n2 = pg.n2;
#pragma unroll
  for (int t = 0; t < kProjTaps; ++t) {
    a.C[t] = pg.coef[t * n2 + c];
    a.z[t] = r[q[t]] * pg.inv[q[t]];
    a.p[t] = first ? 0.0 : po[q[t]];
  }
}

// stopping test and beta from the previous launch, then p_new = M r + beta
// p_old (beta = 0 on the first iteration), Ap and the p.Ap partials
__global__ __launch_bounds__(kProjBlock) void k_proj_A(ProjGeo pg, ProjWork w, int it, int mode,
                                                       const double* __restrict__ po, double* __restrict__ pn) {
  __shared__ double red[4 * 3];
  ProjScalars* sc = w.sc;
  if (sc->done) return;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const int c0 = first_column(pg.nblocks), stride = pg.nblocks * kProjBlock;
  ColA a;
  if (c0 < pg.n2) load_col_a(pg, w.r, po, it == 0, c0, a);
  double beta = 0.0;
  bool stop;
  if (it == 0) {
    double t[3];  // b.b, b.M b, transport norm^2
    fold_partials<3>(w.part_rhs, 4, pg.nblocks, t, red);
    const double bb = t[0];
    bool active;
    double atol2;
    if (mode == 0) {
      active = true;
      atol2 = sc->tol2 * bb;
      if (lead) {
        sc->floor2 = sc->tol2 * bb;
        sc->threshold = sc->ttol2 * t[2];
      }
    } else {
      active = t[2] > sc->threshold;
      atol2 = fmax(sc->tol2 * bb, sc->floor2);
    }
    stop = !active || !(bb > atol2 && sc->maxiter > 0);
    if (lead) {
      sc->active = active ? 1 : 0;
      sc->atol2 = atol2;
      sc->gam[0] = t[1];
    }
  } else {
    double t[2];  // r.M r, r.r after the previous iteration
    fold_partials<2>(w.part_b, 2, pg.nblocks, t, red);
    stop = !(t[1] > sc->atol2 && it < sc->maxiter);
    beta = t[0] / sc->gam[(it - 1) & 1];
    if (lead) sc->gam[it & 1] = t[0];
  }
  if (stop) {
    if (lead) sc->done = 1;
    return;
  }
  double val[1] = {0.0};
  for (int c = c0; c < pg.n2; c += stride) {
    if (c != c0) load_col_a(pg, w.r, po, it == 0, c, a);
    double pv[kProjTaps];
#pragma unroll
    for (int t = 0; t < kProjTaps; ++t) pv[t] = it == 0 ? a.z[t] : a.z[t] + beta * a.p[t];
    double acc = a.C[0] * pv[0];
#pragma unroll
    for (int t = 1; t < kProjTaps; ++t) acc += a.C[t] * pv[t];
    pn[c] = pv[0];
    w.Ap[c] = acc;
    val[0] += pv[0] * acc;
  }
  block_sum<1>(val, red);
  if (threadIdx.x == 0) w.part_a[blockIdx.x] = val[0];
}

// alpha from the p.Ap partials, x += alpha p, r -= alpha Ap and the r.M r, r.r partials
__global__ __launch_bounds__(kProjBlock) void k_proj_B(ProjGeo pg, ProjWork w, int it, int slot,
                                                       const double* __restrict__ pn) {
  __shared__ double red[4 * 2];
  ProjScalars* sc = w.sc;
  if (sc->done) return;
  const int c0 = first_column(pg.nblocks), stride = pg.nblocks * kProjBlock;
  double xv = 0.0, pv = 0.0, rv = 0.0, av = 0.0, mv = 0.0;
  if (c0 < pg.n2) {
    xv = w.x[c0];
    pv = pn[c0];
    rv = w.r[c0];
    av = w.Ap[c0];
    mv = pg.inv[c0];
  }
  double t[1];
  fold_partials<1>(w.part_a, 1, pg.nblocks, t, red);
  const double alpha = sc->gam[it & 1] / t[0];
  double val[2] = {0.0, 0.0};
  for (int c = c0; c < pg.n2; c += stride) {
    if (c != c0) {
      xv = w.x[c];
      pv = pn[c];
      rv = w.r[c];
      av = w.Ap[c];
      mv = pg.inv[c];
    }
    w.x[c] = xv + alpha * pv;
    const double rn = rv - alpha * av;
    w.r[c] = rn;
    val[0] += rn * (rn * mv);
    val[1] += rn * rn;
  }
  block_sum<2>(val, red);
  if (threadIdx.x == 0) {
    w.part_b[blockIdx.x * 2 + 0] = val[0];
    w.part_b[blockIdx.x * 2 + 1] = val[1];
    if (blockIdx.x == 0) sc->its[slot] = it + 1;
  }
}

// ud = us - dtg * grad_c3d(x) (horizontal.py:126-151), one thread per node
template <typename real>
__global__ __launch_bounds__(256) void k_proj_correct(Geo<real> g, ProjGeo pg, ProjWork w, double dtg,
                                                      const real* us, const real* vs, real* ud, real* vd) {
  if (!w.sc->active) return;
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= g.n3) return;
  const int k = (int)(q / pg.n2);
  const int c = (int)(q - (long long)k * pg.n2);
  const int j = c / pg.nx, i = c - j * pg.nx, row = j * pg.nx;
  double gx = 0.0, gy = 0.0;
  if (k < g.nwet[c]) {
    const double* __restrict__ x = w.x;
    const int e = row + wrap_x(i + 1, pg.nx), wc = row + wrap_x(i - 1, pg.nx);
    const double xc = x[c];
    const double dxp = k < g.nwet[e] ? x[e] - xc : 0.0;
    const double dxm = k < g.nwet[wc] ? xc - x[wc] : 0.0;
    gx = pg.idx[j] * 0.5 * (dxp + dxm);
    const double dyp = (j + 1 < pg.ny && k < g.nwet[c + pg.nx]) ? (x[c + pg.nx] - xc) * pg.cfn[j] : 0.0;
    const double dym = (j > 0 && k < g.nwet[c - pg.nx]) ? (xc - x[c - pg.nx]) * pg.cfs[j] : 0.0;
    gy = pg.idyc[j] * 0.5 * (dyp + dym);
  }
  const double un = (double)us[q] - dtg * gx;
  const double vn = (double)vs[q] - dtg * gy;
  ud[q] = (real)un;
  vd[q] = (real)vn;
}

// ---------------------------------------------------------------------------
template <typename real>
void launch_proj_rhs(const Geo<real>& g, const ProjGeo& pg, const ProjWork& w, const real* u, const real* v,
                     double dtg, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_proj_rhs<real>, dim3(pg.nblocks), dim3(kProjBlock), 0, s, g, pg, w, u, v, dtg, mode);
}

void launch_proj_cg(const ProjGeo& pg, const ProjWork& w, int maxiter, int mode, int slot, hipStream_t s) {
  for (int it = 0; it < maxiter; ++it) {
    const double* po = (it & 1) ? w.p1 : w.p0;
    double* pn = (it & 1) ? w.p0 : w.p1;
    hipLaunchKernelGGL(k_proj_A, dim3(pg.nblocks), dim3(kProjBlock), 0, s, pg, w, it, mode, po, pn);
    hipLaunchKernelGGL(k_proj_B, dim3(pg.nblocks), dim3(kProjBlock), 0, s, pg, w, it, slot, (const double*)pn);
  }
}

template <typename real>
void launch_proj_correct(const Geo<real>& g, const ProjGeo& pg, const ProjWork& w, double dtg, const real* us,
                         const real* vs, real* ud, real* vd, hipStream_t s) {
  const unsigned blocks = (unsigned)((g.n3 + 255) / 256);
  hipLaunchKernelGGL(k_proj_correct<real>, dim3(blocks), dim3(256), 0, s, g, pg, w, dtg, us, vs, ud, vd);
}

// ---------------------------------------------------------------------------
// Host: taps of the column operator and the Jacobi diagonal.
//
// Per wet layer the x part of A is -area dz div_x(grad_x psi) with
//   D(t) = O(t) (psi(t+1) - psi(t)),  O(t) = m(i+t) m(i+t+1)  (open east face),
//   div_x(0) = a^2/4 [O(0) (D(-1) + 2 D(0) + D(1)) - O(-1) (D(-2) + 2 D(-1) + D(0))],
// a = 1/dx; the y part has the same shape with the face cosines,
//   E(t) = P(t) cfn(j+t) (psi(t+1) - psi(t)),  gy(t) = b/(2 cos_{j+t}) (E(t) + E(t-1)),
//   div_y(0) = b/cos_j [P(0) cfn(j) (gy(0) + gy(1)) - P(-1) cfn(j-1) (gy(-1) + gy(0))] / 2,
// b = 1/dy and P(t) = 0 for faces outside 0 <= j+t < ny-1.
// The diagonal follows _column_projection_diagonal (projection.py:10-50).
void proj_stencil_host(int nx, int ny, int nz, const int* nwet, const double* wet2, const double* dx_row, double dy,
                       const double* cos_lat, const double* dzn, double* coef, double* diag) {
  const long long n2 = (long long)nx * ny;
  const double b = 1.0 / dy;
  auto cfnr = [&](int r) { return 0.5 * (cos_lat[r] + cos_lat[(r + 1) % ny]); };
  for (int j = 0; j < ny; ++j) {
    const double a = 1.0 / dx_row[j], area = dx_row[j] * dy, cj = cos_lat[j];
    const double cn = cos_lat[(j + 1) % ny], cs = cos_lat[(j + ny - 1) % ny];
    for (int i = 0; i < nx; ++i) {
      const long long c = (long long)j * nx + i;
      double cx[5] = {0, 0, 0, 0, 0}, cy[5] = {0, 0, 0, 0, 0}, d = 0.0;
      const double s2 = -area * wet2[c];
      int nwx[5], nwy[5];
      for (int t = -2; t <= 2; ++t) {
        nwx[t + 2] = nwet[(long long)j * nx + (((i + t) % nx) + nx) % nx];
        nwy[t + 2] = (j + t >= 0 && j + t < ny) ? nwet[(long long)(j + t) * nx + i] : 0;
      }
      for (int k = 0; k < nwet[c]; ++k) {
        const double dz = dzn[k];
        auto mx = [&](int t) { return k < nwx[t + 2] ? 1.0 : 0.0; };
        auto my = [&](int t) { return k < nwy[t + 2] ? 1.0 : 0.0; };
        auto O = [&](int t) { return mx(t) * mx(t + 1); };
        auto Pf = [&](int t) { return (j + t >= 0 && j + t + 1 <= ny - 1) ? my(t) * my(t + 1) : 0.0; };
        auto addD = [&](int t, double wgt) {
          const double o = O(t);
          cx[t + 3] += wgt * o;
          cx[t + 2] -= wgt * o;
        };
        auto addE = [&](int t, double wgt) {
          const double pm = Pf(t);
          if (pm == 0.0) return;
          const double cf = cfnr(j + t);
          cy[t + 3] += wgt * pm * cf;
          cy[t + 2] -= wgt * pm * cf;
        };
        const double f = 0.25 * a * a * dz * s2;
        if (O(0) != 0.0) {
          addD(-1, f);
          addD(0, 2.0 * f);
          addD(1, f);
        }
        if (O(-1) != 0.0) {
          addD(-2, -f);
          addD(-1, -2.0 * f);
          addD(0, -f);
        }
        const double F = (b / cj) * dz * s2;
        if (Pf(0) != 0.0) {
          const double A0 = 0.5 * F * cfnr(j);
          const double g0 = b / (2.0 * cj), g1 = b / (2.0 * cos_lat[j + 1]);
          addE(0, A0 * g0);
          addE(-1, A0 * g0);
          addE(1, A0 * g1);
          addE(0, A0 * g1);
        }
        if (Pf(-1) != 0.0) {
          const double A1 = -0.5 * F * cfnr(j - 1);
          const double gm = b / (2.0 * cos_lat[j - 1]), g0 = b / (2.0 * cj);
          addE(-1, A1 * gm);
          addE(-2, A1 * gm);
          addE(0, A1 * g0);
          addE(-1, A1 * g0);
        }
        // diagonal (closed form of the reference)
        const double pos = mx(0) * mx(1), neg = mx(0) * mx(-1);
        const double dxp = nx <= 2 ? 0.0 : 0.25 * a * a * ((pos - neg) * (pos - neg) + pos * pos + neg * neg);
        const double py = (j < ny - 1 ? my(0) * my(1) : 0.0) * 0.5 * (cj + cn);
        const double ny_ = (j > 0 ? my(0) * my(-1) : 0.0) * 0.5 * (cj + cs);
        const double dyp = 0.25 * b * b *
                           ((py - ny_) * (py - ny_) / (cj * cj) + py * py / (cj * cn) + ny_ * ny_ / (cj * cs));
        d += (dxp + dyp) * dz;
      }
      // taps: centre, east, west, east+2, west+2, north, south, north+2, south+2
      coef[0 * n2 + c] = cx[2] + cy[2];
      coef[1 * n2 + c] = cx[3];
      coef[2 * n2 + c] = cx[1];
      coef[3 * n2 + c] = cx[4];
      coef[4 * n2 + c] = cx[0];
      coef[5 * n2 + c] = cy[3];
      coef[6 * n2 + c] = cy[1];
      coef[7 * n2 + c] = cy[4];
      coef[8 * n2 + c] = cy[0];
      diag[c] = area * wet2[c] * d;
    }
  }
}

#define ZM_PROJ_INSTANTIATE(real)                                                                              \
  template void launch_proj_rhs<real>(const Geo<real>&, const ProjGeo&, const ProjWork&, const real*,           \
                                      const real*, double, int, hipStream_t);                             \
  template void launch_proj_correct<real>(const Geo<real>&, const ProjGeo&, const ProjWork&, double, const real*, \
                                          const real*, real*, real*, hipStream_t);

ZM_PROJ_INSTANTIATE(float)
ZM_PROJ_INSTANTIATE(double)

}  // namespace zm
