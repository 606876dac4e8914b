// Mode-split step kernels for gfx950 (MI355X), templated on the compute type.
//
// Each kernel restates one stage of the reference's legacy mode-split step
// (src/zhenmode/model/solver/timestepping/step.py:341-458); the reference
// function it follows is named above each kernel.  All face gates of the
// reference (wet * neighbour-wet products) are evaluated as selects on the
// wet-level count, which equals the reference's 0/1 products for finite data.
#include <type_traits>
#include <stdexcept>

#include "zm_common.h"
#include "zm_kernels.h"

namespace zm {

// ---------------------------------------------------------------------------
// neighbourhood of a column (i, j)
// ---------------------------------------------------------------------------
struct Nb {
  int i, j, c2;
  int e2, w2;       // east / west (periodic)
  int n2e, s2e;     // north / south, edge-clamped (mirror ghost)
  bool hasN, hasS;  // j < ny-1, j > 0
  int nwC, nwE, nwW, nwN, nwS;  // wet-level counts (N/S edge-clamped)
};

template <typename real>
__device__ __forceinline__ Nb make_nb(const Geo<real>& g, int i, int j) {
  Nb n;
  n.i = i;
  n.j = j;
  n.c2 = j * g.nx + i;
  n.e2 = j * g.nx + wrap_e(i, g.nx);
  n.w2 = j * g.nx + wrap_w(i, g.nx);
  n.hasN = j + 1 < g.ny;
  n.hasS = j > 0;
  n.n2e = n.hasN ? n.c2 + g.nx : n.c2;
  n.s2e = n.hasS ? n.c2 - g.nx : n.c2;
  n.nwC = g.nwet[n.c2];
  n.nwE = g.nwet[n.e2];
  n.nwW = g.nwet[n.w2];
  n.nwN = g.nwet[n.n2e];
  n.nwS = g.nwet[n.s2e];
  return n;
}

// Combination of the Laplacian's three terms and the nu_h update with
// explicit fused multiply-adds: the single-substep and the register-strip
// kernels then round identically whatever the compiler's contraction choices,
// so the strip blocking never changes results.
template <typename real>
__device__ __forceinline__ real lap_sum(real dx, real dy, real dyc, real id2, real idy2, real hidy, real tr) {
  return fma(dx, id2, fma(dy, idy2, -(dyc * hidy) * tr));
}

// One explicit nu_h substep x + nu_h dts lap(x) in neighbour differences
// d = neighbour - x and row coefficients that carry nd = nu_h dts:
//   kE = nd/dx^2, kN = nd (1/dy^2 - a), kS = nd (1/dy^2 + a), kY = nd/dy^2,
//   a = tan(lat)/(2 R dy);
// x + nd lap = x + kE (dE + dW) + kN dN + kS dS.  A closed face drops its
// flux; the metric term keeps the raw north/south values, so a closed north
// or south face is removed by -kY d.  With every gate open that correction is
// fma(-kY, +0, inc) == inc bit for bit, so nu_open equals nu_gated with all
// gates true, and every kernel of the sub-cycle (single substep, register
// strips, streaming) rounds identically whatever the row's class.  No
// multiply here feeds an add, so contraction cannot change the rounding.
template <typename real>
struct NuRow {
  real kE, kN, kS, kY;
};

template <typename real>
__device__ __forceinline__ real nu_open(real x, real xe, real xw, real xn, real xs, const NuRow<real>& k) {
  const real dE = xe - x, dW = xw - x, dN = xn - x, dS = xs - x;
  return x + fma(k.kE, dE + dW, fma(k.kN, dN, k.kS * dS));
}

template <typename real>
__device__ __forceinline__ real nu_gated(real x, real xe, real xw, real xn, real xs, bool ge, bool gw, bool gn,
                                         bool gs, const NuRow<real>& k) {
  const real dE = ge ? xe - x : real(0), dW = gw ? xw - x : real(0);
  const real dN = xn - x, dS = xs - x;
  const real inc = fma(k.kE, dE + dW, fma(k.kN, dN, k.kS * dS));
  return x + fma(-k.kY, (gn ? real(0) : dN) + (gs ? real(0) : dS), inc);
}

// _laplacian_h legacy branch (numerics/horizontal.py:181-215).
template <typename real>
__device__ __forceinline__ real lap_h(const real* __restrict__ f, long long plane, const Nb& n,
                                      int k, const Geo<real>& g) {
  const real u = f[plane + n.c2];
  const real ue = f[plane + n.e2], uw = f[plane + n.w2];
  const real un = f[plane + n.n2e], us = f[plane + n.s2e];
  const bool wc = k < n.nwC;
  const real fe = (wc && k < n.nwE) ? (ue - u) : real(0);
  const real fw = (wc && k < n.nwW) ? (u - uw) : real(0);
  const real fn = (wc && k < n.nwN) ? (un - u) : real(0);
  const real fs = (wc && k < n.nwS) ? (u - us) : real(0);
  return lap_sum(fe - fw, fn - fs, un - us, g.idx2[n.j], g.idy2, g.hidy, g.tanR[n.j]);
}

// _horizontal_diffusion_flux (numerics/horizontal.py:217-238); kap == nullptr
// means the unit coefficient of _horizontal_biharmonic_tracer.
template <typename real>
__device__ __forceinline__ real d1_flux(const real* __restrict__ t, long long plane, const Nb& n,
                                        int k, const Geo<real>& g, const real* __restrict__ kap) {
  const real c = t[plane + n.c2];
  const bool wc = k < n.nwC;
  const real kc = kap ? kap[n.c2] : real(1);
  real zE = 0, zW = 0, mN = 0, mS = 0;
  if (wc && k < n.nwE) {
    const real ke = kap ? kap[n.e2] : real(1);
    zE = (real(0.5) * (kc + ke)) * (t[plane + n.e2] - c) * g.idx[n.j];
  }
  if (wc && k < n.nwW) {
    const real kw = kap ? kap[n.w2] : real(1);
    zW = (real(0.5) * (kw + kc)) * (c - t[plane + n.w2]) * g.idx[n.j];
  }
  if (n.hasN && wc && k < n.nwN) {
    const real kn = kap ? kap[n.n2e] : real(1);
    mN = (real(0.5) * (kc + kn)) * (t[plane + n.n2e] - c) * g.idy * g.cfn[n.j];
  }
  if (n.hasS && wc && k < n.nwS) {
    const real ks = kap ? kap[n.s2e] : real(1);
    mS = (real(0.5) * (ks + kc)) * (c - t[plane + n.s2e]) * g.idy * g.cfn[n.j - 1];
  }
  return (zE - zW) * g.idx[n.j] + (mN - mS) * g.idyc[n.j];
}

// _d2_dz2 with _fill_ghost_bottom (numerics/vertical.py:122-135, 275-293).
template <typename real>
__device__ __forceinline__ real d2dz2(const real* __restrict__ f, int c2, int n2, int k, int nw,
                                      const Geo<real>& g) {
  const int kb = nw > 0 ? nw - 1 : 0;
  auto val = [&](int kk) { return f[(long long)(kk < kb ? kk : kb) * n2 + c2]; };
  if (k == 0) return real(2) * (val(1) - val(0)) * g.ih0top2;
  if (k == g.nz - 1) return real(2) * (val(g.nz - 2) - val(g.nz - 1)) * g.ih0bot2;
  return (val(k + 1) * ldc(g.hm, k - 1) + val(k - 1) * ldc(g.hp, k - 1) - val(k) * ldc(g.hmp, k - 1)) * ldc(g.iden, k - 1);
}

// _gradient_conservative_3d (numerics/horizontal.py:126-151) of a field given
// at the five points (per-level face gates, wrapped y neighbour closed at walls).
template <typename real>
__device__ __forceinline__ void grad_cons(real pc, real pe, real pw, real pn, real ps, const Nb& n,
                                          int k, const Geo<real>& g, real& gx, real& gy) {
  const bool wc = k < n.nwC;
  const real dxp = (wc && k < n.nwE) ? (pe - pc) : real(0);
  const real dxm = (wc && k < n.nwW) ? (pc - pw) : real(0);
  gx = g.idx[n.j] * real(0.5) * (dxp + dxm);
  const real dyp = (n.hasN && wc && k < n.nwN) ? (pn - pc) * g.cfn[n.j] : real(0);
  const real dym = (n.hasS && wc && k < n.nwS) ? (pc - ps) * g.cfs[n.j] : real(0);
  gy = g.idyc[n.j] * real(0.5) * (dyp + dym);
}

__device__ __forceinline__ void cell_of(long long c, int n2, int nx, int& i, int& j, int& k) {
  k = (int)(c / n2);
  const int r = (int)(c - (long long)k * n2);
  j = r / nx;
  i = r - j * nx;
}

// Pointwise 3-D kernels run on 64 x 4 (i x j) tiles of one level, so the
// north/south neighbour rows of a tile's interior come from the same
// workgroup's L1; tiles are numbered i-fastest and mapped XCD-contiguously.
constexpr int kTileI = 64, kTileJ = 4;

// Levels a column-scan kernel loads ahead of the level it sums.
constexpr int kColAhead = 6;

#define ZM_TILE_COORDS                                                      \
  const int ntx_ = (g.nx + kTileI - 1) / kTileI;                            \
  const int nty_ = (g.ny + kTileJ - 1) / kTileJ;                            \
  long long t_ = xcd_tile(blockIdx.x, gridDim.x);                           \
  const int tx_ = (int)(t_ % ntx_);                                         \
  t_ /= ntx_;                                                               \
  const int ty_ = (int)(t_ % nty_);                                         \
  const int k = (int)(t_ / nty_);                                           \
  const int i = tx_ * kTileI + (int)(threadIdx.x % kTileI);                 \
  const int j = ty_ * kTileJ + (int)(threadIdx.x / kTileI);                 \
  if (i >= g.nx || j >= g.ny) return;                                       \
  const long long plane = (long long)k * g.n2;                              \
  const long long c = plane + (long long)j * g.nx + i;

// The same cells with the level fastest: each XCD's contiguous tile range is
// a band of (i, j) tiles over all their levels, so a per-cell kernel that
// reads 2-D column data keeps that band's share in its L2 instead of
// streaming every 2-D array once per level.
#define ZM_TILE_COORDS_LEVEL_FASTEST                                        \
  const int ntx_ = (g.nx + kTileI - 1) / kTileI;                            \
  long long t_ = xcd_tile(blockIdx.x, gridDim.x);                           \
  const int k = (int)(t_ % g.nz);                                           \
  t_ /= g.nz;                                                               \
  const int tx_ = (int)(t_ % ntx_);                                         \
  const int ty_ = (int)(t_ / ntx_);                                         \
  const int i = tx_ * kTileI + (int)(threadIdx.x % kTileI);                 \
  const int j = ty_ * kTileJ + (int)(threadIdx.x / kTileI);                 \
  if (i >= g.nx || j >= g.ny) return;                                       \
  const long long plane = (long long)k * g.n2;                              \
  const long long c = plane + (long long)j * g.nx + i;

#define ZM_CELL_PROLOGUE \
  ZM_TILE_COORDS         \
  const Nb n = make_nb(g, i, j);

// ---------------------------------------------------------------------------
// nu_h sub-cycle: one explicit Laplacian substep of u and v
// (timestepping/step.py:94-107).  Optionally also stores lap(u), lap(v).
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(256) void k_nu_substep(Geo<real> g, const real* __restrict__ u,
                                                    const real* __restrict__ v, real* __restrict__ uo,
                                                    real* __restrict__ vo, real* __restrict__ lu,
                                                    real* __restrict__ lv) {
  ZM_CELL_PROLOGUE
  const real* __restrict__ row = g.nu_met + kNuMetStride * (n.j + kNuMetPad);
  const NuRow<real> kr{row[0], row[1], row[2], row[3]};
  const bool wc = k < n.nwC;
  const bool ge = wc && k < n.nwE, gw = wc && k < n.nwW, gn = wc && k < n.nwN, gs = wc && k < n.nwS;
  uo[c] = nu_gated(u[c], u[plane + n.e2], u[plane + n.w2], u[plane + n.n2e], u[plane + n.s2e], ge, gw, gn, gs, kr);
  vo[c] = nu_gated(v[c], v[plane + n.e2], v[plane + n.w2], v[plane + n.n2e], v[plane + n.s2e], ge, gw, gn, gs, kr);
  if (lu) {
    lu[c] = lap_h(u, plane, n, k, g);
    lv[c] = lap_h(v, plane, n, k, g);
  }
}

// Cross-lane moves (DPP wave_shl:1 / wave_shr:1 with bound_ctrl): lane l
// receives lane l+1 (from_right) or lane l-1 (from_left); lanes shifted in
// from outside the wave read 0 (only halo lanes, never stored).  A source lane
// that is inactive also reads 0, so these are called under wave-uniform
// conditions only.
__device__ __forceinline__ int dpp_from_right(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_from_left(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ float dpp_from_right(float x) { return __int_as_float(dpp_from_right(__float_as_int(x))); }
__device__ __forceinline__ float dpp_from_left(float x) { return __int_as_float(dpp_from_left(__float_as_int(x))); }
__device__ __forceinline__ double dpp_from_right(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_from_right((int)(b & 0xffffffffll)), hi = dpp_from_right((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double dpp_from_left(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_from_left((int)(b & 0xffffffffll)), hi = dpp_from_left((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ unsigned long long dpp_from_right(unsigned long long x) {
  const int lo = dpp_from_right((int)(x & 0xffffffffull)), hi = dpp_from_right((int)(x >> 32));
  return ((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo;
}
__device__ __forceinline__ unsigned long long dpp_from_left(unsigned long long x) {
  const int lo = dpp_from_left((int)(x & 0xffffffffull)), hi = dpp_from_left((int)(x >> 32));
  return ((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo;
}

__device__ __forceinline__ float readlane(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double readlane(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Temporally blocked nu_h sub-cycle (register strips).  One wave owns a tile
// of 64 longitudes x H latitude rows of one level; lane l holds its column's
// H values of u and v in registers, east/west neighbours come from the
// adjacent lanes (DPP), north/south from the neighbouring registers.  `nsub`
// (<= HALO) substeps run per launch; each substep invalidates one more cell
// at every tile edge, so the (64 - 2 HALO) x (H - 2 HALO) centre is exact.
// Every node receives exactly the arithmetic of k_nu_substep (nu_gated); rows
// whose 64 nodes have all four faces open use nu_open, which rounds like
// nu_gated with every gate true.  lap_strip and lap_open are lap_h's
// Laplacian for the same node classes (the stored lap(u), lap(v)).
template <typename real>
__device__ __forceinline__ real lap_strip(real x, real xe, real xw, real xn, real xs, bool fe_on, bool fw_on,
                                          bool fn_on, bool fs_on, real id2, real idy2, real hidy, real tr) {
  const real fe = fe_on ? (xe - x) : real(0), fw = fw_on ? (x - xw) : real(0);
  const real fn = fn_on ? (xn - x) : real(0), fs = fs_on ? (x - xs) : real(0);
  return lap_sum(fe - fw, fn - fs, xn - xs, id2, idy2, hidy, tr);
}

template <typename real>
__device__ __forceinline__ real lap_open(real x, real xe, real xw, real xn, real xs, real id2, real idy2,
                                         real hidy, real tr) {
  return lap_sum((xe - x) - (x - xw), (xn - x) - (x - xs), xn - xs, id2, idy2, hidy, tr);
}

// nsub substeps of the strip; WALL = the tile holds a closed-wall row (its
// own mirror neighbour), otherwise the north/south selects are compiled out.
template <typename real, int H, bool WALL>
__device__ __forceinline__ void nu_strip_substeps(real (&U)[H], real (&V)[H], NuRow<real> lane_k,
                                                  unsigned int me, unsigned int mw, unsigned int mn, unsigned int ms,
                                                  unsigned int wallN, unsigned int wallS, unsigned int open_rows,
                                                  int nsub) {
  for (int s = 0; s < nsub; ++s) {
    // keep one copy of the row coefficients instead of 4 H hoisted scalars
    asm volatile("" : "+v"(lane_k.kE), "+v"(lane_k.kN), "+v"(lane_k.kS), "+v"(lane_k.kY), "+v"(me), "+v"(mw),
                 "+v"(mn), "+v"(ms), "+s"(wallN), "+s"(wallS), "+s"(open_rows));
    real pu = U[0], pv = V[0];  // old values of row r-1
#pragma unroll
    for (int r = 0; r < H; ++r) {
      const NuRow<real> kr{readlane(lane_k.kE, r), readlane(lane_k.kN, r), readlane(lane_k.kS, r),
                           readlane(lane_k.kY, r)};
      const bool atN = WALL && ((wallN >> r) & 1u), atS = WALL && ((wallS >> r) & 1u);
      const int rn = r + 1 < H ? r + 1 : r;
      // ordering barrier: keep each row's neighbour moves next to its update
      asm volatile("" : "+v"(U[r]), "+v"(V[r]));
      const real x = U[r], y = V[r];
      const real xn = atN ? x : U[rn], xs = atS ? x : pu;
      const real yn = atN ? y : V[rn], ys = atS ? y : pv;
      const real xe = dpp_from_right(x), xw = dpp_from_left(x);
      const real ye = dpp_from_right(y), yw = dpp_from_left(y);
      pu = x;
      pv = y;
      if ((open_rows >> r) & 1u) {
        U[r] = nu_open(x, xe, xw, xn, xs, kr);
        V[r] = nu_open(y, ye, yw, yn, ys, kr);
      } else {
        const bool fe_on = (me >> r) & 1u, fw_on = (mw >> r) & 1u;
        const bool fn_on = (mn >> r) & 1u, fs_on = (ms >> r) & 1u;
        U[r] = nu_gated(x, xe, xw, xn, xs, fe_on, fw_on, fn_on, fs_on, kr);
        V[r] = nu_gated(y, ye, yw, yn, ys, fe_on, fw_on, fn_on, fs_on, kr);
      }
    }
  }
}

template <typename real, int H, int HALO>
__global__ __launch_bounds__(256) void k_nu_strip(Geo<real> g, const real* __restrict__ u,
                                                  const real* __restrict__ v, real* __restrict__ uo,
                                                  real* __restrict__ vo, int nsub) {
  constexpr int OX = 64 - 2 * HALO, OY = H - 2 * HALO;
  static_assert(H <= 32, "wet bits are kept in one 32-bit word");
  const int lane = threadIdx.x & 63;
  const int ntx = (g.nx + OX - 1) / OX, nty = (g.ny + OY - 1) / OY;
  const long long ntiles = (long long)ntx * nty * g.nz;
  long long t = xcd_tile(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (t >= ntiles) return;  // uniform per wave
  // the tile index is uniform per wave: keep it (and what derives from it) in SGPRs
  const int tx = __builtin_amdgcn_readfirstlane((int)(t % ntx));
  t /= ntx;
  const int ty = __builtin_amdgcn_readfirstlane((int)(t % nty));
  const int k = __builtin_amdgcn_readfirstlane((int)(t / nty));
  const int i0 = tx * OX - HALO, j0 = ty * OY - HALO;
  int i = (i0 + lane) % g.nx;
  if (i < 0) i += g.nx;
  const long long plane = (long long)k * g.n2;
  const real* __restrict__ up = u + plane;  // uniform bases: 32-bit per-lane offsets
  const real* __restrict__ vp = v + plane;
  real U[H], V[H];
  int NW[H];
#pragma unroll
  for (int r = 0; r < H; ++r) {
    int j = j0 + r;
    j = j < 0 ? 0 : (j >= g.ny ? g.ny - 1 : j);  // out-of-domain rows read a copy; never used by valid nodes
    const int c2 = j * g.nx + i;
    U[r] = up[c2];
    V[r] = vp[c2];
    NW[r] = g.nwet[c2];
  }
  unsigned int mc = 0, wallN = 0, wallS = 0, inside = 0;
#pragma unroll
  for (int r = 0; r < H; ++r) {
    const int j = j0 + r;
    const bool in = j >= 0 && j < g.ny;
    if (in && k < NW[r]) mc |= 1u << r;
    if (in) inside |= 1u << r;
    if (j == g.ny - 1) wallN |= 1u << r;
    if (j == 0) wallS |= 1u << r;
  }
  // face gates per row bit: east/west from the neighbouring lanes, north/south
  // from the neighbouring rows (the wall row is its own mirror neighbour)
  unsigned int me = mc & (unsigned int)dpp_from_right((int)mc);
  unsigned int mw = mc & (unsigned int)dpp_from_left((int)mc);
  unsigned int mn = mc & ((mc >> 1) | wallN);
  unsigned int ms = mc & ((mc << 1) | wallS);
  // rows where every interior lane has all four faces open (uniform)
  unsigned int notfull = (lane == 0 || lane == 63) ? 0u : ~(me & mw & mn & ms);
  for (int o = 1; o < 64; o <<= 1) notfull |= (unsigned int)__shfl_xor((int)notfull, o, 64);
  unsigned int open_rows = __builtin_amdgcn_readfirstlane(~notfull) & inside;
  wallN = __builtin_amdgcn_readfirstlane(wallN);
  wallS = __builtin_amdgcn_readfirstlane(wallS);
  // lane r holds the coefficients of strip row r (read back with readlane)
  NuRow<real> lane_k;
  {
    int j = j0 + (lane < H ? lane : H - 1);
    j = j < 0 ? 0 : (j >= g.ny ? g.ny - 1 : j);
    const real* __restrict__ row = g.nu_met + kNuMetStride * (j + kNuMetPad);
    lane_k = NuRow<real>{row[0], row[1], row[2], row[3]};
  }
  nu_strip_substeps<real, H, true>(U, V, lane_k, me, mw, mn, ms, wallN, wallS, open_rows, nsub);
  const int io = tx * OX + (lane - HALO);
  real* __restrict__ uop = uo + plane;
  real* __restrict__ vop = vo + plane;
  if (lane >= HALO && lane < 64 - HALO && io < g.nx) {
#pragma unroll
    for (int r = HALO; r < H - HALO; ++r) {
      const int j = j0 + r;
      if (j < g.ny) {
        const int o = j * g.nx + io;
        uop[o] = U[r];
        vop[o] = V[r];
      }
    }
  }
}

// Streaming form of the temporally blocked nu_h sub-cycle.  A wave owns 64
// longitudes of one level and marches north through a chunk of rows.  Level l
// of the NS substeps keeps a window of three consecutive rows in registers;
// in one march step every level l computes its row of level l + 1 from the
// window as it stood at the start of the step, so the 2 NS computations of a
// step are independent, and the new rows then enter the windows of the level
// above (level 0: the loaded row).  Level l + 1 therefore trails level l by
// two rows: in step t it computes row t - 2 l - 1.  Rows are only recomputed
// at the chunk ends (NS - l extra rows per side at level l) and input rows
// are prefetched kNuAhead rows ahead.  East/west neighbours come from the
// adjacent lanes (DPP), so as in k_nu_strip lanes [NS, 64 - NS) are exact.
// Every node receives the arithmetic of k_nu_substep (nu_open where the row's
// interior lanes have all four faces open, nu_gated otherwise).
//
// Per-wave data prepared at set-up (never written by a step):
//  * nu_wet[(k * nch + ch) * nx + i]: bit b = node (i, jb - kNuWetBase + b, k)
//    is wet (0 outside the domain), jb = ch * chunk; the four face gates of
//    every row follow from it once per wave, and the rows whose interior
//    lanes are all open from one wave-wide AND;
//  * nu_met[kNuMetStride (j + kNuMetPad) + q] = row j's substep coefficients
//    kE, kN, kS, kY (q = 0..3, see nu_open) and 1/dx^2, tan(lat)/R (q = 4, 5),
//    padded so that every march step reads its rows' coefficients
//    unconditionally through scalar loads.
constexpr int kNuAhead = 6;  // a multiple of the period of the three-row window

template <typename real>
using nu_met_ptr = const __attribute__((address_space(4))) real*;  // read through the scalar cache

// One level's window: rows (south, centre, north) of its next computation.
// Named scalars per level (no per-field arrays over the levels, which the
// compiler keeps as register tuples and copies whole at every branch merge).
template <typename real>
struct NuWin {
  real su, cu, nu, sv, cv, nv;
  __device__ __forceinline__ void push(real u, real v) {
    su = cu;
    cu = nu;
    nu = u;
    sv = cv;
    cv = nv;
    nv = v;
  }
};

template <typename real, int NS>
struct NuStream {
  NuWin<real> w[NS];
  real qu[kNuAhead], qv[kNuAhead];  // prefetched input rows
};

// face gates of this lane, bit b = row wb + b (the wall row is its own mirror)
struct NuGates {
  unsigned long long e, w, n, s;
};

// Raw buffer access (gfx9 descriptor word 3 = 0x00020000, stride 0): the base
// address sits in an SGPR descriptor, each access adds a 32-bit per-lane
// byte offset and a uniform byte offset (the row), so no per-lane 64-bit
// pointer is kept.  Offsets stay below 2^31 bytes (one level plane).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ float buf_ld(float*, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ double buf_ld(double*, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void buf_st(float x, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, voff, soff, 0);
}
__device__ __forceinline__ void buf_st(double x, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
  using b64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(b64, x), r, voff, soff, 0);
}

// The bits of |x| (zero exactly when x is +0 or -0).
__device__ __forceinline__ unsigned long long magnitude_bits(float x) { return __float_as_uint(x) & 0x7fffffffu; }
__device__ __forceinline__ unsigned long long magnitude_bits(double x) {
  return (unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffull;
}

// Buffer descriptors of the streaming kernel's input and output level planes.
struct NuBufs {
  __amdgpu_buffer_rsrc_t u, v, uo, vo, lu, lv;
};

// One level-row of the streaming kernel: row r of level l + 1 from the window
// (xs, x, xn) of level l; b = r's bit in the gate words.  cls (uniform): 1 =
// every interior lane has all four faces open (nu_open), 2 = every face is
// closed (nu_gated with every gate false, which needs no east/west values),
// 0 = mixed.
template <typename real>
__device__ __forceinline__ void nu_stream_node(real x, real xn, real xs, real y, real yn, real ys, int cls,
                                               const NuGates& G, int b, const NuRow<real>& kr, real& ou, real& ov) {
  if (cls == 2) {
    ou = nu_gated(x, x, x, xn, xs, false, false, false, false, kr);
    ov = nu_gated(y, y, y, yn, ys, false, false, false, false, kr);
  } else {
    const real xe = dpp_from_right(x), xw = dpp_from_left(x);
    const real ye = dpp_from_right(y), yw = dpp_from_left(y);
    if (cls == 1) {
      ou = nu_open(x, xe, xw, xn, xs, kr);
      ov = nu_open(y, ye, yw, yn, ys, kr);
    } else {
      const bool fe_on = (G.e >> b) & 1ull, fw_on = (G.w >> b) & 1ull;
      const bool fn_on = (G.n >> b) & 1ull, fs_on = (G.s >> b) & 1ull;
      ou = nu_gated(x, xe, xw, xn, xs, fe_on, fw_on, fn_on, fs_on, kr);
      ov = nu_gated(y, ye, yw, yn, ys, fe_on, fw_on, fn_on, fs_on, kr);
    }
  }
}

// The Laplacians of the same node (lap_h's arithmetic per class: lap_open,
// lap_strip, or lap_strip with zero face differences).
template <typename real>
__device__ __forceinline__ void lap_stream_node(real x, real xn, real xs, real y, real yn, real ys, int cls,
                                                const NuGates& G, int b, real id2, real tr, real idy2, real hidy,
                                                real& Lx, real& Ly) {
  if (cls == 2) {
    Lx = lap_sum(real(0), real(0), xn - xs, id2, idy2, hidy, tr);
    Ly = lap_sum(real(0), real(0), yn - ys, id2, idy2, hidy, tr);
  } else {
    const real xe = dpp_from_right(x), xw = dpp_from_left(x);
    const real ye = dpp_from_right(y), yw = dpp_from_left(y);
    if (cls == 1) {
      Lx = lap_open(x, xe, xw, xn, xs, id2, idy2, hidy, tr);
      Ly = lap_open(y, ye, yw, yn, ys, id2, idy2, hidy, tr);
    } else {
      const bool fe_on = (G.e >> b) & 1ull, fw_on = (G.w >> b) & 1ull;
      const bool fn_on = (G.n >> b) & 1ull, fs_on = (G.s >> b) & 1ull;
      Lx = lap_strip(x, xe, xw, xn, xs, fe_on, fw_on, fn_on, fs_on, id2, idy2, hidy, tr);
      Ly = lap_strip(y, ye, yw, yn, ys, fe_on, fw_on, fn_on, fs_on, id2, idy2, hidy, tr);
    }
  }
}

// Uniform per-wave bounds of the streaming kernel: level l computes its row
// at step t iff 0 <= t - c[l] < n[l].
template <int NS>
struct NuRange {
  int wb, qe, jb, je;
  int c[NS], n[NS];
};

// One march step t = t0 + SLOT.  mt = the coefficient row of row t0 - 2 NS + 1
// (every row of the step then sits at a constant offset from it).  LU: level
// 0 also stores the Laplacians of the input rows [jb, je) (the first
// substep's lap(u), lap(v)).
template <typename real, int NS, int SLOT, bool LU>
__device__ __forceinline__ void nu_stream_step(NuStream<real, NS>& st, const Geo<real>& g, nu_met_ptr<real> mt,
                                               const NuGates& G, unsigned long long openw, unsigned long long dryw,
                                               int t, const NuRange<NS>& R,
                                               const NuBufs& B, unsigned ib, unsigned iob, bool store_lane) {
  // row t enters the level-0 window
  st.w[0].push(st.qu[SLOT], st.qv[SLOT]);
  // prefetch row t + kNuAhead into this slot (uniform row bases)
  if (t + kNuAhead < R.qe) {
    const int ro = (t + kNuAhead) * g.nx * (int)sizeof(real);
    st.qu[SLOT] = buf_ld((real*)nullptr, B.u, ib, ro);
    st.qv[SLOT] = buf_ld((real*)nullptr, B.v, ib, ro);
  }
  // levels from the top: level l's new row enters the window of level l + 1,
  // which has already computed its row of this step
#pragma unroll
  for (int l = NS - 1; l >= 0; --l) {
    const int r = t - 2 * l - 1;  // row of level l + 1 computed now
    const nu_met_ptr<real> row = mt + kNuMetStride * (SLOT + 2 * NS - 2 - 2 * l);  // row r
    NuWin<real>& w = st.w[l];
    // a row outside the level's range is never read by a computed row of
    // the level above: it keeps the centre value
    real ou = w.cu, ov = w.cv;
    if ((unsigned)(t - R.c[l]) < (unsigned)R.n[l]) {  // uniform
      // the wall row is its own mirror neighbour; the slot written is not
      // read by another computed row (the south slot leaves the window, the
      // north one holds row ny).  The empty asm keeps these rare (wall-chunk)
      // branches as branches instead of two selects per level and step.
      if (__builtin_expect(r == 0, 0)) {
        asm volatile("");
        w.su = w.cu;
        w.sv = w.cv;
      }
      if (__builtin_expect(r == g.ny - 1, 0)) {
        asm volatile("");
        w.nu = w.cu;
        w.nv = w.cv;
      }
      const int b = r - R.wb;
      const int cls = ((openw >> b) & 1ull) ? 1 : (((dryw >> b) & 1ull) ? 2 : 0);
      const NuRow<real> kr{row[0], row[1], row[2], row[3]};
      nu_stream_node(w.cu, w.nu, w.su, w.cv, w.nv, w.sv, cls, G, b, kr, ou, ov);
      if constexpr (LU) {
        // uniform condition: the Laplacian reads the neighbouring lanes (DPP),
        // which must all be active; only the store is per lane
        if (l == 0 && r >= R.jb && r < R.je) {
          real Lx, Ly;
          lap_stream_node(w.cu, w.nu, w.su, w.cv, w.nv, w.sv, cls, G, b, row[4], row[5], g.idy2, g.hidy, Lx, Ly);
          if (store_lane) {
            const int ro = r * g.nx * (int)sizeof(real);
            buf_st(Lx, B.lu, iob, ro);
            buf_st(Ly, B.lv, iob, ro);
          }
        }
      }
    }
    if (l + 1 < NS) {
      st.w[l + 1].push(ou, ov);
    } else if (r >= R.jb && r < R.je && store_lane) {  // level NS: the output rows
      const int ro = r * g.nx * (int)sizeof(real);
      buf_st(ou, B.uo, iob, ro);
      buf_st(ov, B.vo, iob, ro);
    }
  }
}

template <typename real, int NS, bool LU>
__global__ __launch_bounds__(256) void k_nu_stream(Geo<real> g, const real* __restrict__ u,
                                                   const real* __restrict__ v, real* __restrict__ uo,
                                                   real* __restrict__ vo, real* __restrict__ lu,
                                                   real* __restrict__ lv, int chunk) {
  constexpr int OX = 64 - 2 * NS;
  static_assert(NS <= kNuWetBase && 2 * NS - 1 < kNuMetPad, "wet word and metric padding cover NS substeps");
  const int lane = threadIdx.x & 63;
  const int ntx = (g.nx + OX - 1) / OX, nch = (g.ny + chunk - 1) / chunk;
  const long long ntiles = (long long)ntx * nch * g.nz;
  // groups of 8 blocks (32 waves, about 2.5 chunk rows of one level) dealt
  // over the XCDs: the open-row share falls with depth, so XCD-contiguous
  // level ranges would leave the surface levels' work to one XCD
  long long t = xcd_spread(blockIdx.x, gridDim.x, 8) * 4 + (threadIdx.x >> 6);
  if (t >= ntiles) return;  // uniform per wave
  const int tx = __builtin_amdgcn_readfirstlane((int)(t % ntx));
  t /= ntx;
  const int ch = __builtin_amdgcn_readfirstlane((int)(t % nch));
  const int k = __builtin_amdgcn_readfirstlane((int)(t / nch));
  int i = (tx * OX - NS + lane) % g.nx;
  if (i < 0) i += g.nx;
  const int jb = ch * chunk, je = jb + chunk < g.ny ? jb + chunk : g.ny;
  const int qs = jb - NS > 0 ? jb - NS : 0, qe = je + NS < g.ny ? je + NS : g.ny;  // input rows [qs, qe)
  const int wb = jb - kNuWetBase;
  const long long plane = (long long)k * g.n2;
  NuBufs B;
  B.u = buf_rsrc(u + plane);
  B.v = buf_rsrc(v + plane);
  B.uo = buf_rsrc(uo + plane);
  B.vo = buf_rsrc(vo + plane);
  if constexpr (LU) {
    B.lu = buf_rsrc(lu + plane);
    B.lv = buf_rsrc(lv + plane);
  }
  const int io = tx * OX + (lane - NS);
  const bool store_lane = lane >= NS && lane < 64 - NS && io < g.nx;
  // 32-bit byte offsets of this lane's column (row bases stay uniform)
  const unsigned ib = (unsigned)i * (unsigned)sizeof(real), iob = store_lane ? (unsigned)io * (unsigned)sizeof(real) : 0u;
  if constexpr (!LU) {
    // A wave whose whole input tile (64 columns, rows [qs, qe)) is land
    // first checks whether every input value is +-0 (land velocities are
    // held at rest away from the coasts).  Then every substep of every node
    // gives +0 whatever its face gates: each difference of zeros is +0, so
    // each product, fused multiply-add and final sum is +0 (a -0 centre
    // value plus +0 is +0), and the wave stores +0 without the march.
    const unsigned long long Wt = g.nu_wet[((long long)k * nch + ch) * g.nx + i];
    const int b0 = qs - wb, nb = qe - qs;  // the input rows' bits of the wet word (b0 + nb <= 64)
    const unsigned long long rows = (nb >= 64 ? ~0ull : (1ull << nb) - 1ull) << b0;
    if (__ballot((Wt & rows) != 0ull) == 0ull) {  // uniform
      unsigned long long acc = 0ull;
#pragma unroll 8
      for (int r = qs; r < qe; ++r) {
        const int ro = r * g.nx * (int)sizeof(real);
        acc |= magnitude_bits(buf_ld((real*)nullptr, B.u, ib, ro)) | magnitude_bits(buf_ld((real*)nullptr, B.v, ib, ro));
      }
      if (__ballot(acc != 0ull) == 0ull) {  // uniform
        if (store_lane)
          for (int r = jb; r < je; ++r) {
            const int ro = r * g.nx * (int)sizeof(real);
            buf_st(real(0), B.uo, iob, ro);
            buf_st(real(0), B.vo, iob, ro);
          }
        return;
      }
    }
  }
  NuStream<real, NS> st;
#pragma unroll
  for (int l = 0; l < NS; ++l) st.w[l] = NuWin<real>{real(0), real(0), real(0), real(0), real(0), real(0)};
#pragma unroll
  for (int q = 0; q < kNuAhead; ++q) {
    st.qu[q] = st.qv[q] = real(0);
    if (qs + q < qe) {
      const int ro = (qs + q) * g.nx * (int)sizeof(real);
      st.qu[q] = buf_ld((real*)nullptr, B.u, ib, ro);
      st.qv[q] = buf_ld((real*)nullptr, B.v, ib, ro);
    }
  }
  // face gates: east/west from the adjacent lanes' wet words, north/south
  // from the adjacent bits; the wall rows are their own mirror neighbours
  const unsigned long long W = g.nu_wet[((long long)k * nch + ch) * g.nx + i];
  const int bn = g.ny - 1 - wb, bs = -wb;
  const unsigned long long wallN = bn >= 0 && bn < 64 ? 1ull << bn : 0ull;
  const unsigned long long wallS = bs >= 0 && bs < 64 ? 1ull << bs : 0ull;
  NuGates G;
  G.e = W & dpp_from_right(W);
  G.w = W & dpp_from_left(W);
  G.n = W & ((W >> 1) | wallN);
  G.s = W & ((W << 1) | wallS);
  // rows where every interior lane has all four faces open, or all closed (uniform)
  const bool edge = lane == 0 || lane == 63;
  unsigned long long full = edge ? ~0ull : (G.e & G.w & G.n & G.s);
  unsigned long long some = edge ? 0ull : (G.e | G.w | G.n | G.s);
  for (int o = 1; o < 64; o <<= 1) {
    full &= (unsigned long long)__shfl_xor((long long)full, o, 64);
    some |= (unsigned long long)__shfl_xor((long long)some, o, 64);
  }
  const unsigned long long openw =
      ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(full >> 32)) << 32) |
      (unsigned)__builtin_amdgcn_readfirstlane((int)(full & 0xffffffffull));
  const unsigned long long dryw =
      ~(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(some >> 32)) << 32) |
        (unsigned)__builtin_amdgcn_readfirstlane((int)(some & 0xffffffffull)));
  const nu_met_ptr<real> met = (nu_met_ptr<real>)g.nu_met;
  NuRange<NS> R;
  R.wb = wb;
  R.qe = qe;
  R.jb = jb;
  R.je = je;
#pragma unroll
  for (int l = 0; l < NS; ++l) {  // rows [max(jb - (NS - l - 1), 0), min(je + NS - l - 1, ny)) at step r + 2 l + 1
    const int lo = jb - (NS - l - 1) > 0 ? jb - (NS - l - 1) : 0;
    const int hi = je + NS - l - 1 < g.ny ? je + NS - l - 1 : g.ny;
    R.c[l] = __builtin_amdgcn_readfirstlane(lo + 2 * l + 1);
    R.n[l] = __builtin_amdgcn_readfirstlane(hi - lo);
  }
  const int tend = qe + 2 * NS - 1;
  static_assert(kNuAhead == 6, "the step below is unrolled over the prefetch ring");
#define ZM_NU_STEP(S) \
  nu_stream_step<real, NS, S, LU>(st, g, mt, G, openw, dryw, t0 + S, R, B, ib, iob, store_lane)
  for (int t0 = qs; t0 < tend; t0 += kNuAhead) {
    const nu_met_ptr<real> mt = met + kNuMetStride * (t0 - 2 * NS + 1 + kNuMetPad);
    ZM_NU_STEP(0);
    if (t0 + 1 >= tend) break;
    ZM_NU_STEP(1);
    if (t0 + 2 >= tend) break;
    ZM_NU_STEP(2);
    if (t0 + 3 >= tend) break;
    ZM_NU_STEP(3);
    if (t0 + 4 >= tend) break;
    ZM_NU_STEP(4);
    if (t0 + 5 >= tend) break;
    ZM_NU_STEP(5);
  }
#undef ZM_NU_STEP
}

template <typename real>
struct NuTile;
template <>
struct NuTile<float> {
  static constexpr int H = 32, HALO = 4;
};
template <>
struct NuTile<double> {
  static constexpr int H = 20, HALO = 4;
};

// Unit-coefficient D1 of T and S (first half of _horizontal_biharmonic_tracer).
template <typename real>
__device__ __forceinline__ real d1_vals(real c, real e, real w, real nn, real s, bool ge, bool gw, bool gn,
                                        bool gs, real kfe, real kfw, real kfn, real kfs, real idx, real idy,
                                        real cfN, real cfS, real idyc) {
  const real zE = ge ? kfe * (e - c) * idx : real(0);
  const real zW = gw ? kfw * (c - w) * idx : real(0);
  const real mN = gn ? kfn * (nn - c) * idy * cfN : real(0);
  const real mS = gs ? kfs * (c - s) * idy * cfS : real(0);
  return (zE - zW) * idx + (mN - mS) * idyc;
}

// ---------------------------------------------------------------------------
// Column-marching form.  A thread owns one (i, j) column over a segment of
// g.seg levels; vertical neighbours come from a register window, the column's
// 2-D data (wet counts, coefficients, metrics, forcing) is loaded once, and
// each level only loads its horizontal neighbours.  The operators below take
// explicit neighbour values and evaluate the same expressions as lap_h,
// d1_flux and d2dz2.
// ---------------------------------------------------------------------------

#define ZM_COLUMN_TILE                                                       \
  const int ntx_ = (g.nx + kTileI - 1) / kTileI;                             \
  const int nty_ = (g.ny + kTileJ - 1) / kTileJ;                             \
  long long t_ = xcd_tile(blockIdx.x, gridDim.x);                            \
  const int tx_ = (int)(t_ % ntx_);                                          \
  t_ /= ntx_;                                                                \
  const int ty_ = (int)(t_ % nty_);                                          \
  const int seg_ = (int)(t_ / nty_);                                         \
  const int i = tx_ * kTileI + (int)(threadIdx.x % kTileI);                  \
  const int j = ty_ * kTileJ + (int)(threadIdx.x / kTileI);

#define ZM_COLUMN_SEGMENT                                                    \
  const int k0 = seg_ * g.seg;                                               \
  const int k1 = k0 + g.seg < g.nz ? k0 + g.seg : g.nz;                      \
  const Nb n = make_nb(g, i, j);                                             \
  const int c2 = n.c2;                                                       \
  const long long n2 = g.n2;                                                 \
  const int kb = n.nwC > 0 ? n.nwC - 1 : 0;

#define ZM_COLUMN_PROLOGUE                                                   \
  ZM_COLUMN_TILE                                                             \
  if (i >= g.nx || j >= g.ny) return;                                        \
  ZM_COLUMN_SEGMENT

// Unit-coefficient D1 of T and S (the inner D1 of _horizontal_biharmonic_tracer,
// horizontal.py:240-246), one thread per column marching down its segment;
// the next level's ten neighbour values are loaded before the current level
// is combined (d1_vals: the arithmetic of d1_flux with unit coefficients).
template <typename real>
__global__ __launch_bounds__(256) void k_d1_unit(Geo<real> g, const real* __restrict__ T,
                                                 const real* __restrict__ S, real* __restrict__ dT,
                                                 real* __restrict__ dS) {
  ZM_COLUMN_PROLOGUE
  const real one = real(1);
  const real idx = g.idx[j], idyc = g.idyc[j], cfN = g.cfn[j], cfS = n.hasS ? g.cfn[j - 1] : real(0);
  const long long dE = n.e2 - c2, dW = n.w2 - c2, dN = n.n2e - c2, dS2 = n.s2e - c2;
  long long o = (long long)k0 * n2 + c2;
  real tc = T[o], te = T[o + dE], tw = T[o + dW], tn = T[o + dN], ts = T[o + dS2];
  real sc = S[o], se = S[o + dE], sw = S[o + dW], sn = S[o + dN], ss = S[o + dS2];
  for (int k = k0; k < k1; ++k, o += n2) {
    // the next level's values (the last level reloads itself: in bounds)
    const long long on = k + 1 < k1 ? o + n2 : o;
    const real tc1 = T[on], te1 = T[on + dE], tw1 = T[on + dW], tn1 = T[on + dN], ts1 = T[on + dS2];
    const real sc1 = S[on], se1 = S[on + dE], sw1 = S[on + dW], sn1 = S[on + dN], ss1 = S[on + dS2];
    const bool wc = k < n.nwC;
    const bool ge = wc && k < n.nwE, gw = wc && k < n.nwW;
    const bool gn = n.hasN && wc && k < n.nwN, gs = n.hasS && wc && k < n.nwS;
    dT[o] = d1_vals(tc, te, tw, tn, ts, ge, gw, gn, gs, one, one, one, one, idx, g.idy, cfN, cfS, idyc);
    dS[o] = d1_vals(sc, se, sw, sn, ss, ge, gw, gn, gs, one, one, one, one, idx, g.idy, cfN, cfS, idyc);
    tc = tc1; te = te1; tw = tw1; tn = tn1; ts = ts1;
    sc = sc1; se = se1; sw = sw1; sn = sn1; ss = ss1;
  }
}

template <typename real>
__device__ __forceinline__ real lap_vals(real c, real e, real w, real nn, real s, bool ge, bool gw, bool gn,
                                         bool gs, real idx2, real idy2, real hidy, real tr) {
  const real fe = ge ? (e - c) : real(0), fw = gw ? (c - w) : real(0);
  const real fn = gn ? (nn - c) : real(0), fs = gs ? (c - s) : real(0);
  return (fe - fw) * idx2 + (fn - fs) * idy2 + (-((nn - s) * hidy) * tr);
}


// node-form d2/dz2 with the seafloor ghost fill, from the window (k-1, k, k+1);
// ghost nodes below the deepest wet level (k > kb) return 0.
template <typename real>
__device__ __forceinline__ real d2z_win(real xm, real xc, real xp, int k, int kb, const Geo<real>& g) {
  if (k == 0) return kb >= 1 ? real(2) * (xp - xc) * g.ih0top2 : real(0);
  if (k == g.nz - 1) return kb >= g.nz - 1 ? real(2) * (xm - xc) * g.ih0bot2 : real(0);
  if (k > kb) return real(0);
  const real up = k < kb ? xp : xc;
  return (up * ldc(g.hm, k - 1) + xm * ldc(g.hp, k - 1) - xc * ldc(g.hmp, k - 1)) * ldc(g.iden, k - 1);
}

// one-sided/centered d/dz of the ghost-filled column (numerics/vertical.py:267-272)
template <typename real>
__device__ __forceinline__ real ddz_win(real xm, real xc, real xp, int k, int kb, const Geo<real>& g) {
  if (k == 0) return ((kb >= 1 ? xp : xc) - xc) * g.idz_top;
  if (k == g.nz - 1) return kb >= g.nz - 1 ? (xc - xm) * g.idz_bot : real(0);
  if (k > kb) return real(0);
  return ((k < kb ? xp : xc) - xm) * ldc(g.iddz, k - 1);
}

// kappa_v vertical diffusion of a tracer at level k (physics/vertical.py:45-53):
// node form kappa d2/dz2 by default; with conservative_kv the interface-flux
// form of _d2_dz2_flux (numerics/vertical.py:317-332), F[k+1/2] =
// -kappa (X[k+1] - X[k]) / dz_iface[k] on wet interfaces, (F[k-1/2] - F[k+1/2]) / dz_node.
template <typename real>
__device__ __forceinline__ real vdiff_win(real xm, real xc, real xp, int k, int nw, int kb, real kv, bool cons,
                                          const Geo<real>& g) {
  if (!cons) return kv * d2z_win(xm, xc, xp, k, kb, g);
  if (k >= nw) return real(0);
  const real fup = k > 0 ? -kv * (xc - xm) * ldc(g.idzi, k - 1) : real(0);
  const real fdn = k + 1 < nw ? -kv * (xp - xc) * ldc(g.idzi, k) : real(0);
  return (fup - fdn) * ldc(g.idzn, k);
}

// minmod slope from wet neighbours (transport.py:199-213)
template <typename real>
__device__ __forceinline__ real minmod_slope(real xm, real x, real xp, bool wm, bool w, bool wp) {
  const real l = x - xm, r = xp - x;
  const real sg = real((l > real(0)) - (l < real(0)) + (r > real(0)) - (r < real(0)));
  return (wm && w && wp) ? real(0.5) * sg * fmin(fabs(l), fabs(r)) : real(0);
}

template <typename real>
static inline unsigned segs3(const Geo<real>& g) {
  return (unsigned)((long long)((g.nx + kTileI - 1) / kTileI) * ((g.ny + kTileJ - 1) / kTileJ) *
                    ((g.nz + g.seg - 1) / g.seg));
}

// LDS halo tiles of the column kernels.  A workgroup of the 64 x 4 column
// tile marches down the levels in step; per level it stages each stencil
// field's tile with a one-cell halo ((64 + 2) x (4 + 2) cells) in LDS with
// coalesced loads, periodic in i and edge-clamped in j (the cells make_nb's
// neighbour indices name), so that each thread reads its four horizontal
// neighbours from LDS instead of issuing four more loads.  The next level's
// cells are loaded into registers while the current level computes and are
// stored into the second of two LDS buffers (one barrier per level).
constexpr int kHaloW = kTileI + 2, kHaloH = kTileJ + 2, kHaloN = kHaloW * kHaloH;
constexpr int kHaloSlots = (kHaloN + 255) / 256;

template <typename real, int NF>
struct Halo {
  int off[kHaloSlots];  // 2-D offsets of this thread's fill cells (-1: none)
  real pre[NF][kHaloSlots];
  __device__ Halo(const Geo<real>& g, int i0, int j0) {
#pragma unroll
    for (int s = 0; s < kHaloSlots; ++s) {
      const int q = (int)threadIdx.x + 256 * s;
      if (q < kHaloN) {
        const int r = q / kHaloW, c = q - r * kHaloW;
        int gj = j0 - 1 + r;
        gj = gj < 0 ? 0 : (gj >= g.ny ? g.ny - 1 : gj);
        int gi = i0 - 1 + c;
        gi = gi < 0 ? gi + g.nx : (gi >= g.nx ? gi - g.nx : gi);
        gi = gi >= g.nx ? gi % g.nx : gi;
        off[s] = gj * g.nx + gi;
      } else {
        off[s] = -1;
      }
    }
  }
  // registers <- level plane `pl` of the NF fields (null fields are skipped)
  __device__ __forceinline__ void load(const real* const (&f)[NF], long long pl) {
#pragma unroll
    for (int s = 0; s < kHaloSlots; ++s)
#pragma unroll
      for (int q = 0; q < NF; ++q) pre[q][s] = (off[s] >= 0 && f[q]) ? f[q][pl + off[s]] : real(0);
  }
  // LDS buffer <- registers; buffer layout [NF][kHaloN]
  __device__ __forceinline__ void store(real* buf) const {
#pragma unroll
    for (int s = 0; s < kHaloSlots; ++s) {
      const int q = (int)threadIdx.x + 256 * s;
      if (q < kHaloN) {
#pragma unroll
        for (int f = 0; f < NF; ++f) buf[f * kHaloN + q] = pre[f][s];
      }
    }
  }
};

// Remainder of _linear_half_step after the nu_h sub-cycle (step.py:111-173):
// tracer D1, biharmonics, vertical diffusion, land/ghost hold, Coriolis rotation.
// Lu0, Lv0, T0, S0, DT0 and DS0 are read at the four horizontal neighbours
// through the LDS halo tiles; every value equals the one a global load would
// give, so the arithmetic is unchanged.
template <typename real>
__global__ __launch_bounds__(256) void k_linear_finish(
    Geo<real> g, Phys<real> P, const real* __restrict__ u0, const real* __restrict__ v0,
    const real* __restrict__ T0, const real* __restrict__ S0, const real* __restrict__ unu,
    const real* __restrict__ vnu, const real* __restrict__ Lu0, const real* __restrict__ Lv0,
    const real* __restrict__ DT0, const real* __restrict__ DS0, real* __restrict__ uo,
    real* __restrict__ vo, real* __restrict__ To, real* __restrict__ So) {
  constexpr int NF = 6;
  __shared__ real hl[2][NF * kHaloN];
  ZM_COLUMN_TILE
  // every thread of the tile takes part in the halo loads and barriers; the
  // ones outside the grid work on a clamped column and store nothing
  const bool active = i < g.nx && j < g.ny;
  const int ic = i < g.nx ? i : g.nx - 1, jc = j < g.ny ? j : g.ny - 1;
  const int k0 = seg_ * g.seg;
  const int k1 = k0 + g.seg < g.nz ? k0 + g.seg : g.nz;
  const Nb n = make_nb(g, ic, jc);
  const int c2 = n.c2;
  const long long n2 = g.n2;
  const int kb = n.nwC > 0 ? n.nwC - 1 : 0;
  const real dth = P.dth;
  const real idx = g.idx[jc], idx2 = g.idx2[jc], tr = g.tanR[jc], idyc = g.idyc[jc];
  const real cfN = g.cfn[jc], cfS = n.hasS ? g.cfn[jc - 1] : real(0);
  const real kc = P.kh2[c2];
  const real kfe = real(0.5) * (kc + P.kh2[n.e2]), kfw = real(0.5) * (P.kh2[n.w2] + kc);
  const real kfn = real(0.5) * (kc + P.kh2[n.n2e]), kfs = real(0.5) * (P.kh2[n.s2e] + kc);
  const real kv = P.kv2[c2];
  const real cr = P.rot_c[jc], sr = P.rot_s[jc];
  const bool bu = P.nu_bi > real(0), bt = P.kappa_bi > real(0);
  const real* const fields[NF] = {bu ? Lu0 : nullptr, bu ? Lv0 : nullptr, T0, S0, bt ? DT0 : nullptr,
                                  bt ? DS0 : nullptr};
  Halo<real, NF> H(g, tx_ * kTileI, ty_ * kTileJ);
  const int me = ((int)(threadIdx.x / kTileI) + 1) * kHaloW + (int)(threadIdx.x % kTileI) + 1;
  H.load(fields, (long long)k0 * n2);
  H.store(hl[0]);
  __syncthreads();
  long long o = (long long)k0 * n2 + c2;
  real um = k0 > 0 ? u0[o - n2] : real(0), uc = u0[o];
  real vm = k0 > 0 ? v0[o - n2] : real(0), vc = v0[o];
  real Tm = k0 > 0 ? T0[o - n2] : real(0), Tc = T0[o];
  real Sm = k0 > 0 ? S0[o - n2] : real(0), Sc = S0[o];
  for (int k = k0; k < k1; ++k, o += n2) {
    const real* L = hl[(k - k0) & 1];
    if (k + 1 < k1) H.load(fields, (long long)(k + 1) * n2);
    const bool more = k + 1 < g.nz;
    const long long onext = more ? o + n2 : o;  // in bounds at the last level
    const real un_ = u0[onext], vn_ = v0[onext], Tn_ = T0[onext], Sn_ = S0[onext];
    const real up = more ? un_ : uc, vp = more ? vn_ : vc;
    const real Tp = more ? Tn_ : Tc, Sp = more ? Sn_ : Sc;
    const bool wc = k < n.nwC;
    const bool ge = wc && k < n.nwE, gw = wc && k < n.nwW;
    const bool gnE = wc && k < n.nwN, gsE = wc && k < n.nwS;  // mirror ghost rows
    const bool gn = n.hasN && gnE, gs = n.hasS && gsE;        // closed walls
    auto at = [&](int f, int d) { return L[f * kHaloN + me + d]; };
    real u = unu[o], v = vnu[o];
    if (bu) {
      u = u - P.nu_bi * lap_vals(at(0, 0), at(0, 1), at(0, -1), at(0, kHaloW), at(0, -kHaloW), ge, gw, gnE, gsE, idx2,
                                 g.idy2, g.hidy, tr) * dth;
      v = v - P.nu_bi * lap_vals(at(1, 0), at(1, 1), at(1, -1), at(1, kHaloW), at(1, -kHaloW), ge, gw, gnE, gsE, idx2,
                                 g.idy2, g.hidy, tr) * dth;
    }
    u = u + P.nu_v * d2z_win(um, uc, up, k, kb, g) * dth;
    v = v + P.nu_v * d2z_win(vm, vc, vp, k, kb, g) * dth;
    real T = Tc + d1_vals(Tc, at(2, 1), at(2, -1), at(2, kHaloW), at(2, -kHaloW), ge, gw, gn, gs, kfe, kfw, kfn, kfs,
                          idx, g.idy, cfN, cfS, idyc) * dth;
    real S = Sc + d1_vals(Sc, at(3, 1), at(3, -1), at(3, kHaloW), at(3, -kHaloW), ge, gw, gn, gs, kfe, kfw, kfn, kfs,
                          idx, g.idy, cfN, cfS, idyc) * dth;
    if (bt) {
      const real one = real(1);
      T = T - P.kappa_bi * d1_vals(at(4, 0), at(4, 1), at(4, -1), at(4, kHaloW), at(4, -kHaloW), ge, gw, gn, gs, one,
                                   one, one, one, idx, g.idy, cfN, cfS, idyc) * dth;
      S = S - P.kappa_bi * d1_vals(at(5, 0), at(5, 1), at(5, -1), at(5, kHaloW), at(5, -kHaloW), ge, gw, gn, gs, one,
                                   one, one, one, idx, g.idy, cfN, cfS, idyc) * dth;
    }
    T = T + vdiff_win(Tm, Tc, Tp, k, n.nwC, kb, kv, P.cons_kv != 0, g) * dth;
    S = S + vdiff_win(Sm, Sc, Sp, k, n.nwC, kb, kv, P.cons_kv != 0, g) * dth;
    if (!wc) {
      u = uc;
      v = vc;
      T = Tc;
      S = Sc;
    }
    if (active) {
      uo[o] = cr * u + sr * v;
      vo[o] = -sr * u + cr * v;
      To[o] = T;
      So[o] = S;
    }
    um = uc; uc = up; vm = vc; vc = vp;
    Tm = Tc; Tc = Tp; Sm = Sc; Sc = Sp;
    if (k + 1 < k1) H.store(hl[(k + 1 - k0) & 1]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fz: _divergence_h (horizontal.py:255-281) accumulated from the seafloor up
// (_vertical_transport_iface, transport.py:110-133).  Fz has nz+1 levels.
// ---------------------------------------------------------------------------
// _divergence_h of one level from its six loaded velocities: face gates per
// level, the wrapped y neighbour closed at the walls.
template <typename real>
__device__ __forceinline__ real div_vals(real uc, real ue, real uw, real vc, real vn, real vs, const Nb& n, int k,
                                         real idx, real idyc, real cfN, real cfS) {
  const bool wc = k < n.nwC;
  const real fxe = (wc && k < n.nwE) ? real(0.5) * (uc + ue) : real(0);
  const real fxw = (wc && k < n.nwW) ? real(0.5) * (uw + uc) : real(0);
  const real fyn = (n.hasN && wc && k < n.nwN) ? real(0.5) * (vc + vn) * cfN : real(0);
  const real fys = (n.hasS && wc && k < n.nwS) ? real(0.5) * (vs + vc) * cfS : real(0);
  return (fxe - fxw) * idx + (fyn - fys) * idyc;
}

template <typename real>
__global__ __launch_bounds__(256) void k_fz(Geo<real> g, const real* __restrict__ u,
                                            const real* __restrict__ v, real* __restrict__ Fz) {
  const int c2 = (int)(xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
  if (c2 >= g.n2) return;
  const int j = c2 / g.nx, i = c2 - j * g.nx;
  const Nb n = make_nb(g, i, j);
  real acc = 0;
  const real idx = g.idx[j], idyc = g.idyc[j], cfN = g.cfn[j], cfS = n.hasS ? g.cfn[j - 1] : real(0);
  Fz[(long long)g.nz * g.n2 + c2] = real(0);
  // dz of the next level is read before this level's store (the store could
  // alias the metric array, which would otherwise serialise every level)
  // the six velocities of each level are loaded kColAhead levels ahead
  // (clamped to level 0), so later levels are in flight during the sum
  constexpr int D = kColAhead;
  real qc[D], qe[D], qw[D], qvc[D], qn[D], qs[D];
  auto fetch = [&](int k, int d) {
    const long long pl = (long long)(k >= 0 ? k : 0) * g.n2;
    qc[d] = u[pl + n.c2];
    qe[d] = u[pl + n.e2];
    qw[d] = u[pl + n.w2];
    qvc[d] = v[pl + n.c2];
    qn[d] = v[pl + n.n2e];
    qs[d] = v[pl + n.s2e];
  };
#pragma unroll
  for (int d = 0; d < D; ++d) fetch(g.nz - 1 - d, d);
  for (int k0 = g.nz - 1; k0 >= 0; k0 -= D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int k = k0 - d;
      if (k >= 0) {  // uniform
        const real dv = div_vals(qc[d], qe[d], qw[d], qvc[d], qn[d], qs[d], n, k, idx, idyc, cfN, cfS);
        fetch(k - D, d);
        const long long plane = (long long)k * g.n2;
        acc = acc + dv * ldc(g.dzn, k);
        Fz[plane + c2] = acc;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Convective mixing: column gate (_convective_mask, physics/vertical.py:63-82)
// and the legacy sub-cycle of _conv_flux_tendency (tendencies.py:149-165,
// physics/vertical.py:13-37).  One column per lane; the working column sits in
// LDS as [k][lane].  conv buffers are written only where the gate is set.
// ---------------------------------------------------------------------------
template <typename real>
__device__ __forceinline__ real rho_anom(real T, real S, const Phys<real>& P) {
  return real(kRho0) * (real(-kAlpha) * (T - P.T_ref) + real(kBeta) * (S - P.S_ref));
}

// Levels of the column kernels split over the four waves of a block: wave q
// of a 256-thread block serves levels [q Ls, min((q + 1) Ls, nz)) of the
// block's 64 columns, Ls = ceil(nz / 4), so each column's chain of dependent
// loads is a quarter as long and four times as many waves are in flight.
constexpr int kColSegs = 4;

template <typename real>
__global__ __launch_bounds__(256) void k_conv_gate(Geo<real> g, Phys<real> P, const real* __restrict__ T,
                                                   const real* __restrict__ S, int* __restrict__ gate,
                                                   unsigned long long* __restrict__ ubits) {
  // pass 1: the column gate and the unstable interfaces k+1/2 (k < 128), one
  // column per lane and no LDS but the combine: wave q tests the interfaces
  // above levels [max(a, 1), b) of its segment and the four waves' flags are
  // OR-ed (exact) in LDS
  __shared__ unsigned long long sb[kColSegs][2][64];
  __shared__ int sun[kColSegs][64];
  const int lane = (int)threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int c2 = (int)(xcd_tile(blockIdx.x, gridDim.x) * 64 + lane);
  const bool live = c2 < g.n2;
  const int Ls = (g.nz + kColSegs - 1) / kColSegs;
  const int a = q * Ls > 1 ? q * Ls : 1, b = (q + 1) * Ls < g.nz ? (q + 1) * Ls : g.nz;  // interfaces above [a, b)
  bool unstable = false;
  unsigned long long ub0 = 0, ub1 = 0;
  if (live && a < b) {  // a, b uniform
    const int nw = g.nwet[c2];
    const long long n2 = g.n2;
    real rprev = rho_anom(T[(long long)(a - 1) * n2 + c2], S[(long long)(a - 1) * n2 + c2], P);
    // T, S loaded kColAhead levels ahead (clamped to the segment's last
    // level) so later levels are in flight while this one is tested; levels
    // at or below nw never set a bit
    constexpr int D = kColAhead;
    real Tq[D], Sq[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const long long o = (long long)(a + d < b ? a + d : b - 1) * n2 + c2;
      Tq[d] = T[o];
      Sq[d] = S[o];
    }
    for (int k0 = a; k0 < b; k0 += D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int k = k0 + d;
        if (k < b) {  // uniform
          const real Tk = Tq[d], Sk = Sq[d];
          const long long on = (long long)(k + D < b ? k + D : b - 1) * n2 + c2;
          Tq[d] = T[on];
          Sq[d] = S[on];
          const real r = rho_anom(Tk, Sk, P);
          const bool un = k < nw && rprev > r;
          unstable |= un;
          if (k - 1 < 64) ub0 |= (unsigned long long)un << (k - 1);
          else ub1 |= (unsigned long long)un << ((k - 65) & 63);
          rprev = r;
        }
      }
    }
  }
  sun[q][lane] = unstable ? 1 : 0;
  sb[q][0][lane] = ub0;
  sb[q][1][lane] = ub1;
  __syncthreads();
  if (q == 0 && live) {
    int any = 0;
    unsigned long long w0 = 0, w1 = 0;
#pragma unroll
    for (int r = 0; r < kColSegs; ++r) {
      any |= sun[r][lane];
      w0 |= sb[r][0][lane];
      w1 |= sb[r][1][lane];
    }
    const bool on = any && P.kappa_conv > real(0);
    gate[c2] = on ? 1 : 0;
    if (on && P.local_conv) {
      ubits[2 * (long long)c2] = w0;
      ubits[2 * (long long)c2 + 1] = w1;
    }
  }
}

template <typename real>
__global__ __launch_bounds__(64) void k_conv(Geo<real> g, Phys<real> P, const real* __restrict__ T,
                                             const real* __restrict__ S, real* __restrict__ cT,
                                             real* __restrict__ cS, const int* __restrict__ gate,
                                             const unsigned long long* __restrict__ ubits) {
  // pass 2, gated columns only: the working column sits in LDS as [k][lane]
  extern __shared__ unsigned char zm_lds[];
  real* tT = reinterpret_cast<real*>(zm_lds);
  real* tS = tT + (size_t)g.nz * 64;
  real* aT = tS + (size_t)g.nz * 64;
  real* aS = aT + (size_t)g.nz * 64;
  const int lane = threadIdx.x;
  const int c2 = blockIdx.x * 64 + lane;
  if (c2 >= g.n2 || !gate[c2]) return;
  const int nw = g.nwet[c2];
  const int n2 = g.n2;
  // stage the wet part of the column, eight levels per batch of loads (the
  // indices clamped to the deepest wet level, nw >= 1 on a gated column)
  for (int k0 = 0; k0 < nw; k0 += 8) {
    real tq[8], sq[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      const long long o = (long long)(k0 + d < nw ? k0 + d : nw - 1) * n2 + c2;
      tq[d] = T[o];
      sq[d] = S[o];
    }
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      if (k0 + d < nw) {
        tT[(k0 + d) * 64 + lane] = tq[d];
        tS[(k0 + d) * 64 + lane] = sq[d];
      }
    }
  }
  unsigned long long ub0 = 0, ub1 = 0;  // unstable interfaces, frozen over the sub-cycle
  if (P.local_conv) {
    ub0 = ubits[2 * (long long)c2];
    ub1 = ubits[2 * (long long)c2 + 1];
  }
  const int nc = P.conv_nsub;
  const real kap = nc > 1 ? P.kappa_conv / real(nc) : P.kappa_conv;
  const real h = P.dt / real(nc);
  for (int s = 0; s < nc; ++s) {
    real fTprev = 0, fSprev = 0;
    for (int k = 0; k < g.nz; ++k) {
      real fT = 0, fS = 0, et = 0, es = 0;
      if (k < nw) {
        const real ot = tT[k * 64 + lane], os = tS[k * 64 + lane];
        const bool open = !P.local_conv || ((k < 64 ? (ub0 >> k) : (ub1 >> (k - 64))) & 1ull);
        if (k + 1 < nw && open) {
          fT = (-kap * (tT[(k + 1) * 64 + lane] - ot)) * ldc(g.idzi, k);
          fS = (-kap * (tS[(k + 1) * 64 + lane] - os)) * ldc(g.idzi, k);
        }
        et = (fTprev - fT) * ldc(g.idzn, k);
        es = (fSprev - fS) * ldc(g.idzn, k);
        tT[k * 64 + lane] = ot + et * h;
        tS[k * 64 + lane] = os + es * h;
      }
      aT[k * 64 + lane] = s == 0 ? et : aT[k * 64 + lane] + et;
      aS[k * 64 + lane] = s == 0 ? es : aS[k * 64 + lane] + es;
      fTprev = fT;
      fSprev = fS;
    }
  }
  for (int k = 0; k < g.nz; ++k) {
    const long long o = (long long)k * n2 + c2;
    cT[o] = nc > 1 ? aT[k * 64 + lane] / real(nc) : aT[k * 64 + lane];
    cS[o] = nc > 1 ? aS[k * 64 + lane] / real(nc) : aS[k * 64 + lane];
  }
}

// Pass 2 for a sub-cycle of NC <= 4 substeps, in registers: substep s updates
// level L from substep s-1's values at L-1, L, L+1 only, so one pass down the
// column runs substep s NC-1 ... 0 levels behind substep 0 (at step m, substep
// s is at level m - s).  Each level receives the arithmetic of k_conv, with
// its tendencies summed in substep order; no LDS, so the kernel runs at full
// occupancy.  The levels are split over the block's four waves (kColSegs):
// wave q outputs levels [a, b) and starts its pass NC - 1 levels higher, at
// s0 = a - (NC - 1) with the exact substep-0 flux above s0; substep s is then
// exact from level s0 + s down, so every output level is (the values above
// are finite and never stored).  T and S are loaded kColAhead levels ahead,
// clamped to the deepest wet level.  Blocks map to columns in launch order,
// which the hardware deals round-robin over the 8 XCDs: the gated columns
// cluster in latitude, and XCD-contiguous tiles would leave most of them to
// one or two XCDs.
template <typename real, int NC>
__global__ __launch_bounds__(256) void k_conv_stream(Geo<real> g, Phys<real> P, const real* __restrict__ T,
                                                     const real* __restrict__ S, real* __restrict__ cT,
                                                     real* __restrict__ cS, const int* __restrict__ gate,
                                                     const unsigned long long* __restrict__ ubits) {
  const int lane = (int)threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int c2 = (int)blockIdx.x * 64 + lane;
  const int Ls = (g.nz + kColSegs - 1) / kColSegs;
  const int a = q * Ls, b = (q + 1) * Ls < g.nz ? (q + 1) * Ls : g.nz;  // uniform
  if (c2 >= g.n2 || a >= b || !gate[c2]) return;
  const int nw = g.nwet[c2];  // >= 2 on a gated column
  const long long n2 = g.n2;
  if (a >= nw) {  // a dry segment: every tendency is 0
    for (int L = a; L < b; ++L) {
      cT[(long long)L * n2 + c2] = real(0);
      cS[(long long)L * n2 + c2] = real(0);
    }
    return;
  }
  unsigned long long ub0 = ~0ull, ub1 = ~0ull;  // unstable interfaces, frozen over the sub-cycle
  if (P.local_conv) {
    ub0 = ubits[2 * (long long)c2];
    ub1 = ubits[2 * (long long)c2 + 1];
  }
  auto open_at = [&](int L) { return ((L < 64 ? (ub0 >> L) : (ub1 >> ((L - 64) & 63))) & 1ull) != 0; };
  const real kap = NC > 1 ? P.kappa_conv / real(NC) : P.kappa_conv;
  const real h = P.dt / real(NC);
  const int s0 = a - (NC - 1) > 0 ? a - (NC - 1) : 0;  // uniform; s0 < nw
  // per substep s, before step m: cur = t^s[m-s], fp = F^s[m-s-1] (0 above the
  // surface), acc = tendency sum of level m-s over substeps 0..s
  real curT[NC], curS[NC], fpT[NC], fpS[NC], acT[NC], acS[NC];
#pragma unroll
  for (int s = 0; s < NC; ++s) {
    curT[s] = curS[s] = fpT[s] = fpS[s] = acT[s] = acS[s] = real(0);
  }
  curT[0] = T[(long long)s0 * n2 + c2];
  curS[0] = S[(long long)s0 * n2 + c2];
  if (s0 > 0 && open_at(s0 - 1)) {  // substep 0's flux above s0 (s0 < nw)
    const real tu = T[(long long)(s0 - 1) * n2 + c2], su = S[(long long)(s0 - 1) * n2 + c2];
    fpT[0] = (-kap * (curT[0] - tu)) * ldc(g.idzi, s0 - 1);
    fpS[0] = (-kap * (curS[0] - su)) * ldc(g.idzi, s0 - 1);
  }
  constexpr int D = kColAhead;
  real Tq[D], Sq[D];  // t^0 of levels m+1 .. m+D
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const long long o = (long long)(s0 + 1 + d < nw ? s0 + 1 + d : nw - 1) * n2 + c2;
    Tq[d] = T[o];
    Sq[d] = S[o];
  }
  const int mend = b + NC - 1;
  for (int m0 = s0; m0 < mend; m0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int m = m0 + d;
      if (m < mend) {  // uniform
        real upT = Tq[d], upS = Sq[d];
        const long long on = (long long)(m + 1 + D < nw ? m + 1 + D : nw - 1) * n2 + c2;
        Tq[d] = T[on];
        Sq[d] = S[on];
        real aoT[NC], aoS[NC];
#pragma unroll
        for (int s = 0; s < NC; ++s) {
          aoT[s] = acT[s];
          aoS[s] = acS[s];
        }
#pragma unroll
        for (int s = 0; s < NC; ++s) {
          const int L = m - s;  // uniform
          const int Lz = L < 0 ? 0 : (L < g.nz - 1 ? L : g.nz - 1);
          const int Li = Lz < g.nz - 2 ? Lz : (g.nz >= 2 ? g.nz - 2 : 0);
          const real ot = curT[s], os = curS[s];
          real fT = 0, fS = 0, et = 0, es = 0;
          if (L >= 0 && L < nw) {
            if (L + 1 < nw && open_at(L)) {
              fT = (-kap * (upT - ot)) * ldc(g.idzi, Li);
              fS = (-kap * (upS - os)) * ldc(g.idzi, Li);
            }
            et = (fpT[s] - fT) * ldc(g.idzn, Lz);
            es = (fpS[s] - fS) * ldc(g.idzn, Lz);
          }
          const real at = s == 0 ? et : aoT[s - 1] + et;
          const real as = s == 0 ? es : aoS[s - 1] + es;
          acT[s] = at;
          acS[s] = as;
          fpT[s] = fT;
          fpS[s] = fS;
          curT[s] = upT;
          curS[s] = upS;
          upT = ot + et * h;  // t^{s+1}[L], the level below substep s+1's
          upS = os + es * h;
          if (s == NC - 1 && L >= a && L < b) {
            const long long o = (long long)L * n2 + c2;
            cT[o] = NC > 1 ? at / real(NC) : at;
            cS[o] = NC > 1 ? as / real(NC) : as;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Tracer residual (tendencies.py:94-245, legacy branches) at one node, fused
// with the RK2 update of _explicit_full_step (step.py:258-307).
//   mode 0: dT1 -> dT1buf, T_pred = T + dT1 dt
//   mode 1: T_new = Tbase + 0.5 (dT1 + dT2) dt
// ---------------------------------------------------------------------------
template <typename real>
__device__ __forceinline__ real adv_scalar(const real* __restrict__ X, const real* __restrict__ u,
                                           const real* __restrict__ v, const real* __restrict__ Fz,
                                           long long plane, const Nb& n, int k, const Geo<real>& g) {
  // _legacy_horizontal_tracer_fluxes (transport.py:224-290) + donor-cell
  // vertical flux (transport.py:444-475).
  const bool wc = k < n.nwC;
  if (!wc) return real(0);
  const real xc = X[plane + n.c2], uc = u[plane + n.c2], vc = v[plane + n.c2];
  real fxe = 0, fxw = 0, fyn = 0, fys = 0;
  if (k < n.nwE) fxe = (real(0.5) * (uc + u[plane + n.e2])) * (real(0.5) * (xc + X[plane + n.e2]));
  if (k < n.nwW) fxw = (real(0.5) * (u[plane + n.w2] + uc)) * (real(0.5) * (X[plane + n.w2] + xc));
  if (n.hasN && k < n.nwN)
    fyn = (real(0.5) * (vc + v[plane + n.n2e])) * (real(0.5) * (xc + X[plane + n.n2e])) * g.cfn[n.j];
  if (n.hasS && k < n.nwS)
    fys = (real(0.5) * (v[plane + n.s2e] + vc)) * (real(0.5) * (X[plane + n.s2e] + xc)) * g.cfn[n.j - 1];
  const real hdiv = (fxe - fxw) * g.idx[n.j] + (fyn - fys) * g.idyc[n.j];
  real up, dn = 0;
  const real fzk = Fz[plane + n.c2];
  if (k == 0) {
    up = fzk * xc;
  } else {
    // both k-1 and k are wet here (k < nwC)
    up = fzk * (fzk > real(0) ? X[plane - g.n2 + n.c2] : xc);
  }
  if (k + 1 < n.nwC) {
    const real fzn = Fz[plane + g.n2 + n.c2];
    dn = fzn * (fzn > real(0) ? xc : X[plane + g.n2 + n.c2]);
  }
  const real divz = (dn - up) * ldc(g.idzn, k);
  return -(hdiv + divz);
}

//
// LED = true is the ledger instantiation (diagnostics/budgets.py:141-186 fed by
// tendencies.py:211-214): it writes no state but the per-block float64 sums
// of the surface sources, the advection and convection tendencies (signed and
// absolute, times the node volume) and the surface advective flux Fz_top*T_top
// times the cell area, kLedgerTracerSums values per block in lpart.
// TERMS = true is the diagnostic instantiation of _tracer_terms
// (tendencies.py:74-92): it writes the temperature's advection, horizontal
// diffusion, vertical diffusion and convection tendencies, masked to wet
// nodes, into dT1, dS1, To and So (in that order) and nothing else; cT must
// then hold the single-pass convective tendency (conv_nsub = 1, kappa_conv).
template <typename real, int ADV, bool LED, bool TERMS = false>
__global__ __launch_bounds__(256) void k_tracer_residual(
    Geo<real> g, Phys<real> P, int mode, const real* __restrict__ u, const real* __restrict__ v,
    const real* __restrict__ T, const real* __restrict__ S, const real* __restrict__ Fz,
    const real* __restrict__ cT, const real* __restrict__ cS, const int* __restrict__ gate,
    real* __restrict__ dT1, real* __restrict__ dS1, const real* __restrict__ Tb,
    const real* __restrict__ Sb, real* __restrict__ To, real* __restrict__ So,
    const double* __restrict__ area_row, const double* __restrict__ dzn64, double* __restrict__ lpart) {
  double lacc[LED ? kLedgerTracerSums : 1];
  if constexpr (LED)
    for (int q = 0; q < kLedgerTracerSums; ++q) lacc[q] = 0.0;
  ZM_COLUMN_TILE
  const bool inside = i < g.nx && j < g.ny;
  if constexpr (!LED) {
    if (!inside) return;
  }
  if (inside) {
  ZM_COLUMN_SEGMENT
  const real idx = g.idx[j], idyc = g.idyc[j];
  const real cfN = g.cfn[j], cfS = n.hasS ? g.cfn[j - 1] : real(0);
  const real kc = P.kh2[c2];
  const real kfe = real(0.5) * (kc + P.kh2[n.e2]), kfw = real(0.5) * (P.kh2[n.w2] + kc);
  const real kfn = real(0.5) * (kc + P.kh2[n.n2e]), kfs = real(0.5) * (P.kh2[n.s2e] + kc);
  const real kv = P.kv2[c2];
  const bool gated = gate[c2] != 0;
  const real h = real(0.5);
  // second neighbours for the minmod reconstruction (x periodic, y edge-clamped)
  int ee2 = 0, ww2 = 0, nn2 = 0, ss2 = 0, nwEE = 0, nwWW = 0, nwNN = 0, nwSS = 0;
  if (ADV == 2) {
    ee2 = j * g.nx + wrap_e(wrap_e(i, g.nx), g.nx);
    ww2 = j * g.nx + wrap_w(wrap_w(i, g.nx), g.nx);
    nn2 = (j + 2 < g.ny ? j + 2 : g.ny - 1) * g.nx + i;
    ss2 = (j >= 2 ? j - 2 : 0) * g.nx + i;
    nwEE = g.nwet[ee2];
    nwWW = g.nwet[ww2];
    nwNN = g.nwet[nn2];
    nwSS = g.nwet[ss2];
  }
  long long o = (long long)k0 * n2 + c2;
  real Tm = k0 > 0 ? T[o - n2] : real(0), Tc = T[o];
  real Sm = k0 > 0 ? S[o - n2] : real(0), Sc = S[o];
  real fz0 = Fz[o];
  // the horizontal loads of a level (all in bounds) are issued one level
  // ahead of their use; the last level of a segment reloads itself
  struct Lv {
    real TE, TW, TN, TS, SE, SW, SN, SS, uc, vc, uE, uW, vN, vS, cT, cS;
    real TEE, TWW, TNN, TSS, SEE, SWW, SNN, SSS, Tb, Sb, dT1, dS1;
  };
  auto load = [&](long long oo, Lv& L) {
    const long long pl = oo - c2;
    L.TE = T[pl + n.e2], L.TW = T[pl + n.w2], L.TN = T[pl + n.n2e], L.TS = T[pl + n.s2e];
    L.SE = S[pl + n.e2], L.SW = S[pl + n.w2], L.SN = S[pl + n.n2e], L.SS = S[pl + n.s2e];
    L.uc = u[oo], L.vc = v[oo];
    L.uE = u[pl + n.e2], L.uW = u[pl + n.w2], L.vN = v[pl + n.n2e], L.vS = v[pl + n.s2e];
    L.cT = cT[oo], L.cS = cS[oo];
    if (ADV == 2) {
      L.TEE = T[pl + ee2], L.TWW = T[pl + ww2], L.TNN = T[pl + nn2], L.TSS = T[pl + ss2];
      L.SEE = S[pl + ee2], L.SWW = S[pl + ww2], L.SNN = S[pl + nn2], L.SSS = S[pl + ss2];
    }
    if (!TERMS && !LED && mode != 0) {  // the second stage's inputs (never written by it)
      L.Tb = Tb[oo], L.Sb = Sb[oo], L.dT1 = dT1[oo], L.dS1 = dS1[oo];
    }
  };
  Lv cur, nxt;
  load(o, cur);
  for (int k = k0; k < k1; ++k, o += n2) {
    const bool more = k + 1 < g.nz;
    const long long onext = more ? o + n2 : o;  // in bounds at the last level
    const real Tn_ = T[onext], Sn_ = S[onext];
    const real Tp = more ? Tn_ : Tc, Sp = more ? Sn_ : Sc;
    const real fz1 = Fz[o + n2];
    load(k + 1 < k1 ? o + n2 : o, nxt);
    const bool wc = k < n.nwC;
    const bool ge = wc && k < n.nwE, gw = wc && k < n.nwW;
    const bool gn = n.hasN && wc && k < n.nwN, gs = n.hasS && wc && k < n.nwS;
    const real TE = cur.TE, TW = cur.TW, TN = cur.TN, TS = cur.TS;
    const real SE = cur.SE, SW = cur.SW, SN = cur.SN, SS = cur.SS;
    const real dhT = d1_vals(Tc, TE, TW, TN, TS, ge, gw, gn, gs, kfe, kfw, kfn, kfs, idx, g.idy, cfN, cfS, idyc);
    const real dhS = d1_vals(Sc, SE, SW, SN, SS, ge, gw, gn, gs, kfe, kfw, kfn, kfs, idx, g.idy, cfN, cfS, idyc);
    const real dvT = vdiff_win(Tm, Tc, Tp, k, n.nwC, kb, kv, P.cons_kv != 0, g);
    const real dvS = vdiff_win(Sm, Sc, Sp, k, n.nwC, kb, kv, P.cons_kv != 0, g);
    // the velocity and convection loads of the level are in bounds: issued
    // unconditionally, used on wet nodes
    const real uc = cur.uc, vc = cur.vc;
    const real uE = cur.uE, uW = cur.uW, vN = cur.vN, vS = cur.vS;
    const real cTo = cur.cT, cSo = cur.cS;
    real dTdt = 0, dSdt = 0;
    real term_adv = 0, term_conv = 0;
    if (wc) {
      // centred face-gated horizontal fluxes (transport.py:224-290)
      const real uxe = h * (uc + uE), uxw = h * (uW + uc);
      const real vyn = h * (vc + vN), vys = h * (vS + vc);
      const real ufe = ge ? uxe : real(0);
      const real ufw = gw ? uxw : real(0);
      const real vfn = gn ? vyn : real(0);
      const real vfs = gs ? vys : real(0);
      real tE, tW, tN, tS, sE_, sW_, sN_, sS_;
      if (ADV == 0) {  // centred face values
        tE = h * (Tc + TE), tW = h * (TW + Tc), tN = h * (Tc + TN), tS = h * (TS + Tc);
        sE_ = h * (Sc + SE), sW_ = h * (SW + Sc), sN_ = h * (Sc + SN), sS_ = h * (SS + Sc);
      } else if (ADV == 1) {  // donor cell (monotone_adv)
        tE = uxe >= real(0) ? Tc : TE, tW = uxw >= real(0) ? TW : Tc;
        tN = vyn >= real(0) ? Tc : TN, tS = vys >= real(0) ? TS : Tc;
        sE_ = uxe >= real(0) ? Sc : SE, sW_ = uxw >= real(0) ? SW : Sc;
        sN_ = vyn >= real(0) ? Sc : SN, sS_ = vys >= real(0) ? SS : Sc;
      } else {  // minmod-limited upwind reconstruction (fct_adv)
        const bool wC = true, wE = k < n.nwE, wW = k < n.nwW, wN = k < n.nwN, wS = k < n.nwS;
        const bool wEE = k < nwEE, wWW = k < nwWW, wNN = k < nwNN, wSS = k < nwSS;
        const real TEE = cur.TEE, TWW = cur.TWW, TNN = cur.TNN, TSS = cur.TSS;
        const real SEE = cur.SEE, SWW = cur.SWW, SNN = cur.SNN, SSS_ = cur.SSS;
        const real tsx = minmod_slope(TW, Tc, TE, wW, wC, wE), tsE = minmod_slope(Tc, TE, TEE, wC, wE, wEE);
        const real tsW = minmod_slope(TWW, TW, Tc, wWW, wW, wC);
        const real tsy = minmod_slope(TS, Tc, TN, wS, wC, wN), tsN = minmod_slope(Tc, TN, TNN, wC, wN, wNN);
        const real tsS = minmod_slope(TSS, TS, Tc, wSS, wS, wC);
        tE = uxe >= real(0) ? Tc + h * tsx : TE - h * tsE;
        tW = uxw >= real(0) ? TW + h * tsW : Tc - h * tsx;
        tN = vyn >= real(0) ? Tc + h * tsy : TN - h * tsN;
        tS = vys >= real(0) ? TS + h * tsS : Tc - h * tsy;
        const real ssx = minmod_slope(SW, Sc, SE, wW, wC, wE), ssE = minmod_slope(Sc, SE, SEE, wC, wE, wEE);
        const real ssW = minmod_slope(SWW, SW, Sc, wWW, wW, wC);
        const real ssy = minmod_slope(SS, Sc, SN, wS, wC, wN), ssN = minmod_slope(Sc, SN, SNN, wC, wN, wNN);
        const real ssS = minmod_slope(SSS_, SS, Sc, wSS, wS, wC);
        sE_ = uxe >= real(0) ? Sc + h * ssx : SE - h * ssE;
        sW_ = uxw >= real(0) ? SW + h * ssW : Sc - h * ssx;
        sN_ = vyn >= real(0) ? Sc + h * ssy : SN - h * ssN;
        sS_ = vys >= real(0) ? SS + h * ssS : Sc - h * ssy;
      }
      const real hT = (ufe * tE - ufw * tW) * idx + (vfn * tN * cfN - vfs * tS * cfS) * idyc;
      const real hS = (ufe * sE_ - ufw * sW_) * idx + (vfn * sN_ * cfN - vfs * sS_ * cfS) * idyc;
      // donor-cell vertical flux on Fz (transport.py:444-475)
      const real upT = k == 0 ? fz0 * Tc : fz0 * (fz0 > real(0) ? Tm : Tc);
      const real upS = k == 0 ? fz0 * Sc : fz0 * (fz0 > real(0) ? Sm : Sc);
      real dnT = 0, dnS = 0;
      if (k + 1 < n.nwC) {
        dnT = fz1 * (fz1 > real(0) ? Tc : Tp);
        dnS = fz1 * (fz1 > real(0) ? Sc : Sp);
      }
      const real advT = -(hT + (dnT - upT) * ldc(g.idzn, k));
      const real advS = -(hS + (dnS - upS) * ldc(g.idzn, k));
      const real convT = gated ? cTo : real(0);
      const real convS = gated ? cSo : real(0);
      term_adv = advT;
      term_conv = convT;
      dTdt = advT + dhT + dvT;
      dSdt = advS + dhS + dvS;
      if constexpr (LED) {
        const double vol = area_row[j] * dzn64[k];
        const double aT = (double)advT * vol, aS = (double)advS * vol;
        const double qT = (double)convT * vol, qS = (double)convS * vol;
        lacc[6] += aT;
        lacc[7] += aS;
        lacc[8] += fabs(aT);
        lacc[9] += fabs(aS);
        lacc[10] += qT;
        lacc[11] += qS;
        lacc[12] += fabs(qT);
        lacc[13] += fabs(qS);
      }
      if (k == 0) {
        const real hf = P.heat_factor;
        const real ta = P.t_atm[c2];
        if constexpr (LED) {
          const double vol = area_row[j] * dzn64[0];
          lacc[0] += (double)(P.q_heat[c2] * hf) * vol;
          lacc[1] += (double)(P.lambda_bulk * (ta - Tc) * hf) * vol;
          lacc[2] += (double)(P.cbl[c2] * (ta - Tc) * hf) * vol;
          lacc[3] += (double)(P.crc[c2] * (P.crT[c2] - Tc)) * vol;
          lacc[4] += (double)(P.restore_coef_S * (P.s_ref2[c2] - Sc)) * vol;
          lacc[5] += (double)(P.ice_salt_flux * (Tc <= P.ice_freeze ? real(1) : real(0))) * vol;
          lacc[14] += (double)(fz0 * Tc) * area_row[j];
          lacc[15] += (double)(fz0 * Sc) * area_row[j];
        }
        dTdt = dTdt + P.q_heat[c2] * hf;
        dTdt = dTdt + P.lambda_bulk * (ta - Tc) * hf;
        dTdt = dTdt + P.cbl[c2] * (ta - Tc) * hf;
        dTdt = dTdt + convT;
        dTdt = dTdt + P.crc[c2] * (P.crT[c2] - Tc);
        dSdt = dSdt + convS;
        dSdt = dSdt + P.restore_coef_S * (P.s_ref2[c2] - Sc);
        dSdt = dSdt + P.ice_salt_flux * (Tc <= P.ice_freeze ? real(1) : real(0));
      } else {
        dTdt = dTdt + convT;
        dSdt = dSdt + convS;
      }
    }
    dTdt = dTdt - dhT;
    dSdt = dSdt - dhS;
    dTdt = dTdt - dvT;
    dSdt = dSdt - dvS;
    if constexpr (TERMS) {
      dT1[o] = term_adv;
      dS1[o] = wc ? dhT : real(0);
      To[o] = wc ? dvT : real(0);
      So[o] = term_conv;
    } else if constexpr (!LED) {
      if (mode == 0) {
        dT1[o] = dTdt;
        dS1[o] = dSdt;
        To[o] = Tc + dTdt * P.dt;
        So[o] = Sc + dSdt * P.dt;
      } else {
        To[o] = cur.Tb + real(0.5) * (cur.dT1 + dTdt) * P.dt;
        So[o] = cur.Sb + real(0.5) * (cur.dS1 + dSdt) * P.dt;
      }
    }
    Tm = Tc; Tc = Tp; Sm = Sc; Sc = Sp;
    fz0 = fz1;
    cur = nxt;
  }
  }
  if constexpr (LED) {
    __shared__ double red[4][kLedgerTracerSums];
    for (int q = 0; q < kLedgerTracerSums; ++q)
      for (int off = 32; off > 0; off >>= 1) lacc[q] += __shfl_down(lacc[q], off, 64);
    if ((threadIdx.x & 63) == 0)
      for (int q = 0; q < kLedgerTracerSums; ++q) red[threadIdx.x >> 6][q] = lacc[q];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int q = 0; q < kLedgerTracerSums; ++q)
        lpart[(long long)blockIdx.x * kLedgerTracerSums + q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
  }
}

// ---------------------------------------------------------------------------
// Baroclinic pressure: trapezoid prefix sum of the masked density anomaly
// from the surface down (_baroclinic_pressure, dynamics/pressure.py:32-42).
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(256) void k_pbc(Geo<real> g, Phys<real> P, const real* __restrict__ T,
                                             const real* __restrict__ S, real* __restrict__ pbc) {
  const int c2 = (int)(xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
  if (c2 >= g.n2) return;
  const int nw = g.nwet[c2];
  real rprev = 0 < nw ? rho_anom(T[c2], S[c2], P) : real(0);
  real acc = 0;
  pbc[c2] = real(0);
  // T, S of every level are loaded kColAhead levels ahead (clamped to the
  // last level), so the loads of later levels are in flight while this one
  // is summed and no load sits behind a wet-level branch
  constexpr int D = kColAhead;
  real Tq[D], Sq[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const long long o = (long long)(1 + d < g.nz ? 1 + d : g.nz - 1) * g.n2 + c2;
    Tq[d] = T[o];
    Sq[d] = S[o];
  }
  for (int k0 = 1; k0 < g.nz; k0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int k = k0 + d;
      if (k < g.nz) {  // uniform
        const real To = Tq[d], So = Sq[d];
        const long long on = (long long)(k + D < g.nz ? k + D : g.nz - 1) * g.n2 + c2;
        Tq[d] = T[on];
        Sq[d] = S[on];
        const long long o = (long long)k * g.n2 + c2;
        const real r = k < nw ? rho_anom(To, So, P) : real(0);
        acc = acc + real(kG) * (real(0.5) * (rprev + r)) * ldc(g.dz3, k - 1);
        pbc[o] = acc;
        rprev = r;
      }
    }
  }
}

// Depth-weighted baroclinic pressure gradient F_rho (_compute_bt_rho_pgf legacy,
// pressure.py:170-180); with ubt0 != nullptr also the legacy depth mean
// velocity (_barotropic_velocity, transport.py:135-143).
template <typename real>
__global__ __launch_bounds__(256) void k_bt_rho(Geo<real> g, Phys<real> P, const real* __restrict__ pbc,
                                                real* __restrict__ Fx, real* __restrict__ Fy,
                                                const real* __restrict__ u, const real* __restrict__ v,
                                                real* __restrict__ ubt0, real* __restrict__ vbt0) {
  const int c2 = (int)(xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
  if (c2 >= g.n2) return;
  const int j = c2 / g.nx, i = c2 - j * g.nx;
  const Nb n = make_nb(g, i, j);
  // the five pbc values of each level are loaded kColAhead levels ahead
  // (clamped to the last level; the north/south rows edge-clamped and unused
  // at the walls), so later levels' loads are in flight during the sum
  constexpr int D = kColAhead;
  real qc[D], qe[D], qw[D], qn[D], qs[D];
  auto fetch = [&](int k, int d) {
    const long long pl = (long long)(k < g.nz ? k : g.nz - 1) * g.n2;
    qc[d] = pbc[pl + n.c2];
    qe[d] = pbc[pl + n.e2];
    qw[d] = pbc[pl + n.w2];
    qn[d] = pbc[pl + n.n2e];
    qs[d] = pbc[pl + n.s2e];
  };
#pragma unroll
  for (int d = 0; d < D; ++d) fetch(d, d);
  real gxp = 0, gyp = 0;
  real sx = 0, sy = 0;
  for (int k0 = 0; k0 < g.nz; k0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int k = k0 + d;
      if (k < g.nz) {  // uniform
        real gxn, gyn;
        grad_cons(qc[d], qe[d], qw[d], qn[d], qs[d], n, k, g, gxn, gyn);
        fetch(k + D, d);
        if (k > 0 && k < n.nwC) {
          sx = sx + real(0.5) * (gxp + gxn) * ldc(g.dz3, k - 1);
          sy = sy + real(0.5) * (gyp + gyn) * ldc(g.dz3, k - 1);
        }
        gxp = gxn;
        gyp = gyn;
      }
    }
  }
  Fx[c2] = -sx * P.inv_rho0H;
  Fy[c2] = -sy * P.inv_rho0H;
  if (ubt0) {
    real su = 0, sv = 0;
    real up = u[c2], vp = v[c2];
#pragma unroll 8
    for (int k = 0; k + 1 < g.nz; ++k) {
      const long long o = (long long)(k + 1) * g.n2 + c2;
      const real un = u[o], vn = v[o];
      su = su + real(0.5) * (up + un) * ldc(g.dznorm, k);
      sv = sv + real(0.5) * (vp + vn) * ldc(g.dznorm, k);
      up = un;
      vp = vn;
    }
    ubt0[c2] = su;
    vbt0[c2] = sv;
  }
}

// ---------------------------------------------------------------------------
// Momentum advection before de-aliasing (_advection_flux_form,
// transport.py:169-197): -(u d/dx + v d/dy + w d/dz) with the face-gated
// gradient (horizontal.py:331-360) and w = -Fz * wet (transport.py:149-166).
// ---------------------------------------------------------------------------
template <typename real>
__device__ __forceinline__ void grad_gated(const real* __restrict__ f, long long plane, const Nb& n,
                                           int k, const Geo<real>& g, real& gx, real& gy) {
  const real fc = f[plane + n.c2];
  const bool wc = k < n.nwC;
  const real dxp = (wc && k < n.nwE) ? (f[plane + n.e2] - fc) : real(0);
  const real dxm = (wc && k < n.nwW) ? (fc - f[plane + n.w2]) : real(0);
  gx = g.idx[n.j] * real(0.5) * (dxp + dxm);
  const real dyp = (wc && k < n.nwN) ? (f[plane + n.n2e] - fc) : real(0);
  const real dym = (wc && k < n.nwS) ? (fc - f[plane + n.s2e]) : real(0);
  gy = g.idy * real(0.5) * (dyp + dym);
}

template <typename real>
__device__ __forceinline__ real d_dz_filled(const real* __restrict__ f, int c2, int n2, int k, int nw,
                                            const Geo<real>& g) {
  const int kb = nw > 0 ? nw - 1 : 0;
  auto val = [&](int kk) { return f[(long long)(kk < kb ? kk : kb) * n2 + c2]; };
  if (k == 0) return (val(1) - val(0)) * g.idz_top;
  if (k == g.nz - 1) return (val(g.nz - 1) - val(g.nz - 2)) * g.idz_bot;
  return (val(k + 1) - val(k - 1)) * ldc(g.iddz, k - 1);
}

template <typename real>
__global__ __launch_bounds__(256) void k_mom_adv(Geo<real> g, const real* __restrict__ u,
                                                 const real* __restrict__ v, const real* __restrict__ Fz,
                                                 real* __restrict__ au, real* __restrict__ av) {
  ZM_COLUMN_PROLOGUE
  const real hidx = real(0.5) * g.idx[j], hidy = real(0.5) * g.idy;
  long long o = (long long)k0 * n2 + c2;
  real um = k0 > 0 ? u[o - n2] : real(0), uc = u[o];
  real vm = k0 > 0 ? v[o - n2] : real(0), vc = v[o];
  for (int k = k0; k < k1; ++k, o += n2) {
    // all loads in bounds (the level below is clamped at the last level):
    // issued unconditionally, then selected
    const bool more = k + 1 < g.nz;
    const long long onext = more ? o + n2 : o;
    const real un_ = u[onext], vn_ = v[onext];
    const real up = more ? un_ : uc, vp = more ? vn_ : vc;
    const long long pl = o - c2;
    const real fz = Fz[o];
    const real uE = u[pl + n.e2], uW = u[pl + n.w2], uN = u[pl + n.n2e], uS = u[pl + n.s2e];
    const real vE = v[pl + n.e2], vW = v[pl + n.w2], vN = v[pl + n.n2e], vS = v[pl + n.s2e];
    const bool wc = k < n.nwC;
    const bool ge = wc && k < n.nwE, gw = wc && k < n.nwW;
    const bool gn = wc && k < n.nwN, gs = wc && k < n.nwS;
    const real w = wc ? -fz : real(0);
    const real dudx = hidx * ((ge ? uE - uc : real(0)) + (gw ? uc - uW : real(0)));
    const real dudy = hidy * ((gn ? uN - uc : real(0)) + (gs ? uc - uS : real(0)));
    const real dvdx = hidx * ((ge ? vE - vc : real(0)) + (gw ? vc - vW : real(0)));
    const real dvdy = hidy * ((gn ? vN - vc : real(0)) + (gs ? vc - vS : real(0)));
    const real dudz = ddz_win(um, uc, up, k, kb, g);
    const real dvdz = ddz_win(vm, vc, vp, k, kb, g);
    au[o] = -(uc * dudx + vc * dudy + w * dudz);
    av[o] = -(uc * dvdx + vc * dvdy + w * dvdz);
    um = uc; uc = up; vm = vc; vc = vp;
  }
}

// ---------------------------------------------------------------------------
// x part of _dealias_h_fd (horizontal.py:297-300): the 2/3-rule projection
// Re(IFFT(mask * FFT(row))).  The projection is real, so one complex row
// z = a + i b carries the a row of u-advection and the b row of v-advection.
// DFT of length N = N1 N2 by the four-step factorisation in LDS; the
// twiddle table holds exp(-2 pi i e / N), e in [0, N); kept[k] != 0 marks the
// retained wavenumbers (the mask of factory.py:529-533, computed on the host).
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(256) void k_dealias_x(int N, int N1, int N2, const int* __restrict__ kept, int nrows,
                                                   const real* __restrict__ twr,
                                                   const real* __restrict__ twi,
                                                   const real* __restrict__ ina,
                                                   const real* __restrict__ inb,
                                                   real* __restrict__ outa, real* __restrict__ outb) {
  extern __shared__ unsigned char zm_lds[];
  real* ar = reinterpret_cast<real*>(zm_lds);
  real* ai = ar + N;
  real* br = ai + N;
  real* bi = br + N;
  real* wr = bi + N;
  real* wi = wr + N;
  const int row = blockIdx.x;
  if (row >= nrows) return;
  const long long off = (long long)row * N;
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    ar[t] = ina[off + t];
    ai[t] = inb[off + t];
    wr[t] = twr[t];
    wi[t] = twi[t];
  }
  __syncthreads();
  // forward, stage 1: Y[k1 N2 + n2] = W^{n2 k1} sum_n1 a[N2 n1 + n2] W^{N2 n1 k1}
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    const int k1 = t / N2, n2 = t - k1 * N2;
    const int step = N2 * k1;  // < N
    int e = 0;
    real sr = 0, si = 0;
    for (int n1 = 0; n1 < N1; ++n1) {
      const real xr = ar[N2 * n1 + n2], xi = ai[N2 * n1 + n2];
      sr += xr * wr[e] - xi * wi[e];
      si += xr * wi[e] + xi * wr[e];
      e += step;
      if (e >= N) e -= N;
    }
    const int e2 = (n2 * k1) % N;
    br[t] = sr * wr[e2] - si * wi[e2];
    bi[t] = sr * wi[e2] + si * wr[e2];
  }
  __syncthreads();
  // forward, stage 2 + mask: X[k1 + N1 k2] = sum_n2 Y[k1 N2 + n2] W^{N1 n2 k2}
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    const int k2 = t / N1, k1 = t - k2 * N1;
    real sr = 0, si = 0;
    if (kept[t]) {
      const int step = N1 * k2;
      int e = 0;
      for (int n2 = 0; n2 < N2; ++n2) {
        const real yr = br[k1 * N2 + n2], yi = bi[k1 * N2 + n2];
        sr += yr * wr[e] - yi * wi[e];
        si += yr * wi[e] + yi * wr[e];
        e += step;
        if (e >= N) e -= N;
      }
    }
    ar[t] = sr;
    ai[t] = si;
  }
  __syncthreads();
  // inverse, stage 1: Z[k1 N2 + n2] = conj(W)^{k1 n2} sum_k2 X[k1 + N1 k2] conj(W)^{N1 k2 n2}
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    const int k1 = t / N2, n2 = t - k1 * N2;
    const int step = N1 * n2 % N;
    int e = 0;
    real sr = 0, si = 0;
    for (int k2 = 0; k2 < N2; ++k2) {
      const real xr = ar[k1 + N1 * k2], xi = ai[k1 + N1 * k2];
      sr += xr * wr[e] + xi * wi[e];
      si += xi * wr[e] - xr * wi[e];
      e += step;
      if (e >= N) e -= N;
    }
    const int e2 = (k1 * n2) % N;
    br[t] = sr * wr[e2] + si * wi[e2];
    bi[t] = si * wr[e2] - sr * wi[e2];
  }
  __syncthreads();
  // inverse, stage 2: x[n2 + N2 n1] = (1/N) sum_k1 Z[k1 N2 + n2] conj(W)^{N2 k1 n1}
  const real invN = real(1) / real(N);
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    const int n1 = t / N2, n2 = t - n1 * N2;
    const int step = N2 * n1;
    int e = 0;
    real sr = 0, si = 0;
    for (int k1 = 0; k1 < N1; ++k1) {
      const real zr = br[k1 * N2 + n2], zi = bi[k1 * N2 + n2];
      sr += zr * wr[e] + zi * wi[e];
      si += zi * wr[e] - zr * wi[e];
      e += step;
      if (e >= N) e -= N;
    }
    outa[off + t] = sr * invN;
    outb[off + t] = si * invN;
  }
}

// ---------------------------------------------------------------------------
// Same projection with a mixed-radix (2, 3, 4, 5) Stockham FFT in LDS (the
// autosorting formulation of Govindaraju et al., "High Performance Discrete
// Fourier Transforms on Graphics Processors", SC 2008): stage s with radix R
// reads R points at stride N/R, twiddles them by exp(-+2 pi i r k / (Ns R)),
// does an R-point DFT and writes them at stride Ns, Ns *= R.  A workgroup
// transforms kFftRowsPerBlock complex rows in turn with the twiddle table
// held in LDS.
// ---------------------------------------------------------------------------
constexpr int kFftRowsPerBlock = 4;

// cos and sin of 2 pi m / R for the butterfly radices (mathematical constants)
template <int R>
struct RootsOfUnity;
template <>
struct RootsOfUnity<2> {
  static constexpr double c[2] = {1.0, -1.0}, s[2] = {0.0, 0.0};
};
template <>
struct RootsOfUnity<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
};
template <>
struct RootsOfUnity<4> {
  static constexpr double c[4] = {1.0, 0.0, -1.0, 0.0}, s[4] = {0.0, 1.0, 0.0, -1.0};
};
template <>
struct RootsOfUnity<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                                  0.30901699437494742410};
  static constexpr double s[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                                  -0.95105651629515357212};
};

// one radix-R butterfly of a Stockham stage: gather, twiddle, R-point DFT
// with constant roots (exp(-2 pi i m / R), conjugated for the inverse), scatter
template <typename real, int R>
__device__ __forceinline__ void fft_butterfly(int bf, int N, int Ns, int tstep, bool inverse,
                                              const real* __restrict__ src, real* __restrict__ dst,
                                              const real* __restrict__ wr, const real* __restrict__ wi) {
  const int T = N / R;
  const int k = bf % Ns;
  real vr[R], vi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const real xr = src[2 * (bf + r * T)], xi = src[2 * (bf + r * T) + 1];
    if (r == 0) {
      vr[0] = xr;
      vi[0] = xi;
    } else {
      const int e = (r * k * tstep) % N;
      const real c = wr[e], sn = inverse ? -wi[e] : wi[e];
      vr[r] = xr * c - xi * sn;
      vi[r] = xr * sn + xi * c;
    }
  }
  const real sg = inverse ? real(1) : real(-1);
  const int base = (bf / Ns) * Ns * R + k;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    real sr = vr[0], si = vi[0];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      const int m = (r * q) % R;
      const real c = real(RootsOfUnity<R>::c[m]), sn = sg * real(RootsOfUnity<R>::s[m]);
      sr += vr[r] * c - vi[r] * sn;
      si += vr[r] * sn + vi[r] * c;
    }
    dst[2 * (base + q * Ns)] = sr;
    dst[2 * (base + q * Ns) + 1] = si;
  }
}

template <typename real>
__device__ __forceinline__ void fft_stages(int N, const FftPlan& plan, bool inverse, real* __restrict__ a,
                                           real* __restrict__ b, const real* __restrict__ wr,
                                           const real* __restrict__ wi, real*& out) {
  real* src = a;
  real* dst = b;
  int Ns = 1;
  for (int st = 0; st < plan.nst; ++st) {
    const int R = plan.rad[st];
    const int T = N / R;
    const int tstep = N / (Ns * R);
    for (int bf = threadIdx.x; bf < T; bf += blockDim.x) {
      switch (R) {
        case 2: fft_butterfly<real, 2>(bf, N, Ns, tstep, inverse, src, dst, wr, wi); break;
        case 3: fft_butterfly<real, 3>(bf, N, Ns, tstep, inverse, src, dst, wr, wi); break;
        case 4: fft_butterfly<real, 4>(bf, N, Ns, tstep, inverse, src, dst, wr, wi); break;
        default: fft_butterfly<real, 5>(bf, N, Ns, tstep, inverse, src, dst, wr, wi); break;
      }
    }
    __syncthreads();
    real* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  out = src;
}

template <typename real>
__global__ __launch_bounds__(256) void k_dealias_fft(int N, FftPlan plan, const int* __restrict__ kept, int nrows,
                                                     const real* __restrict__ twr, const real* __restrict__ twi,
                                                     const real* __restrict__ ina, const real* __restrict__ inb,
                                                     real* __restrict__ outa, real* __restrict__ outb) {
  extern __shared__ unsigned char zm_lds[];
  real* wr = reinterpret_cast<real*>(zm_lds);
  real* wi = wr + N;
  real* A = wi + N;      // interleaved complex row, 2N
  real* B = A + 2 * N;   // ping-pong buffer, 2N
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    wr[t] = twr[t];
    wi[t] = twi[t];
  }
  const real invN = real(1) / real(N);
  for (int q = 0; q < kFftRowsPerBlock; ++q) {
    const int row = blockIdx.x * kFftRowsPerBlock + q;
    if (row >= nrows) break;  // uniform
    const long long off = (long long)row * N;
    __syncthreads();
    for (int t = threadIdx.x; t < N; t += blockDim.x) {
      A[2 * t] = ina[off + t];
      A[2 * t + 1] = inb[off + t];
    }
    __syncthreads();
    real* X = nullptr;
    fft_stages(N, plan, false, A, B, wr, wi, X);
    for (int t = threadIdx.x; t < N; t += blockDim.x) {
      if (!kept[t]) {
        X[2 * t] = 0;
        X[2 * t + 1] = 0;
      }
    }
    __syncthreads();
    real* Y = nullptr;
    fft_stages(N, plan, true, X, X == A ? B : A, wr, wi, Y);
    for (int t = threadIdx.x; t < N; t += blockDim.x) {
      outa[off + t] = Y[2 * t] * invN;
      outb[off + t] = Y[2 * t + 1] * invN;
    }
  }
}

// ---------------------------------------------------------------------------
// The same projection for the production row lengths (720, 800, 480, 360,
// 180, 90) with the radix plan fixed at compile time: every index division is
// by a constant, the twiddle exponent r k N / (Ns R) needs no reduction
// (r < R, k < Ns), the 2-, 3-, 4- and 5-point DFTs are written out (no
// products with the roots 0 and +-1), complex values move through LDS as one
// 64- or 128-bit word, and the rows of a workgroup go through each stage
// together (one barrier per stage for all of them).
// ---------------------------------------------------------------------------
template <typename real>
struct __attribute__((aligned(2 * sizeof(real)))) Cx {
  real re, im;
};

template <int N>
struct FixedFft {
  // radices in the order used: 4s, then 5s, 3s and 2s
  static constexpr int radix(int s) {
    int n = N, c = 0;
    const int order[4] = {4, 5, 3, 2};
    for (int q = 0; q < 4; ++q) {
      while (n % order[q] == 0) {
        if (c == s) return order[q];
        n /= order[q];
        ++c;
      }
    }
    return 0;
  }
  static constexpr int stages() {
    int c = 0;
    while (radix(c) != 0) ++c;
    return c;
  }
  static constexpr int span(int s) {  // Ns before stage s
    int p = 1;
    for (int q = 0; q < s; ++q) p *= radix(q);
    return p;
  }
  static constexpr bool complete() { return span(stages()) == N; }
};

template <typename real>
__device__ __forceinline__ Cx<real> cmul(Cx<real> a, Cx<real> w) {
  return Cx<real>{a.re * w.re - a.im * w.im, a.re * w.im + a.im * w.re};
}
template <typename real>
__device__ __forceinline__ Cx<real> cadd(Cx<real> a, Cx<real> b) {
  return Cx<real>{a.re + b.re, a.im + b.im};
}
template <typename real>
__device__ __forceinline__ Cx<real> csub(Cx<real> a, Cx<real> b) {
  return Cx<real>{a.re - b.re, a.im - b.im};
}
// (-i) z for the forward transform, (+i) z for the inverse
template <typename real, bool INV>
__device__ __forceinline__ Cx<real> rot(Cx<real> z) {
  return INV ? Cx<real>{-z.im, z.re} : Cx<real>{z.im, -z.re};
}

// y_q = sum_r v_r exp(-+2 pi i r q / R), in place
template <typename real, int R, bool INV>
__device__ __forceinline__ void small_dft(Cx<real> (&v)[R]) {
  if constexpr (R == 2) {
    const Cx<real> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const real s = real(0.86602540378443864676);  // sin(2 pi / 3)
    const Cx<real> a = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    const Cx<real> t{v[0].re - real(0.5) * a.re, v[0].im - real(0.5) * a.im};
    const Cx<real> u = rot<real, INV>(Cx<real>{s * d.re, s * d.im});
    v[0] = cadd(v[0], a);
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
  } else if constexpr (R == 4) {
    const Cx<real> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const Cx<real> t2 = cadd(v[1], v[3]), t3 = rot<real, INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else {
    static_assert(R == 5, "radix");
    const real c1 = real(0.30901699437494742410), c2 = real(-0.80901699437494742410);  // cos(2 pi/5), cos(4 pi/5)
    const real s1 = real(0.95105651629515357212), s2 = real(0.58778525229247312917);   // sin(2 pi/5), sin(4 pi/5)
    const Cx<real> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const Cx<real> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const Cx<real> t1{v[0].re + c1 * a1.re + c2 * a2.re, v[0].im + c1 * a1.im + c2 * a2.im};
    const Cx<real> t2{v[0].re + c2 * a1.re + c1 * a2.re, v[0].im + c2 * a1.im + c1 * a2.im};
    const Cx<real> u1 = rot<real, INV>(Cx<real>{s1 * b1.re + s2 * b2.re, s1 * b1.im + s2 * b2.im});
    const Cx<real> u2 = rot<real, INV>(Cx<real>{s2 * b1.re - s1 * b2.re, s2 * b1.im - s1 * b2.im});
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(t1, u1);
    v[4] = csub(t1, u1);
    v[2] = cadd(t2, u2);
    v[3] = csub(t2, u2);
  }
}

// one Stockham stage over `rows` rows of N points (row-major in src / dst)
template <typename real, int N, int S, bool INV>
__device__ __forceinline__ void fixed_stage(const Cx<real>* __restrict__ src, Cx<real>* __restrict__ dst,
                                            const Cx<real>* __restrict__ tw, int rows) {
  constexpr int R = FixedFft<N>::radix(S), Ns = FixedFft<N>::span(S), T = N / R, tstep = N / (Ns * R);
  for (int id = (int)threadIdx.x; id < rows * T; id += (int)blockDim.x) {
    const int row = id / T, bf = id - row * T;
    const int k = bf % Ns;
    const Cx<real>* in = src + row * N;
    Cx<real> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = in[bf + r * T];
      if (Ns > 1 && r > 0) {
        Cx<real> w = tw[r * k * tstep];  // < N: no reduction
        if (INV) w.im = -w.im;
        v[r] = cmul(v[r], w);
      }
    }
    small_dft<real, R, INV>(v);
    Cx<real>* out = dst + row * N + (bf - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) out[q * Ns] = v[q];
  }
}

// all stages; returns the buffer holding the result
template <typename real, int N, int S, bool INV>
__device__ __forceinline__ Cx<real>* fixed_stages(Cx<real>* src, Cx<real>* dst, const Cx<real>* tw, int rows) {
  if constexpr (S == FixedFft<N>::stages()) {
    return src;
  } else {
    fixed_stage<real, N, S, INV>(src, dst, tw, rows);
    __syncthreads();
    return fixed_stages<real, N, S + 1, INV>(dst, src, tw, rows);
  }
}

template <typename real>
constexpr int fixed_fft_rows() {
  return sizeof(real) == 4 ? 4 : 2;  // keeps the LDS of N = 800 rows within 64 KiB
}

template <typename real, int N>
__global__ __launch_bounds__(256) void k_dealias_fixed(const int* __restrict__ kept, int nrows,
                                                       const real* __restrict__ twr, const real* __restrict__ twi,
                                                       const real* __restrict__ ina, const real* __restrict__ inb,
                                                       real* __restrict__ outa, real* __restrict__ outb) {
  static_assert(FixedFft<N>::complete(), "N must factor into 2, 3, 4 and 5");
  constexpr int RB = fixed_fft_rows<real>();
  extern __shared__ unsigned char zm_lds[];
  Cx<real>* tw = reinterpret_cast<Cx<real>*>(zm_lds);
  Cx<real>* A = tw + N;
  Cx<real>* B = A + RB * N;
  const int row0 = (int)blockIdx.x * RB;
  const int rows = nrows - row0 < RB ? nrows - row0 : RB;
  const long long off = (long long)row0 * N;
  for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = Cx<real>{twr[t], twi[t]};
  for (int t = threadIdx.x; t < rows * N; t += blockDim.x) A[t] = Cx<real>{ina[off + t], inb[off + t]};
  __syncthreads();
  Cx<real>* X = fixed_stages<real, N, 0, false>(A, B, tw, rows);
  for (int t = threadIdx.x; t < rows * N; t += blockDim.x) {
    const int i = t % N;
    if (!kept[i]) X[t] = Cx<real>{real(0), real(0)};
  }
  __syncthreads();
  Cx<real>* Y = fixed_stages<real, N, 0, true>(X, X == A ? B : A, tw, rows);
  const real invN = real(1) / real(N);
  for (int t = threadIdx.x; t < rows * N; t += blockDim.x) {
    const Cx<real> y = Y[t];
    outa[off + t] = y.re * invN;
    outb[off + t] = y.im * invN;
  }
}

// ---------------------------------------------------------------------------
// Two-factor form of the same projection for float32 rows, N = N1 N2 with
// n = n1 + N1 n2 and k = k2 + N2 k1:
//   X[k2 + N2 k1] = sum_n1 w_N1^(n1 k1) w_N^(n1 k2) sum_n2 x[n1 + N1 n2] w_N2^(n2 k2)
// and the inverse with the conjugate roots in the reverse order.  A thread
// holds one n1 (N2 values) or one k2 (N1 values) of a row in registers: the
// N2-point DFT, the twiddle w_N^(n1 k2) and an LDS transpose; the N1-point
// DFT, the mask, the inverse N1-point DFT and the conjugate twiddle; another
// transpose and the inverse N2-point DFT, stored straight to HBM.  Two
// transposes and three barriers per row instead of a pass through LDS per
// radix stage; the in-register DFTs are Stockham stages unrolled at compile
// time with their roots as constants (quarter turns without products).
// ---------------------------------------------------------------------------
constexpr double kTwoPi = 6.28318530717958647692;

// cos and sin by their Taylor series, |x| <= pi (double precision at 30 terms)
constexpr double ct_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 30; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
constexpr double ct_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 30; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
// 2 pi m / L reduced to (-pi, pi]
constexpr double ct_angle(int m, int L) {
  const double a = kTwoPi * (double)m / (double)L;
  return a > kTwoPi / 2.0 ? a - kTwoPi : a;
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// v exp(-+2 pi i m / L) for a compile-time m (+ for the inverse)
template <typename real, int M, int L, bool INV>
__device__ __forceinline__ Cx<real> twiddle_ct(Cx<real> v) {
  constexpr int m = M % L;
  if constexpr (m == 0) {
    return v;
  } else if constexpr (4 * m == L) {
    return rot<real, INV>(v);
  } else if constexpr (2 * m == L) {
    return Cx<real>{-v.re, -v.im};
  } else if constexpr (4 * m == 3 * L) {
    return rot<real, !INV>(v);
  } else {
    constexpr double a = ct_angle(m, L);
    constexpr real c = (real)ct_cos(a);
    constexpr real sn = (real)(INV ? ct_sin(a) : -ct_sin(a));
    return cmul(v, Cx<real>{c, sn});
  }
}

// one Stockham stage of an L-point DFT held in registers (fixed_stage with
// compile-time indices and roots)
template <typename real, int L, int S, bool INV>
__device__ __forceinline__ void reg_stage(const Cx<real> (&in)[L], Cx<real> (&out)[L]) {
  constexpr int R = FixedFft<L>::radix(S), Ns = FixedFft<L>::span(S), T = L / R, tstep = L / (Ns * R);
  static_for<0, T>([&](auto bfc) {
    constexpr int bf = decltype(bfc)::value, k = bf % Ns;
    Cx<real> v[R];
    static_for<0, R>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      v[r] = twiddle_ct<real, r * k * tstep, L, INV>(in[bf + r * T]);
    });
    small_dft<real, R, INV>(v);
    static_for<0, R>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      out[(bf - k) * R + k + q * Ns] = v[q];
    });
  });
}

template <typename real, int L, int S, bool INV>
__device__ __forceinline__ void reg_stages(Cx<real> (&a)[L], Cx<real> (&b)[L]) {
  if constexpr (S < FixedFft<L>::stages()) {
    reg_stage<real, L, S, INV>(a, b);
    reg_stages<real, L, S + 1, INV>(b, a);
  }
}

// y_q = sum_r v_r exp(-+2 pi i r q / L), in place in registers
template <typename real, int L, bool INV>
__device__ __forceinline__ void reg_dft(Cx<real> (&v)[L]) {
  static_assert(FixedFft<L>::complete(), "L must factor into 2, 3, 4 and 5");
  Cx<real> w[L];
  reg_stages<real, L, 0, INV>(v, w);
  if constexpr (FixedFft<L>::stages() % 2 == 1)
    static_for<0, L>([&](auto ic) { v[decltype(ic)::value] = w[decltype(ic)::value]; });
}

template <int N>
struct TwoFactor;
template <>
struct TwoFactor<720> {
  static constexpr int N1 = 24, N2 = 30;
};
template <>
struct TwoFactor<800> {
  static constexpr int N1 = 25, N2 = 32;
};
template <>
struct TwoFactor<480> {
  static constexpr int N1 = 20, N2 = 24;
};
template <>
struct TwoFactor<360> {
  static constexpr int N1 = 18, N2 = 20;
};
template <>
struct TwoFactor<180> {
  static constexpr int N1 = 12, N2 = 15;
};
template <>
struct TwoFactor<90> {
  static constexpr int N1 = 9, N2 = 10;
};

template <int N>
struct TwoFactorShape {
  static constexpr int N1 = TwoFactor<N>::N1, N2 = TwoFactor<N>::N2;
  static_assert(N1 * N2 == N, "factorisation");
  static constexpr int RB = 256 / (N1 > N2 ? N1 : N2);  // rows per 256-thread block
  // odd strides of the transposes: 64-bit LDS accesses of consecutive threads
  // at stride P1 (P2) complex values fall in distinct banks
  static constexpr int P1 = N1 % 2 ? N1 : N1 + 1, P2 = N2 % 2 ? N2 : N2 + 1;
  static constexpr int RS = N2 * P1 > N1 * P2 ? N2 * P1 : N1 * P2;  // LDS complex values per row
  static constexpr size_t lds() { return (size_t)(N + RB * RS) * sizeof(Cx<float>); }
};

template <int N>
__global__ __launch_bounds__(256) void k_dealias_2f(const int* __restrict__ kept, int nrows,
                                                    const float* __restrict__ twr, const float* __restrict__ twi,
                                                    const float* __restrict__ ina, const float* __restrict__ inb,
                                                    float* __restrict__ outa, float* __restrict__ outb) {
  using Sh = TwoFactorShape<N>;
  constexpr int N1 = Sh::N1, N2 = Sh::N2, RB = Sh::RB, P1 = Sh::P1, P2 = Sh::P2, RS = Sh::RS;
  extern __shared__ unsigned char zm_lds[];
  Cx<float>* tw = reinterpret_cast<Cx<float>*>(zm_lds);  // exp(-2 pi i e / N)
  Cx<float>* buf = tw + N;
  const int tid = (int)threadIdx.x, row0 = (int)blockIdx.x * RB;
  for (int t = tid; t < N; t += 256) tw[t] = Cx<float>{twr[t], twi[t]};
  // n1 threads (transforms over n2) and k2 threads (transforms over k1)
  const int ra = tid / N1, n1 = tid - ra * N1;
  const bool actA = ra < RB && row0 + ra < nrows;
  const int rb = tid / N2, k2 = tid - rb * N2;
  const bool actB = rb < RB && row0 + rb < nrows;
  Cx<float> x[N2];
  const long long baseA = (long long)(row0 + ra) * N + n1;
  if (actA)
    static_for<0, N2>([&](auto c) {
      constexpr int n2 = decltype(c)::value;
      x[n2] = Cx<float>{ina[baseA + N1 * n2], inb[baseA + N1 * n2]};
    });
  __syncthreads();  // roots in LDS
  if (actA) {
    reg_dft<float, N2, false>(x);
    Cx<float>* y = buf + ra * RS + n1;
    static_for<0, N2>([&](auto c) {
      constexpr int q = decltype(c)::value;
      y[q * P1] = q == 0 ? x[0] : cmul(x[q], tw[n1 * q]);
    });
  }
  __syncthreads();
  Cx<float> z[N1];
  if (actB) {
    const Cx<float>* y = buf + rb * RS + k2 * P1;
    static_for<0, N1>([&](auto c) { z[decltype(c)::value] = y[decltype(c)::value]; });
  }
  __syncthreads();  // every row of Y read before the inverse overwrites it
  if (actB) {
    reg_dft<float, N1, false>(z);  // z[k1] = X[k2 + N2 k1]
    static_for<0, N1>([&](auto c) {
      constexpr int k1 = decltype(c)::value;
      if (!kept[k2 + N2 * k1]) z[k1] = Cx<float>{0.0f, 0.0f};
    });
    reg_dft<float, N1, true>(z);
    Cx<float>* zo = buf + rb * RS + k2;
    static_for<0, N1>([&](auto c) {
      constexpr int q = decltype(c)::value;
      if constexpr (q == 0) {
        zo[0] = z[0];
      } else {
        Cx<float> w = tw[q * k2];
        w.im = -w.im;
        zo[q * P2] = cmul(z[q], w);
      }
    });
  }
  __syncthreads();
  if (actA) {
    const Cx<float>* zi = buf + ra * RS + n1 * P2;
    static_for<0, N2>([&](auto c) { x[decltype(c)::value] = zi[decltype(c)::value]; });
    reg_dft<float, N2, true>(x);
    const float invN = 1.0f / (float)N;
    static_for<0, N2>([&](auto c) {
      constexpr int n2 = decltype(c)::value;
      outa[baseA + N1 * n2] = x[n2].re * invN;
      outb[baseA + N1 * n2] = x[n2].im * invN;
    });
  }
}

// ---------------------------------------------------------------------------
// Momentum residual (tendencies.py:39-72 and 247-319, legacy mode-split
// branch) fused with the RK2 velocity update (step.py:263-319).
//   mode 0: du1 -> buffers, u_pred = u + du1 dt
//   mode 1: u_new = (u + 0.5 (du1 + du2) dt) * exp(-r_bot dt bm)
// ax/av hold the x-de-aliased advection; the y binomial of _dealias_h_fd
// (horizontal.py:301-305) is applied here.
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(256) void k_mom_residual(
    Geo<real> g, Phys<real> P, int mode, const real* __restrict__ ax, const real* __restrict__ ay,
    const real* __restrict__ u, const real* __restrict__ v, const real* __restrict__ pbc,
    const real* __restrict__ eta, const real* __restrict__ Frx, const real* __restrict__ Fry,
    real* __restrict__ du1, real* __restrict__ dv1, real* __restrict__ uo, real* __restrict__ vo,
    real decay_bot) {
  ZM_COLUMN_PROLOGUE
  const real f = g.frow[j];
  const real idx = g.idx[j], idx2 = g.idx2[j], tr = g.tanR[j], idyc = g.idyc[j];
  const real cfn = g.cfn[j], cfs = g.cfs[j];
  // y binomial [1,4,6,4,1]/16 with 2-row edge padding
  const int jm2 = j >= 2 ? j - 2 : 0, jm1 = j >= 1 ? j - 1 : 0;
  const int jp1 = j + 1 < g.ny ? j + 1 : g.ny - 1, jp2 = j + 2 < g.ny ? j + 2 : g.ny - 1;
  const int rm2 = (jm2 - j) * g.nx, rm1 = (jm1 - j) * g.nx, rp1 = (jp1 - j) * g.nx, rp2 = (jp2 - j) * g.nx;
  // column constants: eta at the five points, 2-D forcing
  const real rg = real(kRho0 * kG);
  const real eC = eta[c2], eE = eta[n.e2], eW = eta[n.w2];
  const real eNl = eta[n.n2e], eSl = eta[n.s2e];  // edge-clamped, selected below
  const real eN = n.hasN ? eNl : real(0), eS = n.hasS ? eSl : real(0);
  const real frx = Frx[c2], fry = Fry[c2];
  const real tx = P.tau_x[c2], ty = P.tau_y[c2];
  const real sfx = tx * P.wind_factor, sfy = ty * P.wind_factor;
  const real bfx = tx * P.inv_rho0H, bfy = ty * P.inv_rho0H;
  const real half = real(0.5);
  long long o = (long long)k0 * n2 + c2;
  real um = k0 > 0 ? u[o - n2] : real(0), uc = u[o];
  real vm = k0 > 0 ? v[o - n2] : real(0), vc = v[o];
  for (int k = k0; k < k1; ++k, o += n2) {
    const bool more = k + 1 < g.nz;
    const long long onext = more ? o + n2 : o;  // in bounds at the last level
    const real un_ = u[onext], vn_ = v[onext];
    const real up = more ? un_ : uc, vp = more ? vn_ : vc;
    const long long pl = o - c2;
    const bool wc = k < n.nwC;
    const bool ge = wc && k < n.nwE, gw = wc && k < n.nwW;
    const bool gnE = wc && k < n.nwN, gsE = wc && k < n.nwS;
    const bool gn = n.hasN && gnE, gs = n.hasS && gsE;
    // all loads of the level are in bounds (rows clamped): issued unconditionally
    const real a0 = ax[o + rm2], a1 = ax[o + rm1], a2 = ax[o], a3 = ax[o + rp1], a4 = ax[o + rp2];
    const real b0 = ay[o + rm2], b1 = ay[o + rm1], b2 = ay[o], b3 = ay[o + rp1], b4 = ay[o + rp2];
    const real pbC = pbc[o], pbE = pbc[pl + n.e2], pbW = pbc[pl + n.w2], pbN = pbc[pl + n.n2e], pbS = pbc[pl + n.s2e];
    const real advu_ = (a0 + real(4) * a1 + real(6) * a2 + real(4) * a3 + a4) / real(16);
    const real advv_ = (b0 + real(4) * b1 + real(6) * b2 + real(4) * b3 + b4) / real(16);
    const real advu = wc ? advu_ : real(0), advv = wc ? advv_ : real(0);
    // hydrostatic pressure p = rho0 g eta + p_bc at the five points
    const real pc = rg * eC + pbC;
    const real pe = rg * eE + pbE;
    const real pw = rg * eW + pbW;
    const real pn = n.hasN ? rg * eN + pbN : real(0);
    const real ps = n.hasS ? rg * eS + pbS : real(0);
    const real gpx = idx * half * ((ge ? pe - pc : real(0)) + (gw ? pc - pw : real(0)));
    const real gpy = idyc * half * ((gn ? (pn - pc) * cfn : real(0)) + (gs ? (pc - ps) * cfs : real(0)));
    const real pgfx = -gpx / real(kRho0), pgfy = -gpy / real(kRho0);
    const real lapu = lap_vals(uc, u[pl + n.e2], u[pl + n.w2], u[pl + n.n2e], u[pl + n.s2e], ge, gw, gnE, gsE, idx2,
                               g.idy2, g.hidy, tr);
    const real lapv = lap_vals(vc, v[pl + n.e2], v[pl + n.w2], v[pl + n.n2e], v[pl + n.s2e], ge, gw, gnE, gsE, idx2,
                               g.idy2, g.hidy, tr);
    const real dvu = P.nu_v * d2z_win(um, uc, up, k, kb, g);
    const real dvv = P.nu_v * d2z_win(vm, vc, vp, k, kb, g);
    real dudt = 0, dvdt = 0;
    if (wc) {
      dudt = advu + f * vc + pgfx + P.nu_h * lapu + dvu;
      dvdt = advv + (-f) * uc + pgfy + P.nu_h * lapv + dvv;
      if (k == 0) {
        dudt = dudt + sfx;
        dvdt = dvdt + sfy;
      }
      if (k == g.nz - 1) {
        dudt = dudt + (-P.r_bot) * uc;
        dvdt = dvdt + (-P.r_bot) * vc;
      }
    }
    dudt = dudt - P.nu_h * lapu;
    dvdt = dvdt - P.nu_h * lapv;
    dudt = dudt - dvu;
    dvdt = dvdt - dvv;
    dudt = dudt - f * vc;
    dvdt = dvdt + f * uc;
    // cancel the eta part with the same layer-gated operator
    const real gex = idx * half * ((ge ? eE - eC : real(0)) + (gw ? eC - eW : real(0)));
    const real gey = idyc * half * ((gn ? (eN - eC) * cfn : real(0)) + (gs ? (eC - eS) * cfs : real(0)));
    dudt = dudt + real(kG) * gex;
    dvdt = dvdt + real(kG) * gey;
    dudt = dudt - frx;
    dvdt = dvdt - fry;
    dudt = dudt - bfx;
    dvdt = dvdt - bfy;
    if (!wc) {
      dudt = 0;
      dvdt = 0;
    }
    const bool bm = wc && k == g.nz - 1;
    if (bm) {
      dudt = dudt + P.r_bot * uc;
      dvdt = dvdt + P.r_bot * vc;
    }
    if (mode == 0) {
      du1[o] = dudt;
      dv1[o] = dvdt;
      uo[o] = uc + dudt * P.dt;
      vo[o] = vc + dvdt * P.dt;
    } else {
      real un = uc + real(0.5) * (du1[o] + dudt) * P.dt;
      real vn = vc + real(0.5) * (dv1[o] + dvdt) * P.dt;
      if (bm) {
        un = un * decay_bot;
        vn = vn * decay_bot;
      }
      uo[o] = un;
      vo[o] = vn;
    }
    um = uc; uc = up; vm = vc; vc = vp;
  }
}

// ---------------------------------------------------------------------------
// Barotropic forward-backward substep, legacy mode-split branch of
// _free_surface_step_fd (dynamics/barotropic.py:86-254).  One workgroup per
// latitude row; rows of the polar-cap bands reduce their zonal means in LDS
// (_apply_polar_cap legacy, horizontal.py:387-451).
// ---------------------------------------------------------------------------
// Sums over the 256 threads of one row group (t = thread within the group,
// red = the group's LDS scratch of 4 entries per value), broadcast to the
// group; each value is summed in the same order (wave tree, then the four
// wave sums in turn), NV values sharing one pair of barriers.  Both barriers
// are block-wide: every thread of the block must make the same calls.
template <typename real, int NV>
__device__ __forceinline__ void group_sums(real (&x)[NV], real* red, int t) {
#pragma unroll
  for (int q = 0; q < NV; ++q)
    for (int o = 32; o > 0; o >>= 1) x[q] += __shfl_down(x[q], o, 64);
  const int lane = t & 63, wid = t >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) red[4 * q + wid] = x[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    real s = 0;
    for (int w = 0; w < 4; ++w) s += red[4 * q + w];
    x[q] = s;
  }
}


template <typename real>
__device__ __forceinline__ real cap_weight(const Phys<real>& P, int j, int ny) {
  if (j < P.nb) return P.capw[j];
  if (j >= ny - P.nb) return P.capw[ny - 1 - j];
  return real(-1);
}

constexpr int kRowChunk = 4;  // cells per thread per row (nx <= 1024)

// Loads and stores of the barotropic state.  COH = true makes them agent-scope
// relaxed atomics (cache-coherent across the XCDs without flushing or
// invalidating the L2: the persistent sub-cycle exchanges halo rows this way);
// COH = false is a plain access.  The value moved is the same either way.
template <bool COH, typename real>
__device__ __forceinline__ real bt_load(const real* p) {
  if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *p;
}
template <bool COH, typename real>
__device__ __forceinline__ void bt_store(real* p, real v) {
  if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

// eta' = eta - dt H div(ubt) on row j with the polar-cap blend of band rows;
// reads eta_row[i] (row j of eta), writes out[i] (and out2[i] unless null).
// out may alias eta_row.  t: thread within the 256-thread row group; ALLRED
// (or allred, block-uniform at run time): reduce on every row so that the
// block-wide barriers match across groups (the zonal mean is only used on
// band rows).
template <typename real, bool ALLRED = false, bool COH = false, int NQ = kRowChunk>
__device__ __forceinline__ void bt_eta_row(const Geo<real>& g, const Phys<real>& P, int j, int t,
                                           const real* __restrict__ eta_row, const real* __restrict__ ubt,
                                           const real* __restrict__ vbt, real dtH, real* out, real* red,
                                           real* out2 = nullptr, bool allred = false) {
  const real* wet = P.wet2;
  real en[NQ];
  real ssum = 0, wsum = 0;
  // row metrics in registers: the selects below then only mix loaded values
  const bool hasN = j + 1 < g.ny, hasS = j > 0;
  const real idx = g.idx[j], idyc = g.idyc[j], cfN = g.cfn[j], cfS = g.cfn[hasS ? j - 1 : 0];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    // lanes past the row end repeat its last column and book nothing
    const bool act = t + q * 256 < g.nx;
    const int i = act ? t + q * 256 : g.nx - 1;
    const int c2 = j * g.nx + i, e2 = j * g.nx + wrap_e(i, g.nx), w2 = j * g.nx + wrap_w(i, g.nx);
    const int n2 = hasN ? c2 + g.nx : c2, s2 = hasS ? c2 - g.nx : c2;  // clamped: unused at the walls
    // every index is in bounds: load unconditionally, then select
    const real wC = wet[c2], wE = wet[e2], wW = wet[w2], wN = wet[n2], wS = wet[s2];
    const real uC = bt_load<COH>(ubt + c2), uE = bt_load<COH>(ubt + e2), uW = bt_load<COH>(ubt + w2);
    const real vC = bt_load<COH>(vbt + c2), vN = bt_load<COH>(vbt + n2), vS = bt_load<COH>(vbt + s2);
    const real eC = bt_load<COH>(eta_row + i);
    const bool wc = wC != real(0);
    const real fe = (wc && wE != real(0)) ? real(0.5) * (uC + uE) : real(0);
    const real fw = (wW != real(0) && wc) ? real(0.5) * (uW + uC) : real(0);
    const real divx = (fe - fw) * idx;
    const real fn = (hasN && wc && wN != real(0)) ? real(0.5) * (vC + vN) * cfN : real(0);
    const real fs = (hasS && wc && wS != real(0)) ? real(0.5) * (vS + vC) * cfS : real(0);
    const real divy = (fn - fs) * idyc;
    const real e = wc ? eC - dtH * (divx + divy) : real(0);
    en[q] = act ? e : real(0);
    ssum += act ? e : real(0);
    wsum += act ? wC : real(0);
  }
  const real wb = cap_weight(P, j, g.ny);
  real zm = 0;
  if (ALLRED || allred || wb >= real(0)) {
    real sw2[2] = {ssum, wsum};
    group_sums<real, 2>(sw2, red, t);
    zm = sw2[0] / (sw2[1] > real(1) ? sw2[1] : real(1));
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int i = t + q * 256;
    if (i >= g.nx) continue;
    const int c2 = j * g.nx + i;
    real e = en[q];
    if (wb >= real(0)) e = wet[c2] != real(0) ? wb * zm + (real(1) - wb) * e : real(0);
    out[i] = e;
    if (out2) bt_store<COH>(out2 + i, e);
  }
}

// The inputs of a row's ubt, vbt update that do not depend on the new eta
// (masks, old velocities, forcing), per column chunk; a kernel may load them
// before its eta phase so that their latency overlaps it.
template <typename real, int NQ>
struct BtVelIn {
  real wC[NQ], wE[NQ], wW[NQ], wN[NQ], wS[NQ], u0[NQ], v0[NQ], fx[NQ], fy[NQ], tx[NQ], ty[NQ];
};

template <typename real, bool COH = false, int NQ = kRowChunk>
__device__ __forceinline__ void bt_vel_load(const Geo<real>& g, const Phys<real>& P, int j, int t,
                                            const real* __restrict__ ubt, const real* __restrict__ vbt,
                                            const real* __restrict__ Frx, const real* __restrict__ Fry,
                                            BtVelIn<real, NQ>& in) {
  const real* wet = P.wet2;
  const bool hasN = j + 1 < g.ny, hasS = j > 0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    // lanes past the row end repeat its last column (their results are not kept)
    const bool act = t + q * 256 < g.nx;
    const int i = act ? t + q * 256 : g.nx - 1;
    const int c2 = j * g.nx + i;
    // every index is in bounds (rows j +- 1 clamped to row j at the walls)
    in.wC[q] = wet[c2];
    in.wE[q] = wet[j * g.nx + wrap_e(i, g.nx)];
    in.wW[q] = wet[j * g.nx + wrap_w(i, g.nx)];
    in.wN[q] = wet[hasN ? c2 + g.nx : c2];
    in.wS[q] = wet[hasS ? c2 - g.nx : c2];
    in.u0[q] = bt_load<COH>(ubt + c2);
    in.v0[q] = bt_load<COH>(vbt + c2);
    in.fx[q] = Frx[c2];
    in.fy[q] = Fry[c2];
    in.tx[q] = P.tau_x[c2];
    in.ty[q] = P.tau_y[c2];
  }
}

// ubt, vbt update of row j from the new eta (rows j-1, j, j+1 given as row
// pointers; en / es are only read when that row exists) and the row's loaded
// inputs: gradient, forcing, legacy implicit Coriolis (iden = 1 / (1 + (f
// dt_bt)^2) of the row), drag, mask, polar-cap blend.  Outputs may alias the
// inputs the row was loaded from.
template <typename real, bool ALLRED = false, bool COH = false, int NQ = kRowChunk>
__device__ __forceinline__ void bt_vel_finish(const Geo<real>& g, const Phys<real>& P, int j, int t,
                                              const BtVelIn<real, NQ>& in, const real* ec_row,
                                              const real* en_row, const real* es_row, real* uo, real* vo, real fd,
                                              real iden, real dt_bt, real drag, real* red, bool store = true,
                                              bool allred = false) {
  real un[NQ], vn[NQ];
  real su = 0, sv = 0, sw = 0;
  // row data in registers: the selects below then only mix loaded values
  const bool hasN = j + 1 < g.ny, hasS = j > 0;
  const real idx = g.idx[j], idyc = g.idyc[j], cfN = g.cfn[j], cfS = g.cfs[j];
  const real* en_safe = hasN ? en_row : ec_row;  // rows past the walls are never read
  const real* es_safe = hasS ? es_row : ec_row;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    // lanes past the row end repeat its last column and book nothing
    const bool act = t + q * 256 < g.nx;
    const int i = act ? t + q * 256 : g.nx - 1;
    const real wC = in.wC[q], wE = in.wE[q], wW = in.wW[q], wN = in.wN[q], wS = in.wS[q];
    const real ec = ec_row[i], eE = ec_row[wrap_e(i, g.nx)], eW = ec_row[wrap_w(i, g.nx)];
    const real eN = en_safe[i], eS = es_safe[i];
    const real u0 = in.u0[q], v0 = in.v0[q];
    const bool wc = wC != real(0);
    const real dxp = (wc && wE != real(0)) ? eE - ec : real(0);
    const real dxm = (wc && wW != real(0)) ? ec - eW : real(0);
    const real gx = idx * real(0.5) * (dxp + dxm);
    const real dyp = (hasN && wc && wN != real(0)) ? (eN - ec) * cfN : real(0);
    const real dym = (hasS && wc && wS != real(0)) ? (ec - eS) * cfS : real(0);
    const real gy = idyc * real(0.5) * (dyp + dym);
    const real Fx = in.fx[q] + in.tx[q] * P.inv_rho0H;
    const real Fy = in.fy[q] + in.ty[q] * P.inv_rho0H;
    const real us = u0 + dt_bt * (real(-kG) * gx + Fx);
    const real vs = v0 + dt_bt * (real(-kG) * gy + Fy);
    real a = (us + fd * vs) * iden;
    real b = (vs - fd * us) * iden;
    a = wc ? a * drag : real(0);
    b = wc ? b * drag : real(0);
    un[q] = act ? a : real(0);
    vn[q] = act ? b : real(0);
    su += act ? a : real(0);
    sv += act ? b : real(0);
    sw += act ? wC : real(0);
  }
  const real wb = cap_weight(P, j, g.ny);
  real zu = 0, zv = 0;
  if (ALLRED || allred || wb >= real(0)) {
    real s3[3] = {su, sv, sw};
    group_sums<real, 3>(s3, red, t);
    const real d = s3[2] > real(1) ? s3[2] : real(1);
    zu = s3[0] / d;
    zv = s3[1] / d;
  }
  if (!store) return;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int i = t + q * 256;
    if (i >= g.nx) continue;
    const int c2 = j * g.nx + i;
    real a = un[q], b = vn[q];
    if (wb >= real(0)) {
      const bool wc = in.wC[q] != real(0);
      a = wc ? wb * zu + (real(1) - wb) * a : real(0);
      b = wc ? wb * zv + (real(1) - wb) * b : real(0);
    }
    bt_store<COH>(uo + c2, a);
    bt_store<COH>(vo + c2, b);
  }
}

// The whole row update: bt_vel_load, then bt_vel_finish.
template <typename real, bool ALLRED = false, bool COH = false, int NQ = kRowChunk>
__device__ __forceinline__ void bt_vel_row(const Geo<real>& g, const Phys<real>& P, int j, int t,
                                           const real* ec_row, const real* en_row, const real* es_row,
                                           const real* __restrict__ ubt, const real* __restrict__ vbt, real* uo,
                                           real* vo, const real* __restrict__ Frx, const real* __restrict__ Fry,
                                           real fd, real iden, real dt_bt, real drag, real* red,
                                           bool store = true, bool allred = false) {
  BtVelIn<real, NQ> in;
  bt_vel_load<real, COH, NQ>(g, P, j, t, ubt, vbt, Frx, Fry, in);
  bt_vel_finish<real, ALLRED, COH, NQ>(g, P, j, t, in, ec_row, en_row, es_row, uo, vo, fd, iden, dt_bt, drag, red,
                                       store, allred);
}

template <typename real>
__global__ __launch_bounds__(256) void k_bt_eta(Geo<real> g, Phys<real> P, real* __restrict__ eta,
                                                const real* __restrict__ ubt, const real* __restrict__ vbt,
                                                real dtH) {
  __shared__ real red[8];
  const int j = blockIdx.x;
  bt_eta_row(g, P, j, (int)threadIdx.x, eta + (long long)j * g.nx, ubt, vbt, dtH, eta + (long long)j * g.nx, red);
}

template <typename real>
__global__ __launch_bounds__(256) void k_bt_vel(Geo<real> g, Phys<real> P, const real* __restrict__ eta,
                                                real* __restrict__ ubt, real* __restrict__ vbt,
                                                const real* __restrict__ Frx, const real* __restrict__ Fry,
                                                const real* __restrict__ fdt, const real* __restrict__ fiden,
                                                real dt_bt, real drag) {
  __shared__ real red[16];
  const int j = blockIdx.x;
  const real* ec = eta + (long long)j * g.nx;
  bt_vel_row(g, P, j, (int)threadIdx.x, ec, ec + g.nx, ec - g.nx, ubt, vbt, ubt, vbt, Frx, Fry, fdt[j], fiden[j], dt_bt,
             drag, red);
}

// One whole forward-backward substep per launch (the arithmetic of k_bt_eta
// followed by k_bt_vel).  A block of 4 row groups owns rows j0, j0 + 1: group
// r computes eta' of row j0 - 1 + r into LDS (the two owned rows also to
// eta_out), then groups 0 and 1 update ubt, vbt of the owned rows from the
// LDS rows.  Inputs and outputs are separate buffers (ping-pong across
// substeps), so a block's halo rows of eta' equal those computed by their
// owners.  In blocks that touch a polar-cap band row every group reduces on
// every row (the zonal means are used on the band rows only), so that all
// groups meet the same block-wide barriers, and groups without a row of their
// own work on a clamped row and store nothing.
template <typename real, int NQ>
__global__ __launch_bounds__(1024) void k_bt_sub(Geo<real> g, Phys<real> P, const real* __restrict__ eta_in,
                                                 const real* __restrict__ u_in, const real* __restrict__ v_in,
                                                 real* __restrict__ eta_out, real* __restrict__ u_out,
                                                 real* __restrict__ v_out, const real* __restrict__ Frx,
                                                 const real* __restrict__ Fry, const real* __restrict__ fdt,
                                                 const real* __restrict__ fiden, real dtH, real dt_bt, real drag) {
  __shared__ real erow[4][kRowChunk * 256];
  __shared__ real red[4][16];
  const int grp = (int)threadIdx.x >> 8, t = (int)threadIdx.x & 255;
  const int j0 = (int)blockIdx.x * 2;
  // blocks whose rows (j0 - 1 .. j0 + 2) all lie outside the polar-cap bands
  // need no zonal means: no group reduces, so no group meets a barrier inside
  // the row functions, and the groups without an owned row skip the
  // velocity phase (block-uniform test)
  const bool band = P.nb > 0 && (j0 - 1 < P.nb || j0 + 2 >= g.ny - P.nb);
  const int jv = j0 + (grp & 1);
  const bool own = grp < 2 && jv < g.ny;
  const int jc = jv < g.ny ? jv : g.ny - 1;
  const bool vel = band || own;  // uniform per row group
  // float32: the velocity phase's inputs are loaded first, so that their
  // latency overlaps the eta phase (float64 keeps them after the barrier: the
  // extra registers would spill)
  constexpr bool kPre = sizeof(real) == 4;
  BtVelIn<real, NQ> vin;
  if (kPre && vel) bt_vel_load<real, false, NQ>(g, P, jc, t, u_in, v_in, Frx, Fry, vin);
  {
    const int je = j0 - 1 + grp;
    const bool in = je >= 0 && je < g.ny;
    const int jc = in ? je : (je < 0 ? 0 : g.ny - 1);
    const bool own = in && (grp == 1 || grp == 2);
    bt_eta_row<real, false, false, NQ>(g, P, jc, t, eta_in + (long long)jc * g.nx, u_in, v_in, dtH, &erow[grp][0],
                                       &red[grp][0], own ? eta_out + (long long)jc * g.nx : (real*)nullptr, band);
  }
  __syncthreads();
  const int sl = jc - (j0 - 1);  // LDS slot of row jc (1 or 2)
  if (vel) {
    if (!kPre) bt_vel_load<real, false, NQ>(g, P, jc, t, u_in, v_in, Frx, Fry, vin);
    bt_vel_finish<real, false, false, NQ>(g, P, jc, t, vin, &erow[sl][0], &erow[sl + 1][0], &erow[sl - 1][0], u_out,
                                          v_out, fdt[jc], fiden[jc], dt_bt, drag, &red[grp][0], own, band);
  }
}

// The whole sub-cycle of n_sub substeps in one persistent launch.  Block b
// keeps k_bt_sub's work (rows 2b, 2b + 1, the same row arithmetic, so every
// substep is bit-identical to k_bt_sub's) and steps through the substeps on
// its own; the only dependence between substeps is on the two neighbouring
// blocks (their halo rows), so instead of a grid-wide barrier each block
// publishes the number of substeps it has finished in progress[b] and, before
// substep q > 0, waits until both neighbours have finished substep q - 1.
// That one wait also orders the ping-pong buffers: the set a block writes in
// substep q was last read by its neighbours in substep q - 1.
// Coherence without cache flushes: every load and store of eta, ubt, vbt is
// an agent-scope relaxed atomic (coherent across the XCDs' L2 caches), each
// wave waits until its stores have completed (s_waitcnt 0) before the block
// barrier, and thread 0 then stores the progress count, also as an
// agent-scope atomic; a waiting block polls the neighbours' counts the same
// way.  (Agent-scope fences would write back and invalidate the whole L2 of
// the XCD from every wave: ten times slower than one launch per substep.)
// The waits are bounded: after kBtSpinLimit polls a wait records a timeout in
// *timeouts and returns, and every later wait returns at once while
// *timeouts is nonzero, so the kernel always finishes (with wrong values,
// which the host detects from the counter).  All blocks must be resident at
// once: the grid is ceil(ny / 2) blocks of 1024 threads, which the host
// allows only up to the device's CU count, and the kernel runs alone on its
// stream.  progress[] must be zero at launch.
constexpr int kBtSpinLimit = 1 << 20;

__device__ __forceinline__ void bt_wait(int* progress, int b, int nblocks, int q, int* timeouts) {
  const int* ps = progress + (b > 0 ? b - 1 : b);
  const int* pn = progress + (b + 1 < nblocks ? b + 1 : b);
  for (int spins = 0;; ++spins) {
    // both counters in flight together (a missing neighbour reads this block's own count, which is >= q)
    const int s = __hip_atomic_load(ps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int n = __hip_atomic_load(pn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (s >= q && n >= q) return;
    if ((spins & 63) == 63) {
      if (__hip_atomic_load(timeouts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
      if (spins >= kBtSpinLimit) {
        __hip_atomic_fetch_add(timeouts, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
    }
  }
}

template <typename real>
__global__ __launch_bounds__(1024) void k_bt_persist(Geo<real> g, Phys<real> P, const real* __restrict__ eta0,
                                                     const real* __restrict__ u0, const real* __restrict__ v0,
                                                     real* ea, real* ua, real* va, real* eb, real* ub, real* vb,
                                                     const real* __restrict__ Frx, const real* __restrict__ Fry,
                                                     const real* __restrict__ fdt, const real* __restrict__ fiden,
                                                     real dtH, real dt_bt, real drag, int nsub, int* progress,
                                                     int* timeouts) {
  __shared__ real erow[4][kRowChunk * 256];
  __shared__ real red[4][16];
  // The rows, wet mask, forcing and wind are the same in every substep; left
  // to itself the compiler hoists their index arithmetic and loads out of the
  // substep loop, which keeps them live across it and spills (a 1024-thread
  // block has 128 VGPRs per thread).  The first row and the base pointers are
  // therefore read from LDS after a barrier in every substep, which the
  // compiler cannot move out of the loop.
  __shared__ const real* invariant[5];
  __shared__ int first_row;
  if (threadIdx.x == 0) {
    invariant[0] = P.wet2;
    invariant[1] = P.tau_x;
    invariant[2] = P.tau_y;
    invariant[3] = Frx;
    invariant[4] = Fry;
    first_row = (int)blockIdx.x * 2;
  }
  const int grp = (int)threadIdx.x >> 8, t = (int)threadIdx.x & 255;
  const int b = (int)blockIdx.x, nblocks = (int)gridDim.x;
  const real *ei = eta0, *ui = u0, *vi = v0;
  for (int q = 0; q < nsub; ++q) {
    // the last substep lands in set a, the sets alternate backwards from there
    const bool last_half = ((nsub - 1 - q) & 1) == 0;
    real* eo = last_half ? ea : eb;
    real* uo = last_half ? ua : ub;
    real* vo = last_half ? va : vb;
    if (q > 0 && threadIdx.x == 0) bt_wait(progress, b, nblocks, q, timeouts);
    __syncthreads();
    Phys<real> Pq = P;
    Pq.wet2 = invariant[0];
    Pq.tau_x = invariant[1];
    Pq.tau_y = invariant[2];
    const real* Frq = invariant[3];
    const real* Fyq = invariant[4];
    const int j0 = first_row;
    {
      const int je = j0 - 1 + grp;
      const bool in = je >= 0 && je < g.ny;
      const int jc = in ? je : (je < 0 ? 0 : g.ny - 1);
      const bool own = in && (grp == 1 || grp == 2);
      bt_eta_row<real, true, true>(g, Pq, jc, t, ei + (long long)jc * g.nx, ui, vi, dtH, &erow[grp][0],
                                   &red[grp][0], own ? eo + (long long)jc * g.nx : (real*)nullptr);
    }
    __syncthreads();
    {
      const int jv = j0 + (grp & 1);
      const bool own = grp < 2 && jv < g.ny;
      const int jc = jv < g.ny ? jv : g.ny - 1;
      const int sl = jc - (j0 - 1);
      bt_vel_row<real, true, true>(g, Pq, jc, t, &erow[sl][0], &erow[sl + 1][0], &erow[sl - 1][0], ui, vi, uo,
                                   vo, Frq, Fyq, fdt[jc], fiden[jc], dt_bt, drag, &red[grp][0], own);
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have completed
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(progress + b, q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ei = eo;
    ui = uo;
    vi = vo;
  }
}

__global__ __launch_bounds__(256) void k_zero_int(int* __restrict__ a, int n) {
  const int i = (int)(blockIdx.x * 256 + threadIdx.x);
  if (i < n) a[i] = 0;
}

// The substep's two kernels for many independent grids in one launch each:
// workgroup b serves row b - row0[m] of member m (row0 ascending, a member
// table in device memory), with exactly the arithmetic of k_bt_eta / k_bt_vel.
template <typename real>
__device__ __forceinline__ int bt_member_of(const BtMember<real>* __restrict__ mem, int nmem, int b) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {  // last member with row0 <= b
    const int mid = (lo + hi + 1) >> 1;
    if (mem[mid].row0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <typename real>
__global__ __launch_bounds__(256) void k_bt_eta_multi(const BtMember<real>* __restrict__ mem, int nmem) {
  __shared__ real red[8];
  const BtMember<real>& m = mem[bt_member_of(mem, nmem, (int)blockIdx.x)];
  const int j = (int)blockIdx.x - m.row0;
  real* row = m.eta + (long long)j * m.g.nx;
  bt_eta_row(m.g, m.P, j, (int)threadIdx.x, row, m.ubt, m.vbt, m.dtH, row, red);
}

template <typename real>
__global__ __launch_bounds__(256) void k_bt_vel_multi(const BtMember<real>* __restrict__ mem, int nmem) {
  __shared__ real red[16];
  const BtMember<real>& m = mem[bt_member_of(mem, nmem, (int)blockIdx.x)];
  const int j = (int)blockIdx.x - m.row0;
  const real* ec = m.eta + (long long)j * m.g.nx;
  bt_vel_row(m.g, m.P, j, (int)threadIdx.x, ec, ec + m.g.nx, ec - m.g.nx, m.ubt, m.vbt, m.ubt, m.vbt, m.Frx, m.Fry,
             m.fdt[j], m.fiden[j], m.dt_bt, m.drag, red);
}

// ---------------------------------------------------------------------------
// Final masks (step.py:419-450): project the barotropic delta on every node,
// polar-cap u, v, T, S with the 3-D wet mask, hold land/ghost at the pre-step
// values and zero v on the wall rows.  k_cap_means reduces the band rows.
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(256) void k_cap_means(Geo<real> g, Phys<real> P, const real* __restrict__ u,
                                                   const real* __restrict__ v, const real* __restrict__ T,
                                                   const real* __restrict__ S, const real* __restrict__ ubt,
                                                   const real* __restrict__ ubt0, const real* __restrict__ vbt,
                                                   const real* __restrict__ vbt0, real* __restrict__ means) {
  __shared__ real red[20];
  const int band_row = blockIdx.x % (2 * P.nb);
  const int k = blockIdx.x / (2 * P.nb);
  const int j = band_row < P.nb ? band_row : g.ny - 2 * P.nb + band_row;
  const long long plane = (long long)k * g.n2;
  real su = 0, sv = 0, sT = 0, sS = 0, sw = 0;
  for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
    const int c2 = j * g.nx + i;
    if (k < g.nwet[c2]) {
      su += u[plane + c2] + (ubt[c2] - ubt0[c2]);
      sv += v[plane + c2] + (vbt[c2] - vbt0[c2]);
      sT += T[plane + c2];
      sS += S[plane + c2];
      sw += real(1);
    }
  }
  const int t = (int)threadIdx.x;  // one 256-thread group per block
  real s5[5] = {su, sv, sT, sS, sw};
  group_sums<real, 5>(s5, red, t);
  su = s5[0];
  sv = s5[1];
  sT = s5[2];
  sS = s5[3];
  sw = s5[4];
  if (threadIdx.x == 0) {
    const real d = sw > real(1) ? sw : real(1);
    const long long o = (long long)k * 2 * P.nb + band_row;
    const long long stride = (long long)g.nz * 2 * P.nb;
    means[o] = su / d;
    means[o + stride] = sv / d;
    means[o + 2 * stride] = sT / d;
    means[o + 3 * stride] = sS / d;
  }
}

template <typename real>
__global__ __launch_bounds__(256) void k_final(Geo<real> g, Phys<real> P, const real* __restrict__ u3,
                                               const real* __restrict__ v3, const real* __restrict__ T3,
                                               const real* __restrict__ S3, const real* __restrict__ ubt,
                                               const real* __restrict__ ubt0, const real* __restrict__ vbt,
                                               const real* __restrict__ vbt0, const real* __restrict__ means,
                                               const real* __restrict__ uin, const real* __restrict__ vin,
                                               const real* __restrict__ Tin, const real* __restrict__ Sin,
                                               real* __restrict__ uo, real* __restrict__ vo,
                                               real* __restrict__ To, real* __restrict__ So,
                                               const real* __restrict__ eta_src, real* __restrict__ eta_dst,
                                               const real* __restrict__ ice_src, real* __restrict__ ice_dst) {
  ZM_TILE_COORDS_LEVEL_FASTEST
  const int c2 = j * g.nx + i;
  if (k == 0) {  // the step's 2-D outputs: eta of the last substep, ice unchanged
    eta_dst[c2] = eta_src[c2];
    if (ice_dst) ice_dst[c2] = ice_src[c2];
  }
  const bool wc = k < g.nwet[c2];
  if (!wc) {  // land and ghost nodes hold their pre-step values
    uo[c] = uin[c];
    vo[c] = (j == 0 || j == g.ny - 1) ? real(0) : vin[c];
    To[c] = Tin[c];
    So[c] = Sin[c];
    return;
  }
  real u = u3[c] + (ubt[c2] - ubt0[c2]);
  real v = v3[c] + (vbt[c2] - vbt0[c2]);
  real T = T3[c], S = S3[c];
  const real wb = P.nb > 0 ? cap_weight(P, j, g.ny) : real(-1);
  if (wb >= real(0)) {
    const int band_row = j < P.nb ? j : j - (g.ny - 2 * P.nb);
    const long long o = (long long)k * 2 * P.nb + band_row;
    const long long stride = (long long)g.nz * 2 * P.nb;
    const real om = real(1) - wb;
    u = wb * means[o] + om * u;
    v = wb * means[o + stride] + om * v;
    T = wb * means[o + 2 * stride] + om * T;
    S = wb * means[o + 3 * stride] + om * S;
  }
  if (j == 0 || j == g.ny - 1) v = real(0);
  uo[c] = u;
  vo[c] = v;
  To[c] = T;
  So[c] = S;
}

// Diagnostics (factory.py:752-760): rho = rho0 + rho', p = rho0 g eta + p_bc.
template <typename real>
__global__ __launch_bounds__(256) void k_diag(Geo<real> g, Phys<real> P, const real* __restrict__ T,
                                              const real* __restrict__ S, const real* __restrict__ eta,
                                              const real* __restrict__ pbc, const real* __restrict__ Fz,
                                              real* __restrict__ rho, real* __restrict__ pres,
                                              real* __restrict__ w) {
  ZM_TILE_COORDS
  const int c2 = j * g.nx + i;
  rho[c] = real(kRho0) + rho_anom(T[c], S[c], P);
  pres[c] = real(kRho0 * kG) * eta[c2] + pbc[c];
  w[c] = k < g.nwet[c2] ? -Fz[c] : real(0);
}

template <typename real>
__global__ __launch_bounds__(256) void k_copy(const real* __restrict__ a, real* __restrict__ b, long long n) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) b[t] = a[t];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(long long n, int per) { return (unsigned)((n + per - 1) / per); }

template <typename real>
static inline unsigned tiles3(const Geo<real>& g) {
  return (unsigned)((long long)((g.nx + kTileI - 1) / kTileI) * ((g.ny + kTileJ - 1) / kTileJ) * g.nz);
}

// The row coefficients in g.nu_met carry nu_h dts: a launch for another
// pair would silently use the wrong diffusion.
template <typename real>
static void check_nu_coefficients(const Geo<real>& g, real nu_h, real dts) {
  if (nu_h != g.nu_h_met || dts != g.nu_dts_met)
    throw std::runtime_error("nu_h sub-cycle: the row coefficients were built for nu_h " +
                             std::to_string((double)g.nu_h_met) + ", substep " + std::to_string((double)g.nu_dts_met) +
                             " s, not nu_h " + std::to_string((double)nu_h) + ", substep " +
                             std::to_string((double)dts) + " s");
}

template <typename real>
void launch_nu_substep(const Geo<real>& g, const real* u, const real* v, real* uo, real* vo, real* lu,
                       real* lv, real nu_h, real dts, hipStream_t s) {
  check_nu_coefficients(g, nu_h, dts);
  hipLaunchKernelGGL(k_nu_substep<real>, dim3(tiles3(g)), dim3(256), 0, s, g, u, v, uo, vo, lu, lv);
}

template <typename real>
int nu_block_max() {
  return NuTile<real>::HALO;
}

template <typename real, int H, int HALO>
static void launch_nu_strip(const Geo<real>& g, const real* u, const real* v, real* uo, real* vo, int nsub,
                            hipStream_t s) {
  constexpr int OX = 64 - 2 * HALO, OY = H - 2 * HALO;
  const long long tiles = (long long)((g.nx + OX - 1) / OX) * ((g.ny + OY - 1) / OY) * g.nz;
  hipLaunchKernelGGL((k_nu_strip<real, H, HALO>), dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s, g, u, v, uo,
                     vo, nsub);
}

// nsub substeps per launch on the narrowest halo that covers them: a 3-wide
// halo computes 12 % more useful nodes per tile than a 4-wide one.
template <typename real>
bool nu_block_writes_lap(const Geo<real>& g, int nsub) {
  return nsub >= 2 && nsub <= 4 && g.nu_chunk > 0;
}

template <typename real>
void launch_nu_block(const Geo<real>& g, const real* u, const real* v, real* uo, real* vo, real nu_h, real dts,
                     int nsub, hipStream_t s, real* lu, real* lv) {
  check_nu_coefficients(g, nu_h, dts);
  if (nsub >= 2 && nsub <= 4 && g.nu_chunk > 0) {
    const int chunk = g.nu_chunk;
    auto go = [&](auto kern, int ns) {
      const int OX = 64 - 2 * ns;
      const long long waves = (long long)((g.nx + OX - 1) / OX) * ((g.ny + chunk - 1) / chunk) * g.nz;
      const unsigned blocks = (unsigned)((waves + 3) / 4);
      hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, g, u, v, uo, vo, lu, lv, chunk);
    };
    if (lu) {
      if (nsub == 2) go(k_nu_stream<real, 2, true>, 2);
      else if (nsub == 3) go(k_nu_stream<real, 3, true>, 3);
      else go(k_nu_stream<real, 4, true>, 4);
    } else {
      if (nsub == 2) go(k_nu_stream<real, 2, false>, 2);
      else if (nsub == 3) go(k_nu_stream<real, 3, false>, 3);
      else go(k_nu_stream<real, 4, false>, 4);
    }
    return;
  }
  if (lu) throw std::runtime_error("launch_nu_block: lap output needs the streaming kernel (2-4 substeps)");
  using T = NuTile<real>;
  if (nsub <= 3)
    launch_nu_strip<real, T::H, 3>(g, u, v, uo, vo, nsub, s);
  else
    launch_nu_strip<real, T::H, T::HALO>(g, u, v, uo, vo, nsub, s);
}

template <typename real>
void launch_d1_unit(const Geo<real>& g, const real* T, const real* S, real* dT, real* dS, hipStream_t s) {
  hipLaunchKernelGGL(k_d1_unit<real>, dim3(segs3(g)), dim3(256), 0, s, g, T, S, dT, dS);
}

template <typename real>
void launch_linear_finish(const Geo<real>& g, const Phys<real>& P, const real* u0, const real* v0,
                          const real* T0, const real* S0, const real* unu, const real* vnu, const real* Lu0,
                          const real* Lv0, const real* DT0, const real* DS0, real* uo, real* vo, real* To,
                          real* So, hipStream_t s) {
  hipLaunchKernelGGL(k_linear_finish<real>, dim3(segs3(g)), dim3(256), 0, s, g, P, u0, v0, T0,
                     S0, unu, vnu, Lu0, Lv0, DT0, DS0, uo, vo, To, So);
}

template <typename real>
void launch_fz(const Geo<real>& g, const real* u, const real* v, real* Fz, hipStream_t s) {
  hipLaunchKernelGGL(k_fz<real>, dim3(blocks_for(g.n2, 256)), dim3(256), 0, s, g, u, v, Fz);
}

template <typename real>
void launch_conv(const Geo<real>& g, const Phys<real>& P, const real* T, const real* S, real* cT, real* cS,
                 int* gate, unsigned long long* ubits, hipStream_t s) {
  hipLaunchKernelGGL(k_conv_gate<real>, dim3(blocks_for(g.n2, 64)), dim3(256), 0, s, g, P, T, S, gate, ubits);
  if (P.conv_nsub >= 1 && P.conv_nsub <= 4) {
    const unsigned b = blocks_for(g.n2, 64);  // 64 columns x kColSegs level segments per block
    const auto cg = (const int*)gate;
    const auto cu = (const unsigned long long*)ubits;
    if (P.conv_nsub == 1) hipLaunchKernelGGL((k_conv_stream<real, 1>), dim3(b), dim3(256), 0, s, g, P, T, S, cT, cS, cg, cu);
    else if (P.conv_nsub == 2) hipLaunchKernelGGL((k_conv_stream<real, 2>), dim3(b), dim3(256), 0, s, g, P, T, S, cT, cS, cg, cu);
    else if (P.conv_nsub == 3) hipLaunchKernelGGL((k_conv_stream<real, 3>), dim3(b), dim3(256), 0, s, g, P, T, S, cT, cS, cg, cu);
    else hipLaunchKernelGGL((k_conv_stream<real, 4>), dim3(b), dim3(256), 0, s, g, P, T, S, cT, cS, cg, cu);
    return;
  }
  const size_t lds = (size_t)4 * g.nz * 64 * sizeof(real);
  if (lds > 65536) {
    static bool raised = false;  // once per instantiation (one per real type)
    if (!raised) {
      hipFuncSetAttribute(reinterpret_cast<const void*>(k_conv<real>), hipFuncAttributeMaxDynamicSharedMemorySize,
                          160 * 1024);
      raised = true;
    }
  }
  hipLaunchKernelGGL(k_conv<real>, dim3(blocks_for(g.n2, 64)), dim3(64), lds, s, g, P, T, S, cT, cS,
                     (const int*)gate, (const unsigned long long*)ubits);
}

template <typename real>
void launch_tracer_residual(const Geo<real>& g, const Phys<real>& P, int mode, const real* u, const real* v,
                            const real* T, const real* S, const real* Fz, const real* cT, const real* cS,
                            const int* gate, real* dT1, real* dS1, const real* Tb, const real* Sb, real* To,
                            real* So, hipStream_t s) {
  const double* nd = nullptr;
  double* np_ = nullptr;
  if (P.adv_scheme == 2)
    hipLaunchKernelGGL((k_tracer_residual<real, 2, false>), dim3(segs3(g)), dim3(256), 0, s, g, P, mode, u, v,
                     T, S, Fz, cT, cS, gate, dT1, dS1, Tb, Sb, To, So, nd, nd, np_);
  else if (P.adv_scheme == 1)
    hipLaunchKernelGGL((k_tracer_residual<real, 1, false>), dim3(segs3(g)), dim3(256), 0, s, g, P, mode, u, v,
                     T, S, Fz, cT, cS, gate, dT1, dS1, Tb, Sb, To, So, nd, nd, np_);
  else
    hipLaunchKernelGGL((k_tracer_residual<real, 0, false>), dim3(segs3(g)), dim3(256), 0, s, g, P, mode, u, v,
                     T, S, Fz, cT, cS, gate, dT1, dS1, Tb, Sb, To, So, nd, nd, np_);
}

template <typename real>
void launch_tracer_terms(const Geo<real>& g, const Phys<real>& P, const real* u, const real* v, const real* T,
                         const real* S, const real* Fz, const real* cT, const real* cS, const int* gate, real* adv,
                         real* diff_h, real* diff_v, real* conv, hipStream_t s) {
  const double* nd = nullptr;
  double* np_ = nullptr;
  const real* nr = nullptr;
  if (P.adv_scheme == 2)
    hipLaunchKernelGGL((k_tracer_residual<real, 2, false, true>), dim3(segs3(g)), dim3(256), 0, s, g, P, 0, u, v,
                       T, S, Fz, cT, cS, gate, adv, diff_h, nr, nr, diff_v, conv, nd, nd, np_);
  else if (P.adv_scheme == 1)
    hipLaunchKernelGGL((k_tracer_residual<real, 1, false, true>), dim3(segs3(g)), dim3(256), 0, s, g, P, 0, u, v,
                       T, S, Fz, cT, cS, gate, adv, diff_h, nr, nr, diff_v, conv, nd, nd, np_);
  else
    hipLaunchKernelGGL((k_tracer_residual<real, 0, false, true>), dim3(segs3(g)), dim3(256), 0, s, g, P, 0, u, v,
                       T, S, Fz, cT, cS, gate, adv, diff_h, nr, nr, diff_v, conv, nd, nd, np_);
}

template <typename real>
unsigned tracer_ledger_blocks(const Geo<real>& g) {
  return segs3(g);
}

template <typename real>
void launch_tracer_ledger(const Geo<real>& g, const Phys<real>& P, const real* u, const real* v, const real* T,
                          const real* S, const real* Fz, const real* cT, const real* cS, const int* gate,
                          const double* area_row, const double* dzn64, double* lpart, hipStream_t s) {
  real* nr = nullptr;
  const real* cr = nullptr;
  if (P.adv_scheme == 2)
    hipLaunchKernelGGL((k_tracer_residual<real, 2, true>), dim3(segs3(g)), dim3(256), 0, s, g, P, 0, u, v, T,
                     S, Fz, cT, cS, gate, nr, nr, cr, cr, nr, nr, area_row, dzn64, lpart);
  else if (P.adv_scheme == 1)
    hipLaunchKernelGGL((k_tracer_residual<real, 1, true>), dim3(segs3(g)), dim3(256), 0, s, g, P, 0, u, v, T,
                     S, Fz, cT, cS, gate, nr, nr, cr, cr, nr, nr, area_row, dzn64, lpart);
  else
    hipLaunchKernelGGL((k_tracer_residual<real, 0, true>), dim3(segs3(g)), dim3(256), 0, s, g, P, 0, u, v, T,
                     S, Fz, cT, cS, gate, nr, nr, cr, cr, nr, nr, area_row, dzn64, lpart);
}

template <typename real>
void launch_pbc(const Geo<real>& g, const Phys<real>& P, const real* T, const real* S, real* pbc,
                hipStream_t s) {
  hipLaunchKernelGGL(k_pbc<real>, dim3(blocks_for(g.n2, 256)), dim3(256), 0, s, g, P, T, S, pbc);
}

template <typename real>
void launch_bt_rho(const Geo<real>& g, const Phys<real>& P, const real* pbc, real* Fx, real* Fy, const real* u,
                   const real* v, real* ubt0, real* vbt0, hipStream_t s) {
  hipLaunchKernelGGL(k_bt_rho<real>, dim3(blocks_for(g.n2, 256)), dim3(256), 0, s, g, P, pbc, Fx, Fy, u, v,
                     ubt0, vbt0);
}

template <typename real>
void launch_mom_adv(const Geo<real>& g, const real* u, const real* v, const real* Fz, real* au, real* av,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_mom_adv<real>, dim3(segs3(g)), dim3(256), 0, s, g, u, v, Fz, au, av);
}

template <typename real>
void launch_dealias_x(int N, int N1, int N2, const int* kept, int nrows, const real* twr, const real* twi,
                      const real* ina, const real* inb, real* outa, real* outb, hipStream_t s) {
  const size_t lds = (size_t)6 * N * sizeof(real);
  hipLaunchKernelGGL(k_dealias_x<real>, dim3(nrows), dim3(256), lds, s, N, N1, N2, kept, nrows, twr, twi, ina,
                     inb, outa, outb);
}

template <typename real>
void launch_dealias_fft(int N, const FftPlan& plan, const int* kept, int nrows, const real* twr, const real* twi,
                        const real* ina, const real* inb, real* outa, real* outb, hipStream_t s) {
  if constexpr (sizeof(real) == 4) {
#define ZM_TWO_FACTOR(LEN)                                                                                       if (N == LEN) {                                                                                                 using Sh = TwoFactorShape<LEN>;                                                                               hipLaunchKernelGGL((k_dealias_2f<LEN>), dim3((unsigned)((nrows + Sh::RB - 1) / Sh::RB)), dim3(256),                              Sh::lds(), s, kept, nrows, (const float*)twr, (const float*)twi, (const float*)ina,                           (const float*)inb, (float*)outa, (float*)outb);                                            return;                                                                                                     }
    ZM_TWO_FACTOR(720)
    ZM_TWO_FACTOR(800)
    ZM_TWO_FACTOR(480)
    ZM_TWO_FACTOR(360)
    ZM_TWO_FACTOR(180)
    ZM_TWO_FACTOR(90)
#undef ZM_TWO_FACTOR
  }
  constexpr int RB = fixed_fft_rows<real>();
  const size_t lds_fixed = (size_t)(1 + 2 * RB) * N * sizeof(Cx<real>);
  const unsigned fixed_blocks = (unsigned)((nrows + RB - 1) / RB);
#define ZM_FIXED_FFT(LEN)                                                                                    if (N == LEN) {                                                                                              hipLaunchKernelGGL((k_dealias_fixed<real, LEN>), dim3(fixed_blocks), dim3(256), lds_fixed, s, kept, nrows,                        twr, twi, ina, inb, outa, outb);                                                        return;                                                                                                  }
  ZM_FIXED_FFT(720)
  ZM_FIXED_FFT(800)
  ZM_FIXED_FFT(480)
  ZM_FIXED_FFT(360)
  ZM_FIXED_FFT(180)
  ZM_FIXED_FFT(90)
#undef ZM_FIXED_FFT
  const size_t lds = (size_t)6 * N * sizeof(real);
  const unsigned blocks = (unsigned)((nrows + kFftRowsPerBlock - 1) / kFftRowsPerBlock);
  hipLaunchKernelGGL(k_dealias_fft<real>, dim3(blocks), dim3(256), lds, s, N, plan, kept, nrows, twr, twi, ina, inb,
                     outa, outb);
}

template <typename real>
void launch_mom_residual(const Geo<real>& g, const Phys<real>& P, int mode, const real* ax, const real* ay,
                         const real* u, const real* v, const real* pbc, const real* eta, const real* Frx,
                         const real* Fry, real* du1, real* dv1, real* uo, real* vo, real decay_bot,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_mom_residual<real>, dim3(segs3(g)), dim3(256), 0, s, g, P, mode, ax, ay, u,
                     v, pbc, eta, Frx, Fry, du1, dv1, uo, vo, decay_bot);
}

template <typename real>
void launch_bt_eta(const Geo<real>& g, const Phys<real>& P, real* eta, const real* ubt, const real* vbt,
                   real dtH, hipStream_t s) {
  hipLaunchKernelGGL(k_bt_eta<real>, dim3(g.ny), dim3(256), 0, s, g, P, eta, ubt, vbt, dtH);
}

template <typename real>
void launch_bt_vel(const Geo<real>& g, const Phys<real>& P, const real* eta, real* ubt, real* vbt,
                   const real* Frx, const real* Fry, const real* fdt, const real* fiden, real dt_bt, real drag,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_bt_vel<real>, dim3(g.ny), dim3(256), 0, s, g, P, eta, ubt, vbt, Frx, Fry, fdt, fiden,
                     dt_bt, drag);
}

template <typename real>
void launch_bt_sub(const Geo<real>& g, const Phys<real>& P, const real* eta_in, const real* u_in,
                   const real* v_in, real* eta_out, real* u_out, real* v_out, const real* Frx, const real* Fry,
                   const real* fdt, const real* fiden, real dtH, real dt_bt, real drag, hipStream_t s) {
  // column chunks of 256 per row: exactly as many as the row needs, so that
  // no thread runs a chunk past the row end
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)((g.ny + 1) / 2)), dim3(1024), 0, s, g, P, eta_in, u_in, v_in, eta_out,
                       u_out, v_out, Frx, Fry, fdt, fiden, dtH, dt_bt, drag);
  };
  switch ((g.nx + 255) / 256) {
    case 1: go(k_bt_sub<real, 1>); break;
    case 2: go(k_bt_sub<real, 2>); break;
    case 3: go(k_bt_sub<real, 3>); break;
    default: go(k_bt_sub<real, kRowChunk>); break;
  }
}

template <typename real>
void launch_bt_persist(const Geo<real>& g, const Phys<real>& P, const real* eta0, const real* u0, const real* v0,
                       real* ea, real* ua, real* va, real* eb, real* ub, real* vb, const real* Frx, const real* Fry,
                       const real* fdt, const real* fiden, real dtH, real dt_bt, real drag, int nsub, int* progress,
                       int* timeouts, hipStream_t s) {
  const int nblocks = (g.ny + 1) / 2;
  hipLaunchKernelGGL(k_zero_int, dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, s, progress, nblocks);
  hipLaunchKernelGGL(k_bt_persist<real>, dim3((unsigned)nblocks), dim3(1024), 0, s, g, P, eta0, u0, v0, ea, ua, va,
                     eb, ub, vb, Frx, Fry, fdt, fiden, dtH, dt_bt, drag, nsub, progress, timeouts);
}

template <typename real>
void launch_bt_multi(const BtMember<real>* table, int nmem, int total_rows, hipStream_t s) {
  hipLaunchKernelGGL(k_bt_eta_multi<real>, dim3(total_rows), dim3(256), 0, s, table, nmem);
  hipLaunchKernelGGL(k_bt_vel_multi<real>, dim3(total_rows), dim3(256), 0, s, table, nmem);
}

template <typename real>
void launch_cap_means(const Geo<real>& g, const Phys<real>& P, const real* u, const real* v, const real* T,
                      const real* S, const real* ubt, const real* ubt0, const real* vbt, const real* vbt0,
                      real* means, hipStream_t s) {
  if (P.nb <= 0) return;
  hipLaunchKernelGGL(k_cap_means<real>, dim3(2 * P.nb * g.nz), dim3(256), 0, s, g, P, u, v, T, S, ubt, ubt0,
                     vbt, vbt0, means);
}

template <typename real>
void launch_final(const Geo<real>& g, const Phys<real>& P, const real* u3, const real* v3, const real* T3,
                  const real* S3, const real* ubt, const real* ubt0, const real* vbt, const real* vbt0,
                  const real* means, const real* uin, const real* vin, const real* Tin, const real* Sin,
                  real* uo, real* vo, real* To, real* So, const real* eta_src, real* eta_dst,
                  const real* ice_src, real* ice_dst, hipStream_t s) {
  hipLaunchKernelGGL(k_final<real>, dim3(tiles3(g)), dim3(256), 0, s, g, P, u3, v3, T3, S3, ubt,
                     ubt0, vbt, vbt0, means, uin, vin, Tin, Sin, uo, vo, To, So, eta_src, eta_dst, ice_src, ice_dst);
}

template <typename real>
void launch_diag(const Geo<real>& g, const Phys<real>& P, const real* T, const real* S, const real* eta,
                 const real* pbc, const real* Fz, real* rho, real* pres, real* w, hipStream_t s) {
  hipLaunchKernelGGL(k_diag<real>, dim3(tiles3(g)), dim3(256), 0, s, g, P, T, S, eta, pbc, Fz, rho,
                     pres, w);
}

template <typename real>
void launch_copy(const real* a, real* b, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy<real>, dim3(blocks_for(n, 256)), dim3(256), 0, s, a, b, n);
}

#define ZM_INSTANTIATE(real)                                                                                 \
  template void launch_nu_substep<real>(const Geo<real>&, const real*, const real*, real*, real*, real*,    \
                                        real*, real, real, hipStream_t);                                      \
  template void launch_d1_unit<real>(const Geo<real>&, const real*, const real*, real*, real*, hipStream_t); \
  template int nu_block_max<real>();                                                                         \
  template bool nu_block_writes_lap<real>(const Geo<real>&, int);                                            \
  template void launch_nu_block<real>(const Geo<real>&, const real*, const real*, real*, real*, real, real,  \
                                      int, hipStream_t, real*, real*);                                       \
  template void launch_linear_finish<real>(const Geo<real>&, const Phys<real>&, const real*, const real*,   \
                                           const real*, const real*, const real*, const real*, const real*, \
                                           const real*, const real*, const real*, real*, real*, real*,      \
                                           real*, hipStream_t);                                              \
  template void launch_fz<real>(const Geo<real>&, const real*, const real*, real*, hipStream_t);            \
  template void launch_conv<real>(const Geo<real>&, const Phys<real>&, const real*, const real*, real*,     \
                                  real*, int*, unsigned long long*, hipStream_t);                            \
  template void launch_tracer_terms<real>(const Geo<real>&, const Phys<real>&, const real*, const real*,     \
                                          const real*, const real*, const real*, const real*, const real*,   \
                                          const int*, real*, real*, real*, real*, hipStream_t);              \
  template void launch_tracer_residual<real>(const Geo<real>&, const Phys<real>&, int, const real*,         \
                                             const real*, const real*, const real*, const real*,            \
                                             const real*, const real*, const int*, real*, real*,            \
                                             const real*, const real*, real*, real*, hipStream_t);           \
  template unsigned tracer_ledger_blocks<real>(const Geo<real>&);                                          \
  template void launch_tracer_ledger<real>(const Geo<real>&, const Phys<real>&, const real*, const real*,   \
                                           const real*, const real*, const real*, const real*, const real*, \
                                           const int*, const double*, const double*, double*, hipStream_t); \
  template void launch_pbc<real>(const Geo<real>&, const Phys<real>&, const real*, const real*, real*,      \
                                 hipStream_t);                                                               \
  template void launch_bt_rho<real>(const Geo<real>&, const Phys<real>&, const real*, real*, real*,         \
                                    const real*, const real*, real*, real*, hipStream_t);                    \
  template void launch_mom_adv<real>(const Geo<real>&, const real*, const real*, const real*, real*, real*,  \
                                     hipStream_t);                                                           \
  template void launch_dealias_x<real>(int, int, int, const int*, int, const real*, const real*, const real*, \
                                       const real*, real*, real*, hipStream_t);                              \
  template void launch_dealias_fft<real>(int, const FftPlan&, const int*, int, const real*, const real*,     \
                                         const real*, const real*, real*, real*, hipStream_t);                \
  template void launch_mom_residual<real>(const Geo<real>&, const Phys<real>&, int, const real*,            \
                                          const real*, const real*, const real*, const real*, const real*,  \
                                          const real*, const real*, real*, real*, real*, real*, real,       \
                                          hipStream_t);                                                      \
  template void launch_bt_eta<real>(const Geo<real>&, const Phys<real>&, real*, const real*, const real*,   \
                                    real, hipStream_t);                                                      \
  template void launch_bt_vel<real>(const Geo<real>&, const Phys<real>&, const real*, real*, real*,         \
                                    const real*, const real*, const real*, const real*, real, real,         \
                                    hipStream_t);                                                            \
  template void launch_bt_multi<real>(const BtMember<real>*, int, int, hipStream_t);                        \
  template void launch_bt_persist<real>(const Geo<real>&, const Phys<real>&, const real*, const real*,        \
                                        const real*, real*, real*, real*, real*, real*, real*, const real*,     \
                                        const real*, const real*, const real*, real, real, real, int, int*,     \
                                        int*, hipStream_t);                                                     \
  template void launch_bt_sub<real>(const Geo<real>&, const Phys<real>&, const real*, const real*, const real*,  \
                                    real*, real*, real*, const real*, const real*, const real*, const real*,   \
                                    real, real, real, hipStream_t);                                             \
  template void launch_cap_means<real>(const Geo<real>&, const Phys<real>&, const real*, const real*,       \
                                       const real*, const real*, const real*, const real*, const real*,     \
                                       const real*, real*, hipStream_t);                                     \
  template void launch_final<real>(const Geo<real>&, const Phys<real>&, const real*, const real*,           \
                                   const real*, const real*, const real*, const real*, const real*,         \
                                   const real*, const real*, const real*, const real*, const real*,         \
                                   const real*, real*, real*, real*, real*, const real*, real*, const real*, \
                                   real*, hipStream_t);                                                      \
  template void launch_diag<real>(const Geo<real>&, const Phys<real>&, const real*, const real*,            \
                                  const real*, const real*, const real*, real*, real*, real*, hipStream_t); \
  template void launch_copy<real>(const real*, real*, long long, hipStream_t);

ZM_INSTANTIATE(float)
ZM_INSTANTIATE(double)

}  // namespace zm
