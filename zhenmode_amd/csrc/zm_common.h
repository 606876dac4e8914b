// Shared device-side definitions for the mode-split step kernels.
//
// Layout: every 3-D field is stored i-innermost, index ((k * ny) + j) * nx + i,
// with i the periodic longitude, j the latitude (0 = south) and k the level
// (0 = surface).  2-D fields are (j * nx + i).  The wet mask is contiguous from
// the surface in every column, so it is carried as nwet[j*nx+i], the number of
// wet levels: level k is wet iff k < nwet.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zm {

constexpr double kG = 9.81;
constexpr double kRho0 = 1025.0;
constexpr double kAlpha = 2.0e-4;
constexpr double kBeta = 7.6e-4;
constexpr double kCp = 3992.0;

// Geometry: metric rows, vertical coefficients and the wet-level count.
template <typename real>
struct Geo {
  int nx, ny, nz;
  int n2;       // nx * ny
  long long n3; // nx * ny * nz
  // per latitude row [ny]
  const real* cosl;   // cos(lat)
  const real* idx;    // 1/dx
  const real* idx2;   // 1/dx^2
  const real* tanR;   // tan(lat)/R (from f of row 0)
  const real* frow;   // f = 2 Omega sin(lat)
  const real* cfn;    // 0.5 (cos_j + cos_{j+1 mod ny})
  const real* cfs;    // 0.5 (cos_j + cos_{j-1 mod ny})
  const real* idyc;   // (1/dy) / cos_j
  real idy, idy2, hidy;   // 1/dy, 1/dy^2, 0.5/dy
  // vertical
  const real* dzn;    // node thickness [nz]
  const real* idzn;   // 1/dz_node [nz]
  const real* dzi;    // interface distance [nz-1]
  const real* idzi;   // 1/dz_iface [nz-1]
  const real* dz3;    // layer thickness dz [nz-1]
  const real* dznorm; // dz/H_sw [nz-1]
  const real* hm;     // |z_{k-1}-z_k| for interior k (index k-1) [nz-2]
  const real* hp;     // |z_{k+1}-z_k| [nz-2]
  const real* hmp;    // hm + hp [nz-2]
  const real* iden;   // 1 / (hm hp (hm+hp)/2) [nz-2]
  const real* iddz;   // 1 / ((z_{k+1}-z_k) + (z_k-z_{k-1})) [nz-2]
  real idz_top, idz_bot;  // 1/(z1-z0), 1/(z_{nz-1}-z_{nz-2}) (negative)
  real ih0top2, ih0bot2;  // 1/h0^2 at top and bottom
  const int* nwet;        // wet levels per column [n2]
  int seg;                // levels per thread in the column-marching kernels
  int nu_chunk;           // rows per wave of the streaming nu_h kernel (0: register-tile kernel)
  const unsigned long long* nu_wet;  // streaming nu_h kernel: wet words [nz][chunks][nx]
  const real* nu_met;                // nu_h substep row coefficients kE, kN, kS, kY, 1/dx^2, tan(lat)/R
                                     // [ny + 2 kNuMetPad][kNuMetStride], for nu_h_met and nu_dts_met
  real nu_h_met, nu_dts_met;
};

// Scalars and 2-D fields of the physics.
template <typename real>
struct Phys {
  real dt, dth;                // baroclinic dt and dt/2
  real nu_h, nu_v, kappa_h, nu_bi, kappa_bi, kappa_conv;
  real T_ref, S_ref, r_bot;
  real H_sw, dz_surface;
  real wind_factor;            // 1 / (rho0 dz_surface)
  real inv_rho0H;              // 1 / (rho0 H_sw)
  real lambda_bulk, restore_coef_S, ice_freeze, ice_salt_flux;
  real heat_factor;            // 1 / (dz_surface rho0 cp)
  int conv_nsub;
  int cons_kv;                 // conservative_kv: interface-flux vertical tracer diffusion
  int local_conv;              // localize_conv: convective flux needs its own interface unstable
  int adv_scheme;              // horizontal tracer face values: 0 centred, 1 donor cell, 2 minmod (fct_adv)
  int ncap, nb;                // polar cap rows and band width (ncap + taper)
  const real* capw;            // pole-inward blend weights [nb]
  const real* wet2;            // wet_mask [n2]
  const real* kh2;             // kappa_h + coastal kappa_h [n2]
  const real* kv2;             // kappa_v + coastal kappa_v [n2]
  const real* tau_x;           // [n2]
  const real* tau_y;
  const real* q_heat;
  const real* t_atm;
  const real* s_ref2;          // SSS target [n2]
  const real* crc;             // coastal restore coefficient [n2]
  const real* crT;             // coastal restore target [n2]
  const real* cbl;             // coastal bulk lambda [n2]
  const real* rot_c;           // cos(f dt/2) per row
  const real* rot_s;           // sin(f dt/2) per row
};

__device__ __forceinline__ int wrap_e(int i, int nx) { return i + 1 == nx ? 0 : i + 1; }
__device__ __forceinline__ int wrap_w(int i, int nx) { return i == 0 ? nx - 1 : i - 1; }

// Read-only set-up table element through the scalar cache (constant address
// space): uniform indices become scalar loads that the compiler may hoist past
// the kernel's own stores.  Only for tables no kernel writes.
template <typename T>
__device__ __forceinline__ T ldc(const T* p, int i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Map a launch-order block id onto a tile id so that each of the 8 XCDs gets
// one contiguous range of tiles (neighbouring tiles share halos through the
// same L2).  Bijective for any block count.
__device__ __forceinline__ long long xcd_tile(long long L, long long B) {
  const long long nxcd = 8;
  long long x = L % nxcd, q = L / nxcd;
  long long per = B / nxcd, rem = B % nxcd;
  return x < rem ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

// Map a launch-order block id onto a tile id so that groups of G consecutive
// tiles are dealt round-robin over the 8 XCDs: work that varies with depth or
// latitude (tile ranges) spreads evenly over the XCDs while the tiles of a
// group still share their halos through one L2.  Bijective for any block
// count (blocks past the last whole round of groups keep their own id).
__device__ __forceinline__ long long xcd_spread(long long L, long long B, long long G) {
  const long long nxcd = 8;
  const long long full = (B / (nxcd * G)) * nxcd * G;
  if (L >= full) return L;
  const long long x = L % nxcd, i = L / nxcd;
  return ((i / G) * nxcd + x) * G + (i % G);
}

}  // namespace zm
