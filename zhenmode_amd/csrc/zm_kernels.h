// Host-side launchers for the mode-split kernels (explicitly instantiated for
// float and double in zm_kernels.hip).
#pragma once

#include "zm_common.h"

namespace zm {

// Radix sequence of a mixed-radix FFT of length n (radices 2, 3, 4, 5).
struct FftPlan {
  int nst;
  int rad[16];
};

// Set-up data of the nu_h sub-cycle kernels (Geo::nu_wet, Geo::nu_met): the
// wet word of (level k, chunk ch, column i) holds bit b = node (i, ch * chunk -
// kNuWetBase + b, k) is wet; row j's coefficients sit at kNuMetStride (j +
// kNuMetPad): kE, kN, kS, kY for nu_h dts (the substep's update), then
// 1/dx^2 and tan(lat)/R (its Laplacian).
constexpr int kNuWetBase = 4;                     // >= the substeps per launch
constexpr int kNuChunkMax = 64 - 2 * kNuWetBase;  // a chunk and its margins fit one 64-bit word
constexpr int kNuMetPad = 8;                      // > 2 * substeps - 1
constexpr int kNuMetStride = 8;

template <typename real>
void launch_nu_substep(const Geo<real>& g, const real* u, const real* v, real* uo, real* vo, real* lu,
                       real* lv, real nu_h, real dts, hipStream_t s);
template <typename real>
int nu_block_max();
// nsub substeps in one launch; lu, lv (optional) receive lap(u), lap(v) of the
// input, which needs nu_block_writes_lap(g, nsub).
template <typename real>
bool nu_block_writes_lap(const Geo<real>& g, int nsub);
template <typename real>
void launch_nu_block(const Geo<real>& g, const real* u, const real* v, real* uo, real* vo, real nu_h, real dts,
                     int nsub, hipStream_t s, real* lu = nullptr, real* lv = nullptr);
template <typename real>
void launch_d1_unit(const Geo<real>& g, const real* T, const real* S, real* dT, real* dS, hipStream_t s);
template <typename real>
void launch_linear_finish(const Geo<real>& g, const Phys<real>& P, const real* u0, const real* v0,
                          const real* T0, const real* S0, const real* unu, const real* vnu, const real* Lu0,
                          const real* Lv0, const real* DT0, const real* DS0, real* uo, real* vo, real* To,
                          real* So, hipStream_t s);
template <typename real>
void launch_fz(const Geo<real>& g, const real* u, const real* v, real* Fz, hipStream_t s);
template <typename real>
// Convective gate (gate[c] = 1 where a wet interface is unstable and
// kappa_conv > 0; ubits[2c..2c+1] = the unstable interfaces when localize_conv)
// and the conv sub-cycle of the gated columns into cT, cS.
void launch_conv(const Geo<real>& g, const Phys<real>& P, const real* T, const real* S, real* cT, real* cS,
                 int* gate, unsigned long long* ubits, hipStream_t s);
// _tracer_terms (tendencies.py:74-92): the temperature's advection,
// horizontal diffusion, vertical diffusion and convection tendencies on wet
// nodes (zero elsewhere); cT, cS, gate from launch_conv with conv_nsub = 1.
template <typename real>
void launch_tracer_terms(const Geo<real>& g, const Phys<real>& P, const real* u, const real* v, const real* T,
                         const real* S, const real* Fz, const real* cT, const real* cS, const int* gate, real* adv,
                         real* diff_h, real* diff_v, real* conv, hipStream_t s);
template <typename real>
void launch_tracer_residual(const Geo<real>& g, const Phys<real>& P, int mode, const real* u, const real* v,
                            const real* T, const real* S, const real* Fz, const real* cT, const real* cS,
                            const int* gate, real* dT1, real* dS1, const real* Tb, const real* Sb, real* To,
                            real* So, hipStream_t s);

// Ledger sums of one tracer residual evaluation (no state is written): per
// block, kLedgerTracerSums float64 values = volume integrals of the six
// surface sources (prescribed heat, bulk heat, coastal bulk heat, coastal
// temperature restoring, salinity restoring, brine), of the advection
// tendency of T and S, their absolute values, of the convection tendency of T
// and S and their absolute values, then the area integrals of Fz_top*T_top
// and Fz_top*S_top.  lpart holds tracer_ledger_blocks(g) * kLedgerTracerSums.
constexpr int kLedgerTracerSums = 16;
template <typename real>
unsigned tracer_ledger_blocks(const Geo<real>& g);
template <typename real>
void launch_tracer_ledger(const Geo<real>& g, const Phys<real>& P, const real* u, const real* v, const real* T,
                          const real* S, const real* Fz, const real* cT, const real* cS, const int* gate,
                          const double* area_row, const double* dzn64, double* lpart, hipStream_t s);
template <typename real>
void launch_pbc(const Geo<real>& g, const Phys<real>& P, const real* T, const real* S, real* pbc,
                hipStream_t s);
template <typename real>
void launch_bt_rho(const Geo<real>& g, const Phys<real>& P, const real* pbc, real* Fx, real* Fy, const real* u,
                   const real* v, real* ubt0, real* vbt0, hipStream_t s);
template <typename real>
void launch_mom_adv(const Geo<real>& g, const real* u, const real* v, const real* Fz, real* au, real* av,
                    hipStream_t s);
template <typename real>
void launch_dealias_x(int N, int N1, int N2, const int* kept, int nrows, const real* twr, const real* twi,
                      const real* ina, const real* inb, real* outa, real* outb, hipStream_t s);
template <typename real>
void launch_dealias_fft(int N, const FftPlan& plan, const int* kept, int nrows, const real* twr, const real* twi,
                        const real* ina, const real* inb, real* outa, real* outb, hipStream_t s);
template <typename real>
void launch_mom_residual(const Geo<real>& g, const Phys<real>& P, int mode, const real* ax, const real* ay,
                         const real* u, const real* v, const real* pbc, const real* eta, const real* Frx,
                         const real* Fry, real* du1, real* dv1, real* uo, real* vo, real decay_bot,
                         hipStream_t s);
template <typename real>
void launch_bt_eta(const Geo<real>& g, const Phys<real>& P, real* eta, const real* ubt, const real* vbt,
                   real dtH, hipStream_t s);
template <typename real>
void launch_bt_vel(const Geo<real>& g, const Phys<real>& P, const real* eta, real* ubt, real* vbt,
                   const real* Frx, const real* Fry, const real* fdt, const real* fiden, real dt_bt, real drag,
                   hipStream_t s);
// A member of a batched barotropic substep: one grid's constants and state.
// row0 = first workgroup of this member in the batched launch.
template <typename real>
struct BtMember {
  Geo<real> g;
  Phys<real> P;
  real* eta;
  real* ubt;
  real* vbt;
  const real* Frx;
  const real* Fry;
  const real* fdt;
  const real* fiden;
  real dtH, dt_bt, drag;
  int row0;
};

// One whole barotropic substep in one launch, from (eta_in, u_in, v_in) into
// separate output buffers (same arithmetic as launch_bt_eta + launch_bt_vel).
// All n_sub substeps in one persistent launch (the arithmetic of n_sub
// launch_bt_sub calls, alternating sets so that the last lands in (ea, ua, va);
// the first reads eta0, u0, v0).  progress: int[(ny + 1) / 2] device scratch,
// cleared by the launcher; timeouts: int device counter, incremented when a
// bounded wait gives up (results are then invalid).
template <typename real>
void launch_bt_persist(const Geo<real>& g, const Phys<real>& P, const real* eta0, const real* u0, const real* v0,
                       real* ea, real* ua, real* va, real* eb, real* ub, real* vb, const real* Frx, const real* Fry,
                       const real* fdt, const real* fiden, real dtH, real dt_bt, real drag, int nsub, int* progress,
                       int* timeouts, hipStream_t s);

template <typename real>
void launch_bt_sub(const Geo<real>& g, const Phys<real>& P, const real* eta_in, const real* u_in,
                   const real* v_in, real* eta_out, real* u_out, real* v_out, const real* Frx, const real* Fry,
                   const real* fdt, const real* fiden, real dtH, real dt_bt, real drag, hipStream_t s);

// One barotropic substep (eta then velocity kernel) for every member of a
// device-resident table.
template <typename real>
void launch_bt_multi(const BtMember<real>* table, int nmem, int total_rows, hipStream_t s);

template <typename real>
void launch_cap_means(const Geo<real>& g, const Phys<real>& P, const real* u, const real* v, const real* T,
                      const real* S, const real* ubt, const real* ubt0, const real* vbt, const real* vbt0,
                      real* means, hipStream_t s);
template <typename real>
void launch_final(const Geo<real>& g, const Phys<real>& P, const real* u3, const real* v3, const real* T3,
                  const real* S3, const real* ubt, const real* ubt0, const real* vbt, const real* vbt0,
                  const real* means, const real* uin, const real* vin, const real* Tin, const real* Sin,
                  real* uo, real* vo, real* To, real* So, const real* eta_src, real* eta_dst,
                  const real* ice_src, real* ice_dst, hipStream_t s);
template <typename real>
void launch_diag(const Geo<real>& g, const Phys<real>& P, const real* T, const real* S, const real* eta,
                 const real* pbc, const real* Fz, real* rho, real* pres, real* w, hipStream_t s);
template <typename real>
void launch_copy(const real* a, real* b, long long n, hipStream_t s);

}  // namespace zm
