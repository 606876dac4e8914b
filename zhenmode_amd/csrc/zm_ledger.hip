// Shadow-step stage ledger on the device (diagnostics/budgets.py:62-270).
//
// Every integral is a float64 sum: per thread over a grid-stride range, then
// per block in a fixed tree, then over the kLedgerBlocks block partials in a
// fixed order by one block that updates the raw accumulator, so a ledger does
// not depend on block scheduling.
#include <hip/hip_runtime.h>

#include "../../include/zhenmode.h"
#include "zm_kernels.h"
#include "zm_ledger.h"

namespace zm {

namespace {

constexpr double kRhoCp = kRho0 * kCp;          // J m^-3 K^-1
constexpr double kRhoSalt = kRho0 / 1000.0;     // kg of salt per m^3 per g/kg
constexpr double kIceLatent = 917.0 * 3.34e5;   // rho_ice * latent heat of fusion

template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double (*red)[NV]) {
  for (int q = 0; q < NV; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < NV; ++q) red[threadIdx.x >> 6][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; ++q) v[q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
}

// sum of part[b][q] over b in a fixed order (valid in thread 0)
template <int NV>
__device__ __forceinline__ void fold(const double* __restrict__ part, int nblocks, double (&t)[NV],
                                     double (*red)[NV]) {
  for (int q = 0; q < NV; ++q) t[q] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
    for (int q = 0; q < NV; ++q) t[q] += part[(long long)b * NV + q];
  block_sums<NV>(t, red);
}

}  // namespace

__global__ void k_ledger_reset(double* raw) {
  for (int q = threadIdx.x; q < kRawSize; q += blockDim.x) raw[q] = 0.0;
}

// acc: 0 heat, 1 |heat|, 2 salt, 3 |salt|, 4 latent, 5 |latent|, 6 displacement,
// 7 |displacement|, 8 eta*T_top change, 9 eta*S_top change
template <typename real>
__global__ __launch_bounds__(256) void k_ledger_diff(Geo<real> g, const double* __restrict__ area_row,
                                                     const double* __restrict__ dzn, const real* __restrict__ Ta,
                                                     const real* __restrict__ Sa, const real* __restrict__ eta_a,
                                                     const real* __restrict__ ice_a, const real* __restrict__ Tb,
                                                     const real* __restrict__ Sb, const real* __restrict__ eta_b,
                                                     const real* __restrict__ ice_b, double* __restrict__ part) {
  __shared__ double red[4][kLedgerDiffSums];
  double acc[kLedgerDiffSums];
  for (int q = 0; q < kLedgerDiffSums; ++q) acc[q] = 0.0;
  const long long n2 = g.n2, n3 = g.n3;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n3; q += stride) {
    const int k = (int)(q / n2);
    const long long c2 = q - (long long)k * n2;
    if (k < g.nwet[c2]) {
      const double vol = area_row[c2 / g.nx] * dzn[k];
      const double h = kRhoCp * ((double)Ta[q] - (double)Tb[q]) * vol;
      const double sl = kRhoSalt * ((double)Sa[q] - (double)Sb[q]) * vol;
      acc[0] += h;
      acc[1] += fabs(h);
      acc[2] += sl;
      acc[3] += fabs(sl);
    }
  }
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n2; q += stride) {
    if (g.nwet[q] > 0) {
      const double sa = area_row[q / g.nx];
      const double lat = kIceLatent * ((double)ice_a[q] - (double)ice_b[q]) * sa;
      const double d = ((double)eta_a[q] - (double)eta_b[q]) * sa;
      acc[4] += lat;
      acc[5] += fabs(lat);
      acc[6] += d;
      acc[7] += fabs(d);
      acc[8] += ((double)eta_a[q] * (double)Ta[q] - (double)eta_b[q] * (double)Tb[q]) * sa;
      acc[9] += ((double)eta_a[q] * (double)Sa[q] - (double)eta_b[q] * (double)Sb[q]) * sa;
    }
  }
  block_sums<kLedgerDiffSums>(acc, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < kLedgerDiffSums; ++q) part[(long long)blockIdx.x * kLedgerDiffSums + q] = acc[q];
}

__global__ __launch_bounds__(256) void k_ledger_diff_fold(const double* __restrict__ part, int nblocks, int slot,
                                                          double* __restrict__ raw) {
  __shared__ double red[4][kLedgerDiffSums];
  double t[kLedgerDiffSums];
  fold<kLedgerDiffSums>(part, nblocks, t, red);
  if (threadIdx.x != 0) return;
  if (slot >= 0) {
    raw[kRawStages + 3 * slot + 0] += t[0] - t[4];
    raw[kRawStages + 3 * slot + 1] += t[2];
    raw[kRawStages + 3 * slot + 2] += t[6];
    raw[kRawScale + 0] += t[1] + t[5];
    raw[kRawScale + 1] += t[3];
    raw[kRawScale + 2] += t[7];
  } else {
    raw[kRawObserved + 0] = t[0] - t[4];
    raw[kRawObserved + 1] = t[2];
    raw[kRawObserved + 2] = t[6];
    raw[kRawSurfDisp + 0] = kRhoCp * t[8];
    raw[kRawSurfDisp + 1] = kRhoSalt * t[9];
    raw[kRawSurfDisp + 2] = 0.0;
  }
}

// tracer sums: 0..5 surface sources, 6/7 advection T/S, 8/9 |advection|,
// 10/11 convection, 12/13 |convection|, 14/15 surface advective flux T/S
__global__ __launch_bounds__(256) void k_ledger_tracer_fold(const double* __restrict__ part, int nblocks,
                                                            double interval, double* __restrict__ raw) {
  __shared__ double red[4][kLedgerTracerSums];
  double t[kLedgerTracerSums];
  fold<kLedgerTracerSums>(part, nblocks, t, red);
  if (threadIdx.x != 0) return;
  for (int q = 0; q < 4; ++q) raw[kRawSources + 3 * q + 0] += t[q] * kRhoCp * interval;
  raw[kRawSources + 3 * 4 + 1] += t[4] * kRhoSalt * interval;
  raw[kRawSources + 3 * 5 + 1] += t[5] * kRhoSalt * interval;
  raw[kRawProcesses + 0] += kRhoCp * interval * t[6];
  raw[kRawProcesses + 1] += kRhoSalt * interval * t[7];
  raw[kRawProcesses + 3] += kRhoCp * interval * t[10];
  raw[kRawProcesses + 4] += kRhoSalt * interval * t[11];
  raw[kRawProcScale + 0] += kRhoCp * interval * (t[8] + t[12]);
  raw[kRawProcScale + 1] += kRhoSalt * interval * (t[9] + t[13]);
  raw[kRawAdvBoundary + 0] += kRhoCp * interval * t[14];
  raw[kRawAdvBoundary + 1] += kRhoSalt * interval * t[15];
}

template <typename real>
__global__ __launch_bounds__(256) void k_ledger_transport(Geo<real> g, const double* __restrict__ area_row,
                                                          const real* __restrict__ Fa, const real* __restrict__ Fb,
                                                          double* __restrict__ part) {
  __shared__ double red[4][2];
  double acc[2] = {0.0, 0.0};
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.n2; q += stride) {
    if (g.nwet[q] > 0) {
      const double sa = area_row[q / g.nx];
      const double a = (double)Fa[q], b = (double)Fb[q];
      acc[0] += a * a * sa;
      acc[1] += b * b * sa;
    }
  }
  block_sums<2>(acc, red);
  if (threadIdx.x == 0) {
    part[(long long)blockIdx.x * 2 + 0] = acc[0];
    part[(long long)blockIdx.x * 2 + 1] = acc[1];
  }
}

__global__ __launch_bounds__(256) void k_ledger_transport_fold(const double* __restrict__ part, int nblocks,
                                                               double* __restrict__ raw) {
  __shared__ double red[4][2];
  double t[2];
  fold<2>(part, nblocks, t, red);
  if (threadIdx.x != 0) return;
  raw[kRawProjNorms + 0] += t[0];
  raw[kRawProjNorms + 1] += t[1];
  const double rel = t[0] > 0.0 ? sqrt(t[1] / t[0]) : (t[1] == 0.0 ? 0.0 : __builtin_huge_val());
  raw[kRawProjMax] = fmax(raw[kRawProjMax], rel);
}

// budgets.py:187-225 (result)
__global__ void k_ledger_finish(const double* __restrict__ raw, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int q = 0; q < ZM_LEDGER_SIZE; ++q) out[q] = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double observed = raw[kRawObserved + c];
    double stages = 0.0, sources = 0.0, sources6 = 0.0, processes = 0.0;
    for (int s = 0; s < 6; ++s) stages += raw[kRawStages + 3 * s + c];
    for (int s = 0; s < 9; ++s) sources += raw[kRawSources + 3 * s + c];
    for (int s = 0; s < 6; ++s) sources6 += raw[kRawSources + 3 * s + c];
    for (int s = 0; s < 4; ++s) processes += raw[kRawProcesses + 3 * s + c];
    const double decomposition = observed - stages;
    const double budget = observed - sources;
    const double boundary = raw[kRawProcesses + c] - raw[kRawAdvBoundary + c];
    const double nonlinear = raw[kRawStages + 3 * kStageNonlinear + c] - sources6 - processes;
    out[ZM_LEDGER_OBSERVED + c] = observed;
    out[ZM_LEDGER_DECOMPOSITION + c] = decomposition;
    out[ZM_LEDGER_BUDGET + c] = budget;
    out[ZM_LEDGER_ABS_DECOMPOSITION + c] = fabs(decomposition);
    out[ZM_LEDGER_ABS_BUDGET + c] = fabs(budget);
    out[ZM_LEDGER_CHANGE_SCALE + c] = raw[kRawScale + c];
    out[ZM_LEDGER_PROCESS_SCALE + c] = raw[kRawProcScale + c];
    out[ZM_LEDGER_ADVECTION_BOUNDARY + c] = raw[kRawAdvBoundary + c];
    out[ZM_LEDGER_ADVECTION_BOUNDARY_RESIDUAL + c] = boundary;
    out[ZM_LEDGER_ABS_ADVECTION_BOUNDARY_RESIDUAL + c] = fabs(boundary);
    out[ZM_LEDGER_NONLINEAR_RESIDUAL + c] = nonlinear;
    out[ZM_LEDGER_ABS_NONLINEAR_RESIDUAL + c] = fabs(nonlinear);
    out[ZM_LEDGER_SURFACE_DISPLACEMENT + c] = raw[kRawSurfDisp + c];
  }
  for (int q = 0; q < 18; ++q) out[ZM_LEDGER_STAGES + q] = raw[kRawStages + q];
  for (int q = 0; q < 27; ++q) out[ZM_LEDGER_SOURCES + q] = raw[kRawSources + q];
  for (int q = 0; q < 12; ++q) out[ZM_LEDGER_PROCESSES + q] = raw[kRawProcesses + q];
  out[ZM_LEDGER_PROJECTION_NORMS + 0] = raw[kRawProjNorms + 0];
  out[ZM_LEDGER_PROJECTION_NORMS + 1] = raw[kRawProjNorms + 1];
  out[ZM_LEDGER_PROJECTION_MAX] = raw[kRawProjMax];
}

// ---------------------------------------------------------------------------
void launch_ledger_reset(double* raw, hipStream_t s) {
  hipLaunchKernelGGL(k_ledger_reset, dim3(1), dim3(128), 0, s, raw);
}

template <typename real>
void launch_ledger_diff(const Geo<real>& g, const double* area_row, const double* dzn, const real* Ta,
                        const real* Sa, const real* eta_a, const real* ice_a, const real* Tb, const real* Sb,
                        const real* eta_b, const real* ice_b, int slot, double* part, double* raw, hipStream_t s) {
  hipLaunchKernelGGL(k_ledger_diff<real>, dim3(kLedgerBlocks), dim3(256), 0, s, g, area_row, dzn, Ta, Sa, eta_a,
                     ice_a, Tb, Sb, eta_b, ice_b, part);
  hipLaunchKernelGGL(k_ledger_diff_fold, dim3(1), dim3(256), 0, s, (const double*)part, kLedgerBlocks, slot, raw);
}

void launch_ledger_tracer_fold(const double* part, int nblocks, double interval, double* raw, hipStream_t s) {
  hipLaunchKernelGGL(k_ledger_tracer_fold, dim3(1), dim3(256), 0, s, part, nblocks, interval, raw);
}

template <typename real>
void launch_ledger_transport(const Geo<real>& g, const double* area_row, const real* Fz_before,
                             const real* Fz_after, double* part, double* raw, hipStream_t s) {
  hipLaunchKernelGGL(k_ledger_transport<real>, dim3(kLedgerBlocks), dim3(256), 0, s, g, area_row, Fz_before,
                     Fz_after, part);
  hipLaunchKernelGGL(k_ledger_transport_fold, dim3(1), dim3(256), 0, s, (const double*)part, kLedgerBlocks, raw);
}

void launch_ledger_finish(const double* raw, double* ledger, hipStream_t s) {
  hipLaunchKernelGGL(k_ledger_finish, dim3(1), dim3(64), 0, s, raw, ledger);
}

#define ZM_LEDGER_INSTANTIATE(real)                                                                             \
  template void launch_ledger_diff<real>(const Geo<real>&, const double*, const double*, const real*,          \
                                         const real*, const real*, const real*, const real*, const real*,      \
                                         const real*, const real*, int, double*, double*, hipStream_t);        \
  template void launch_ledger_transport<real>(const Geo<real>&, const double*, const real*, const real*,       \
                                              double*, double*, hipStream_t);

ZM_LEDGER_INSTANTIATE(float)
ZM_LEDGER_INSTANTIATE(double)

}  // namespace zm
