// Run-time kernels that keep a run loop on the device: the 30-day seasonal
// forcing blend (inputs/forcing/seasonal.py:86-103), the state monitor of
// runtime/monitor.py:20-30 with the first-rejection bookkeeping of
// runtime/run_loop.py:53-96, and the float64 compensated budget sums of
// diagnostics/snapshot.py:67-128.
#include <hip/hip_runtime.h>

#include "zm_runtime.h"

namespace zm {

template <typename real>
__global__ __launch_bounds__(256) void k_seasonal_blend(const real* __restrict__ stack, int nfields, long long n2,
                                                        int mi, real wc, real wa, real wb, real* __restrict__ out) {
  const long long n = (long long)nfields * n2;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const int ni = (mi + 1) % 12, pi = (mi + 11) % 12;
  const long long month = n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    out[q] = wc * stack[mi * month + q] + wa * stack[ni * month + q] + wb * stack[pi * month + q];
}

// Neumaier compensated accumulation
__device__ __forceinline__ void kahan_add(double& s, double& c, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x))
    c += (s - t) + x;
  else
    c += (x - t) + s;
  s = t;
}

__device__ __forceinline__ double block_reduce_sum(double x, double* red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = x;
  __syncthreads();
  double s = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q];
  return s;
}

__device__ __forceinline__ double block_reduce_max(double x, double* red) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_down(x, o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = x;
  __syncthreads();
  double s = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s = fmax(s, red[q]);
  return s;
}

// ---------------------------------------------------------------------------
// monitor: per-block max |u|, max |v|, max |eta| and the count of non-finite
// values over every field of the state, then one block folds the partials
// into the status (classify_state codes 0 accepted, 1 non-finite,
// 3 velocity, 4 surface height).  After the first rejection later states are
// ignored, so the status keeps the first-rejection record of the run loop.
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(256) void k_monitor_partials(long long n3, long long n2, const real* __restrict__ u,
                                                          const real* __restrict__ v, const real* __restrict__ T,
                                                          const real* __restrict__ S, const real* __restrict__ eta,
                                                          const real* __restrict__ ice, double* __restrict__ partials,
                                                          const MonitorStatus* __restrict__ status, long long step) {
  __shared__ double red[4];
  if (status->first_failure_step >= 0 && step > status->first_failure_step) return;  // uniform
  double mu = 0, mv = 0, me = 0, bad = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n3; q += stride) {
    const double a = (double)u[q], b = (double)v[q], t = (double)T[q], s = (double)S[q];
    mu = fmax(mu, fabs(a));
    mv = fmax(mv, fabs(b));
    bad += (!isfinite(a)) + (!isfinite(b)) + (!isfinite(t)) + (!isfinite(s));
  }
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n2; q += stride) {
    const double e = (double)eta[q], c = (double)ice[q];
    me = fmax(me, fabs(e));
    bad += (!isfinite(e)) + (!isfinite(c));
  }
  mu = block_reduce_max(mu, red);
  mv = block_reduce_max(mv, red);
  me = block_reduce_max(me, red);
  bad = block_reduce_sum(bad, red);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 4 + 0] = mu;
    partials[blockIdx.x * 4 + 1] = mv;
    partials[blockIdx.x * 4 + 2] = me;
    partials[blockIdx.x * 4 + 3] = bad;
  }
}

__global__ __launch_bounds__(256) void k_monitor_final(int nblocks, const double* __restrict__ partials,
                                                       MonitorStatus* __restrict__ status, long long step,
                                                       double velocity_limit, double eta_limit, int inclusive) {
  __shared__ double red[4];
  if (status->first_failure_step >= 0 && step > status->first_failure_step) return;  // uniform
  double mu = 0, mv = 0, me = 0, bad = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    mu = fmax(mu, partials[b * 4 + 0]);
    mv = fmax(mv, partials[b * 4 + 1]);
    me = fmax(me, partials[b * 4 + 2]);
    bad += partials[b * 4 + 3];
  }
  mu = block_reduce_max(mu, red);
  mv = block_reduce_max(mv, red);
  me = block_reduce_max(me, red);
  bad = block_reduce_sum(bad, red);
  if (threadIdx.x == 0) {
    const double vel = fmax(mu, mv);
    const bool finite = bad == 0.0;
    const bool bad_v = inclusive ? vel >= velocity_limit : vel > velocity_limit;
    const bool bad_e = inclusive ? me >= eta_limit : me > eta_limit;
    const int failure = !finite ? 1 : (bad_v ? 3 : (bad_e ? 4 : 0));
    status->max_u = mu;
    status->max_velocity = vel;
    status->max_eta = me;
    status->peak_u = fmax(status->peak_u, mu);
    status->peak_velocity = fmax(status->peak_velocity, vel);
    status->peak_eta = fmax(status->peak_eta, me);
    status->finite = finite ? 1 : 0;
    status->failure = failure;
    status->last_step = step;
    if (failure != 0) status->first_failure_step = step;
  }
}

__global__ void k_monitor_reset(MonitorStatus* status) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    MonitorStatus z{};
    z.finite = 1;
    z.first_failure_step = -1;
    z.last_step = -1;
    *status = z;
  }
}

// ---------------------------------------------------------------------------
// budget sums (float64, Neumaier-compensated per thread, tree within a block,
// compensated across blocks)
// ---------------------------------------------------------------------------
constexpr int kBudgetSums = 9;

template <typename real>
__global__ __launch_bounds__(256) void k_budget_partials(Geo<real> g, const double* __restrict__ area_row,
                                                         const double* __restrict__ dzn, const real* __restrict__ T,
                                                         const real* __restrict__ S, const real* __restrict__ eta,
                                                         double* __restrict__ partials) {
  __shared__ double red[4];
  double s[kBudgetSums], c[kBudgetSums];
  for (int q = 0; q < kBudgetSums; ++q) s[q] = c[q] = 0.0;
  const long long n2 = g.n2, n3 = g.n3;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n3; q += stride) {
    const int k = (int)(q / n2);
    const long long c2 = q - (long long)k * n2;
    if (k < g.nwet[c2]) {
      const int j = (int)(c2 / g.nx);
      const double vol = area_row[j] * dzn[k];
      kahan_add(s[0], c[0], vol);
      kahan_add(s[1], c[1], (double)T[q] * vol);
      kahan_add(s[2], c[2], (double)S[q] * vol);
    }
  }
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n2; q += stride) {
    const int j = (int)(q / g.nx);
    const double a = area_row[j];
    kahan_add(s[6], c[6], a);
    if (g.nwet[q] > 0) {
      kahan_add(s[3], c[3], a);
      kahan_add(s[4], c[4], (double)T[q] * a);
      kahan_add(s[5], c[5], (double)S[q] * a);
      kahan_add(s[7], c[7], (double)eta[q] * a);
    }
  }
  for (int q = 0; q < kBudgetSums; ++q) {
    const double tot = block_reduce_sum(s[q] + c[q], red);
    if (threadIdx.x == 0) partials[blockIdx.x * kBudgetSums + q] = tot;
  }
}

__global__ __launch_bounds__(64) void k_budget_final(int nblocks, const double* __restrict__ partials,
                                                     double* __restrict__ out) {
  const int q = threadIdx.x;
  if (q >= kBudgetSums) return;
  double s = 0.0, c = 0.0;
  for (int b = 0; b < nblocks; ++b) kahan_add(s, c, partials[b * kBudgetSums + q]);
  out[q] = s + c;
}

// ---------------------------------------------------------------------------
template <typename real>
void launch_seasonal_blend(const real* stack, int nfields, long long n2, int mi, double w_after, double w_before,
                           real* out, hipStream_t s) {
  const long long n = (long long)nfields * n2;
  const unsigned blocks = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_seasonal_blend<real>, dim3(blocks), dim3(256), 0, s, stack, nfields, n2, mi,
                     (real)(1.0 - w_after - w_before), (real)w_after, (real)w_before, out);
}

template <typename real>
void launch_monitor(const Geo<real>& g, const real* u, const real* v, const real* T, const real* S, const real* eta,
                    const real* ice, long long step, double velocity_limit, double eta_limit, int inclusive,
                    double* partials, MonitorStatus* status, hipStream_t s) {
  hipLaunchKernelGGL(k_monitor_partials<real>, dim3(kReduceBlocks), dim3(256), 0, s, g.n3, (long long)g.n2, u, v, T,
                     S, eta, ice, partials, status, step);
  hipLaunchKernelGGL(k_monitor_final, dim3(1), dim3(256), 0, s, kReduceBlocks, partials, status, step,
                     velocity_limit, eta_limit, inclusive);
}

void launch_monitor_reset(MonitorStatus* status, hipStream_t s) {
  hipLaunchKernelGGL(k_monitor_reset, dim3(1), dim3(64), 0, s, status);
}

template <typename real>
void launch_budget(const Geo<real>& g, const double* area_row, const double* dzn, const real* T, const real* S,
                   const real* eta, double* partials, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_budget_partials<real>, dim3(kReduceBlocks), dim3(256), 0, s, g, area_row, dzn, T, S, eta,
                     partials);
  hipLaunchKernelGGL(k_budget_final, dim3(1), dim3(64), 0, s, kReduceBlocks, partials, out);
}

#define ZM_RT_INSTANTIATE(real)                                                                                  \
  template void launch_seasonal_blend<real>(const real*, int, long long, int, double, double, real*, hipStream_t); \
  template void launch_monitor<real>(const Geo<real>&, const real*, const real*, const real*, const real*,        \
                                     const real*, const real*, long long, double, double, int, double*,           \
                                     MonitorStatus*, hipStream_t);                                                \
  template void launch_budget<real>(const Geo<real>&, const double*, const double*, const real*, const real*,     \
                                    const real*, double*, double*, hipStream_t);

ZM_RT_INSTANTIATE(float)
ZM_RT_INSTANTIATE(double)

}  // namespace zm
