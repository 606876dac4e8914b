// Launchers of the run-time kernels that keep the run loop on the device:
// seasonal forcing blending, the state monitor with first-rejection
// bookkeeping, and float64 compensated budget sums.
#pragma once

#include <hip/hip_runtime.h>

#include "zm_common.h"

namespace zm {

constexpr int kReduceBlocks = 1024;  // grid-stride partial sums per launch

// out[f][c] = (1 - wa - wb) stack[mi][f][c] + wa stack[(mi+1)%12][f][c] + wb stack[(mi+11)%12][f][c]
template <typename real>
void launch_seasonal_blend(const real* stack, int nfields, long long n2, int mi, double w_after, double w_before,
                           real* out, hipStream_t s);

// Device status of a monitored run (layout mirrored by zm_monitor_status in
// include/zhenmode.h).
struct MonitorStatus {
  double max_u, max_velocity, max_eta;  // metrics of the last classified state
  double peak_u, peak_velocity, peak_eta;
  int finite, failure;
  long long first_failure_step;  // -1 while every state was accepted
  long long last_step;           // index of the last classified state
};

// partials: 4 * kReduceBlocks doubles of scratch
template <typename real>
void launch_monitor(const Geo<real>& g, const real* u, const real* v, const real* T, const real* S,
                    const real* eta, const real* ice, long long step, double velocity_limit, double eta_limit,
                    int inclusive, double* partials, MonitorStatus* status, hipStream_t s);

void launch_monitor_reset(MonitorStatus* status, hipStream_t s);

// Budget sums (float64, compensated): out[0..8] = total volume, sum T vol,
// sum S vol, wet surface area, sum T_top area, sum S_top area, total area,
// sum eta area (wet), unused.  partials: 9 * kReduceBlocks doubles.
template <typename real>
void launch_budget(const Geo<real>& g, const double* area_row, const double* dzn, const real* T, const real* S,
                   const real* eta, double* partials, double* out, hipStream_t s);

}  // namespace zm
