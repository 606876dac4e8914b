// Shadow-step stage ledger (diagnostics/budgets.py:62-270): float64 volume
// and area integrals of stage differences, of the tracer residual's sources
// and process tendencies, and of the projection's surface transports, kept in
// a raw device accumulator and expanded into the ledger layout of
// include/zhenmode.h (ZM_LEDGER_*) at the end of the step.
#pragma once

#include <hip/hip_runtime.h>

#include "zm_common.h"

namespace zm {

// raw accumulator layout (float64)
constexpr int kRawStages = 0;       // [6][3] heat J, salt kg, displacement m3
constexpr int kRawScale = 18;       // [3] sum of |stage changes|
constexpr int kRawSources = 21;     // [9][3]
constexpr int kRawProcesses = 48;   // [4][3] advection, convection, gm, redi
constexpr int kRawProcScale = 60;   // [3]
constexpr int kRawAdvBoundary = 63; // [3]
constexpr int kRawProjNorms = 66;   // [2] surface transport norm^2 before / after
constexpr int kRawProjMax = 68;     // worst after/before ratio
constexpr int kRawObserved = 69;    // [3]
constexpr int kRawSurfDisp = 72;    // [3]
constexpr int kRawSize = 75;

constexpr int kLedgerBlocks = 1024;
constexpr int kLedgerDiffSums = 10;

enum LedgerStage { kStageLinearDiffusion = 0, kStageSponge, kStageNonlinear, kStageFreeSurface, kStagePolarCap,
                   kStageDynamicIce };

void launch_ledger_reset(double* raw, hipStream_t s);

// Difference of (T, S, eta, ice) between two states (after minus before),
// integrated with the wet node volume / wet cell area.  slot >= 0 adds it to
// that stage and to the change scale; slot < 0 stores it as the observed
// change together with the surface-displacement tracer change.
// part: kLedgerBlocks * kLedgerDiffSums doubles of scratch.
template <typename real>
void launch_ledger_diff(const Geo<real>& g, const double* area_row, const double* dzn, const real* Ta,
                        const real* Sa, const real* eta_a, const real* ice_a, const real* Tb, const real* Sb,
                        const real* eta_b, const real* ice_b, int slot, double* part, double* raw, hipStream_t s);

// Fold the tracer ledger partials of one residual evaluation (nblocks blocks)
// into the raw accumulator with the interval `interval` (seconds).
void launch_ledger_tracer_fold(const double* part, int nblocks, double interval, double* raw, hipStream_t s);

// Surface transport norm^2 (sum Fz_top^2 * wet area) of two Fz fields, added
// to the projection norms; the worst ratio sqrt(after / before) is kept.
// part: kLedgerBlocks * 2 doubles of scratch.
template <typename real>
void launch_ledger_transport(const Geo<real>& g, const double* area_row, const real* Fz_before,
                             const real* Fz_after, double* part, double* raw, hipStream_t s);

// Expand the raw accumulator into the ledger layout (ZM_LEDGER_SIZE doubles).
void launch_ledger_finish(const double* raw, double* ledger, hipStream_t s);

}  // namespace zm
