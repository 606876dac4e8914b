// Host-side launchers of the column kernels: the implicit vertical-mixing
// Thomas solve, the per-column MLP surrogate (features, BF16 MFMA GEMM with
// fused epilogues, apply with integral correction and residual check) and its
// training kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "zm_common.h"

namespace zm {

// Column-mixing coefficients shared by the Thomas solve and the surrogate's
// residual check: kappa_{k+1/2} = kappa_v + kappa_conv [rho'_k > rho'_{k+1}].
struct MixParams {
  double dt, kappa_v, kappa_conv;
  double T_ref, S_ref;
  const double* dzn;   // dz_node [nz], float64
  const double* dzi;   // dz_iface [nz-1], float64
  const double* idzn;  // 1 / dz_node [nz]
  const double* idzi;  // 1 / dz_iface [nz-1]
};

// work: 3 nz n2 float64 values
template <typename real>
void launch_vmix(const Geo<real>& g, const MixParams& m, const real* T, const real* S, real* To, real* So,
                 const int* only, double* work, hipStream_t s);

// The same solve through the library's static device scratch (no allocation):
// kStaticScratchDoubles float64 per GPU, the grid solved in column ranges of
// static_vmix_columns(nz) columns; throws std::runtime_error when nz is too
// large for one 256-column range or the scratch's address cannot be found.
constexpr long long kStaticScratchDoubles = 4ll << 20;  // 32 MiB
long long static_vmix_columns(int nz);
template <typename real>
void launch_vmix_static(const Geo<real>& g, const MixParams& m, const real* T, const real* S, real* To, real* So,
                        const int* only, hipStream_t s);

// Feature rows X[c][0..255] (bf16, row c = j*nx + i) of every column:
// T, S, u, v, N^2 at nz levels, q_heat, SSS-restoring salt flux, |tau|, zero pad;
// normalised as (x - mean) * inv_std.  ld = row stride in elements (>= 256).
template <typename real>
void launch_col_features(const Geo<real>& g, const MixParams& m, const real* T, const real* S, const real* u,
                         const real* v, const real* q_heat, const real* s_ref2, real restore_coef_S,
                         const real* tau_x, const real* tau_y, const float* mean, const float* inv_std,
                         void* X, long long rows, hipStream_t s);

// X_new = X + e + s (y * out_std + out_mean) per tracer (e: the explicit
// mixing increment over dt, s = max_k |e_k|) with the column-integral
// correction; columns whose implicit-system residual exceeds tol get flag = 1.
template <typename real>
void launch_mlp_apply(const Geo<real>& g, const MixParams& m, const real* T, const real* S, const float* Y,
                      int ldy, const float* out_mean, const float* out_std, double tol_T, double tol_S, real* To,
                      real* So, int* flag, int* count, hipStream_t s);

// Training targets of the surrogate: per column and tracer, the Thomas
// solution (Tt, St) minus the state and minus the explicit mixing increment,
// in units of the column's mixing scale s = max_k |e_k| (0 where s = 0 and on
// dry levels); float32 rows of ldy (nz T values, nz S values, zero padding).
template <typename real>
void launch_mix_targets(const Geo<real>& g, const MixParams& m, const real* T, const real* S, const real* Tt,
                        const real* St, float* Y, int ldy, hipStream_t s);

// The surrogate step on a float32 state in one launch: features, the four
// layers (weights[l]: bf16 [N][256], N = 256, 256, 256, 128; biases[l] fp32
// [N]) and the apply step of launch_mlp_apply; results equal the unfused
// sequence bit for bit.  nz <= 50.
void launch_mlp_fused(const Geo<float>& g, const MixParams& m, const float* T, const float* S, const float* u,
                      const float* v, const float* q_heat, const float* s_ref2, float restore_coef_S,
                      const float* tau_x, const float* tau_y, const float* in_mean, const float* in_inv_std,
                      const void* const* weights, const float* const* biases, const float* out_mean,
                      const float* out_std, double tol_T, double tol_S, float* To, float* So, int* flag, int* count,
                      hipStream_t s);

// C[m][n] = sum_k A[m][k] B[n][k], A (M x K) and B (N x K) bf16 row-major,
// fp32 accumulation on MFMA; M, N multiples of 128, K a multiple of 32.
enum GemmEpilogue {
  kEpiF32 = 0,        // C_f32 = acc
  kEpiBiasF32 = 1,    // C_f32 = acc + bias[n]
  kEpiBiasGelu = 2,   // Z = acc + bias -> C_f32 (if set); C_bf16 = gelu(Z), CT_bf16 = its transpose
  kEpiGeluGrad = 3,   // G = acc * gelu'(aux[m][n]) -> C_bf16, CT_bf16
};
void launch_gemm_bf16_nt(long long M, long long N, long long K, const void* A, const void* B, int epi,
                         const float* bias, const float* aux, float* C_f32, void* C_bf16, void* CT_bf16,
                         hipStream_t s);

// dY = 2 (Y - target) / (rows * n_out) on n < n_out (0 on pads), as bf16 in
// both layouts; loss_sum += sum (Y - target)^2 over the same entries.
void launch_mse_grad(long long rows, int ld, int n_out, const float* Y, const float* target, void* dY,
                     void* dYT, float* loss_sum, hipStream_t s);

// out[r] = sum_c in[r][c] (bf16 rows of length cols), fp32.
void launch_rowsum_bf16(long long rows, long long cols, const void* in, float* out, hipStream_t s);

// out[c][r] = in[r][c] for a rows x cols bf16 matrix.
void launch_transpose_bf16(long long rows, long long cols, const void* in, void* out, hipStream_t s);

// Adam on fp32 master weights w (n_out x n_in, row-major "out x in"): writes the
// bf16 copies w_bf (out x in) and wT_bf (in x out); bias arrays are updated
// the same way when given (b, gb, bm, bv of length n_out).
void launch_adam(int n_out, int n_in, float* w, const float* gw, float* m, float* v, void* w_bf, void* wT_bf,
                 float* b, const float* gb, float* bm, float* bv, float lr, float beta1, float beta2, float eps,
                 float c1, float c2, hipStream_t s);

}  // namespace zm
