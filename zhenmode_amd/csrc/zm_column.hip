// Column kernels for gfx950: the implicit vertical-mixing Thomas solve and
// the per-column MLP surrogate that stands in for it (features, BF16 MFMA
// GEMM with fused epilogues, apply with the exact column-integral correction
// and the residual check that selects the Thomas fallback), plus the
// surrogate's training kernels (loss gradient, bias reduction, transpose,
// Adam).
//
// The reference has no implicit vertical solve and no learned component; the
// mixing operator is the backward-Euler counterpart of its explicit flux-form
// vertical operator (_diffusive_interface_flux / _iface_flux_divergence,
// src/zhenmode/model/solver/numerics/vertical.py:296-332) with the linear EOS
// of physics/eos.py:13-18.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "zm_column.h"

namespace zm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double rho_anom(double T, double S, const MixParams& m) {
  return kRho0 * (-kAlpha * (T - m.T_ref) + kBeta * (S - m.S_ref));
}

__device__ __forceinline__ double mix_kappa(double Tc, double Sc, double Tn, double Sn, const MixParams& m) {
  return m.kappa_v + (rho_anom(Tc, Sc, m) > rho_anom(Tn, Sn, m) ? m.kappa_conv : 0.0);
}

// ---------------------------------------------------------------------------
// Thomas solve of (I - dt D_kappa) X_new = X for X in (T, S), one column per
// lane, float64 elimination.  The forward sweep writes c' and d'_T, d'_S to a
// float64 scratch of 3 nz n2 values laid out [k][column] (coalesced across
// lanes), the back substitution reads them in reverse; one reciprocal of the
// pivot per level.  Dry levels are held.  `only` (optional) restricts the
// solve to flagged columns.  Blocks map to columns in launch order (dealt
// round-robin over the XCDs): flagged columns cluster in latitude.
// ---------------------------------------------------------------------------
//
// A launch solves the columns [c_begin, c_begin + c_count); its elimination
// values go to work[3][nz][c_count], so a bounded buffer can serve any grid in
// column ranges (the arithmetic per column does not depend on the range).
template <typename real>
__global__ __launch_bounds__(256) void k_vmix(Geo<real> g, MixParams m, const real* __restrict__ T,
                                              const real* __restrict__ S, real* __restrict__ To,
                                              real* __restrict__ So, const int* __restrict__ only,
                                              double* __restrict__ work, long long c_begin, long long c_count) {
  const long long off = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long c2 = c_begin + off;
  if (off >= c_count || c2 >= g.n2) return;
  if (only && only[c2] == 0) return;
  const int nz = g.nz, nw = g.nwet[c2];
  const long long n2 = g.n2, plane = c_count * nz;
  double* cpl = work + off;
  double* dTl = cpl + plane;
  double* dSl = dTl + plane;
  for (int k = nw; k < nz; ++k) {
    To[k * n2 + c2] = T[k * n2 + c2];
    So[k * n2 + c2] = S[k * n2 + c2];
  }
  if (nw == 0) return;
  double Tc = T[c2], Sc = S[c2];
  double km = 0.0, cprev = 0.0, dT = 0.0, dS = 0.0;
  for (int k = 0; k < nw; ++k) {
    double Tn = 0.0, Sn = 0.0, kp = 0.0;
    if (k + 1 < nw) {
      Tn = T[(k + 1) * n2 + c2];
      Sn = S[(k + 1) * n2 + c2];
      kp = mix_kappa(Tc, Sc, Tn, Sn, m);
    }
    const double a = k > 0 ? -m.dt * km / (m.dzn[k] * m.dzi[k - 1]) : 0.0;
    const double c = k + 1 < nw ? -m.dt * kp / (m.dzn[k] * m.dzi[k]) : 0.0;
    const double b = 1.0 - a - c;
    const double inv = 1.0 / (b - a * cprev);
    cprev = c * inv;
    dT = (Tc - a * dT) * inv;
    dS = (Sc - a * dS) * inv;
    cpl[k * c_count] = cprev;
    dTl[k * c_count] = dT;
    dSl[k * c_count] = dS;
    Tc = Tn;
    Sc = Sn;
    km = kp;
  }
  double xT = dT, xS = dS;
  To[(nw - 1) * n2 + c2] = (real)xT;
  So[(nw - 1) * n2 + c2] = (real)xS;
  for (int k = nw - 2; k >= 0; --k) {
    const double cp = cpl[k * c_count];
    xT = dTl[k * c_count] - cp * xT;
    xS = dSl[k * c_count] - cp * xS;
    To[k * n2 + c2] = (real)xT;
    So[k * n2 + c2] = (real)xS;
  }
}

// ---------------------------------------------------------------------------
// Feature rows of the surrogate: a workgroup builds the rows of 64 columns in
// LDS (4 waves: T; S; u, v; N^2, surface fluxes and padding) and writes them
// as one contiguous block of 64 x 512 bytes.
// ---------------------------------------------------------------------------
constexpr int kFeat = 256;
constexpr int kFeatRow = kFeat + 2;  // bf16 per LDS row: 129 words, conflict-free row writes

template <typename real>
__global__ __launch_bounds__(256) void k_col_features(Geo<real> g, MixParams m, const real* __restrict__ T,
                                                      const real* __restrict__ S, const real* __restrict__ u,
                                                      const real* __restrict__ v, const real* __restrict__ q_heat,
                                                      const real* __restrict__ s_ref2, real restore_coef_S,
                                                      const real* __restrict__ tau_x,
                                                      const real* __restrict__ tau_y, const float* __restrict__ mean,
                                                      const float* __restrict__ inv_std, unsigned* __restrict__ X,
                                                      long long rows) {
  __shared__ unsigned short rowbuf[64 * kFeatRow];
  const int lane = threadIdx.x & 63, part = threadIdx.x >> 6;
  const long long r0 = (long long)blockIdx.x * 64;
  const long long c2 = r0 + lane;
  const int nz = g.nz;
  const long long n2 = g.n2;
  unsigned short* row = rowbuf + lane * kFeatRow;
  auto put = [&](int f, float x) {
    const __bf16 b = (__bf16)((x - mean[f]) * inv_std[f]);
    row[f] = __builtin_bit_cast(unsigned short, b);
  };
  if (c2 < g.n2) {
    const int nw = g.nwet[c2];
    if (part == 0) {
      for (int k = 0; k < nz; ++k) put(k, k < nw ? (float)T[k * n2 + c2] : 0.f);
    } else if (part == 1) {
      for (int k = 0; k < nz; ++k) put(nz + k, k < nw ? (float)S[k * n2 + c2] : 0.f);
    } else if (part == 2) {
      for (int k = 0; k < nz; ++k) put(2 * nz + k, k < nw ? (float)u[k * n2 + c2] : 0.f);
      for (int k = 0; k < nz; ++k) put(3 * nz + k, k < nw ? (float)v[k * n2 + c2] : 0.f);
    } else {
      double rc = nw > 0 ? rho_anom((double)T[c2], (double)S[c2], m) : 0.0;
      for (int k = 0; k < nz; ++k) {
        float n2k = 0.f;
        if (k + 1 < nw) {
          const double rn = rho_anom((double)T[(k + 1) * n2 + c2], (double)S[(k + 1) * n2 + c2], m);
          n2k = (float)(kG / kRho0 * (rn - rc) / m.dzi[k]);
          rc = rn;
        }
        put(4 * nz + k, n2k);
      }
      const bool wet = nw > 0;
      put(5 * nz, wet ? (float)q_heat[c2] : 0.f);
      put(5 * nz + 1, wet ? (float)(restore_coef_S * (s_ref2[c2] - S[c2])) : 0.f);
      put(5 * nz + 2, wet ? sqrtf((float)tau_x[c2] * (float)tau_x[c2] + (float)tau_y[c2] * (float)tau_y[c2]) : 0.f);
      for (int f = 5 * nz + 3; f < kFeat; ++f) put(f, 0.f);
    }
  } else if (part == 0) {
    for (int f = 0; f < kFeat; ++f) row[f] = 0;
  }
  __syncthreads();
  const long long nrows = rows - r0 < 64 ? rows - r0 : 64;
  const unsigned* words = reinterpret_cast<const unsigned*>(rowbuf);
  for (int q = threadIdx.x; q < nrows * (kFeat / 2); q += 256) {
    const int rr = q / (kFeat / 2), w = q - rr * (kFeat / 2);
    X[(r0 + rr) * (kFeat / 2) + w] = words[rr * (kFeatRow / 2) + w];
  }
}

// ---------------------------------------------------------------------------
// Apply the surrogate's increments: dX_k = dt (y_k out_std_k + out_mean_k) on
// wet levels, shifted uniformly so that sum_k dz_node dX_k = 0 in float64 (the
// column's heat and salt content is conserved exactly before the final
// rounding to the state dtype), X_new = X + dX.  The residual of the implicit
// system, r = X_new - dt D(X_new) - X, bounds |X_new - X_thomas| (the matrix is
// an M-matrix with unit row sums, so ||A^-1||_inf <= 1): a column with
// max|r_T| > tol_T or max|r_S| > tol_S is flagged for the Thomas fallback.
// A workgroup stages the Y rows and T, S columns of 64 columns in LDS with
// coalesced loads (4 waves); one wave then walks the columns.
// ---------------------------------------------------------------------------
constexpr int kApplyMaxNz = 50;

template <typename real>
__global__ __launch_bounds__(256) void k_mlp_apply(Geo<real> g, MixParams m, const real* __restrict__ T,
                                                   const real* __restrict__ S, const float* __restrict__ Y, int ldy,
                                                   const float* __restrict__ omean, const float* __restrict__ ostd,
                                                   double tol_T, double tol_S, real* __restrict__ To,
                                                   real* __restrict__ So, int* __restrict__ flag,
                                                   int* __restrict__ count) {
  constexpr int YW = 2 * kApplyMaxNz + 1;
  __shared__ float ys[64 * YW];
  __shared__ real ts[kApplyMaxNz * 64], ss[kApplyMaxNz * 64];
  const int nz = g.nz;
  const long long n2 = g.n2;
  const long long r0 = (long long)blockIdx.x * 64;
  const int ncol = (int)(n2 - r0 < 64 ? n2 - r0 : 64);
  for (int q = threadIdx.x; q < ncol * 2 * nz; q += 256) {
    const int rr = q / (2 * nz), k = q - rr * 2 * nz;
    ys[rr * YW + k] = Y[(r0 + rr) * ldy + k];
  }
  for (int q = threadIdx.x; q < nz * 64; q += 256) {
    const int k = q >> 6, l = q & 63;
    if (l < ncol) {
      ts[k * 64 + l] = T[k * n2 + r0 + l];
      ss[k * 64 + l] = S[k * n2 + r0 + l];
    }
  }
  __syncthreads();
  if (threadIdx.x >= 64 || (int)threadIdx.x >= ncol) return;
  const int l = threadIdx.x;
  const long long c2 = r0 + l;
  const int nw = g.nwet[c2];
  for (int k = nw; k < nz; ++k) {
    To[k * n2 + c2] = ts[k * 64 + l];
    So[k * n2 + c2] = ss[k * 64 + l];
  }
  if (nw == 0) {
    flag[c2] = 0;
    return;
  }
  const float* y = ys + l * YW;
  auto incT = [&](int k) { return m.dt * ((double)y[k] * (double)ostd[k] + (double)omean[k]); };
  auto incS = [&](int k) { return m.dt * ((double)y[nz + k] * (double)ostd[nz + k] + (double)omean[nz + k]); };
  double sT = 0.0, sS = 0.0, H = 0.0;
  for (int k = 0; k < nw; ++k) {
    sT += m.dzn[k] * incT(k);
    sS += m.dzn[k] * incS(k);
    H += m.dzn[k];
  }
  const double shT = sT / H, shS = sS / H;
  double T0 = ts[l], S0 = ss[l];
  double xT = T0 + (incT(0) - shT), xS = S0 + (incS(0) - shS);
  double xTm = 0.0, xSm = 0.0, km = 0.0, rT = 0.0, rS = 0.0;
  for (int k = 0; k < nw; ++k) {
    double T1 = 0.0, S1 = 0.0, xTp = 0.0, xSp = 0.0, kp = 0.0;
    if (k + 1 < nw) {
      T1 = ts[(k + 1) * 64 + l];
      S1 = ss[(k + 1) * 64 + l];
      xTp = T1 + (incT(k + 1) - shT);
      xSp = S1 + (incS(k + 1) - shS);
      kp = mix_kappa(T0, S0, T1, S1, m);
    }
    const double fm = k > 0 ? km / m.dzi[k - 1] : 0.0;
    const double fp = k + 1 < nw ? kp / m.dzi[k] : 0.0;
    const double qT = -fm * (xT - xTm) + fp * (xTp - xT);
    const double qS = -fm * (xS - xSm) + fp * (xSp - xS);
    rT = fmax(rT, fabs(xT - m.dt * qT / m.dzn[k] - T0));
    rS = fmax(rS, fabs(xS - m.dt * qS / m.dzn[k] - S0));
    To[k * n2 + c2] = (real)xT;
    So[k * n2 + c2] = (real)xS;
    xTm = xT;
    xSm = xS;
    xT = xTp;
    xS = xSp;
    T0 = T1;
    S0 = S1;
    km = kp;
  }
  const int bad = (rT > tol_T || rS > tol_S) ? 1 : 0;
  flag[c2] = bad;
  if (bad) atomicAdd(count, 1);
}

// ---------------------------------------------------------------------------
// BF16 GEMM C = A B^T on MFMA (v_mfma_f32_16x16x32_bf16), fp32 accumulation.
// 128 x 128 block tile, 4 waves of 64 x 64, K step 32 staged in LDS with
// padded rows (40 bf16).  Operand fragments: lane l holds row (l % 16) and the
// 8 consecutive k of group l / 16; the 16 x 16 result holds rows
// 4 (l / 16) + r, column l % 16 in element r.
// ---------------------------------------------------------------------------
constexpr int kGB = 128, kGK = 32, kGPad = 40;

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.7071067811865476f)); }

__device__ __forceinline__ float gelu_grad_f(float z) {
  return 0.5f * (1.0f + erff(z * 0.7071067811865476f)) + z * 0.3989422804014327f * expf(-0.5f * z * z);
}

__global__ __launch_bounds__(256) void k_gemm_bf16_nt(long long M, long long N, long long K,
                                                      const __bf16* __restrict__ A, const __bf16* __restrict__ B,
                                                      int epi, const float* __restrict__ bias,
                                                      const float* __restrict__ aux, float* __restrict__ Cf,
                                                      __bf16* __restrict__ Cb, __bf16* __restrict__ CTb) {
  __shared__ __attribute__((aligned(16))) __bf16 As[kGB * kGPad];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[kGB * kGPad];
  const long long ntn = N / kGB;
  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  const long long m0 = (tile / ntn) * kGB, n0 = (tile % ntn) * kGB;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  f32x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int lr = t >> 1, lh = (t & 1) * 16;
  const __bf16* Ag = A + (m0 + lr) * K + lh;
  const __bf16* Bg = B + (n0 + lr) * K + lh;
  __bf16* Asw = As + lr * kGPad + lh;
  __bf16* Bsw = Bs + lr * kGPad + lh;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (long long k0 = 0; k0 < K; k0 += kGK) {
    const uint4 a0 = *reinterpret_cast<const uint4*>(Ag + k0);
    const uint4 a1 = *reinterpret_cast<const uint4*>(Ag + k0 + 8);
    const uint4 b0 = *reinterpret_cast<const uint4*>(Bg + k0);
    const uint4 b1 = *reinterpret_cast<const uint4*>(Bg + k0 + 8);
    __syncthreads();
    *reinterpret_cast<uint4*>(Asw) = a0;
    *reinterpret_cast<uint4*>(Asw + 8) = a1;
    *reinterpret_cast<uint4*>(Bsw) = b0;
    *reinterpret_cast<uint4*>(Bsw + 8) = b1;
    __syncthreads();
    bf16x8 af[4], bf[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      af[q] = *reinterpret_cast<const bf16x8*>(As + (wm * 64 + q * 16 + fr) * kGPad + fk);
      bf[q] = *reinterpret_cast<const bf16x8*>(Bs + (wn * 64 + q * 16 + fr) * kGPad + fk);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bf[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const long long mb = m0 + wm * 64 + a * 16 + 4 * (lane >> 4);
      const long long n = n0 + wn * 64 + b * 16 + fr;
      float out[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long mm = mb + r;
        float x = acc[a][b][r];
        if (epi == kEpiF32) {
          Cf[mm * N + n] = x;
        } else if (epi == kEpiBiasF32) {
          Cf[mm * N + n] = x + bias[n];
        } else if (epi == kEpiBiasGelu) {
          const float z = x + bias[n];
          if (Cf) Cf[mm * N + n] = z;
          x = gelu_f(z);
          Cb[mm * N + n] = (__bf16)x;
        } else {
          x = x * gelu_grad_f(aux[mm * N + n]);
          Cb[mm * N + n] = (__bf16)x;
        }
        out[r] = x;
      }
      if ((epi == kEpiBiasGelu || epi == kEpiGeluGrad) && CTb) {
        bf16x4 pk = {(__bf16)out[0], (__bf16)out[1], (__bf16)out[2], (__bf16)out[3]};
        *reinterpret_cast<bf16x4*>(CTb + n * M + mb) = pk;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// training helpers
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mse_grad(long long rows, int ld, int n_out, const float* __restrict__ Y,
                                                  const float* __restrict__ target, __bf16* __restrict__ dY,
                                                  __bf16* __restrict__ dYT, float* __restrict__ loss_sum) {
  __shared__ float red[256];
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  float d2 = 0.f;
  if (idx < rows * ld) {
    const long long r = idx / ld;
    const int n = (int)(idx - r * ld);
    const float d = n < n_out ? Y[idx] - target[idx] : 0.f;
    const float gr = 2.0f * d / ((float)rows * (float)n_out);
    dY[idx] = (__bf16)gr;
    dYT[(long long)n * rows + r] = (__bf16)gr;
    d2 = d * d;
  }
  red[threadIdx.x] = d2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss_sum, red[0]);
}

__global__ __launch_bounds__(256) void k_rowsum_bf16(long long cols, const __bf16* __restrict__ in,
                                                     float* __restrict__ out) {
  __shared__ float red[256];
  const __bf16* row = in + (long long)blockIdx.x * cols;
  float s = 0.f;
  for (long long c = threadIdx.x; c < cols; c += 256) s += (float)row[c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_transpose_bf16(long long rows, long long cols, const __bf16* __restrict__ in,
                                                        __bf16* __restrict__ out) {
  __shared__ __bf16 tile[32][33];
  const long long ntc = (cols + 31) / 32;
  const long long br = blockIdx.x / ntc, bc = blockIdx.x % ntc;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int q = ty; q < 32; q += 8) {
    const long long r = br * 32 + q, c = bc * 32 + tx;
    if (r < rows && c < cols) tile[q][tx] = in[r * cols + c];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const long long c = bc * 32 + q, r = br * 32 + tx;
    if (r < rows && c < cols) out[c * rows + r] = tile[tx][q];
  }
}

__global__ __launch_bounds__(256) void k_adam(int n_out, int n_in, float* __restrict__ w, const float* __restrict__ gw,
                                              float* __restrict__ m, float* __restrict__ v, __bf16* __restrict__ w_bf,
                                              __bf16* __restrict__ wT_bf, float* __restrict__ b,
                                              const float* __restrict__ gb, float* __restrict__ bm,
                                              float* __restrict__ bv, float lr, float beta1, float beta2, float eps,
                                              float c1, float c2) {
  const long long nw = (long long)n_out * n_in;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx < nw) {
    const float gr = gw[idx];
    const float mm = beta1 * m[idx] + (1.0f - beta1) * gr;
    const float vv = beta2 * v[idx] + (1.0f - beta2) * gr * gr;
    m[idx] = mm;
    v[idx] = vv;
    const float x = w[idx] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    w[idx] = x;
    const long long o = idx / n_in, i = idx - o * n_in;
    w_bf[idx] = (__bf16)x;
    wT_bf[i * n_out + o] = (__bf16)x;
  } else if (b && idx < nw + n_out) {
    const long long q = idx - nw;
    const float gr = gb[q];
    const float mm = beta1 * bm[q] + (1.0f - beta1) * gr;
    const float vv = beta2 * bv[q] + (1.0f - beta2) * gr * gr;
    bm[q] = mm;
    bv[q] = vv;
    b[q] = b[q] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename real>
void launch_vmix(const Geo<real>& g, const MixParams& m, const real* T, const real* S, real* To, real* So,
                 const int* only, double* work, hipStream_t s) {
  const unsigned blocks = (unsigned)((g.n2 + 255) / 256);
  hipLaunchKernelGGL(k_vmix<real>, dim3(blocks), dim3(256), 0, s, g, m, T, S, To, So, only, work, 0ll, g.n2);
}

// ---------------------------------------------------------------------------
// Allocation-free Thomas solve: the elimination values live in a file-scope
// device array of kStaticScratchDoubles float64 (32 MiB, held by every GPU
// the library runs on until the process ends) and the grid is solved in
// column ranges that fit it, one launch per range, in stream order.  In host
// code `scratch` only names the symbol; its device address on the current
// GPU comes from hipGetSymbolAddress, looked up once per device.  Every call
// on one GPU shares that GPU's array: calls that overlap on one GPU (different
// streams, or threads) race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_lock;
double* scratch_address[kMaxDevices] = {};
}  // namespace

double* static_scratch() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal beyond the static scratch table");
  std::lock_guard<std::mutex> guard(scratch_lock);
  if (!scratch_address[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_address[dev] = static_cast<double*>(p);
  }
  return scratch_address[dev];
}

long long static_vmix_columns(int nz) {
  const long long width = kStaticScratchDoubles / (3ll * nz) / 256 * 256;
  return width;
}

template <typename real>
void launch_vmix_static(const Geo<real>& g, const MixParams& m, const real* T, const real* S, real* To, real* So,
                        const int* only, hipStream_t s) {
  const long long width = static_vmix_columns(g.nz);
  if (width < 256) throw std::runtime_error("too many levels for the static Thomas scratch");
  double* work = static_scratch();
  for (long long c0 = 0; c0 < g.n2; c0 += width) {
    const long long count = g.n2 - c0 < width ? g.n2 - c0 : width;
    const unsigned blocks = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL(k_vmix<real>, dim3(blocks), dim3(256), 0, s, g, m, T, S, To, So, only, work, c0, count);
  }
}

template <typename real>
void launch_col_features(const Geo<real>& g, const MixParams& m, const real* T, const real* S, const real* u,
                         const real* v, const real* q_heat, const real* s_ref2, real restore_coef_S,
                         const real* tau_x, const real* tau_y, const float* mean, const float* inv_std, void* X,
                         long long rows, hipStream_t s) {
  const unsigned blocks = (unsigned)((rows + 63) / 64);
  hipLaunchKernelGGL(k_col_features<real>, dim3(blocks), dim3(256), 0, s, g, m, T, S, u, v, q_heat, s_ref2,
                     restore_coef_S, tau_x, tau_y, mean, inv_std, (unsigned*)X, rows);
}

template <typename real>
void launch_mlp_apply(const Geo<real>& g, const MixParams& m, const real* T, const real* S, const float* Y, int ldy,
                      const float* out_mean, const float* out_std, double tol_T, double tol_S, real* To, real* So,
                      int* flag, int* count, hipStream_t s) {
  const unsigned blocks = (unsigned)((g.n2 + 63) / 64);
  hipLaunchKernelGGL(k_mlp_apply<real>, dim3(blocks), dim3(256), 0, s, g, m, T, S, Y, ldy, out_mean, out_std, tol_T,
                     tol_S, To, So, flag, count);
}

void launch_gemm_bf16_nt(long long M, long long N, long long K, const void* A, const void* B, int epi,
                         const float* bias, const float* aux, float* C_f32, void* C_bf16, void* CT_bf16,
                         hipStream_t s) {
  const unsigned blocks = (unsigned)((M / kGB) * (N / kGB));
  hipLaunchKernelGGL(k_gemm_bf16_nt, dim3(blocks), dim3(256), 0, s, M, N, K, (const __bf16*)A, (const __bf16*)B,
                     epi, bias, aux, C_f32, (__bf16*)C_bf16, (__bf16*)CT_bf16);
}

void launch_mse_grad(long long rows, int ld, int n_out, const float* Y, const float* target, void* dY, void* dYT,
                     float* loss_sum, hipStream_t s) {
  const unsigned blocks = (unsigned)((rows * ld + 255) / 256);
  hipLaunchKernelGGL(k_mse_grad, dim3(blocks), dim3(256), 0, s, rows, ld, n_out, Y, target, (__bf16*)dY,
                     (__bf16*)dYT, loss_sum);
}

void launch_rowsum_bf16(long long rows, long long cols, const void* in, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_rowsum_bf16, dim3((unsigned)rows), dim3(256), 0, s, cols, (const __bf16*)in, out);
}

void launch_transpose_bf16(long long rows, long long cols, const void* in, void* out, hipStream_t s) {
  const unsigned blocks = (unsigned)(((rows + 31) / 32) * ((cols + 31) / 32));
  hipLaunchKernelGGL(k_transpose_bf16, dim3(blocks), dim3(256), 0, s, rows, cols, (const __bf16*)in, (__bf16*)out);
}

void launch_adam(int n_out, int n_in, float* w, const float* gw, float* m, float* v, void* w_bf, void* wT_bf,
                 float* b, const float* gb, float* bm, float* bv, float lr, float beta1, float beta2, float eps,
                 float c1, float c2, hipStream_t s) {
  const long long n = (long long)n_out * n_in + (b ? n_out : 0);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_out, n_in, w, gw, m, v,
                     (__bf16*)w_bf, (__bf16*)wT_bf, b, gb, bm, bv, lr, beta1, beta2, eps, c1, c2);
}

#define ZM_COL_INSTANTIATE(real)                                                                              \
  template void launch_vmix<real>(const Geo<real>&, const MixParams&, const real*, const real*, real*, real*, \
                                  const int*, double*, hipStream_t);                                          \
  template void launch_vmix_static<real>(const Geo<real>&, const MixParams&, const real*, const real*, real*,  \
                                         real*, const int*, hipStream_t);                                     \
  template void launch_col_features<real>(const Geo<real>&, const MixParams&, const real*, const real*,       \
                                          const real*, const real*, const real*, const real*, real,           \
                                          const real*, const real*, const float*, const float*, void*,        \
                                          long long, hipStream_t);                                            \
  template void launch_mlp_apply<real>(const Geo<real>&, const MixParams&, const real*, const real*,          \
                                       const float*, int, const float*, const float*, double, double, real*,  \
                                       real*, int*, int*, hipStream_t);

ZM_COL_INSTANTIATE(float)
ZM_COL_INSTANTIATE(double)

}  // namespace zm
