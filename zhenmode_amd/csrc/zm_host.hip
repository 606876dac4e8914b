// Host orchestration of the mode-split step and the C ABI (include/zhenmode.h).
//
// One step (reference _step_impl, timestepping/step.py:341-458):
//   L(dt/2) -> N(dt) -> L(dt/2) -> barotropic sub-cycle -> final masks,
// issued as a fixed kernel sequence on one stream and, by default, replayed
// from a hipGraph keyed by the state/forcing pointers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/zhenmode.h"
#include "zm_common.h"
#include "zm_column.h"
#include "zm_kernels.h"
#include "zm_ledger.h"
#include "zm_project.h"
#include "zm_runtime.h"

namespace {

thread_local std::string g_thread_error;

struct HipError {
  hipError_t code;
  std::string where;
};

#define ZM_HIP(call)                                                  \
  do {                                                                \
    hipError_t e_ = (call);                                           \
    if (e_ != hipSuccess) throw HipError{e_, std::string(#call)};     \
  } while (0)

struct ConfigError {
  std::string msg;
};

using GraphKey = std::tuple<std::vector<const void*>, int>;

struct SolverBase {
  virtual ~SolverBase() = default;
  virtual void step(const zm_state& in, const zm_state& out, int nsteps, const void* const* forcing,
                    hipStream_t s) = 0;
  virtual void diagnostics(const zm_state& in, void* rho, void* p, void* w, hipStream_t s) = 0;
  virtual void tracer_terms(const zm_state& in, void* terms, hipStream_t s) = 0;
  virtual void stage(int which, const zm_state& in, const zm_state& out, hipStream_t s) = 0;
  virtual double time_kernel(int which, int reps) = 0;
  virtual int nu_substeps_per_launch() const = 0;
  virtual void vertical_mix(const void* T, const void* S, void* To, void* So, double dt, double kv, double kc,
                            const int* only, hipStream_t s) = 0;
  virtual void vertical_mix_static(const void* T, const void* S, void* To, void* So, double dt, double kv,
                                   double kc, const int* only, hipStream_t s) = 0;
  virtual void column_features(const zm_state& in, const void* tau_x, const void* tau_y, const void* q_heat,
                               const float* mean, const float* inv_std, void* X, long long rows,
                               hipStream_t s) = 0;
  virtual void mlp_apply(const zm_state& in, const float* Y, int ldy, const float* omean, const float* ostd,
                         double dt, double kv, double kc, double tol_T, double tol_S, void* To, void* So, int* flags,
                         int* count, hipStream_t s) = 0;
  virtual void mix_targets(const zm_state& in, const void* Tt, const void* St, double dt, double kv, double kc,
                           float* Y, int ldy, hipStream_t s) = 0;
  virtual void mlp_fused(const zm_state& in, const void* tau_x, const void* tau_y, const void* q_heat,
                         const float* in_mean, const float* in_inv_std, const void* const* weights,
                         const float* const* biases, const float* omean, const float* ostd, double dt, double kv,
                         double kc, double tol_T, double tol_S, void* To, void* So, int* flags, int* count,
                         hipStream_t s) = 0;
  virtual void seasonal(const void* stack, int nfields, double day, double blend_days, void* out,
                        hipStream_t s) = 0;
  virtual void monitor(const zm_state& st, long long step, double vlim, double elim, int inclusive, void* status,
                       hipStream_t s) = 0;
  virtual void budget(const zm_state& st, double* sums, hipStream_t s) = 0;
  virtual void project_velocity(const void* u, const void* v, void* uo, void* vo, double dt, int* its,
                                hipStream_t s) = 0;
  virtual void projection_iterations(int* its, hipStream_t s) = 0;
  virtual void budget_step(const zm_state& in, const zm_state& out, const void* const* forcing, double* ledger,
                           hipStream_t s) = 0;
  virtual void set_tiles(int nu_block, int seg) = 0;
  virtual void get_tiles(int* nu_block, int* seg) const = 0;
  virtual void set_bt_persistent(bool on) = 0;
  virtual bool bt_persistent_on() const = 0;
  virtual int bt_timeouts() = 0;
  int nz_ = 0;
  long long n2_ = 0;
  bool use_graph = true;
};

template <typename real>
struct Solver : SolverBase {
  zm::Geo<real> g{};
  zm::Phys<real> P{};
  int nx = 0, ny = 0, nz = 0;
  long long n2 = 0, n3 = 0;
  int n_sub = 0, n_nu = 0, N1 = 1, N2 = 1;
  zm::FftPlan fft{};
  bool use_fft = false;
  int* kept = nullptr;
  real dt_bt = 0, dtH = 0, drag = 0, decay_bot = 0, dts = 0;
  int nu_block = 1;
  std::vector<void*> allocs;
  real *twr = nullptr, *twi = nullptr, *fdt = nullptr, *fiden = nullptr;
  double *dzn64 = nullptr, *dzi64 = nullptr, *idzn64 = nullptr, *idzi64 = nullptr;
  double* vmix_scratch = nullptr;  // Thomas c', d' (3 n3), allocated on first use
  double* area64 = nullptr;        // cell area per latitude row (dx dy), float64
  double* red_partials = nullptr;  // monitor / budget block partials
  double T_ref64 = 0, S_ref64 = 0;
  // column-divergence projection (float64 solver state)
  bool proj_on = false;
  int proj_niter = 1, proj_maxref = 0;
  double dt64 = 0;
  zm::ProjGeo pg{};
  zm::ProjWork pw{};
  real *ua = nullptr, *va = nullptr;  // projected advecting velocity
  // stage ledger (allocated on the first budget step)
  bool ledger_on = false;
  double *led_raw = nullptr, *led_part = nullptr;
  size_t led_part_n = 0;
  // workspace
  real *u1, *v1, *T1, *S1, *nAu, *nAv, *nBu, *nBv, *Lu, *Lv, *DT, *DS;
  real *Fz1, *Fz2, *cT, *cS;
  int* gate;
  unsigned long long* ubits;  // unstable interfaces of the gated columns (2 words per column)
  real *dT1, *dS1, *Tp, *Sp, *du1, *dv1, *up, *vp, *T2, *S2, *u2, *v2, *pbc, *au, *av, *axu, *axv;
  real *u3, *v3, *T3, *S3, *tmpA, *tmpB;
  real *Frx, *Fry, *ubt0, *vbt0, *ubt, *vbt, *etaw, *means;
  // persistent barotropic sub-cycle: per-block progress counters, timeout counter
  bool bt_persist = true, bt_persist_fits = true;
  int* bt_progress = nullptr;
  int* bt_timeout = nullptr;
  real *ubt2, *vbt2, *etaw2;  // the other half of the barotropic ping-pong
  hipStream_t own = nullptr;
  // two side streams for the stages of a step that do not depend on each other
  // (captured as parallel branches of the step graph), joined by events
  hipStream_t sideA = nullptr, sideB = nullptr;
  static constexpr int kEvents = 16;
  // captured step graphs kept per handle, one per (input, output, forcing)
  // pointer set: a run loop stepping through a ring of n device states needs n
  static constexpr size_t kGraphCache = 32;
  // false while an ensemble issues this member's stages: the members then run
  // side by side on the ensemble's own streams and each member's stages stay
  // on the stream it is given
  bool fork_stages = true;
  bool fork_linear = true;     // the tracer biharmonic's D1 beside the nu_h chain
  bool fork_nonlinear = true;  // convection and momentum advection beside each other
  hipStream_t lane_a(hipStream_t s) const { return fork_stages ? sideA : s; }
  hipStream_t lane_b(hipStream_t s) const { return fork_stages ? sideB : s; }
  hipEvent_t ev_pool[kEvents] = {};
  int ev_next = 0;
  std::map<GraphKey, std::pair<hipGraph_t, hipGraphExec_t>> graphs;

  template <typename T>
  T* dev_alloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) throw HipError{e, "hipMalloc"};
    allocs.push_back(p);
    ZM_HIP(hipMemset(p, 0, count * sizeof(T)));
    return static_cast<T*>(p);
  }

  real* upload(const std::vector<double>& h) {
    real* d = dev_alloc<real>(h.size());
    std::vector<real> t(h.begin(), h.end());
    if (!t.empty()) ZM_HIP(hipMemcpy(d, t.data(), t.size() * sizeof(real), hipMemcpyHostToDevice));
    return d;
  }

  real* upload2d(const double* h, double fill) {
    std::vector<double> v(n2, fill);
    if (h) std::copy(h, h + n2, v.begin());
    return upload(v);
  }

  ~Solver() override {
    for (auto& kv : graphs) {
      hipGraphExecDestroy(kv.second.second);
      hipGraphDestroy(kv.second.first);
    }
    if (own) hipStreamDestroy(own);
    if (sideA) hipStreamDestroy(sideA);
    if (sideB) hipStreamDestroy(sideB);
    for (hipEvent_t e : ev_pool)
      if (e) hipEventDestroy(e);
    for (void* p : allocs) hipFree(p);
  }

  // `waiter` continues only after everything issued on `src` so far
  void after(hipStream_t waiter, hipStream_t src) {
    if (waiter == src) return;
    hipEvent_t e = ev_pool[ev_next];
    ev_next = (ev_next + 1) % kEvents;
    ZM_HIP(hipEventRecord(e, src));
    ZM_HIP(hipStreamWaitEvent(waiter, e, 0));
  }

  explicit Solver(const zm_config& c) {
    nx = c.nx;
    ny = c.ny;
    nz = c.nz;
    n2 = (long long)nx * ny;
    n3 = n2 * nz;
    n_sub = c.n_subcyc;
    n_nu = c.n_nu;
    const int ncap = c.polar_cap_rows > 0 ? c.polar_cap_rows : 0;
    const int nb = ncap > 0 ? ncap + c.polar_cap_taper : 0;
    // geometry rows
    std::vector<double> cosl(ny), idx(ny), idx2(ny), tanR(ny), frow(ny), cfn(ny), cfs(ny), idyc(ny);
    std::vector<double> rc(ny), rs(ny), fd(ny), fiden_h(ny);
    const double idy = 1.0 / c.dy;
    for (int j = 0; j < ny; ++j) {
      cosl[j] = c.cos_lat[j];
      idx[j] = 1.0 / c.dx_row[j];
      idx2[j] = 1.0 / (c.dx_row[j] * c.dx_row[j]);
      const double sinl = c.f_row[j] / (2.0 * 7.2921e-5);
      tanR[j] = (sinl / c.cos_lat[j]) / 6371.0e3;
      frow[j] = c.f_row[j];
      cfn[j] = 0.5 * (c.cos_lat[j] + c.cos_lat[(j + 1) % ny]);
      cfs[j] = 0.5 * (c.cos_lat[j] + c.cos_lat[(j + ny - 1) % ny]);
      idyc[j] = idy / c.cos_lat[j];
      rc[j] = std::cos(c.f_row[j] * (c.dt / 2.0));
      rs[j] = std::sin(c.f_row[j] * (c.dt / 2.0));
      fd[j] = c.f_row[j] * c.dt_bt;
      fiden_h[j] = 1.0 / (1.0 + fd[j] * fd[j]);
    }
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.n2 = (int)n2;
    g.n3 = n3;
    g.cosl = upload(cosl);
    g.idx = upload(idx);
    g.idx2 = upload(idx2);
    g.tanR = upload(tanR);
    g.frow = upload(frow);
    g.cfn = upload(cfn);
    g.cfs = upload(cfs);
    g.idyc = upload(idyc);
    g.idy = (real)idy;
    g.idy2 = (real)(idy * idy);
    g.hidy = (real)(0.5 * idy);
    std::vector<double> dzn(c.dz_node, c.dz_node + nz), idzn(nz);
    for (int k = 0; k < nz; ++k) idzn[k] = 1.0 / dzn[k];
    std::vector<double> dzi(c.dz_iface, c.dz_iface + nz - 1), idzi(nz - 1);
    for (int k = 0; k + 1 < nz; ++k) idzi[k] = 1.0 / dzi[k];
    std::vector<double> dz3(c.dz, c.dz + nz - 1), dznorm(c.dz_norm, c.dz_norm + nz - 1);
    const int ni = nz > 2 ? nz - 2 : 0;
    std::vector<double> hm(ni), hp(ni), hmp(ni), iden(ni), iddz(ni);
    for (int k = 0; k < ni; ++k) {
      hm[k] = c.d2z_hm[k];
      hp[k] = c.d2z_hp[k];
      hmp[k] = c.d2z_hm[k] + c.d2z_hp[k];
      iden[k] = 1.0 / c.d2z_denom[k];
      iddz[k] = 1.0 / c.dz_denom_interior[k];
    }
    g.dzn = upload(dzn);
    g.idzn = upload(idzn);
    g.dzi = upload(dzi);
    g.idzi = upload(idzi);
    g.dz3 = upload(dz3);
    g.dznorm = upload(dznorm);
    g.hm = upload(hm);
    g.hp = upload(hp);
    g.hmp = upload(hmp);
    g.iden = upload(iden);
    g.iddz = upload(iddz);
    g.idz_top = (real)(1.0 / c.dz_bnd_top);
    g.idz_bot = (real)(1.0 / c.dz_bnd_bot);
    g.ih0top2 = (real)(1.0 / (c.d2z_h0_top * c.d2z_h0_top));
    g.ih0bot2 = (real)(1.0 / (c.d2z_h0_bot * c.d2z_h0_bot));
    int* nw = dev_alloc<int>(n2);
    ZM_HIP(hipMemcpy(nw, c.nwet, n2 * sizeof(int), hipMemcpyHostToDevice));
    g.nwet = nw;
    // levels per thread of the column kernels: whole columns when there are
    // enough of them to fill the GPU (best for every column kernel at 0.5
    // degrees), else 10 (zm_autotune refines either)
    g.seg = n2 >= 65536 ? nz : std::min(10, nz);
    g.nu_chunk = 28;  // rows per wave chunk of the streaming nu_h kernel (measured best at 0.5 degrees)
    if (const char* env = std::getenv("ZHENMODE_NU_CHUNK")) g.nu_chunk = std::max(0, std::atoi(env));
    // independent stages as parallel branches of the step graph (1, default) or one chain (0)
    if (const char* env = std::getenv("ZHENMODE_FORK_STAGES")) {
      // 0: one chain; 1: both stages forked; 2: the nonlinear stage only; 3: the linear stages only
      const int mode = std::atoi(env);
      fork_stages = mode != 0;
      fork_linear = mode == 1 || mode == 3;
      fork_nonlinear = mode == 1 || mode == 2;
    }
    g.nu_chunk = std::min(g.nu_chunk, zm::kNuChunkMax);
    {
      // nu_h sub-cycle: padded row coefficients of the substep x + nu_h dts
      // lap(x) (the same nu_h and dts as P.nu_h and dts below) and
      // per-(level, chunk, column) wet words of the streaming kernel
      const double nd = c.nu_h * ((c.dt / 2.0) / c.n_nu), idy2 = idy * idy;
      std::vector<double> met((size_t)zm::kNuMetStride * (ny + 2 * zm::kNuMetPad), 0.0);
      for (int j = 0; j < ny; ++j) {
        double* row = met.data() + (size_t)zm::kNuMetStride * (j + zm::kNuMetPad);
        const double a = 0.5 * idy * tanR[j];
        row[0] = nd * idx2[j];
        row[1] = nd * (idy2 - a);
        row[2] = nd * (idy2 + a);
        row[3] = nd * idy2;
        row[4] = idx2[j];
        row[5] = tanR[j];
      }
      g.nu_met = upload(met);
      g.nu_h_met = (real)c.nu_h;
      g.nu_dts_met = (real)((c.dt / 2.0) / c.n_nu);
      const int chunk = g.nu_chunk > 0 ? g.nu_chunk : 32;
      const int nch = (ny + chunk - 1) / chunk;
      std::vector<unsigned long long> wet((size_t)nz * nch * nx, 0ull);
      for (int ch = 0; ch < nch; ++ch)
        for (int b = 0; b < 64; ++b) {
          const int j = ch * chunk - zm::kNuWetBase + b;
          if (j < 0 || j >= ny) continue;
          for (int i = 0; i < nx; ++i) {
            const int nwc = std::min(c.nwet[(size_t)j * nx + i], nz);
            for (int k = 0; k < nwc; ++k) wet[((size_t)k * nch + ch) * nx + i] |= 1ull << b;
          }
        }
      unsigned long long* dw = dev_alloc<unsigned long long>(wet.size());
      ZM_HIP(hipMemcpy(dw, wet.data(), wet.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
      g.nu_wet = dw;
    }
    dzn64 = dev_alloc<double>(nz);
    ZM_HIP(hipMemcpy(dzn64, c.dz_node, nz * sizeof(double), hipMemcpyHostToDevice));
    dzi64 = dev_alloc<double>(nz > 1 ? nz - 1 : 1);
    if (nz > 1) ZM_HIP(hipMemcpy(dzi64, c.dz_iface, (nz - 1) * sizeof(double), hipMemcpyHostToDevice));
    {
      std::vector<double> rn(nz), ri(nz > 1 ? nz - 1 : 1, 0.0);
      for (int k = 0; k < nz; ++k) rn[k] = 1.0 / c.dz_node[k];
      for (int k = 0; k + 1 < nz; ++k) ri[k] = 1.0 / c.dz_iface[k];
      idzn64 = dev_alloc<double>(nz);
      idzi64 = dev_alloc<double>(ri.size());
      ZM_HIP(hipMemcpy(idzn64, rn.data(), rn.size() * sizeof(double), hipMemcpyHostToDevice));
      ZM_HIP(hipMemcpy(idzi64, ri.data(), ri.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    {
      std::vector<double> ar(ny);
      for (int j = 0; j < ny; ++j) ar[j] = c.dx_row[j] * c.dy;
      area64 = dev_alloc<double>(ny);
      ZM_HIP(hipMemcpy(area64, ar.data(), ny * sizeof(double), hipMemcpyHostToDevice));
      red_partials = dev_alloc<double>((size_t)9 * zm::kReduceBlocks);
    }
    T_ref64 = c.T_ref;
    S_ref64 = c.S_ref;
    dt64 = c.dt;
    setup_projection(c);
    nz_ = nz;
    n2_ = n2;

    P.dt = (real)c.dt;
    P.dth = (real)(c.dt / 2.0);
    P.nu_h = (real)c.nu_h;
    P.nu_v = (real)c.nu_v;
    P.kappa_h = (real)c.kappa_h;
    P.nu_bi = (real)c.nu_bi;
    P.kappa_bi = (real)c.kappa_bi;
    P.kappa_conv = (real)c.kappa_conv;
    P.T_ref = (real)c.T_ref;
    P.S_ref = (real)c.S_ref;
    P.r_bot = (real)c.r_bot;
    P.H_sw = (real)c.H_sw;
    P.dz_surface = (real)c.dz_surface;
    P.wind_factor = (real)(1.0 / (1025.0 * c.dz_surface));
    P.inv_rho0H = (real)(1.0 / (1025.0 * c.H_sw));
    P.lambda_bulk = (real)c.lambda_bulk;
    P.restore_coef_S = (real)c.restore_coef_S;
    P.ice_freeze = (real)c.ice_freeze_temp_c;
    P.ice_salt_flux = (real)c.ice_salt_flux;
    P.heat_factor = (real)((1.0 / c.dz_surface) / (1025.0 * 3992.0));
    P.conv_nsub = c.conv_nsub;
    P.cons_kv = c.conservative_kv != 0;
    P.local_conv = c.localize_conv != 0;
    P.adv_scheme = c.fct_adv ? 2 : (c.monotone_adv ? 1 : 0);
    P.ncap = ncap;
    P.nb = nb;
    std::vector<double> capw(nb > 0 ? nb : 1, 0.0);
    for (int q = 0; q < nb; ++q) capw[q] = c.cap_weights[q];
    P.capw = upload(capw);
    P.wet2 = upload2d(c.wet_mask, 0.0);
    P.kh2 = upload2d(c.kappa_h_2d, c.kappa_h);
    P.kv2 = upload2d(c.kappa_v_2d, c.kappa_v);
    P.tau_x = upload2d(c.tau_x, 0.0);
    P.tau_y = upload2d(c.tau_y, 0.0);
    P.q_heat = upload2d(c.q_heat, 0.0);
    P.t_atm = upload2d(c.t_atm, 0.0);
    P.s_ref2 = upload2d(c.s_ref_surf, 0.0);
    P.crc = upload2d(c.coastal_restore_coef, 0.0);
    P.crT = upload2d(c.coastal_restore_T, 0.0);
    P.cbl = upload2d(c.coastal_bulk_lambda, 0.0);
    P.rot_c = upload(rc);
    P.rot_s = upload(rs);
    fdt = upload(fd);
    fiden = upload(fiden_h);

    dt_bt = (real)c.dt_bt;
    dtH = (real)(c.dt_bt * c.H_sw);
    drag = (real)(1.0 / (1.0 + c.r_bot * c.dt_bt));
    decay_bot = (real)std::exp(-c.r_bot * c.dt);
    dts = (real)((c.dt / 2.0) / n_nu);
    // substeps fused per launch of the nu_h sub-cycle (1 = one launch per substep)
    nu_block = zm::nu_block_max<real>();
    if (const char* env = std::getenv("ZHENMODE_NU_BLOCK")) nu_block = std::max(1, std::min(nu_block, std::atoi(env)));

    // DFT factorisation nx = N1 * N2 with N1 the largest divisor <= sqrt(nx)
    N1 = 1;
    for (int d = 1; d * d <= nx; ++d)
      if (nx % d == 0) N1 = d;
    N2 = nx / N1;
    // mixed-radix plan (radices 4, 2, 3, 5); the four-step DFT serves lengths
    // with other prime factors
    {
      int rest = nx;
      fft.nst = 0;
      for (int r : {4, 2, 3, 5})
        while (rest % r == 0 && fft.nst < 16) {
          fft.rad[fft.nst++] = r;
          rest /= r;
        }
      use_fft = rest == 1;
    }
    {
      std::vector<int> km(c.dealias_mask, c.dealias_mask + nx);
      kept = dev_alloc<int>(nx);
      ZM_HIP(hipMemcpy(kept, km.data(), nx * sizeof(int), hipMemcpyHostToDevice));
    }
    std::vector<double> wr(nx), wi(nx);
    for (int e = 0; e < nx; ++e) {
      const double a = -2.0 * M_PI * (double)e / (double)nx;
      wr[e] = std::cos(a);
      wi[e] = std::sin(a);
    }
    twr = upload(wr);
    twi = upload(wi);

    real** f3[] = {&u1, &v1, &T1, &S1, &nAu, &nAv, &nBu, &nBv, &Lu, &Lv, &DT, &DS, &cT, &cS,
                   &dT1, &dS1, &Tp, &Sp, &du1, &dv1, &up, &vp, &T2, &S2, &u2, &v2, &pbc,
                   &au, &av, &axu, &axv, &u3, &v3, &T3, &S3, &tmpA, &tmpB};
    for (real** p : f3) *p = dev_alloc<real>(n3);
    Fz1 = dev_alloc<real>(n3 + n2);
    Fz2 = dev_alloc<real>(n3 + n2);
    gate = dev_alloc<int>(n2);
    ubits = dev_alloc<unsigned long long>((size_t)2 * n2);
    real** f2[] = {&Frx, &Fry, &ubt0, &vbt0, &ubt, &vbt, &etaw, &ubt2, &vbt2, &etaw2};
    for (real** p : f2) *p = dev_alloc<real>(n2);
    means = dev_alloc<real>((size_t)4 * nz * (2 * (nb > 0 ? nb : 1)));
    bt_progress = dev_alloc<int>((size_t)(ny + 1) / 2);
    bt_timeout = dev_alloc<int>(1);
    ZM_HIP(hipMemset(bt_timeout, 0, sizeof(int)));
    {
      // the persistent sub-cycle needs all its blocks resident at once: one
      // 1024-thread block per CU at most
      int dev = 0, cus = 0;
      ZM_HIP(hipGetDevice(&dev));
      ZM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      bt_persist_fits = (ny + 1) / 2 <= cus;
      // one launch per substep stays the default: at 0.5 degrees it takes
      // 0.41 ms per sub-cycle against 0.56 ms for the persistent launch
      bt_persist = false;
    }
    ZM_HIP(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
    ZM_HIP(hipStreamCreateWithFlags(&sideA, hipStreamNonBlocking));
    ZM_HIP(hipStreamCreateWithFlags(&sideB, hipStreamNonBlocking));
    for (hipEvent_t& e : ev_pool) ZM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ZM_HIP(hipDeviceSynchronize());
  }

  // ledger sums of one tracer residual evaluation, booked over dt/2
  // (tendencies.py:211-214 with the recorder's default interval)
  void ledger_tracer(const zm::Phys<real>& Pf, const real* u, const real* v, const real* T, const real* S,
                     const real* Fz, hipStream_t s) {
    zm::launch_tracer_ledger(g, Pf, u, v, T, S, Fz, (const real*)cT, (const real*)cS, (const int*)gate, area64,
                             dzn64, led_part, s);
    zm::launch_ledger_tracer_fold(led_part, (int)zm::tracer_ledger_blocks(g), dt64 / 2.0, led_raw, s);
  }

  // make_budget_step (diagnostics/budgets.py:244-270): the step with the
  // recorder's stage differences (step.py:138-161, 365-368, 426-427, 452-456)
  void budget_step(const zm_state& in, const zm_state& out, const void* const* forcing, double* ledger,
                   hipStream_t s) override {
    const void* ins[] = {in.u, in.v, in.T, in.S, in.eta, in.ice};
    const void* outs[] = {out.u, out.v, out.T, out.S, out.eta, out.ice};
    for (const void* a : ins)
      for (const void* b : outs)
        if (a == b) throw ConfigError{"zm_budget_step output must not alias its input"};
    const size_t need = std::max<size_t>({(size_t)zm::tracer_ledger_blocks(g) * zm::kLedgerTracerSums,
                                          (size_t)zm::kLedgerBlocks * zm::kLedgerDiffSums,
                                          (size_t)zm::kLedgerBlocks * 2});
    if (!led_raw) led_raw = dev_alloc<double>(zm::kRawSize);
    if (led_part_n < need) {
      led_part = dev_alloc<double>(need);
      led_part_n = need;
    }
    const zm::Phys<real> Pf = forced(forcing);
    const real* Ti = (const real*)in.T;
    const real* Si = (const real*)in.S;
    const real* ei = (const real*)in.eta;
    const real* ci = (const real*)in.ice;
    auto diff = [&](const real* Ta, const real* Sa, const real* ea, const real* Tb, const real* Sb,
                    const real* eb, int slot) {
      zm::launch_ledger_diff(g, area64, dzn64, Ta, Sa, ea, ci, Tb, Sb, eb, ci, slot, led_part, led_raw, s);
    };
    zm::launch_ledger_reset(led_raw, s);
    linear((const real*)in.u, (const real*)in.v, Ti, Si, u1, v1, T1, S1, s);
    diff(T1, S1, ei, Ti, Si, ei, zm::kStageLinearDiffusion);
    ledger_on = true;
    try {
      nonlinear(Pf, ei, u2, v2, T2, S2, s);
    } catch (...) {
      ledger_on = false;
      throw;
    }
    ledger_on = false;
    diff(T2, S2, ei, T1, S1, ei, zm::kStageNonlinear);
    linear(u2, v2, T2, S2, u3, v3, T3, S3, s);
    diff(T3, S3, ei, T2, S2, ei, zm::kStageLinearDiffusion);
    barotropic(Pf, ei, s);
    diff(T3, S3, etaw, T3, S3, ei, zm::kStageFreeSurface);
    post_barotropic(in, out, Pf, s);
    diff((const real*)out.T, (const real*)out.S, (const real*)out.eta, T3, S3, etaw, zm::kStagePolarCap);
    zm::launch_ledger_diff(g, area64, dzn64, (const real*)out.T, (const real*)out.S, (const real*)out.eta,
                           (const real*)out.ice, Ti, Si, ei, ci, -1, led_part, led_raw, s);
    zm::launch_ledger_finish(led_raw, ledger, s);
    ZM_HIP(hipGetLastError());
  }

  // operator taps, Jacobi scaling and solver buffers of the projection
  void setup_projection(const zm_config& c) {
    proj_on = c.project_adv_vel != 0;
    proj_niter = c.projection_niter;
    proj_maxref = c.projection_max_refinements;
    std::vector<double> coef((size_t)zm::kProjTaps * n2), diag(n2), inv(n2), wet(c.wet_mask, c.wet_mask + n2);
    zm::proj_stencil_host(nx, ny, nz, c.nwet, c.wet_mask, c.dx_row, c.dy, c.cos_lat, c.dz_node, coef.data(),
                          diag.data());
    for (long long q = 0; q < n2; ++q)
      inv[q] = c.projection_preconditioner == 1 ? 1.0 / (diag[q] > 0.0 ? diag[q] : 1.0) : 1.0;
    std::vector<double> idx(ny), idyc(ny), cfn(ny), cfs(ny);
    for (int j = 0; j < ny; ++j) {
      idx[j] = 1.0 / c.dx_row[j];
      idyc[j] = (1.0 / c.dy) / c.cos_lat[j];
      cfn[j] = 0.5 * (c.cos_lat[j] + c.cos_lat[(j + 1) % ny]);
      cfs[j] = 0.5 * (c.cos_lat[j] + c.cos_lat[(j + ny - 1) % ny]);
    }
    auto up64 = [&](const std::vector<double>& h) {
      double* d = dev_alloc<double>(h.size());
      ZM_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
      return d;
    };
    pg.nx = nx;
    pg.ny = ny;
    pg.n2 = (int)n2;
    pg.nblocks = (int)std::min<long long>((n2 + zm::kProjBlock - 1) / zm::kProjBlock, zm::kProjMaxBlocks);
    pg.coef = up64(coef);
    pg.inv = up64(inv);
    pg.wet = up64(wet);
    pg.idx = up64(idx);
    pg.idyc = up64(idyc);
    pg.cfn = up64(cfn);
    pg.cfs = up64(cfs);
    pg.area = area64;
    pg.dzn = dzn64;
    pw.x = dev_alloc<double>(n2);
    pw.r = dev_alloc<double>(n2);
    pw.p0 = dev_alloc<double>(n2);
    pw.p1 = dev_alloc<double>(n2);
    pw.Ap = dev_alloc<double>(n2);
    pw.part_rhs = dev_alloc<double>((size_t)4 * pg.nblocks);
    pw.part_a = dev_alloc<double>((size_t)pg.nblocks);
    pw.part_b = dev_alloc<double>((size_t)2 * pg.nblocks);
    pw.sc = dev_alloc<zm::ProjScalars>(1);
    zm::ProjScalars sc{};
    sc.tol2 = c.projection_rtol * c.projection_rtol;
    sc.ttol2 = c.projection_transport_rtol * c.projection_transport_rtol;
    sc.maxiter = proj_niter;
    sc.done = 1;
    ZM_HIP(hipMemcpy(pw.sc, &sc, sizeof(sc), hipMemcpyHostToDevice));
    if (proj_on) {
      ua = dev_alloc<real>(n3);
      va = dev_alloc<real>(n3);
    }
  }

  // _project_column_divergence(u, v, p, dt) (dynamics/projection.py:64-127)
  void project(const real* u, const real* v, real* uo, real* vo, double dt, hipStream_t s) {
    const double dtg = dt * zm::kG;
    zm::launch_proj_rhs(g, pg, pw, u, v, dtg, 0, s);
    zm::launch_proj_cg(pg, pw, proj_niter, 0, 0, s);
    zm::launch_proj_correct(g, pg, pw, dtg, u, v, uo, vo, s);
    for (int q = 0; q < proj_maxref; ++q) {
      zm::launch_proj_rhs(g, pg, pw, (const real*)uo, (const real*)vo, dtg, 1, s);
      zm::launch_proj_cg(pg, pw, proj_niter, 1, q + 1, s);
      zm::launch_proj_correct(g, pg, pw, dtg, (const real*)uo, (const real*)vo, uo, vo, s);
    }
  }

  void project_velocity(const void* u, const void* v, void* uo, void* vo, double dt, int* its,
                        hipStream_t s) override {
    if (proj_niter < 1) throw ConfigError{"projection_niter must be a positive integer"};
    project((const real*)u, (const real*)v, (real*)uo, (real*)vo, dt, s);
    if (its) projection_iterations(its, s);
    ZM_HIP(hipGetLastError());
  }

  void projection_iterations(int* its, hipStream_t s) override {
    ZM_HIP(hipMemcpyAsync(its, pw.sc->its, 3 * sizeof(int), hipMemcpyDeviceToDevice, s));
  }

  // x part of the de-alias of both advection fields (one complex row each)
  void dealias(const real* a, const real* b, real* oa, real* ob, hipStream_t s) {
    if (use_fft)
      zm::launch_dealias_fft(nx, fft, kept, (int)(ny * nz), twr, twi, a, b, oa, ob, s);
    else
      zm::launch_dealias_x(nx, N1, N2, kept, (int)(ny * nz), twr, twi, a, b, oa, ob, s);
  }

  // ---- stages -----------------------------------------------------------
  void linear(const real* u0, const real* v0, const real* T0, const real* S0, real* uo, real* vo, real* To,
              real* So, hipStream_t s) {
    const bool bih_u = P.nu_bi > real(0), bih_t = P.kappa_bi > real(0);
    // the inner D1 of the tracer biharmonic (memory-bound) beside the nu_h chain (VALU-bound)
    const hipStream_t A = fork_linear ? lane_a(s) : s;
    if (bih_t) {
      after(A, s);
      zm::launch_d1_unit(g, T0, S0, DT, DS, A);
    }
    const real* su = u0;
    const real* sv = v0;
    int n = 0, flip = 0;
    while (n < n_nu) {
      real* du = flip ? nBu : nAu;
      real* dv = flip ? nBv : nAv;
      const bool lap = n == 0 && bih_u;  // the first substep also stores lap(u0), lap(v0)
      // blocks of nu_block substeps; a last single substep would cost nearly
      // a whole block, so a remainder of nu_block + 1 is split in two
      int h = std::min(nu_block, n_nu - n);
      if (n_nu - n - h == 1 && h > 2) h -= 1;
      if (nu_block <= 1 || h <= 1 || (lap && !zm::nu_block_writes_lap(g, h))) {
        zm::launch_nu_substep(g, su, sv, du, dv, lap ? Lu : (real*)nullptr, lap ? Lv : (real*)nullptr, P.nu_h,
                              dts, s);
        n += 1;
      } else {
        zm::launch_nu_block(g, su, sv, du, dv, P.nu_h, dts, h, s, lap ? Lu : (real*)nullptr,
                            lap ? Lv : (real*)nullptr);
        n += h;
      }
      su = du;
      sv = dv;
      flip ^= 1;
    }
    if (bih_t) after(s, A);
    zm::launch_linear_finish(g, P, u0, v0, T0, S0, su, sv, Lu, Lv, DT, DS, uo, vo, To, So, s);
  }

  void nonlinear(const zm::Phys<real>& Pf, const real* eta, real* uo, real* vo, real* To, real* So,
                 hipStream_t s) {
    // tracer residual #1 at the N-step start (u1, v1, T1, S1); the convective
    // gate and sub-cycle (side stream A) need only T1, S1, the momentum
    // advection and its de-aliasing (side stream B) only u1, v1 and Fz1
    const hipStream_t A = fork_nonlinear ? lane_a(s) : s, B = fork_nonlinear ? lane_b(s) : s;
    after(A, s);
    zm::launch_conv(g, Pf, T1, S1, cT, cS, gate, ubits, A);
    zm::launch_fz(g, u1, v1, Fz1, s);
    after(B, s);
    zm::launch_mom_adv(g, u1, v1, Fz1, au, av, B);
    dealias(au, av, axu, axv, B);
    after(s, A);
    zm::launch_tracer_residual(g, Pf, 0, u1, v1, T1, S1, Fz1, cT, cS, gate, dT1, dS1, T1, S1, Tp, Sp, s);
    if (ledger_on) ledger_tracer(Pf, u1, v1, T1, S1, Fz1, s);
    // the second convective pass needs only the predicted tracers (and the
    // convective buffers free again): side stream A, beside momentum residual #1
    after(A, s);
    zm::launch_conv(g, Pf, Tp, Sp, cT, cS, gate, ubits, A);
    // momentum residual #1 at (u1, v1, T_pred, S_pred)
    zm::launch_pbc(g, Pf, Tp, Sp, pbc, s);
    zm::launch_bt_rho(g, Pf, pbc, Frx, Fry, (const real*)nullptr, (const real*)nullptr, (real*)nullptr,
                      (real*)nullptr, s);
    after(s, B);
    zm::launch_mom_residual(g, Pf, 0, axu, axv, u1, v1, pbc, eta, Frx, Fry, du1, dv1, up, vp, decay_bot, s);
    // tracer residual #2 at (u_adv, v_adv, T_pred, S_pred) -> T_new, with
    // (u_adv, v_adv) the predictor velocity or its column-divergence projection
    const real* uad = up;
    const real* vad = vp;
    if (proj_on) {
      project(up, vp, ua, va, dt64, s);
      uad = ua;
      vad = va;
    }
    zm::launch_fz(g, uad, vad, Fz2, s);
    if (ledger_on && proj_on) {
      // surface transports of the predictor (into the free Fz1) and of its projection
      zm::launch_fz(g, up, vp, Fz1, s);
      zm::launch_ledger_transport(g, area64, (const real*)Fz1, (const real*)Fz2, led_part, led_raw, s);
    }
    after(s, A);
    zm::launch_tracer_residual(g, Pf, 1, uad, vad, Tp, Sp, Fz2, cT, cS, gate, dT1, dS1, T1, S1, To, So, s);
    if (ledger_on) ledger_tracer(Pf, uad, vad, Tp, Sp, Fz2, s);
    // momentum residual #2 at (u1, v1, T_new, S_new); advection of (u1, v1) is reused
    zm::launch_pbc(g, Pf, To, So, pbc, s);
    zm::launch_bt_rho(g, Pf, pbc, Frx, Fry, (const real*)nullptr, (const real*)nullptr, (real*)nullptr,
                      (real*)nullptr, s);
    zm::launch_mom_residual(g, Pf, 1, axu, axv, u1, v1, pbc, eta, Frx, Fry, du1, dv1, uo, vo, decay_bot, s);
  }

  // barotropic entry: F_rho and ubt0 from the post-L2 state
  void bt_prepare(const zm::Phys<real>& Pf, const real* eta_in, hipStream_t s) {
    zm::launch_pbc(g, Pf, T3, S3, pbc, s);
    zm::launch_bt_rho(g, Pf, pbc, Frx, Fry, u3, v3, ubt0, vbt0, s);
    zm::launch_copy(eta_in, etaw, n2, s);
    zm::launch_copy((const real*)ubt0, ubt, n2, s);
    zm::launch_copy((const real*)vbt0, vbt, n2, s);
  }

  // n_sub forward-backward substeps, one launch each, alternating between
  // (etaw, ubt, vbt) and (etaw2, ubt2, vbt2) so that the last substep lands in
  // the former; the first reads eta_in, ubt0, vbt0 directly
  void barotropic(const zm::Phys<real>& Pf, const real* eta_in, hipStream_t s) {
    zm::launch_pbc(g, Pf, T3, S3, pbc, s);
    zm::launch_bt_rho(g, Pf, pbc, Frx, Fry, u3, v3, ubt0, vbt0, s);
    if (bt_persist) {
      zm::launch_bt_persist(g, Pf, eta_in, ubt0, vbt0, etaw, ubt, vbt, etaw2, ubt2, vbt2, Frx, Fry, fdt, fiden, dtH,
                            dt_bt, drag, n_sub, bt_progress, bt_timeout, s);
      return;
    }
    const real *ei = eta_in, *ui = ubt0, *vi = vbt0;
    for (int q = 0; q < n_sub; ++q) {
      const bool last_half = ((n_sub - 1 - q) & 1) == 0;
      real* eo = last_half ? etaw : etaw2;
      real* uo = last_half ? ubt : ubt2;
      real* vo = last_half ? vbt : vbt2;
      zm::launch_bt_sub(g, Pf, ei, ui, vi, eo, uo, vo, Frx, Fry, fdt, fiden, dtH, dt_bt, drag, s);
      ei = eo;
      ui = uo;
      vi = vo;
    }
  }

  // the member record of a batched barotropic substep (frozen forcing)
  zm::BtMember<real> bt_member(int row0) const {
    zm::BtMember<real> m{};
    m.g = g;
    m.P = P;
    m.eta = etaw;
    m.ubt = ubt;
    m.vbt = vbt;
    m.Frx = Frx;
    m.Fry = Fry;
    m.fdt = fdt;
    m.fiden = fiden;
    m.dtH = dtH;
    m.dt_bt = dt_bt;
    m.drag = drag;
    m.row0 = row0;
    return m;
  }

  // everything before / after the barotropic substeps of one step
  void pre_barotropic(const zm_state& in, const zm::Phys<real>& Pf, hipStream_t s) {
    const real* ui = (const real*)in.u;
    const real* vi = (const real*)in.v;
    const real* Ti = (const real*)in.T;
    const real* Si = (const real*)in.S;
    linear(ui, vi, Ti, Si, u1, v1, T1, S1, s);
    nonlinear(Pf, (const real*)in.eta, u2, v2, T2, S2, s);
    linear(u2, v2, T2, S2, u3, v3, T3, S3, s);
  }

  void post_barotropic(const zm_state& in, const zm_state& out, const zm::Phys<real>& Pf, hipStream_t s) {
    zm::launch_cap_means(g, Pf, u3, v3, T3, S3, ubt, ubt0, vbt, vbt0, means, s);
    zm::launch_final(g, Pf, u3, v3, T3, S3, ubt, ubt0, vbt, vbt0, means, (const real*)in.u, (const real*)in.v,
                     (const real*)in.T, (const real*)in.S, (real*)out.u, (real*)out.v, (real*)out.T, (real*)out.S,
                     (const real*)etaw, (real*)out.eta, (const real*)in.ice,
                     out.ice != in.ice ? (real*)out.ice : (real*)nullptr, s);
  }

  void one_step(const zm_state& in, const zm_state& out, const zm::Phys<real>& Pf, hipStream_t s) {
    pre_barotropic(in, Pf, s);
    barotropic(Pf, (const real*)in.eta, s);
    post_barotropic(in, out, Pf, s);
  }

  void seasonal(const void* stack, int nfields, double day, double blend_days, void* out,
                hipStream_t s) override {
    // the 30-day calendar of interp_monthly_field_jit (inputs/forcing/seasonal.py:86-103)
    const double month_len = 30.0;
    double mpos = std::fmod(day, month_len);
    if (mpos < 0) mpos += month_len;
    long long mi = (long long)std::floor(day / month_len) % 12;
    if (mi < 0) mi += 12;
    const double half = blend_days / 2.0;
    double wa = 0.0, wb = 0.0;
    if (blend_days > 0.0) {
      wa = std::min(1.0, std::max(0.0, (mpos - (month_len - half)) / blend_days));
      wb = std::min(1.0, std::max(0.0, (half - mpos) / blend_days));
    }
    zm::launch_seasonal_blend((const real*)stack, nfields, n2, (int)mi, wa, wb, (real*)out, s);
    ZM_HIP(hipGetLastError());
  }

  void monitor(const zm_state& st, long long step, double vlim, double elim, int inclusive, void* status,
               hipStream_t s) override {
    zm::launch_monitor(g, (const real*)st.u, (const real*)st.v, (const real*)st.T, (const real*)st.S,
                       (const real*)st.eta, (const real*)st.ice, step, vlim, elim, inclusive, red_partials,
                       (zm::MonitorStatus*)status, s);
    ZM_HIP(hipGetLastError());
  }

  void budget(const zm_state& st, double* sums, hipStream_t s) override {
    zm::launch_budget(g, area64, dzn64, (const real*)st.T, (const real*)st.S, (const real*)st.eta, red_partials, sums,
                      s);
    ZM_HIP(hipGetLastError());
  }

  void set_tiles(int nu, int seg) override {
    if (nu < 1 || nu > zm::nu_block_max<real>()) throw ConfigError{"nu_block out of range"};
    if (seg < 1 || seg > nz) throw ConfigError{"column segment out of range"};
    nu_block = nu;
    g.seg = seg;
    clear_graphs();
  }

  void get_tiles(int* nu, int* seg) const override {
    *nu = nu_block;
    *seg = g.seg;
  }

  void set_bt_persistent(bool on) override {
    if (on && !bt_persist_fits)
      throw ConfigError{"the persistent barotropic sub-cycle needs (ny + 1) / 2 <= the device's CU count"};
    bt_persist = on;
    clear_graphs();
  }

  bool bt_persistent_on() const override { return bt_persist; }

  // waits of the persistent sub-cycle that gave up since the handle was made
  int bt_timeouts() override {
    int v = 0;
    ZM_HIP(hipStreamSynchronize(own));
    ZM_HIP(hipMemcpy(&v, bt_timeout, sizeof(int), hipMemcpyDeviceToHost));
    return v;
  }

  void clear_graphs() {
    for (auto& kv : graphs) {
      hipGraphExecDestroy(kv.second.second);
      hipGraphDestroy(kv.second.first);
    }
    graphs.clear();
  }

  zm::Phys<real> forced(const void* const* forcing) const {
    zm::Phys<real> Pf = P;
    if (forcing) {
      if (forcing[0]) Pf.tau_x = (const real*)forcing[0];
      if (forcing[1]) Pf.tau_y = (const real*)forcing[1];
      if (forcing[2]) Pf.q_heat = (const real*)forcing[2];
      if (forcing[3]) Pf.t_atm = (const real*)forcing[3];
    }
    return Pf;
  }

  void issue(const zm_state& in, const zm_state& out, int nsteps, const zm::Phys<real>& Pf, hipStream_t s) {
    one_step(in, out, Pf, s);
    for (int q = 1; q < nsteps; ++q) one_step(out, out, Pf, s);
  }

  void step(const zm_state& in, const zm_state& out, int nsteps, const void* const* forcing,
            hipStream_t s) override {
    const zm::Phys<real> Pf = forced(forcing);
    if (!use_graph) {
      issue(in, out, nsteps, Pf, s);
      ZM_HIP(hipGetLastError());
      return;
    }
    std::vector<const void*> key = {in.u, in.v, in.T, in.S, in.eta, in.ice,
                                    out.u, out.v, out.T, out.S, out.eta, out.ice,
                                    Pf.tau_x, Pf.tau_y, Pf.q_heat, Pf.t_atm};
    GraphKey gk{key, nsteps};
    auto it = graphs.find(gk);
    if (it == graphs.end()) {
      if (graphs.size() >= kGraphCache) {
        auto first = graphs.begin();
        ZM_HIP(hipStreamSynchronize(s));
        hipGraphExecDestroy(first->second.second);
        hipGraphDestroy(first->second.first);
        graphs.erase(first);
      }
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      ZM_HIP(hipStreamBeginCapture(own, hipStreamCaptureModeThreadLocal));
      issue(in, out, nsteps, Pf, own);
      hipError_t e = hipStreamEndCapture(own, &graph);
      if (e != hipSuccess) throw HipError{e, "hipStreamEndCapture"};
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        hipGraphDestroy(graph);
        throw HipError{e, "hipGraphInstantiate"};
      }
      it = graphs.emplace(gk, std::make_pair(graph, exec)).first;
    }
    ZM_HIP(hipGraphLaunch(it->second.second, s));
  }

  // _tracer_terms (tendencies.py:74-92) into terms[6][nz][ny][nx]: advection,
  // horizontal diffusion, vertical diffusion and single-pass convection of T
  // on wet nodes, then zero GM and Redi planes (neither closure is supported)
  void tracer_terms(const zm_state& in, void* terms, hipStream_t s) override {
    const real *u = (const real*)in.u, *v = (const real*)in.v, *T = (const real*)in.T, *S = (const real*)in.S;
    zm::launch_fz(g, u, v, Fz1, s);
    zm::Phys<real> Pc = P;
    Pc.conv_nsub = 1;
    zm::launch_conv(g, Pc, T, S, cT, cS, gate, ubits, s);
    real* o = (real*)terms;
    zm::launch_tracer_terms(g, P, u, v, T, S, Fz1, cT, cS, gate, o, o + n3, o + 2 * n3, o + 3 * n3, s);
    ZM_HIP(hipMemsetAsync(o + 4 * n3, 0, (size_t)2 * n3 * sizeof(real), s));
    ZM_HIP(hipGetLastError());
  }

  void diagnostics(const zm_state& in, void* rho, void* pres, void* w, hipStream_t s) override {
    zm::launch_pbc(g, P, (const real*)in.T, (const real*)in.S, pbc, s);
    zm::launch_fz(g, (const real*)in.u, (const real*)in.v, Fz1, s);
    zm::launch_diag(g, P, (const real*)in.T, (const real*)in.S, (const real*)in.eta, pbc, Fz1, (real*)rho,
                    (real*)pres, (real*)w, s);
    ZM_HIP(hipGetLastError());
  }

  void stage(int which, const zm_state& in, const zm_state& out, hipStream_t s) override {
    const real* ui = (const real*)in.u;
    const real* vi = (const real*)in.v;
    const real* Ti = (const real*)in.T;
    const real* Si = (const real*)in.S;
    const real* ei = (const real*)in.eta;
    switch (which) {
      case ZM_STAGE_LINEAR:
        linear(ui, vi, Ti, Si, (real*)out.u, (real*)out.v, (real*)out.T, (real*)out.S, s);
        break;
      case ZM_STAGE_NONLINEAR:
        zm::launch_copy(ui, u1, n3, s);
        zm::launch_copy(vi, v1, n3, s);
        zm::launch_copy(Ti, T1, n3, s);
        zm::launch_copy(Si, S1, n3, s);
        nonlinear(P, ei, (real*)out.u, (real*)out.v, (real*)out.T, (real*)out.S, s);
        break;
      case ZM_STAGE_BAROTROPIC:
        zm::launch_copy(ui, u3, n3, s);
        zm::launch_copy(vi, v3, n3, s);
        zm::launch_copy(Ti, T3, n3, s);
        zm::launch_copy(Si, S3, n3, s);
        barotropic(P, ei, s);
        zm::launch_cap_means(g, P, u3, v3, T3, S3, ubt, ubt0, vbt, vbt0, means, s);
        zm::launch_final(g, P, u3, v3, T3, S3, ubt, ubt0, vbt, vbt0, means, ui, vi, Ti, Si, (real*)out.u,
                         (real*)out.v, (real*)out.T, (real*)out.S, (const real*)etaw, (real*)out.eta,
                         (const real*)nullptr, (real*)nullptr, s);
        break;
      case ZM_STAGE_TRACER_RES:
        zm::launch_fz(g, ui, vi, Fz1, s);
        zm::launch_conv(g, P, Ti, Si, cT, cS, gate, ubits, s);
        zm::launch_tracer_residual(g, P, 0, ui, vi, Ti, Si, Fz1, cT, cS, gate, (real*)out.T, (real*)out.S, Ti,
                                   Si, tmpA, tmpB, s);
        break;
      case ZM_STAGE_MOMENTUM_RES:
        zm::launch_fz(g, ui, vi, Fz1, s);
        zm::launch_pbc(g, P, Ti, Si, pbc, s);
        zm::launch_bt_rho(g, P, pbc, Frx, Fry, (const real*)nullptr, (const real*)nullptr, (real*)nullptr,
                          (real*)nullptr, s);
        zm::launch_mom_adv(g, ui, vi, Fz1, au, av, s);
        dealias(au, av, axu, axv, s);
        zm::launch_mom_residual(g, P, 0, axu, axv, ui, vi, pbc, ei, Frx, Fry, (real*)out.u, (real*)out.v, tmpA,
                                tmpB, decay_bot, s);
        break;
      default:
        throw ConfigError{"unknown stage"};
    }
    ZM_HIP(hipGetLastError());
  }

  int nu_substeps_per_launch() const override { return nu_block; }

  // ---- column operators (implicit mixing, surrogate) ----------------------
  zm::MixParams mix(double dt, double kv, double kc) const {
    zm::MixParams m{};
    m.dt = dt;
    m.kappa_v = kv;
    m.kappa_conv = kc;
    m.T_ref = T_ref64;
    m.S_ref = S_ref64;
    m.dzn = dzn64;
    m.dzi = dzi64;
    m.idzn = idzn64;
    m.idzi = idzi64;
    return m;
  }

  void vertical_mix(const void* T, const void* S, void* To, void* So, double dt, double kv, double kc,
                    const int* only, hipStream_t s) override {
    if (!vmix_scratch) vmix_scratch = dev_alloc<double>((size_t)3 * n3);
    zm::launch_vmix(g, mix(dt, kv, kc), (const real*)T, (const real*)S, (real*)To, (real*)So, only, vmix_scratch, s);
    ZM_HIP(hipGetLastError());
  }

  void vertical_mix_static(const void* T, const void* S, void* To, void* So, double dt, double kv, double kc,
                           const int* only, hipStream_t s) override {
    zm::launch_vmix_static(g, mix(dt, kv, kc), (const real*)T, (const real*)S, (real*)To, (real*)So, only, s);
    ZM_HIP(hipGetLastError());
  }

  void column_features(const zm_state& in, const void* tau_x, const void* tau_y, const void* q_heat,
                       const float* mean, const float* inv_std, void* X, long long rows, hipStream_t s) override {
    zm::launch_col_features(g, mix(P.dt, 0.0, 0.0), (const real*)in.T, (const real*)in.S, (const real*)in.u,
                            (const real*)in.v, q_heat ? (const real*)q_heat : P.q_heat, P.s_ref2,
                            P.restore_coef_S, tau_x ? (const real*)tau_x : P.tau_x,
                            tau_y ? (const real*)tau_y : P.tau_y, mean, inv_std, X, rows, s);
    ZM_HIP(hipGetLastError());
  }

  void mlp_apply(const zm_state& in, const float* Y, int ldy, const float* omean, const float* ostd, double dt,
                 double kv, double kc, double tol_T, double tol_S, void* To, void* So, int* flags, int* count,
                 hipStream_t s) override {
    zm::launch_mlp_apply(g, mix(dt, kv, kc), (const real*)in.T, (const real*)in.S, Y, ldy, omean, ostd, tol_T, tol_S,
                         (real*)To, (real*)So, flags, count, s);
    ZM_HIP(hipGetLastError());
  }

  void mix_targets(const zm_state& in, const void* Tt, const void* St, double dt, double kv, double kc, float* Y,
                   int ldy, hipStream_t s) override {
    zm::launch_mix_targets(g, mix(dt, kv, kc), (const real*)in.T, (const real*)in.S, (const real*)Tt,
                           (const real*)St, Y, ldy, s);
    ZM_HIP(hipGetLastError());
  }

  void mlp_fused(const zm_state& in, const void* tau_x, const void* tau_y, const void* q_heat, const float* in_mean,
                 const float* in_inv_std, const void* const* weights, const float* const* biases, const float* omean,
                 const float* ostd, double dt, double kv, double kc, double tol_T, double tol_S, void* To, void* So,
                 int* flags, int* count, hipStream_t s) override {
    if constexpr (std::is_same_v<real, float>) {
      zm::launch_mlp_fused(g, mix(dt, kv, kc), (const float*)in.T, (const float*)in.S, (const float*)in.u,
                           (const float*)in.v, q_heat ? (const float*)q_heat : P.q_heat, P.s_ref2, P.restore_coef_S,
                           tau_x ? (const float*)tau_x : P.tau_x, tau_y ? (const float*)tau_y : P.tau_y, in_mean,
                           in_inv_std, weights, biases, omean, ostd, tol_T, tol_S, (float*)To, (float*)So, flags,
                           count, s);
      ZM_HIP(hipGetLastError());
    } else {
      throw ConfigError{"the fused surrogate step runs on float32 handles"};
    }
  }

  double time_kernel(int which, int reps) override {
    hipEvent_t a, b;
    ZM_HIP(hipEventCreate(&a));
    ZM_HIP(hipEventCreate(&b));
    hipStream_t s = own;
    auto body = [&](int r) {
      switch (which) {
        case 0:
          zm::launch_nu_substep(g, (r % 2) ? nBu : nAu, (r % 2) ? nBv : nAv, (r % 2) ? nAu : nBu,
                                (r % 2) ? nAv : nBv, (real*)nullptr, (real*)nullptr, P.nu_h, dts, s);
          break;
        case 1:
          zm::launch_tracer_residual(g, P, 0, u1, v1, T1, S1, Fz1, cT, cS, gate, dT1, dS1, T1, S1, Tp, Sp, s);
          break;
        case 2:
          zm::launch_mom_residual(g, P, 0, axu, axv, u1, v1, pbc, etaw, Frx, Fry, du1, dv1, up, vp, decay_bot,
                                  s);
          break;
        case 3:
          dealias(au, av, axu, axv, s);
          break;
        case 5:
          zm::launch_nu_block(g, (r % 2) ? nBu : nAu, (r % 2) ? nBv : nAv, (r % 2) ? nAu : nBu,
                              (r % 2) ? nAv : nBv, P.nu_h, dts, std::max(1, nu_block), s);
          break;
        case 6:
          zm::launch_copy((const real*)((r % 2) ? tmpB : tmpA), (r % 2) ? tmpA : tmpB, n3, s);
          break;
        case 7:
          zm::launch_linear_finish(g, P, u1, v1, T1, S1, nAu, nAv, Lu, Lv, DT, DS, u3, v3, T3, S3, s);
          break;
        case 8:
          zm::launch_conv(g, P, T1, S1, cT, cS, gate, ubits, s);
          break;
        case 9:
          zm::launch_fz(g, u1, v1, Fz1, s);
          break;
        case 10:
          zm::launch_d1_unit(g, T1, S1, DT, DS, s);
          break;
        case 11:
          zm::launch_mom_adv(g, u1, v1, Fz1, au, av, s);
          break;
        case 12:
          zm::launch_final(g, P, u3, v3, T3, S3, ubt, ubt0, vbt, vbt0, means, u1, v1, T1, S1, tmpA, tmpB, u2, v2,
                           (const real*)etaw, etaw2, (const real*)nullptr, (real*)nullptr, s);
          break;
        case 4:
          zm::launch_bt_sub(g, P, etaw, ubt, vbt, etaw2, ubt2, vbt2, Frx, Fry, fdt, fiden, dtH, dt_bt, drag, s);
          break;
        case 15:  // the whole persistent sub-cycle (n_sub substeps)
          if (!bt_persist_fits) throw ConfigError{"the persistent sub-cycle does not fit this device"};
          zm::launch_bt_persist(g, P, (r % 2) ? etaw2 : etaw, (r % 2) ? ubt2 : ubt, (r % 2) ? vbt2 : vbt,
                                (r % 2) ? etaw : etaw2, (r % 2) ? ubt : ubt2, (r % 2) ? vbt : vbt2, (real*)tmpA,
                                (real*)tmpA + n2, (real*)tmpB, Frx, Fry, fdt, fiden, dtH, dt_bt, drag, n_sub,
                                bt_progress, bt_timeout, s);
          break;
        case 14:
          zm::launch_dealias_x(nx, N1, N2, kept, (int)(ny * nz), twr, twi, au, av, axu, axv, s);
          break;
        default:
          throw ConfigError{"unknown kernel selector"};
      }
    };
    body(0);
    ZM_HIP(hipEventRecord(a, s));
    for (int r = 0; r < reps; ++r) body(r);
    ZM_HIP(hipEventRecord(b, s));
    ZM_HIP(hipEventSynchronize(b));
    float ms = 0;
    ZM_HIP(hipEventElapsedTime(&ms, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    ZM_HIP(hipGetLastError());
    return (double)ms / reps;
  }
};

void validate(const zm_config& c) {
  auto bad = [](const std::string& m) { throw ConfigError{m}; };
  if (c.nx < 2 || c.ny < 2 || c.nz < 2) bad("nx, ny, nz must be >= 2");
  if (c.nx > 1024) bad("nx > 1024 is not supported by the row kernels");
  if (c.dtype_bytes != 4 && c.dtype_bytes != 8) bad("dtype must be float32 or float64");
  if (!(c.dt > 0) || !(c.dt_bt > 0)) bad("dt and dt_bt must be positive");
  if (c.n_subcyc < 1 || c.n_nu < 1 || c.conv_nsub < 1) bad("substep counts must be >= 1");
  if (c.adv_nsub != 1) bad("adv_nsub > 1 (vertical advection sub-cycling) is not supported");
  if ((long long)4 * c.nz * 64 * c.dtype_bytes > 160 * 1024)
    bad("too many levels for the convective column kernel (nz <= 160 in float32, 80 in float64)");
  if (c.localize_conv && c.nz > 129) bad("localize_conv supports at most 129 levels");
  if (c.project_adv_vel) {
    if (c.projection_niter < 1) bad("projection_niter must be a positive integer");
    if (c.projection_max_refinements < 0 || c.projection_max_refinements > 2)
      bad("projection_max_refinements must be an integer from 0 to 2");
    if (c.projection_preconditioner != 0 && c.projection_preconditioner != 1)
      bad("projection_preconditioner must be none or jacobi");
    if (!(c.projection_rtol > 0) || !std::isfinite(c.projection_rtol)) bad("projection_rtol must be finite and positive");
    if (!(c.projection_transport_rtol > 0) || !std::isfinite(c.projection_transport_rtol))
      bad("projection_transport_rtol must be finite and positive");
  }
  if (c.polar_cap_rows < 0 || c.polar_cap_taper < 0) bad("polar cap rows must be >= 0");
  if (c.polar_cap_rows > 0 && 2 * (c.polar_cap_rows + c.polar_cap_taper) > c.ny)
    bad("polar cap bands must not overlap");
  const void* req[] = {c.cos_lat, c.dx_row, c.f_row, c.dz_node, c.dz_iface, c.dz, c.dz_norm, c.nwet, c.wet_mask,
                       c.dealias_mask};
  for (const void* p : req)
    if (!p) bad("missing required array");
  if (c.nz > 2 && (!c.d2z_hm || !c.d2z_hp || !c.d2z_denom || !c.dz_denom_interior)) bad("missing vertical arrays");
  if (c.polar_cap_rows > 0 && !c.cap_weights) bad("missing polar cap weights");
  const long long n2 = (long long)c.nx * c.ny;
  for (long long q = 0; q < n2; ++q)
    if (c.nwet[q] < 0 || c.nwet[q] > c.nz) bad("nwet out of range");
  for (int j = 0; j < c.ny; ++j)
    if (!(c.dx_row[j] > 0) || !(c.cos_lat[j] > 0)) bad("dx and cos(lat) must be positive");
}

int fail(zm_handle* h, const std::string& m, int code);

int zm_bt_row_max() { return 1024; }

}  // namespace

struct zm_handle {
  SolverBase* solver = nullptr;
  int dtype_bytes = 0;
  std::string err;
};

// Uneven float32 members advanced together: the 3-D stages of the members on
// kEnsembleLanes side streams (member i on lane i mod kEnsembleLanes, so the
// kernels of different members can run at once; captured as parallel graph
// branches), the barotropic sub-cycle as one batched launch per substep over
// a member table on the main stream, then the final masks of the members on
// the side streams again; the whole ensemble step is one hipGraph launch.
constexpr int kEnsembleLanes = 4;

struct zm_ensemble {
  std::vector<Solver<float>*> mem;
  zm::BtMember<float>* table = nullptr;
  int total_rows = 0, n_sub = 0;
  hipStream_t own = nullptr;
  hipStream_t lane[kEnsembleLanes] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kEnsembleLanes] = {};
  bool use_graph = true;
  std::map<std::vector<const void*>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
  std::string err;

  ~zm_ensemble() {
    for (auto& kv : graphs) {
      hipGraphExecDestroy(kv.second.second);
      hipGraphDestroy(kv.second.first);
    }
    for (int k = 0; k < kEnsembleLanes; ++k) {
      if (lane[k]) hipStreamDestroy(lane[k]);
      if (join_ev[k]) hipEventDestroy(join_ev[k]);
    }
    if (fork_ev) hipEventDestroy(fork_ev);
    if (own) hipStreamDestroy(own);
    if (table) hipFree(table);
  }

  void make_lanes() {
    for (int k = 0; k < kEnsembleLanes; ++k) {
      ZM_HIP(hipStreamCreateWithFlags(&lane[k], hipStreamNonBlocking));
      ZM_HIP(hipEventCreateWithFlags(&join_ev[k], hipEventDisableTiming));
    }
    ZM_HIP(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
  }

  // the side streams start after everything issued on s so far
  void fork(hipStream_t s) {
    ZM_HIP(hipEventRecord(fork_ev, s));
    for (int k = 0; k < kEnsembleLanes; ++k) ZM_HIP(hipStreamWaitEvent(lane[k], fork_ev, 0));
  }

  // s continues after everything issued on the side streams
  void join(hipStream_t s) {
    for (int k = 0; k < kEnsembleLanes; ++k) {
      ZM_HIP(hipEventRecord(join_ev[k], lane[k]));
      ZM_HIP(hipStreamWaitEvent(s, join_ev[k], 0));
    }
  }

  void issue(const zm_state* in, const zm_state* out, hipStream_t s) {
    const int n = (int)mem.size();
    // Each member's stages stay on its lane: forking them further inside the
    // ensemble graph (stage streams nested in lane streams) ended the first
    // capture of a 24-member ensemble in a segmentation fault in the runtime.
    struct InStream {
      std::vector<Solver<float>*>& m;
      explicit InStream(std::vector<Solver<float>*>& v) : m(v) {
        for (auto* x : m) x->fork_stages = false;
      }
      ~InStream() {
        for (auto* x : m) x->fork_stages = true;
      }
    } in_stream(mem);
    fork(s);
    for (int i = 0; i < n; ++i) {
      hipStream_t ls = lane[i % kEnsembleLanes];
      mem[i]->pre_barotropic(in[i], mem[i]->P, ls);
      mem[i]->bt_prepare(mem[i]->P, (const float*)in[i].eta, ls);
    }
    join(s);
    for (int q = 0; q < n_sub; ++q) zm::launch_bt_multi(table, n, total_rows, s);
    fork(s);
    for (int i = 0; i < n; ++i) mem[i]->post_barotropic(in[i], out[i], mem[i]->P, lane[i % kEnsembleLanes]);
    join(s);
  }

  void step(const zm_state* in, const zm_state* out, hipStream_t s) {
    const int n = (int)mem.size();
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        if (i != j && (out[i].u == in[j].u || out[i].u == out[j].u))
          throw ConfigError{"ensemble members must not share state buffers"};
    if (!use_graph) {
      issue(in, out, s);
      ZM_HIP(hipGetLastError());
      return;
    }
    std::vector<const void*> key;
    for (int i = 0; i < n; ++i) {
      const void* p[] = {in[i].u, in[i].v, in[i].T, in[i].S, in[i].eta, in[i].ice,
                         out[i].u, out[i].v, out[i].T, out[i].S, out[i].eta, out[i].ice};
      key.insert(key.end(), p, p + 12);
    }
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      if (graphs.size() >= 4) {
        auto first = graphs.begin();
        ZM_HIP(hipStreamSynchronize(s));
        hipGraphExecDestroy(first->second.second);
        hipGraphDestroy(first->second.first);
        graphs.erase(first);
      }
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      ZM_HIP(hipStreamBeginCapture(own, hipStreamCaptureModeThreadLocal));
      issue(in, out, own);
      hipError_t e = hipStreamEndCapture(own, &graph);
      if (e != hipSuccess) throw HipError{e, "hipStreamEndCapture"};
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        hipGraphDestroy(graph);
        throw HipError{e, "hipGraphInstantiate"};
      }
      it = graphs.emplace(key, std::make_pair(graph, exec)).first;
    }
    ZM_HIP(hipGraphLaunch(it->second.second, s));
  }
};

namespace {
int fail(zm_handle* h, const std::string& m, int code) {
  if (h)
    h->err = m;
  else
    g_thread_error = m;
  return code;
}

template <typename F>
int guarded(zm_handle* h, F&& f) {
  try {
    f();
    return ZM_OK;
  } catch (const ConfigError& e) {
    return fail(h, e.msg, ZM_EINVAL);
  } catch (const HipError& e) {
    return fail(h, e.where + ": " + hipGetErrorString(e.code),
                e.code == hipErrorOutOfMemory ? ZM_ENOMEM : ZM_EHIP);
  } catch (const std::exception& e) {
    return fail(h, e.what(), ZM_EHIP);
  }
}
}  // namespace

extern "C" {

int32_t zm_version(void) { return 1; }

int zm_create(const zm_config* cfg, zm_handle** out) {
  if (!cfg || !out) return fail(nullptr, "null argument", ZM_EINVAL);
  *out = nullptr;
  zm_handle* h = new zm_handle();
  int rc = guarded(nullptr, [&] {
    validate(*cfg);
    if (cfg->dtype_bytes == 4)
      h->solver = new Solver<float>(*cfg);
    else
      h->solver = new Solver<double>(*cfg);
    h->dtype_bytes = cfg->dtype_bytes;
  });
  if (rc != ZM_OK) {
    delete h->solver;
    delete h;
    return rc;
  }
  *out = h;
  return ZM_OK;
}

int zm_destroy(zm_handle* h) {
  if (!h) return ZM_OK;
  delete h->solver;
  delete h;
  return ZM_OK;
}

const char* zm_last_error(const zm_handle* h) { return h ? h->err.c_str() : g_thread_error.c_str(); }

int zm_step(zm_handle* h, const zm_state* in, const zm_state* out, int32_t nsteps, void* stream) {
  if (!h || !in || !out || nsteps < 1) return fail(h, "invalid zm_step arguments", ZM_EINVAL);
  return guarded(h, [&] { h->solver->step(*in, *out, nsteps, nullptr, (hipStream_t)stream); });
}

int zm_step_dyn(zm_handle* h, const zm_state* in, const zm_state* out, const void* tau_x, const void* tau_y,
                const void* q_heat, const void* t_atm, void* stream) {
  if (!h || !in || !out || !tau_x || !tau_y || !q_heat) return fail(h, "invalid zm_step_dyn arguments", ZM_EINVAL);
  const void* forcing[4] = {tau_x, tau_y, q_heat, t_atm};
  return guarded(h, [&] { h->solver->step(*in, *out, 1, forcing, (hipStream_t)stream); });
}

int zm_diagnostics(zm_handle* h, const zm_state* in, void* rho, void* pressure, void* w, void* stream) {
  if (!h || !in || !rho || !pressure || !w) return fail(h, "invalid zm_diagnostics arguments", ZM_EINVAL);
  return guarded(h, [&] { h->solver->diagnostics(*in, rho, pressure, w, (hipStream_t)stream); });
}

int zm_tracer_terms(zm_handle* h, const zm_state* in, void* terms, void* stream) {
  if (!h || !in || !terms) return fail(h, "invalid zm_tracer_terms arguments", ZM_EINVAL);
  return guarded(h, [&] { h->solver->tracer_terms(*in, terms, (hipStream_t)stream); });
}

int zm_stage(zm_handle* h, int32_t which, const zm_state* in, const zm_state* out, void* stream) {
  if (!h || !in || !out) return fail(h, "invalid zm_stage arguments", ZM_EINVAL);
  return guarded(h, [&] { h->solver->stage(which, *in, *out, (hipStream_t)stream); });
}

int zm_set_graph(zm_handle* h, int32_t enabled) {
  if (!h) return fail(h, "null handle", ZM_EINVAL);
  h->solver->use_graph = enabled != 0;
  return ZM_OK;
}

int32_t zm_nu_substeps_per_launch(const zm_handle* h) { return h ? h->solver->nu_substeps_per_launch() : 0; }

int zm_vertical_mix(zm_handle* h, const void* T, const void* S, void* T_out, void* S_out, double dt, double kappa_v,
                    double kappa_conv, const int32_t* only_columns, void* stream) {
  if (!h || !T || !S || !T_out || !S_out || !(dt >= 0.0) || !(kappa_v >= 0.0) || !(kappa_conv >= 0.0))
    return fail(h, "invalid zm_vertical_mix arguments", ZM_EINVAL);
  if (T == T_out || S == S_out) return fail(h, "zm_vertical_mix output must not alias its input", ZM_EINVAL);
  return guarded(h, [&] {
    h->solver->vertical_mix(T, S, T_out, S_out, dt, kappa_v, kappa_conv, only_columns, (hipStream_t)stream);
  });
}

int zm_vertical_mix_noalloc(zm_handle* h, const void* T, const void* S, void* T_out, void* S_out, double dt,
                            double kappa_v, double kappa_conv, const int32_t* only_columns, void* stream) {
  if (!h || !T || !S || !T_out || !S_out || !(dt >= 0.0) || !(kappa_v >= 0.0) || !(kappa_conv >= 0.0))
    return fail(h, "invalid zm_vertical_mix_noalloc arguments", ZM_EINVAL);
  if (T == T_out || S == S_out) return fail(h, "zm_vertical_mix_noalloc output must not alias its input", ZM_EINVAL);
  if (zm::static_vmix_columns(h->solver->nz_) < 256)
    return fail(h, "zm_vertical_mix_noalloc: nz too large for the 32 MiB static scratch (3 nz x 256 float64 "
                   "per column range)", ZM_EINVAL);
  return guarded(h, [&] {
    h->solver->vertical_mix_static(T, S, T_out, S_out, dt, kappa_v, kappa_conv, only_columns, (hipStream_t)stream);
  });
}

int zm_column_features(zm_handle* h, const zm_state* in, const void* tau_x, const void* tau_y, const void* q_heat,
                       const float* mean, const float* inv_std, void* X, int64_t rows, void* stream) {
  if (!h || !in || !mean || !inv_std || !X) return fail(h, "invalid zm_column_features arguments", ZM_EINVAL);
  if (5 * h->solver->nz_ + 3 > 256) return fail(h, "column features need nz <= 50", ZM_EINVAL);
  if (rows < h->solver->n2_) return fail(h, "feature rows must cover every column", ZM_EINVAL);
  return guarded(h, [&] {
    h->solver->column_features(*in, tau_x, tau_y, q_heat, mean, inv_std, X, rows, (hipStream_t)stream);
  });
}

int zm_mlp_apply(zm_handle* h, const zm_state* in, const float* Y, int32_t ldy, const float* out_mean,
                 const float* out_std, double dt, double kappa_v, double kappa_conv, double tol_T, double tol_S,
                 void* T_out, void* S_out, int32_t* flags, int32_t* n_flagged, void* stream) {
  if (!h || !in || !Y || !out_mean || !out_std || !T_out || !S_out || !flags || !n_flagged)
    return fail(h, "invalid zm_mlp_apply arguments", ZM_EINVAL);
  if (ldy < 2 * h->solver->nz_) return fail(h, "ldy must hold 2 nz outputs per column", ZM_EINVAL);
  if (h->solver->nz_ > 50) return fail(h, "the surrogate apply step takes at most 50 levels", ZM_EINVAL);
  if (T_out == in->T || S_out == in->S) return fail(h, "zm_mlp_apply output must not alias its input", ZM_EINVAL);
  return guarded(h, [&] {
    h->solver->mlp_apply(*in, Y, ldy, out_mean, out_std, dt, kappa_v, kappa_conv, tol_T, tol_S, T_out, S_out,
                         flags, n_flagged, (hipStream_t)stream);
  });
}

int zm_mix_targets(zm_handle* h, const zm_state* in, const void* T_thomas, const void* S_thomas, double dt,
                   double kappa_v, double kappa_conv, float* Y, int32_t ldy, void* stream) {
  if (!h || !in || !T_thomas || !S_thomas || !Y || !(dt >= 0.0) || !(kappa_v >= 0.0) || !(kappa_conv >= 0.0))
    return fail(h, "invalid zm_mix_targets arguments", ZM_EINVAL);
  if (ldy < 2 * h->solver->nz_) return fail(h, "ldy must hold 2 nz targets per column", ZM_EINVAL);
  return guarded(h, [&] {
    h->solver->mix_targets(*in, T_thomas, S_thomas, dt, kappa_v, kappa_conv, Y, ldy, (hipStream_t)stream);
  });
}

int zm_mlp_mix(zm_handle* h, const zm_state* in, const void* tau_x, const void* tau_y, const void* q_heat,
               const float* in_mean, const float* in_inv_std, const void* const* weights, const float* const* biases,
               const float* out_mean, const float* out_std, double dt, double kappa_v, double kappa_conv,
               double tol_T, double tol_S, void* T_out, void* S_out, int32_t* flags, int32_t* n_flagged,
               void* stream) {
  if (!h || !in || !in_mean || !in_inv_std || !weights || !biases || !out_mean || !out_std || !T_out || !S_out ||
      !flags || !n_flagged || !(dt >= 0.0) || !(kappa_v >= 0.0) || !(kappa_conv >= 0.0))
    return fail(h, "invalid zm_mlp_mix arguments", ZM_EINVAL);
  for (int l = 0; l < 4; ++l)
    if (!weights[l] || !biases[l]) return fail(h, "zm_mlp_mix needs the four layers' weights and biases", ZM_EINVAL);
  if (h->dtype_bytes != 4) return fail(h, "zm_mlp_mix runs on float32 handles", ZM_EINVAL);
  if (5 * h->solver->nz_ + 3 > 256) return fail(h, "the surrogate takes at most 50 levels", ZM_EINVAL);
  if (T_out == in->T || S_out == in->S) return fail(h, "zm_mlp_mix output must not alias its input", ZM_EINVAL);
  return guarded(h, [&] {
    h->solver->mlp_fused(*in, tau_x, tau_y, q_heat, in_mean, in_inv_std, weights, biases, out_mean, out_std, dt,
                         kappa_v, kappa_conv, tol_T, tol_S, T_out, S_out, flags, n_flagged, (hipStream_t)stream);
  });
}

int zm_gemm_bf16_nt(int64_t M, int64_t N, int64_t K, const void* A, const void* B, int32_t epilogue,
                    const float* bias, const float* aux, float* C_f32, void* C_bf16, void* CT_bf16, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0 || M % 128 || N % 128 || K % 32 || !A || !B)
    return fail(nullptr, "zm_gemm_bf16_nt: M, N must be positive multiples of 128 and K of 32", ZM_EINVAL);
  const bool ok = (epilogue == zm::kEpiF32 && C_f32) || (epilogue == zm::kEpiBiasF32 && C_f32 && bias) ||
                  (epilogue == zm::kEpiBiasGelu && bias && C_bf16) || (epilogue == zm::kEpiGeluGrad && aux && C_bf16);
  if (!ok) return fail(nullptr, "zm_gemm_bf16_nt: missing operand for the epilogue", ZM_EINVAL);
  return guarded(nullptr, [&] {
    zm::launch_gemm_bf16_nt(M, N, K, A, B, epilogue, bias, aux, C_f32, C_bf16, CT_bf16, (hipStream_t)stream);
    ZM_HIP(hipGetLastError());
  });
}

int zm_mse_grad(int64_t rows, int32_t ld, int32_t n_out, const float* Y, const float* target, void* dY, void* dYT,
                float* loss_sum, void* stream) {
  if (rows <= 0 || ld <= 0 || n_out <= 0 || n_out > ld || !Y || !target || !dY || !dYT || !loss_sum)
    return fail(nullptr, "invalid zm_mse_grad arguments", ZM_EINVAL);
  return guarded(nullptr, [&] {
    zm::launch_mse_grad(rows, ld, n_out, Y, target, dY, dYT, loss_sum, (hipStream_t)stream);
    ZM_HIP(hipGetLastError());
  });
}

int zm_rowsum_bf16(int64_t rows, int64_t cols, const void* in, float* out, void* stream) {
  if (rows <= 0 || cols <= 0 || !in || !out) return fail(nullptr, "invalid zm_rowsum_bf16 arguments", ZM_EINVAL);
  return guarded(nullptr, [&] {
    zm::launch_rowsum_bf16(rows, cols, in, out, (hipStream_t)stream);
    ZM_HIP(hipGetLastError());
  });
}

int zm_transpose_bf16(int64_t rows, int64_t cols, const void* in, void* out, void* stream) {
  if (rows <= 0 || cols <= 0 || !in || !out || in == out)
    return fail(nullptr, "invalid zm_transpose_bf16 arguments", ZM_EINVAL);
  return guarded(nullptr, [&] {
    zm::launch_transpose_bf16(rows, cols, in, out, (hipStream_t)stream);
    ZM_HIP(hipGetLastError());
  });
}

int zm_adam(int32_t n_out, int32_t n_in, float* w, const float* gw, float* m, float* v, void* w_bf, void* wT_bf,
            float* b, const float* gb, float* bm, float* bv, float lr, float beta1, float beta2, float eps, int32_t t,
            void* stream) {
  if (n_out <= 0 || n_in <= 0 || !w || !gw || !m || !v || !w_bf || !wT_bf || t < 1 ||
      (b && (!gb || !bm || !bv)))
    return fail(nullptr, "invalid zm_adam arguments", ZM_EINVAL);
  const float c1 = 1.0f - std::pow(beta1, (float)t), c2 = 1.0f - std::pow(beta2, (float)t);
  return guarded(nullptr, [&] {
    zm::launch_adam(n_out, n_in, w, gw, m, v, w_bf, wT_bf, b, gb, bm, bv, lr, beta1, beta2, eps, c1, c2,
                    (hipStream_t)stream);
    ZM_HIP(hipGetLastError());
  });
}

int zm_time_kernel(zm_handle* h, int32_t which, int32_t reps, double* ms_per_launch) {
  if (!h || reps < 1 || !ms_per_launch) return fail(h, "invalid zm_time_kernel arguments", ZM_EINVAL);
  return guarded(h, [&] { *ms_per_launch = h->solver->time_kernel(which, reps); });
}

int zm_set_tiles(zm_handle* h, int32_t nu_block, int32_t seg) {
  if (!h) return fail(h, "null handle", ZM_EINVAL);
  return guarded(h, [&] { h->solver->set_tiles(nu_block, seg); });
}

int zm_set_barotropic_persistent(zm_handle* h, int32_t enabled) {
  if (!h) return fail(h, "null handle", ZM_EINVAL);
  return guarded(h, [&] { h->solver->set_bt_persistent(enabled != 0); });
}

int zm_barotropic_timeouts(zm_handle* h, int32_t* count) {
  if (!h || !count) return fail(h, "invalid zm_barotropic_timeouts arguments", ZM_EINVAL);
  return guarded(h, [&] { *count = h->solver->bt_timeouts(); });
}

int zm_get_tiles(const zm_handle* h, int32_t* nu_block, int32_t* seg) {
  if (!h || !nu_block || !seg) return fail(nullptr, "invalid zm_get_tiles arguments", ZM_EINVAL);
  int a = 0, b = 0;
  h->solver->get_tiles(&a, &b);
  *nu_block = a;
  *seg = b;
  return ZM_OK;
}

// The autotuner times eagerly issued steps (no graph capture: an earlier
// version that captured a step graph for every candidate on a fresh stream
// aborted once without a message, cause not identified).
constexpr bool kAutotuneEnabled = true;

int zm_autotune(zm_handle* h, const zm_state* sample, int32_t reps, int32_t* nu_block, int32_t* seg,
                double* best_ms) {
  if (!h || !sample || reps < 1 || !nu_block || !seg || !best_ms)
    return fail(h, "invalid zm_autotune arguments", ZM_EINVAL);
  if (!kAutotuneEnabled) return fail(h, "zm_autotune is switched off: set launch shapes with zm_set_tiles", ZM_EINVAL);
  return guarded(h, [&] {
    SolverBase* sv = h->solver;
    const size_t w = (size_t)h->dtype_bytes;
    const size_t b3 = (size_t)sv->n2_ * sv->nz_ * w, b2 = (size_t)sv->n2_ * w;
    // two scratch states for ping-pong timing, freed before returning
    std::vector<void*> bufs;
    auto alloc = [&](size_t bytes) {
      void* p = nullptr;
      hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess) {
        for (void* q : bufs) hipFree(q);
        throw HipError{e, "hipMalloc (autotune scratch)"};
      }
      bufs.push_back(p);
      return p;
    };
    zm_state st[2];
    for (auto& x : st) {
      x.u = alloc(b3);
      x.v = alloc(b3);
      x.T = alloc(b3);
      x.S = alloc(b3);
      x.eta = alloc(b2);
      x.ice = alloc(b2);
    }
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ZM_HIP(hipStreamCreate(&s));
    ZM_HIP(hipEventCreate(&e0));
    ZM_HIP(hipEventCreate(&e1));
    int best_nu = 1, best_seg = 1;
    double best = 1e30;
    std::vector<int> segs;
    for (int c : {4, 5, 7, 10, 13, 17, 25, 50})
      if (c <= sv->nz_) segs.push_back(c);
    if (segs.empty() || segs.back() != sv->nz_) segs.push_back(sv->nz_);
    int nu_max = 1, cur_seg = 1;
    sv->get_tiles(&nu_max, &cur_seg);
    const bool graph = sv->use_graph;
    sv->use_graph = false;  // every candidate's steps are issued eagerly
    try {
      for (int nu = 1; nu <= nu_max; ++nu) {
        for (int sg : segs) {
          sv->set_tiles(nu, sg);
          const void* src[] = {sample->u, sample->v, sample->T, sample->S, sample->eta, sample->ice};
          void* dst[] = {st[0].u, st[0].v, st[0].T, st[0].S, st[0].eta, st[0].ice};
          for (int f = 0; f < 6; ++f)
            ZM_HIP(hipMemcpyAsync(dst[f], src[f], f < 4 ? b3 : b2, hipMemcpyDeviceToDevice, s));
          sv->step(st[0], st[1], 1, nullptr, s);
          sv->step(st[1], st[0], 1, nullptr, s);
          ZM_HIP(hipEventRecord(e0, s));
          for (int r = 0; r < reps; ++r) sv->step(st[r % 2], st[(r + 1) % 2], 1, nullptr, s);
          ZM_HIP(hipEventRecord(e1, s));
          ZM_HIP(hipEventSynchronize(e1));
          float ms = 0;
          ZM_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (ms / reps < best) {
            best = ms / reps;
            best_nu = nu;
            best_seg = sg;
          }
        }
      }
    } catch (...) {
      sv->use_graph = graph;
      sv->set_tiles(nu_max, cur_seg);
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      hipStreamDestroy(s);
      for (void* q : bufs) hipFree(q);
      throw;
    }
    sv->use_graph = graph;
    sv->set_tiles(best_nu, best_seg);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(s);
    for (void* q : bufs) hipFree(q);
    *nu_block = best_nu;
    *seg = best_seg;
    *best_ms = best;
  });
}

int zm_seasonal_forcing(zm_handle* h, const void* stack, int32_t nfields, double day, double blend_days, void* out,
                        void* stream) {
  if (!h || !stack || !out || nfields < 1 || !std::isfinite(day) || !std::isfinite(blend_days) ||
      blend_days > 30.0)
    return fail(h, "invalid zm_seasonal_forcing arguments", ZM_EINVAL);
  return guarded(h, [&] { h->solver->seasonal(stack, nfields, day, blend_days, out, (hipStream_t)stream); });
}

int zm_monitor_reset(zm_handle* h, zm_monitor_status* status, void* stream) {
  if (!h || !status) return fail(h, "invalid zm_monitor_reset arguments", ZM_EINVAL);
  return guarded(h, [&] {
    zm::launch_monitor_reset((zm::MonitorStatus*)status, (hipStream_t)stream);
    ZM_HIP(hipGetLastError());
  });
}

int zm_monitor(zm_handle* h, const zm_state* st, int64_t step, double velocity_limit, double eta_limit,
               int32_t inclusive, zm_monitor_status* status, void* stream) {
  if (!h || !st || !status || step < 0) return fail(h, "invalid zm_monitor arguments", ZM_EINVAL);
  return guarded(h, [&] {
    h->solver->monitor(*st, step, velocity_limit, eta_limit, inclusive, status, (hipStream_t)stream);
  });
}

int zm_budget(zm_handle* h, const zm_state* st, double* sums, void* stream) {
  if (!h || !st || !sums) return fail(h, "invalid zm_budget arguments", ZM_EINVAL);
  return guarded(h, [&] { h->solver->budget(*st, sums, (hipStream_t)stream); });
}

int zm_budget_step(zm_handle* h, const zm_state* in, const zm_state* out, const void* tau_x, const void* tau_y,
                   const void* q_heat, const void* t_atm, double* ledger, void* stream) {
  if (!h || !in || !out || !ledger) return fail(h, "invalid zm_budget_step arguments", ZM_EINVAL);
  const void* forcing[4] = {tau_x, tau_y, q_heat, t_atm};
  return guarded(h, [&] { h->solver->budget_step(*in, *out, forcing, ledger, (hipStream_t)stream); });
}

int zm_project_velocity(zm_handle* h, const void* u, const void* v, void* u_out, void* v_out, double dt,
                        int32_t* iterations, void* stream) {
  if (!h || !u || !v || !u_out || !v_out || !(dt > 0.0) || !std::isfinite(dt))
    return fail(h, "invalid zm_project_velocity arguments", ZM_EINVAL);
  return guarded(h, [&] {
    h->solver->project_velocity(u, v, u_out, v_out, dt, iterations, (hipStream_t)stream);
  });
}

int zm_projection_operator(const zm_config* cfg, double* taps, double* diagonal) {
  if (!cfg || !cfg->nwet || !cfg->wet_mask || !cfg->dx_row || !cfg->cos_lat || !cfg->dz_node || cfg->nx < 2 ||
      cfg->ny < 2 || cfg->nz < 1 || !(cfg->dy > 0))
    return fail(nullptr, "invalid zm_projection_operator arguments", ZM_EINVAL);
  return guarded(nullptr, [&] {
    const long long n2 = (long long)cfg->nx * cfg->ny;
    std::vector<double> t((size_t)zm::kProjTaps * n2), d(n2);
    zm::proj_stencil_host(cfg->nx, cfg->ny, cfg->nz, cfg->nwet, cfg->wet_mask, cfg->dx_row, cfg->dy, cfg->cos_lat,
                          cfg->dz_node, t.data(), d.data());
    if (taps) std::copy(t.begin(), t.end(), taps);
    if (diagonal) std::copy(d.begin(), d.end(), diagonal);
  });
}

int zm_projection_iterations(zm_handle* h, int32_t* iterations, void* stream) {
  if (!h || !iterations) return fail(h, "invalid zm_projection_iterations arguments", ZM_EINVAL);
  return guarded(h, [&] { h->solver->projection_iterations(iterations, (hipStream_t)stream); });
}

int zm_ensemble_create(zm_handle* const* members, int32_t n, zm_ensemble** out) {
  if (!members || n < 1 || !out) return fail(nullptr, "invalid zm_ensemble_create arguments", ZM_EINVAL);
  *out = nullptr;
  zm_ensemble* e = new zm_ensemble();
  int rc = guarded(nullptr, [&] {
    std::vector<zm::BtMember<float>> host;
    for (int i = 0; i < n; ++i) {
      if (!members[i]) throw ConfigError{"null member handle"};
      auto* sv = dynamic_cast<Solver<float>*>(members[i]->solver);
      if (!sv) throw ConfigError{"ensemble members must be float32 handles"};
      for (auto* other : e->mem)
        if (other == sv) throw ConfigError{"a handle may appear only once in an ensemble"};
      if (i == 0) e->n_sub = sv->n_sub;
      if (sv->n_sub != e->n_sub) throw ConfigError{"ensemble members must share the barotropic substep count"};
      if (sv->nx > zm_bt_row_max()) throw ConfigError{"nx too large for the batched barotropic substep"};
      host.push_back(sv->bt_member(e->total_rows));
      e->total_rows += sv->ny;
      e->mem.push_back(sv);
    }
    ZM_HIP(hipMalloc(&e->table, host.size() * sizeof(zm::BtMember<float>)));
    ZM_HIP(hipMemcpy(e->table, host.data(), host.size() * sizeof(zm::BtMember<float>), hipMemcpyHostToDevice));
    ZM_HIP(hipStreamCreateWithFlags(&e->own, hipStreamNonBlocking));
    e->make_lanes();
  });
  if (rc != ZM_OK) {
    delete e;
    return rc;
  }
  *out = e;
  return ZM_OK;
}

int zm_ensemble_step(zm_ensemble* e, const zm_state* in, const zm_state* out, void* stream) {
  if (!e || !in || !out) return fail(nullptr, "invalid zm_ensemble_step arguments", ZM_EINVAL);
  int rc = guarded(nullptr, [&] { e->step(in, out, (hipStream_t)stream); });
  if (rc != ZM_OK) e->err = g_thread_error;
  return rc;
}

int zm_ensemble_set_graph(zm_ensemble* e, int32_t enabled) {
  if (!e) return fail(nullptr, "null ensemble", ZM_EINVAL);
  e->use_graph = enabled != 0;
  return ZM_OK;
}

int32_t zm_ensemble_size(const zm_ensemble* e) { return e ? (int32_t)e->mem.size() : 0; }

int zm_ensemble_destroy(zm_ensemble* e) {
  delete e;
  return ZM_OK;
}

}  // extern "C"
