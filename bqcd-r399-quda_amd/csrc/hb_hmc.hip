// hb_hmc.hip - one HMC trajectory on device fields: the pseudofermion set-up
// of hmc_init_phi.F90, the Hamiltonian of hmc.F90 / action.F90, the force
// groups of hmc_integrator.F90 (dsf1.F90, dsf2.F90, dsd.F90, dsg.F90), the
// multi-time-scale leapfrog of hmc_leap_frog.F90 or an Omelyan integrator on
// the same groups, the link update of hmc_u.F90 and the chronological initial
// guesses of mre.F90.  Only scalars (actions, dot products, the small MRE
// system) cross to the host.
#include <chrono>
#include <cmath>
#include <complex>

#include "hb_blas.h"
#include "hb_ops.h"
#include "hb_su3.h"

namespace hb {

int dsf_dev(Context *c, double *P, double step, double kappa, double csw_kappa, const double2 *a,
            const double2 *b);
int dsd_dev(Context *c, double *P, double step, double csw_kappa);
int dsg_dev(Context *c, double *P, double step, double beta);
int hmc_u_dev(Context *c, const double *P, double step, double csw_kappa);
int clover_action_dev(Context *c, double *res_d);
int build_clover_dev(Context *c, double csw_kappa);
int plaquette_dev(Context *c, double *res_d);
int cg_mixed(Context *c, int op, double kappa, double rho, double2 *x, const double2 *b, float rest,
             int maxiter, int log, int prec, double delta, int max_inner, int *iters, int *n_rel,
             double *true_rtr);
int cg_multi(Context *c, int op, int nrhs, const double *kappa, const double *rho, double csw,
             double2 *const *x, const double2 *const *b, const float *rest, int maxiter, int log,
             int prec, int block, int *iters, double *true_rtr);

namespace {

constexpr int kRedGrid = 512;  // blocks of the vector reductions (fits the context's partials)
enum {
  SLOT_HMC_P = SLOT_HMC,
  SLOT_HMC_PHI1,
  SLOT_HMC_PHI2,
  SLOT_HMC_A,
  SLOT_HMC_B,
  SLOT_HMC_A2,
  SLOT_HMC_B2,
  SLOT_HMC_S,
  SLOT_HMC_RES,
  SLOT_HMC_U0,     // links at the start of the trajectory (restored on rejection)
  SLOT_HMC_LAST
};
static_assert(SLOT_HMC_LAST <= SLOT_HMC_MV, "HMC fields overlap the MRE products");

// ---------------------------------------------------------------------------
// vector kernels on n complex (double2) numbers
// ---------------------------------------------------------------------------
// Re, Im of sum conj(x) y (sc_cdotc, sc.F90:268-287) or, with y == x, |x|^2
__global__ __launch_bounds__(kBlock) void k_cdot(const double2 *__restrict__ x,
                                                 const double2 *__restrict__ y, size_t n,
                                                 double *partials) {
  double re = 0.0, im = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double2 a = x[i], b = y[i];
    re += a.x * b.x + a.y * b.y;
    im += a.x * b.y - a.y * b.x;
  }
  block_partials<2>({re, im}, partials);
}

// sum of squares of n doubles (sp, action.F90:126-143, as 2 x sum)
__global__ __launch_bounds__(kBlock) void k_sumsq(const double *__restrict__ x, size_t n,
                                                  double *partials) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    s += x[i] * x[i];
  block_partials<1>({s}, partials);
}

// y += (ar + i ai) x  (sc_axpy / sc_caxpy, sc.F90)
__global__ void k_caxpy(double2 *__restrict__ y, const double2 *__restrict__ x, double ar, double ai,
                        size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double2 v = x[i];
    double2 w = y[i];
    w.x += ar * v.x - ai * v.y;
    w.y += ar * v.y + ai * v.x;
    y[i] = w;
  }
}

__global__ void k_scale(double2 *__restrict__ x, double s, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double2 v = x[i];
    v.x *= s;
    v.y *= s;
    x[i] = v;
  }
}

__global__ void k_gauss(double *__restrict__ x, size_t n, uint64_t seed, double sigma) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; 2 * i < n; i += (size_t)gridDim.x * blockDim.x) {
    Rng r(seed, i);
    double a, b;
    r.gauss2(a, b);
    x[2 * i] = sigma * a;
    if (2 * i + 1 < n) x[2 * i + 1] = sigma * b;
  }
}

struct FinCopy2 {
  double *out;
  __device__ void operator()(const double (&t)[2]) const { out[0] = t[0]; out[1] = t[1]; }
};

int grid_for(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + kBlock - 1) / kBlock, kRedGrid)); }

// the HMC's scalar reductions: device result, read back synchronously
struct Vec {
  Context *c;
  double *res_d;
  size_t n;  // complex numbers per spinor
  double host[4];

  int cdot(const double2 *x, const double2 *y, std::complex<double> *out) {
    const int g = grid_for(n);
    k_cdot<<<g, kBlock, 0, c->stream>>>(x, y, n, c->red);
    k_finalize<2><<<1, kFinThreads, 0, c->stream>>>(c->red, g, nullptr, FinCopy2{res_d});
    HB_HIP(hipGetLastError());
    HB_HIP(hipMemcpyAsync(host, res_d, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HB_HIP(hipStreamSynchronize(c->stream));
    *out = {host[0], host[1]};
    return HB_OK;
  }
  int norm2(const double2 *x, double *out) {
    std::complex<double> z;
    HB_TRY(cdot(x, x, &z));
    *out = z.real();
    return HB_OK;
  }
  int sumsq(const double *x, size_t m, double *out) {
    const int g = grid_for(m);
    k_sumsq<<<g, kBlock, 0, c->stream>>>(x, m, c->red);
    k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(c->red, g, nullptr, FinStore1{res_d});
    HB_HIP(hipGetLastError());
    HB_HIP(hipMemcpyAsync(host, res_d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HB_HIP(hipStreamSynchronize(c->stream));
    *out = host[0];
    return HB_OK;
  }
  int caxpy(double2 *y, const double2 *x, std::complex<double> a) {
    k_caxpy<<<grid_for(n) * 4, kBlock, 0, c->stream>>>(y, x, a.real(), a.imag(), n);
    HB_HIP(hipGetLastError());
    return HB_OK;
  }
  int scale(double2 *x, double s) {
    k_scale<<<grid_for(n) * 4, kBlock, 0, c->stream>>>(x, s, n);
    HB_HIP(hipGetLastError());
    return HB_OK;
  }
};

// ---------------------------------------------------------------------------
// trajectory state
// ---------------------------------------------------------------------------
struct Traj {
  Context *c;
  hb_hmc_params prm;
  Vec v;
  double ck;       // csw * kappa
  bool clover, hasenbusch;
  double *P;
  double2 *phi1, *phi2, *a, *b, *a2, *b2, *s;
  double sf1 = 0.0, sf2 = 0.0;
  int ncall = 0, niter_tot = 0, niter_max = 0;
  int n_ir = 0, n_uv = 0;
  std::chrono::steady_clock::time_point t0;

  void progress(int it) {
    if (!prm.verbose) return;
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fprintf(stderr, "[hb_hmc] outer step %d/%d: %d solves, %d CG iterations, %.1f s\n", it, prm.ntau,
            ncall, niter_tot, s);
    fflush(stderr);
  }

  // cg_stat bookkeeping (cg.F90:126-131)
  void count(int it) {
    ncall++;
    niter_tot += it;
    niter_max = std::max(niter_max, it);
  }

  // solve op x = src from the guess in x
  int solve(int op, double rho, double2 *x, const double2 *src) {
    int it = 0;
    double rtr = 0.0;
    if (prm.precision == HB_PREC_F64) {
      HB_TRY(cg_f64(c, op, prm.kappa, rho, x, src, prm.rest, prm.maxiter, prm.log, &it, &rtr));
    } else {
      int nrel = 0;
      HB_TRY(cg_mixed(c, op, prm.kappa, rho, x, src, prm.rest, prm.maxiter, prm.log, prm.precision,
                      prm.delta, prm.max_inner, &it, &nrel, &rtr));
    }
    count(it);
    return HB_OK;
  }

  double2 *mre_vec(int basis, int k) {
    int rc = 0;
    return spinor_scratch(c, c->mre_slot[basis][k], &rc);
  }

  // mre_get (mre.F90:68-146): trial solution of the system whose operator
  // M (w_mult for basis 0, mtil for basis 1) maps the basis vectors
  int mre_get(int basis, double rho, double2 *trial, const double2 *phi) {
    const int rank = c->mre_rank[basis];
    const size_t n = v.n;
    if (c->mre_nvec == 0 || rank == 0) return copy_f64(c, trial, phi, n);
    if (rank == 1) return copy_f64(c, trial, mre_vec(basis, 0), n);
    // modified Gram-Schmidt with real inner products (mre.F90:170-195)
    for (int k = 0; k < rank; k++) {
      double r_kk;
      HB_TRY(v.norm2(mre_vec(basis, k), &r_kk));
      HB_TRY(v.scale(mre_vec(basis, k), 1.0 / std::sqrt(r_kk)));
      for (int j = k + 1; j < rank; j++) {
        std::complex<double> r_kj;
        HB_TRY(v.cdot(mre_vec(basis, k), mre_vec(basis, j), &r_kj));
        HB_TRY(v.caxpy(mre_vec(basis, j), mre_vec(basis, k), -r_kj.real()));
      }
    }
    // b_i = <v_i, phi>, G_ij = <M v_i, M v_j>
    std::complex<double> g[kMaxMre][kMaxMre], bb[kMaxMre];
    int rc = 0;
    double2 *t = spinor_scratch(c, SLOT_T, &rc); HB_TRY(rc);
    double2 *y = spinor_scratch(c, SLOT_Y, &rc); HB_TRY(rc);
    for (int i = 0; i < rank; i++) {
      HB_TRY(v.cdot(mre_vec(basis, i), phi, &bb[i]));
      double2 *mv = spinor_scratch(c, SLOT_HMC_MV + i, &rc); HB_TRY(rc);
      HB_TRY(apply_f64(c, basis == 0 ? HB_OP_W : HB_OP_MTIL, prm.kappa, rho, mv, mre_vec(basis, i), t, y,
                       nullptr, nullptr));
    }
    for (int i = 0; i < rank; i++) {
      double2 *mi = spinor_scratch(c, SLOT_HMC_MV + i, &rc); HB_TRY(rc);
      HB_TRY(v.cdot(mi, mi, &g[i][i]));
      g[i][i] = {g[i][i].real(), 0.0};
      for (int j = i + 1; j < rank; j++) {
        double2 *mj = spinor_scratch(c, SLOT_HMC_MV + j, &rc); HB_TRY(rc);
        HB_TRY(v.cdot(mi, mj, &g[i][j]));
        g[j][i] = std::conj(g[i][j]);
      }
    }
    HB_TRY(gauss_jordan(g, bb, rank));
    // trial = sum_j b_j v_j (sc_cax2 / sc_caxpy2 / sc_caxpy)
    HB_TRY(zero_f64(c, trial, n));
    for (int j = 0; j < rank; j++) HB_TRY(v.caxpy(trial, mre_vec(basis, j), bb[j]));
    return HB_OK;
  }

  // Gauss-Jordan elimination with full pivoting (mre_gauss_jordan, mre.F90:199-279)
  static int gauss_jordan(std::complex<double> (&a)[kMaxMre][kMaxMre], std::complex<double> (&b)[kMaxMre],
                          int n) {
    int ipiv[kMaxMre] = {0};
    for (int i = 0; i < n; i++) {
      double big = 0.0;
      int irow = 0, icol = 0;
      for (int j = 0; j < n; j++) {
        if (ipiv[j] == 1) continue;
        for (int k = 0; k < n; k++) {
          if (ipiv[k] == 0) {
            const double t = std::abs(a[j][k]);
            if (t >= big) {
              big = t;
              irow = j;
              icol = k;
            }
          } else if (ipiv[k] > 1) {
            set_error("mre: singular matrix");
            return HB_E_STATE;
          }
        }
      }
      ipiv[icol]++;
      if (irow != icol) {
        for (int l = 0; l < n; l++) std::swap(a[irow][l], a[icol][l]);
        std::swap(b[irow], b[icol]);
      }
      HB_CHECK(a[icol][icol] != std::complex<double>(0.0), HB_E_STATE, "mre: singular matrix");
      const std::complex<double> pivinv = 1.0 / a[icol][icol];
      for (int l = 0; l < n; l++) a[icol][l] *= pivinv;
      b[icol] *= pivinv;
      for (int ll = 0; ll < n; ll++) {
        if (ll == icol) continue;
        const std::complex<double> dum = a[ll][icol];
        a[ll][icol] = 0.0;
        for (int l = 0; l < n; l++) a[ll][l] -= a[icol][l] * dum;
        b[ll] -= b[icol] * dum;
      }
    }
    return HB_OK;
  }

  // mre_put (mre.F90:28-66): reset, or push x as the newest basis vector
  int mre_put(int basis, const double2 *x, bool reset) {
    if (c->mre_nvec == 0) return HB_OK;
    if (reset) {
      c->mre_rank[basis] = 0;
      return HB_OK;
    }
    const int nv = c->mre_nvec;
    const int oldest = c->mre_slot[basis][nv - 1];
    for (int i = nv - 1; i > 0; i--) c->mre_slot[basis][i] = c->mre_slot[basis][i - 1];
    c->mre_slot[basis][0] = oldest;
    HB_TRY(copy_f64(c, mre_vec(basis, 0), x, v.n));
    if (c->mre_rank[basis] < nv) c->mre_rank[basis]++;
    return HB_OK;
  }

  // dsf1.F90:28-82: a = (W^+W)^-1 phi1, b = W a, force; S_f1 = |b|^2
  int dsf1(double step, bool calc) {
    HB_TRY(mre_get(0, prm.rho, a, phi1));
    HB_TRY(solve(HB_OP_WDAGW, prm.rho, a, phi1));
    return dsf1_tail(step, calc);
  }
  int dsf1_tail(double step, bool calc) {
    int rc = 0;
    double2 *t = spinor_scratch(c, SLOT_T, &rc); HB_TRY(rc);
    HB_TRY(mre_put(0, a, calc));
    HB_TRY(apply_f64(c, HB_OP_W, prm.kappa, prm.rho, b, a, t, nullptr, nullptr, nullptr));
    if (calc) HB_TRY(v.norm2(b, &sf1));
    return dsf_dev(c, P, step, prm.kappa, ck, a, b);
  }

  // dsf2.F90:28-80: a = (M~^+M~)^-1 W^+ phi2, b = M~ a - phi2, force; S_f2 = |M~ a|^2
  int dsf2_src() {
    int rc = 0;
    double2 *t = spinor_scratch(c, SLOT_T, &rc); HB_TRY(rc);
    HB_TRY(apply_f64(c, HB_OP_W_DAG, prm.kappa, prm.rho, s, phi2, t, nullptr, nullptr, nullptr));
    return mre_get(1, 0.0, a2, s);
  }
  int dsf2(double step, bool calc) {
    HB_TRY(dsf2_src());
    HB_TRY(solve(HB_OP_MTDAGMT, 0.0, a2, s));
    return dsf2_tail(step, calc);
  }
  int dsf2_tail(double step, bool calc) {
    int rc = 0;
    double2 *t = spinor_scratch(c, SLOT_T, &rc); HB_TRY(rc);
    HB_TRY(mre_put(1, a2, calc));
    HB_TRY(apply_f64(c, HB_OP_MTIL, prm.kappa, 0.0, b2, a2, t, nullptr, nullptr, nullptr));
    if (calc) HB_TRY(v.norm2(b2, &sf2));
    HB_TRY(v.caxpy(b2, phi2, -1.0));
    return dsf_dev(c, P, step, prm.kappa, ck, a2, b2);
  }

  // both pseudofermion systems in one multi-RHS solve (model B: both on the
  // outer steps; model C: at the points where an inner and an outer kick meet),
  // S_f1's force with step1, S_f2's with step2
  int dsf12_batched(double step1, double step2, bool calc) {
    HB_TRY(mre_get(0, prm.rho, a, phi1));
    HB_TRY(dsf2_src());
    double kap[2] = {prm.kappa, prm.kappa}, rh[2] = {prm.rho, 0.0};
    double2 *xs[2] = {a, a2};
    const double2 *bs[2] = {phi1, s};
    float rest[2] = {prm.rest, prm.rest};
    int it[2] = {0, 0};
    double tr[2];
    HB_TRY(cg_multi(c, HB_OP_WDAGW, 2, kap, rh, clover ? -1.0 : 0.0, xs, bs, rest, prm.maxiter, prm.log,
                    prm.precision, 2, it, tr));
    count(it[0]);
    count(it[1]);
    HB_TRY(dsf1_tail(step1, calc));
    return dsf2_tail(step2, calc);
  }
  bool batched() const { return prm.batch && prm.precision != HB_PREC_F64; }

  int dsd(double step) { return clover ? dsd_dev(c, P, step, ck) : HB_OK; }

  // hmc_integrator_p_ir / _p_uv / _q (hmc_integrator.F90:28-103)
  int p_ir(double step, bool calc) {
    n_ir++;
    switch (prm.model) {
      case 'A':
        HB_TRY(dsf1(step, calc));
        return dsd(step);
      case 'B':
        if (batched()) {
          HB_TRY(dsf12_batched(step, step, calc));
        } else {
          HB_TRY(dsf1(step, calc));
          HB_TRY(dsf2(step, calc));
        }
        return dsd(step);
      default:  // 'C'
        return dsf2(step, calc);
    }
  }
  int p_uv(double step, bool calc) {
    n_uv++;
    HB_TRY(dsg_dev(c, P, step, prm.beta));
    if (prm.model == 'C') {
      HB_TRY(dsf1(step, calc));
      return dsd(step);
    }
    return HB_OK;
  }
  int q(double step) { return hmc_u_dev(c, P, step, clover ? ck : 0.0); }

  // an inner and an outer kick at the same links (momentum kicks commute):
  // under model C with the batched solver the inner S_f1 system and the outer
  // S_f2 system are solved together as one 2-RHS block
  int p_both(double s_uv, double s_ir, bool calc) {
    if (prm.model == 'C' && batched()) {
      n_uv++;
      n_ir++;
      HB_TRY(dsg_dev(c, P, s_uv, prm.beta));
      HB_TRY(dsf12_batched(s_uv, s_ir, calc));
      return dsd(s_uv);
    }
    HB_TRY(p_uv(s_uv, calc));
    return p_ir(s_ir, calc);
  }

  // hmc_leap_frog.F90:28-79
  int leapfrog() {
    const int m = prm.m_scale;
    double s_ir = 0.5 * prm.tau, s_uv = 0.5 * prm.tau / m;
    HB_TRY(p_ir(s_ir, false));
    HB_TRY(p_uv(s_uv, false));
    s_ir = prm.tau;
    s_uv = prm.tau / m;
    bool calc = false;
    for (int it = 1; it <= prm.ntau; it++) {
      for (int is = 1; is <= m; is++) {
        HB_TRY(q(s_uv));
        if (it == prm.ntau && is == m) {
          s_ir *= 0.5;
          s_uv *= 0.5;
          calc = true;
        }
        HB_TRY(p_uv(s_uv, calc));
      }
      HB_TRY(p_ir(s_ir, calc));
      progress(it);
    }
    return HB_OK;
  }

  // Omelyan (second-order minimum norm) on both scales: per outer step
  // P_ir(l h) U_uv(h/2) P_ir((1-2l) h) U_uv(h/2) P_ir(l h), with U_uv(h) made
  // of m_scale inner steps P_uv(l e) Q(e/2) P_uv((1-2l) e) Q(e/2) P_uv(l e).
  // Kicks between two link updates act on the same links and are merged:
  // P_uv(2 l e) between inner steps, and where one inner sequence ends and the
  // next begins the inner P_uv(2 l e) and the outer kick together (p_both)
  int omelyan() {
    const double l = prm.lambda > 0.0 ? prm.lambda : 0.1931833275037836;
    const int m = prm.m_scale;
    const double h = prm.tau, e = 0.5 * h / m;
    HB_TRY(p_both(l * e, l * h, false));
    for (int it = 1; it <= prm.ntau; it++) {
      for (int half = 0; half < 2; half++) {
        for (int is = 1; is <= m; is++) {
          HB_TRY(q(0.5 * e));
          HB_TRY(p_uv((1.0 - 2.0 * l) * e, false));
          HB_TRY(q(0.5 * e));
          if (is < m) {
            HB_TRY(p_uv(2.0 * l * e, false));
          } else if (half == 0) {
            HB_TRY(p_both(2.0 * l * e, (1.0 - 2.0 * l) * h, false));
          } else if (it == prm.ntau) {
            HB_TRY(p_both(l * e, l * h, true));
          } else {
            HB_TRY(p_both(2.0 * l * e, 2.0 * l * h, false));
          }
        }
      }
      progress(it);
    }
    return HB_OK;
  }

  int sg(double *out) {
    double *r = v.res_d + 2;
    HB_TRY(plaquette_dev(c, r));
    HB_HIP(hipMemcpyAsync(v.host, r, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HB_HIP(hipStreamSynchronize(c->stream));
    *out = 12.0 * c->g.vh - v.host[0] / 3.0;  // 6 V - sum Re Tr U_p / 3
    return HB_OK;
  }
  int sd(double *out) {
    *out = 0.0;
    if (!clover) return HB_OK;
    double *r = v.res_d + 2;
    HB_TRY(clover_action_dev(c, r));
    HB_HIP(hipMemcpyAsync(v.host, r, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HB_HIP(hipStreamSynchronize(c->stream));
    *out = v.host[0];
    return HB_OK;
  }
  int sp(double *out) {
    double s2;
    HB_TRY(v.sumsq(P, (size_t)64 * c->g.vh, &s2));
    *out = 0.5 * s2;
    return HB_OK;
  }
};

// hmc.F90:28-215.  dir = 0: a trajectory; dir = +1 / -1: the reversibility test
// modes HMC_TEST_FORWARDS / HMC_TEST_BACKWARDS (hmc.F90:124-210): the backward
// one negates the step and skips the draws and H before, integrating with the
// phi fields left by the preceding trajectory.  p_out (may be null) receives
// the final momenta.
int trajectory(Context *c, const hb_hmc_params *prm, const double *p_d, const double *eta1_d,
               const double *eta2_d, double accept_random, hb_hmc_out *out, int dir = 0,
               double *p_out = nullptr) {
  const bool backward = dir < 0;
  HB_CHECK(prm && out && p_d && (eta1_d || backward), HB_E_ARG, "hmc: null argument");
  HB_CHECK(prm->model == 'A' || prm->model == 'B' || prm->model == 'C', HB_E_ARG, "hmc: model '%c'",
           prm->model);
  HB_CHECK(prm->model == 'A' || eta2_d || backward, HB_E_ARG, "hmc: models B and C need eta2");
  HB_CHECK(!backward || c->hmc_phi_model == prm->model, HB_E_STATE,
           "hmc: a backward test trajectory needs a trajectory of model %c before it", prm->model);
  HB_CHECK(prm->ntau >= 1 && prm->m_scale >= 1, HB_E_ARG, "hmc: ntau, m_scale must be >= 1");
  HB_CHECK(prm->integrator == HB_INT_LEAPFROG || prm->integrator == HB_INT_OMELYAN, HB_E_ARG,
           "hmc: integrator %d", prm->integrator);
  HB_CHECK(prm->precision == HB_PREC_F64 || prm->precision == HB_PREC_F32 || prm->precision == HB_PREC_F16,
           HB_E_ARG, "hmc: precision %d", prm->precision);
  HB_CHECK(prm->mre_vectors >= 0 && prm->mre_vectors <= kMaxMre, HB_E_ARG, "hmc: mre_vectors %d (0..%d)",
           prm->mre_vectors, kMaxMre);
  HB_CHECK(prm->maxiter >= 1, HB_E_ARG, "hmc: maxiter must be >= 1");
  HB_CHECK(prm->csw != 0.0 || !use_clover(c), HB_E_STATE,
           "hmc: csw = 0 but the context holds a clover term (use a context without one)");
  const auto t0 = std::chrono::steady_clock::now();
  const int vh = c->g.vh;
  int rc = 0;
  Traj T{};
  T.c = c;
  T.prm = *prm;
  if (backward) T.prm.tau = -prm->tau;  // hmc_test.F90:45
  T.t0 = t0;
  if (T.prm.delta <= 0.0) T.prm.delta = 0.1;
  if (T.prm.max_inner <= 0) T.prm.max_inner = 1000;
  T.ck = prm->csw * prm->kappa;
  T.clover = prm->csw != 0.0;
  T.hasenbusch = prm->model != 'A';
  T.v.c = c;
  T.v.n = (size_t)12 * vh;
  T.v.res_d = static_cast<double *>(scratch(c, SLOT_HMC_RES, 8 * sizeof(double), &rc)); HB_TRY(rc);
  T.P = static_cast<double *>(scratch(c, SLOT_HMC_P, sizeof(double) * 64 * (size_t)vh, &rc)); HB_TRY(rc);
  T.phi1 = spinor_scratch(c, SLOT_HMC_PHI1, &rc); HB_TRY(rc);
  T.phi2 = spinor_scratch(c, SLOT_HMC_PHI2, &rc); HB_TRY(rc);
  T.a = spinor_scratch(c, SLOT_HMC_A, &rc); HB_TRY(rc);
  T.b = spinor_scratch(c, SLOT_HMC_B, &rc); HB_TRY(rc);
  T.a2 = spinor_scratch(c, SLOT_HMC_A2, &rc); HB_TRY(rc);
  T.b2 = spinor_scratch(c, SLOT_HMC_B2, &rc); HB_TRY(rc);
  T.s = spinor_scratch(c, SLOT_HMC_S, &rc); HB_TRY(rc);
  double2 *t = spinor_scratch(c, SLOT_T, &rc); HB_TRY(rc);

  // MRE bases: (re)initialise when the vector count changes (mre_allocate)
  if (c->mre_nvec != prm->mre_vectors) {
    c->mre_nvec = prm->mre_vectors;
    for (int bsel = 0; bsel < 2; bsel++) {
      c->mre_rank[bsel] = 0;
      for (int k = 0; k < kMaxMre; k++) c->mre_slot[bsel][k] = SLOT_HMC_MRE + kMaxMre * bsel + k;
    }
  }

  // start: backups and the clover term of the current links (hmc.F90:104-113)
  // (the library's own slot: a caller's hb_gauge_backup slots are not touched)
  double2 *u0 = nullptr;
  if (accept_random >= 0.0) {
    u0 = static_cast<double2 *>(scratch(c, SLOT_HMC_U0, sizeof(double2) * 72 * (size_t)vh, &rc)); HB_TRY(rc);
    HB_HIP(hipMemcpyAsync(u0, c->gauge, sizeof(double2) * 72 * (size_t)vh, hipMemcpyDeviceToDevice, c->stream));
  }
  if (T.clover) HB_TRY(build_clover_dev(c, T.ck));
  HB_HIP(hipMemcpyAsync(T.P, p_d, sizeof(double) * 64 * (size_t)vh, hipMemcpyDeviceToDevice, c->stream));

  double sf1_old = 0.0, sf2_old = 0.0, sd_old = 0.0, sg_old = 0.0, sp_old = 0.0, h_old = 0.0;
  if (!backward) {
    // hmc_init_phi.F90:28-79 (no boundary flip: the kernels apply the sign)
    c->hmc_phi_model = 0;
    const double2 *eta1 = reinterpret_cast<const double2 *>(eta1_d);
    HB_TRY(T.v.norm2(eta1, &sf1_old));
    HB_TRY(apply_f64(c, HB_OP_W_DAG, prm->kappa, prm->rho, T.phi1, eta1, t, nullptr, nullptr, nullptr));
    if (T.hasenbusch) {
      const double2 *eta2 = reinterpret_cast<const double2 *>(eta2_d);
      HB_TRY(T.v.norm2(eta2, &sf2_old));
      HB_TRY(apply_f64(c, HB_OP_MTIL_DAG, prm->kappa, 0.0, T.phi2, eta2, t, nullptr, nullptr, nullptr));
      HB_TRY(copy_f64(c, T.a, eta2, T.v.n));  // the noise is the initial guess (hmc_init_phi.F90:64-70)
      HB_TRY(T.solve(HB_OP_WDAGW, prm->rho, T.a, T.phi2));
      HB_TRY(apply_f64(c, HB_OP_W, prm->kappa, prm->rho, T.phi2, T.a, t, nullptr, nullptr, nullptr));
    }
    c->hmc_phi_model = prm->model;

    // H before (hmc.F90:127-134)
    HB_TRY(T.sd(&sd_old));
    HB_TRY(T.sg(&sg_old));
    HB_TRY(T.sp(&sp_old));
    h_old = sd_old + sp_old + prm->beta * sg_old + sf1_old + sf2_old;
  }

  HB_TRY(prm->integrator == HB_INT_LEAPFROG ? T.leapfrog() : T.omelyan());

  // H after (hmc.F90:145-154)
  double sd_new, sg_new, sp_new;
  HB_TRY(T.sd(&sd_new));
  HB_TRY(T.sg(&sg_new));
  HB_TRY(T.sp(&sp_new));
  const double h_new = sd_new + sp_new + prm->beta * sg_new + T.sf1 + T.sf2;

  hb_hmc_out o{};
  o.h_old = h_old;
  o.h_new = h_new;
  o.exp_dh = backward ? 0.0 : std::exp(h_old - h_new);
  o.sg_old = sg_old;
  o.sg_new = sg_new;
  o.sp_old = sp_old;
  o.sp_new = sp_new;
  o.sd_old = sd_old;
  o.sd_new = sd_new;
  o.sf1_old = sf1_old;
  o.sf1_new = T.sf1;
  o.sf2_old = sf2_old;
  o.sf2_new = T.sf2;
  o.accepted = (accept_random < 0.0 || accept_random < o.exp_dh) ? 1 : 0;
  if (!o.accepted) {
    HB_HIP(hipMemcpyAsync(c->gauge, u0, sizeof(double2) * 72 * (size_t)vh, hipMemcpyDeviceToDevice, c->stream));
    c->gauge_version++;
    if (T.clover) HB_TRY(build_clover_dev(c, T.ck));
    HB_TRY(T.sg(&o.plaquette));
  } else {
    o.plaquette = sg_new;
  }
  o.plaquette = 1.0 - o.plaquette / (12.0 * vh);  // <Re Tr U_p>/3
  o.cg_ncall = T.ncall;
  o.cg_niter_tot = T.niter_tot;
  o.cg_niter_max = T.niter_max;
  o.n_force_ir = T.n_ir;
  o.n_force_uv = T.n_uv;
  if (p_out)
    HB_HIP(hipMemcpyAsync(p_out, T.P, sizeof(double) * 64 * (size_t)vh, hipMemcpyDeviceToDevice, c->stream));
  HB_HIP(hipStreamSynchronize(c->stream));
  o.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = o;
  return HB_OK;
}

}  // namespace
}  // namespace hb

using namespace hb;

extern "C" {

int hb_hmc_trajectory_d(hb_context *ctx, const hb_hmc_params *prm, const double *p_d, const double *eta1_d,
                        const double *eta2_d, double accept_random, hb_hmc_out *out) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_HIP(hipSetDevice(c->device));
  return trajectory(c, prm, p_d, eta1_d, eta2_d, accept_random, out);
}

int hb_hmc_test_d(hb_context *ctx, const hb_hmc_params *prm, int direction, double *p_d,
                  const double *eta1_d, const double *eta2_d, hb_hmc_out *out) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(direction == 1 || direction == -1, HB_E_ARG, "hb_hmc_test_d: direction %d (+1 or -1)", direction);
  HB_HIP(hipSetDevice(c->device));
  // hmc_test.F90:41-47: force-accepted in both directions
  return trajectory(c, prm, p_d, eta1_d, eta2_d, -1.0, out, direction, p_d);
}

int hb_hmc_trajectory(hb_context *ctx, const hb_hmc_params *prm, const double *p, int site_stride,
                      const double *eta1, const double *eta2, double accept_random, hb_hmc_out *out) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p && eta1 && site_stride >= c->g.vh, HB_E_ARG, "hb_hmc_trajectory: bad arguments");
  HB_HIP(hipSetDevice(c->device));
  const int vh = c->g.vh;
  int rc = 0;
  // staging outside the trajectory's own slots
  double *pd = static_cast<double *>(scratch(c, SLOT_MOM, sizeof(double) * 64 * (size_t)vh, &rc)); HB_TRY(rc);
  double2 *e1 = spinor_scratch(c, SLOT_X, &rc); HB_TRY(rc);
  double2 *e2 = spinor_scratch(c, SLOT_B, &rc); HB_TRY(rc);
  HB_TRY(gen_upload(c, pd, p, site_stride));
  HB_TRY(spinor_upload(c, e1, eta1));
  if (eta2) HB_TRY(spinor_upload(c, e2, eta2));
  return trajectory(c, prm, pd, reinterpret_cast<double *>(e1), eta2 ? reinterpret_cast<double *>(e2) : nullptr,
                    accept_random, out);
}

int hb_hmc_test(hb_context *ctx, const hb_hmc_params *prm, int direction, double *p, int site_stride,
                const double *eta1, const double *eta2, hb_hmc_out *out) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(direction == 1 || direction == -1, HB_E_ARG, "hb_hmc_test: direction %d (+1 or -1)", direction);
  HB_CHECK(p && site_stride >= c->g.vh && (eta1 || direction < 0), HB_E_ARG, "hb_hmc_test: bad arguments");
  HB_HIP(hipSetDevice(c->device));
  const int vh = c->g.vh;
  int rc = 0;
  double *pd = static_cast<double *>(scratch(c, SLOT_MOM, sizeof(double) * 64 * (size_t)vh, &rc)); HB_TRY(rc);
  double2 *e1 = nullptr, *e2 = nullptr;
  HB_TRY(gen_upload(c, pd, p, site_stride));
  if (direction > 0) {
    e1 = spinor_scratch(c, SLOT_X, &rc); HB_TRY(rc);
    HB_TRY(spinor_upload(c, e1, eta1));
    if (eta2) {
      e2 = spinor_scratch(c, SLOT_B, &rc); HB_TRY(rc);
      HB_TRY(spinor_upload(c, e2, eta2));
    }
  }
  // hmc_test.F90:41-47: force-accepted in both directions
  HB_TRY(trajectory(c, prm, pd, reinterpret_cast<double *>(e1), reinterpret_cast<double *>(e2), -1.0, out,
                    direction, pd));
  return gen_download(c, p, pd, site_stride);
}

int hb_gaussian_d(hb_context *ctx, unsigned long long seed, double var, double *x_d, long long n) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c && x_d && n >= 0 && var >= 0.0, HB_E_ARG, "hb_gaussian_d: bad arguments");
  HB_HIP(hipSetDevice(c->device));
  if (n == 0) return HB_OK;
  const size_t pairs = ((size_t)n + 1) / 2;
  const int grid = (int)std::min<size_t>((pairs + kBlock - 1) / kBlock, 65536);
  k_gauss<<<grid, kBlock, 0, c->stream>>>(x_d, (size_t)n, seed, std::sqrt(var));
  HB_HIP(hipGetLastError());
  return HB_OK;
}

}  // extern "C"
