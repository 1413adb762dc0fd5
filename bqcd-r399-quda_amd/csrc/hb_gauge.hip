// hb_gauge.hip - gauge-field kernels: clover term construction (clover_init),
// plaquette sum, Wilson gauge force (dsg), link update (hmc_u), and seeded
// synthetic gauge fields (exp(i s H) and Haar + APE smearing).
#include "hb_ops.h"
#include "hb_su3.h"

namespace hb {

// ---------------------------------------------------------------------------
// clover_f_mu_nu (clover/clover_f_mu_nu.F90:37-103): f = -i (Q - Q^+) with Q the
// sum of the four plaquette leaves in plane (mu, nu) around x (parity e)
// ---------------------------------------------------------------------------
__device__ M3 f_mu_nu(const double2 *__restrict__ U, const Geom &g, int e, int i, const int (&x)[4],
                      int mu, int nu) {
  const int o = 1 - e, vh = g.vh;
  const int xp_ = nb(g, x[0], x[1], x[2], x[3], mu, 1);
  const int xm_ = nb(g, x[0], x[1], x[2], x[3], mu, -1);
  const int x_p = nb(g, x[0], x[1], x[2], x[3], nu, 1);
  const int x_m = nb(g, x[0], x[1], x[2], x[3], nu, -1);
  const int xmp = nb2(g, x[0], x[1], x[2], x[3], mu, -1, nu, 1);
  const int xmm = nb2(g, x[0], x[1], x[2], x[3], mu, -1, nu, -1);
  const int xpm = nb2(g, x[0], x[1], x[2], x[3], mu, 1, nu, -1);
  const M3 umu_x = link(U, mu, e, i, vh);
  const M3 unu_x = link(U, nu, e, i, vh);
  const M3 umu_xm = link(U, mu, o, xm_, vh);
  const M3 unu_xm = link(U, nu, o, x_m, vh);
  // leaf 1: U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
  M3 q = m3_muld(m3_muld(m3_mul(umu_x, link(U, nu, o, xp_, vh)), link(U, mu, o, x_p, vh)), unu_x);
  // leaf 2: U_nu(x) U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu)
  m3_acc(q, m3_mul(m3_muld(m3_muld(unu_x, link(U, mu, e, xmp, vh)), link(U, nu, o, xm_, vh)), umu_xm));
  // leaf 3: U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu)
  m3_acc(q, m3_mul(m3_dmul(m3_mul(link(U, nu, e, xmm, vh), umu_xm), link(U, mu, e, xmm, vh)), unu_xm));
  // leaf 4: U_nu(x-nu)^+ U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^+
  m3_acc(q, m3_muld(m3_mul(m3_dmul(unu_xm, link(U, mu, o, x_m, vh)), link(U, nu, e, xpm, vh)), umu_x));
  M3 f;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      double2 d = csub(q.m[3 * a + b], cconj(q.m[3 * b + a]));
      f.m[3 * a + b] = make_double2(d.y, -d.x);  // -i d
    }
#pragma unroll
  for (int a = 0; a < 3; a++) f.m[4 * a].y = 0.0;
  return f;
}

// LDL^+ of a Hermitian 6x6 block given as a full matrix; writes the
// reference's stored forms: A/2 (packed), A^-1/2 (packed), L and 1/(2 d_k)
// (clover/clover_inv.F90:29-220)
__device__ void clover_block_store(double2 (&A)[6][6], double2 *__restrict__ Ca,
                                   double2 *__restrict__ Ci, double2 *__restrict__ Cb, size_t base,
                                   int vh) {
  double d[6];
  double2 L[6][6];
#pragma unroll
  for (int r = 0; r < 6; r++) {
#pragma unroll
    for (int c = 0; c < r; c++) {
      double2 acc = cconj(A[c][r]);  // A_rc
#pragma unroll
      for (int k = 0; k < c; k++) acc = csub(acc, cscale<double2>(d[k], cmul(L[r][k], cconj(L[c][k]))));
      L[r][c] = cscale<double2>(1.0 / d[c], acc);
    }
    double dr = A[r][r].x;
#pragma unroll
    for (int k = 0; k < r; k++) dr -= d[k] * (L[r][k].x * L[r][k].x + L[r][k].y * L[r][k].y);
    d[r] = dr;
  }
  // B words: L_rc at r(r-1)/2 + c, then (1/(2 d)) pairs
#pragma unroll
  for (int r = 1; r < 6; r++)
#pragma unroll
    for (int c = 0; c < r; c++) Cb[base + (size_t)(r * (r - 1) / 2 + c) * vh] = L[r][c];
  Cb[base + (size_t)15 * vh] = make_double2(0.5 / d[0], 0.5 / d[1]);
  Cb[base + (size_t)16 * vh] = make_double2(0.5 / d[2], 0.5 / d[3]);
  Cb[base + (size_t)17 * vh] = make_double2(0.5 / d[4], 0.5 / d[5]);
  // A/2 words
  Ca[base] = make_double2(0.5 * A[0][0].x, 0.5 * A[1][1].x);
  Ca[base + (size_t)10 * vh] = make_double2(0.5 * A[2][2].x, 0.5 * A[3][3].x);
  Ca[base + (size_t)16 * vh] = make_double2(0.5 * A[4][4].x, 0.5 * A[5][5].x);
#pragma unroll
  for (int r = 0; r < 6; r++)
#pragma unroll
    for (int c = r + 1; c < 6; c++) Ca[base + (size_t)a_off_word(r, c) * vh] = cscale<double2>(0.5, A[r][c]);
  // inverse column by column: x = A^-1 e_col; row col of A^-1 = conj(x)
  double2 Ainv[6][6];
#pragma unroll
  for (int col = 0; col < 6; col++) {
    double2 y[6], x[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
      double2 acc = make_double2(r == col ? 1.0 : 0.0, 0.0);
#pragma unroll
      for (int c = 0; c < r; c++) acc = csub(acc, cmul(L[r][c], y[c]));
      y[r] = acc;
    }
#pragma unroll
    for (int r = 5; r >= 0; r--) {
      double2 acc = cscale<double2>(1.0 / d[r], y[r]);
#pragma unroll
      for (int c = r + 1; c < 6; c++) acc = csub(acc, cmulc(L[c][r], x[c]));
      x[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < 6; r++) Ainv[col][r] = cconj(x[r]);
  }
  Ci[base] = make_double2(0.5 * Ainv[0][0].x, 0.5 * Ainv[1][1].x);
  Ci[base + (size_t)10 * vh] = make_double2(0.5 * Ainv[2][2].x, 0.5 * Ainv[3][3].x);
  Ci[base + (size_t)16 * vh] = make_double2(0.5 * Ainv[4][4].x, 0.5 * Ainv[5][5].x);
#pragma unroll
  for (int r = 0; r < 6; r++)
#pragma unroll
    for (int c = r + 1; c < 6; c++)
      Ci[base + (size_t)a_off_word(r, c) * vh] = cscale<double2>(0.5, Ainv[r][c]);
}

// clover_init (clover/clover_init.F90:28-189): one thread per site of one parity
__global__ __launch_bounds__(128) void k_clover_build(const double2 *__restrict__ U, Geom g, int e,
                                                      double factor, double2 *__restrict__ Ca,
                                                      double2 *__restrict__ Ci,
                                                      double2 *__restrict__ Cb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.vh) return;
  int x[4];
  site_coords(g, e, i, x[0], x[1], x[2], x[3]);
  double2 P[6][6], Q[6][6];
  {
    M3 f = f_mu_nu(U, g, e, i, x, 1, 0);  // f_21
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        P[r][c] = f.m[3 * r + c];
        P[r + 3][c + 3] = make_double2(-f.m[3 * r + c].x, -f.m[3 * r + c].y);
      }
    M3 f32 = f_mu_nu(U, g, e, i, x, 2, 1);
    M3 f31 = f_mu_nu(U, g, e, i, x, 2, 0);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) P[r][c + 3] = cadd(f32.m[3 * r + c], ci(f31.m[3 * r + c]));
  }
  {
    M3 f = f_mu_nu(U, g, e, i, x, 2, 3);  // f_34
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        Q[r][c] = f.m[3 * r + c];
        Q[r + 3][c + 3] = make_double2(-f.m[3 * r + c].x, -f.m[3 * r + c].y);
      }
    M3 f14 = f_mu_nu(U, g, e, i, x, 0, 3);
    M3 f42 = f_mu_nu(U, g, e, i, x, 3, 1);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) Q[r][c + 3] = cadd(f14.m[3 * r + c], ci(f42.m[3 * r + c]));
  }
  // A_1 = 1 + s (P + Q), A_2 = 1 + s (P - Q) on the upper triangle
  double2 A[6][6];
#pragma unroll
  for (int J = 0; J < 2; J++) {
    const double sg = J == 0 ? 1.0 : -1.0;
#pragma unroll
    for (int r = 0; r < 6; r++)
#pragma unroll
      for (int c = r; c < 6; c++) {
        double2 v = cscale<double2>(factor, make_double2(P[r][c].x + sg * Q[r][c].x, P[r][c].y + sg * Q[r][c].y));
        if (r == c) v = make_double2(v.x + 1.0, 0.0);
        A[r][c] = v;
      }
    const size_t base = (size_t)(e * 36 + J * 18) * g.vh + i;
    clover_block_store(A, Ca, Ci, Cb, base, g.vh);
  }
}

// ---------------------------------------------------------------------------
// plaquette sum: sum over sites and planes mu<nu of Re Tr U_p
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_plaquette(const double2 *__restrict__ U, Geom g,
                                                      double *partials) {
  int tid = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (tid < 2 * g.vh) {
    int e = tid >= g.vh, i = tid - e * g.vh, o = 1 - e;
    int x[4];
    site_coords(g, e, i, x[0], x[1], x[2], x[3]);
    for (int mu = 0; mu < 4; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        int j1 = nb(g, x[0], x[1], x[2], x[3], mu, 1), j2 = nb(g, x[0], x[1], x[2], x[3], nu, 1);
        M3 p = m3_muld(m3_muld(m3_mul(link(U, mu, e, i, g.vh), link(U, nu, o, j1, g.vh)),
                               link(U, mu, o, j2, g.vh)),
                       link(U, nu, e, i, g.vh));
        s += m3_retr(p);
      }
  }
  double vals[1] = {s};
  block_partials<1>(vals, partials);
}

// ---------------------------------------------------------------------------
// synthetic gauge fields
// ---------------------------------------------------------------------------
// scale > 0: U = exp(i scale sum_j w_j lambda_j), w_j ~ N(0,1);
// scale <= 0: Haar-random SU(3) (Gram-Schmidt of a complex Gaussian matrix)
__global__ void k_gauge_random(double2 *__restrict__ U, Geom g, unsigned long long seed, double scale) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int me = blockIdx.y;
  if (i >= g.vh) return;
  Rng rng(seed, (uint64_t)me * g.vh + i);
  M3 u;
  if (scale > 0.0) {
    double h[8];
    for (int j = 0; j < 8; j += 2) {
      rng.gauss2(h[j], h[j + 1]);
      h[j] *= scale;
      h[j + 1] *= scale;
    }
    u = m3_expi(gellmann(h));
  } else {
    for (int k = 0; k < 9; k++) {
      double a, b;
      rng.gauss2(a, b);
      u.m[k] = make_double2(a, b);
    }
    m3_reunitarize(u);
  }
  store_link(U, me >> 1, me & 1, i, g.vh, u);
}

// one APE sweep: U' = P_SU3[(1 - alpha) U + alpha/6 sum_nu (staples along U)]
__global__ __launch_bounds__(128) void k_ape(double2 *__restrict__ Unew, const double2 *__restrict__ U,
                                             Geom g, double alpha) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int me = blockIdx.y;
  if (i >= g.vh) return;
  const int mu = me >> 1, e = me & 1, o = 1 - e, vh = g.vh;
  int x[4];
  site_coords(g, e, i, x[0], x[1], x[2], x[3]);
  M3 s = m3_zero();
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    int xpn = nb(g, x[0], x[1], x[2], x[3], nu, 1), xpm = nb(g, x[0], x[1], x[2], x[3], mu, 1);
    int xmn = nb(g, x[0], x[1], x[2], x[3], nu, -1);
    int xmnpm = nb2(g, x[0], x[1], x[2], x[3], nu, -1, mu, 1);
    // U_nu(x) U_mu(x+nu) U_nu(x+mu)^+
    m3_acc(s, m3_muld(m3_mul(link(U, nu, e, i, vh), link(U, mu, o, xpn, vh)), link(U, nu, o, xpm, vh)));
    // U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu)
    m3_acc(s, m3_mul(m3_dmul(link(U, nu, o, xmn, vh), link(U, mu, o, xmn, vh)), link(U, nu, e, xmnpm, vh)));
  }
  M3 u = link(U, mu, e, i, vh);
  m3_scale(u, 1.0 - alpha);
  m3_scale(s, alpha / 6.0);
  m3_acc(u, s);
  m3_reunitarize(u);
  store_link(Unew, mu, e, i, vh, u);
}

int build_clover_into(Context *c, double csw_kappa, double2 *Ca, double2 *Ci, double2 *Cb) {
  for (int e = 0; e < 2; e++) {
    k_clover_build<<<nblocks(c->g.vh, 128), 128, 0, c->stream>>>(c->gauge, c->g, e, -csw_kappa / 8.0,
                                                                 Ca, Ci, Cb);
    HB_HIP(hipGetLastError());
  }
  return HB_OK;
}

int build_clover_dev(Context *c, double csw_kappa) {
  HB_TRY(alloc_fp64_fields(c));
  if (csw_kappa == 0.0) {
    for (int k = 0; k < 3; k++) c->have_clover[k] = false;
    c->clover_version++;
    return HB_OK;
  }
  HB_TRY(build_clover_into(c, csw_kappa, c->clov[HB_CLOVER_A], c->clov[HB_CLOVER_AINV],
                           c->clov[HB_CLOVER_B]));
  for (int k = 0; k < 3; k++) c->have_clover[k] = true;
  c->clover_version++;
  return HB_OK;
}

int plaquette_dev(Context *c, double *res_d) {
  int n = 2 * c->g.vh;
  k_plaquette<<<nblocks(n), kBlock, 0, c->stream>>>(c->gauge, c->g, c->red);
  k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(c->red, nblocks(n), nullptr, FinStore1{res_d});
  HB_HIP(hipGetLastError());
  return HB_OK;
}

}  // namespace hb

using namespace hb;

extern "C" {

int hb_build_clover(hb_context *ctx, double csw_kappa) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_HIP(hipSetDevice(c->device));
  HB_TRY(build_clover_dev(c, csw_kappa));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

int hb_plaquette(hb_context *ctx, double *re_tr_sum) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(re_tr_sum, HB_E_ARG, "hb_plaquette: null pointer");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double *r = static_cast<double *>(scratch(c, SLOT_STATE, sizeof(CgState), &rc));
  HB_TRY(rc);
  HB_TRY(plaquette_dev(c, r));
  HB_HIP(hipMemcpyAsync(re_tr_sum, r, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

int hb_gauge_random(hb_context *ctx, unsigned long long seed, double scale) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c, HB_E_ARG, "hb_gauge_random: null context");
  HB_HIP(hipSetDevice(c->device));
  HB_TRY(alloc_fp64_fields(c));
  k_gauge_random<<<dim3(nblocks(c->g.vh), 8), kBlock, 0, c->stream>>>(c->gauge, c->g, seed, scale);
  HB_HIP(hipGetLastError());
  HB_HIP(hipStreamSynchronize(c->stream));
  c->have_gauge = true;
  c->gauge_version++;
  return HB_OK;
}

int hb_gauge_smear(hb_context *ctx, int sweeps, double alpha) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_HIP(hipSetDevice(c->device));
  double2 *tmp = nullptr;
  HB_HIP(hipMalloc(&tmp, sizeof(double2) * 72 * (size_t)c->g.vh));
  for (int s = 0; s < sweeps; s++) {
    k_ape<<<dim3(nblocks(c->g.vh, 128), 8), 128, 0, c->stream>>>(tmp, c->gauge, c->g, alpha);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(c->gauge, tmp, sizeof(double2) * 72 * (size_t)c->g.vh,
                                            hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) {
      (void)hipFree(tmp);
      set_error("hb_gauge_smear: %s", hipGetErrorString(e));
      return HB_E_HIP;
    }
  }
  HB_HIP(hipStreamSynchronize(c->stream));
  HB_HIP(hipFree(tmp));
  c->gauge_version++;
  return HB_OK;
}

}  // extern "C"
