// hb_plugin.hip - entry points behind BQCD's link-time plugin libraries, one
// per routine the reference links from them, with the reference's host
// layouts:
//   LIBD      (d/Makefile:113-129): the per-direction hopping-force pieces
//             dsf_{x,y,z,t}{f,b} (d/DSFxyzt.F90:77-117); d and d_dag are
//             hb_dslash on links loaded as given
//   LIBCLOVER (clover/Makefile:44-71): clover_mult_a / _ao / _b on one parity
//             of a clover field, clover_dsf, clover_dsd, clover_action
//             (clover_init is hb_load_gauge + hb_build_clover + hb_save_clover)
// Every call uploads its operands and downloads its result; the fields stay
// owned by the caller.
#include "hb_ops.h"

namespace hb {

namespace {

// one parity of a BQCD clover field (18 words, 2 blocks, volh) into parity 0
// of a device field [2][36][vh]
__global__ void k_clover_parity_to_soa(double2 *__restrict__ dst, const double2 *__restrict__ src, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const double2 *s = src + (size_t)i * 36;
  for (int w = 0; w < 36; w++) dst[(size_t)w * vh + i] = s[w];
}

__device__ __forceinline__ double2 ld_aos(const double2 *f, int site, int spin, int col) {
  return f[((size_t)site * 3 + col) * 4 + spin];
}

// The 2x2 blocks of gamma_mu = [[0, G], [G^+, 0]] for x, y, z in the gamma
// basis of the hopping matrix (d/D21.F90:79-182): G_x = [[0, i], [i, 0]],
// G_y = [[0, 1], [-1, 0]], G_z = [[i, 0], [0, -i]].
__device__ __forceinline__ double2 gblock(int mu, int r, int c) {
  if (mu == 0) return (r != c) ? make_double2(0.0, 1.0) : make_double2(0.0, 0.0);
  if (mu == 1) return r == c ? make_double2(0.0, 0.0) : make_double2(r == 0 ? 1.0 : -1.0, 0.0);
  return r != c ? make_double2(0.0, 0.0) : make_double2(0.0, r == 0 ? 1.0 : -1.0);
}

// (1 - sg gamma_mu) a at one site, all four spin components, one colour;
// for t the projector keeps spins 3, 4 (sg = +1) or 1, 2 (sg = -1), doubled
__device__ __forceinline__ void proj_full(double2 (&ga)[4], const double2 *a, int site, int col, int mu,
                                          double sg) {
  if (mu == 3) {
    const int up = sg > 0 ? 2 : 0;
    for (int s = 0; s < 4; s++) ga[s] = make_double2(0.0, 0.0);
    ga[up] = cscale<double2>(2.0, ld_aos(a, site, up, col));
    ga[up + 1] = cscale<double2>(2.0, ld_aos(a, site, up + 1, col));
    return;
  }
  const double2 a2 = ld_aos(a, site, 2, col), a3 = ld_aos(a, site, 3, col);
  double2 h[2];
  for (int k = 0; k < 2; k++) {
    double2 g = cmul(gblock(mu, k, 0), a2);
    cfma(g, gblock(mu, k, 1), a3);
    h[k] = csub(ld_aos(a, site, k, col), cscale<double2>(sg, g));
  }
  ga[0] = h[0];
  ga[1] = h[1];
  for (int k = 0; k < 2; k++) {  // lower half: -sg G^+ h
    double2 g = cmulc(gblock(mu, 0, k), h[0]);
    cfmac(g, gblock(mu, 1, k), h[1]);
    ga[2 + k] = cscale<double2>(-sg, g);
  }
}

// dsf_{mu}{f,b}: for each site i of the parity, j = nn[i] - 1;
//   forward:  w_ab = sum_s ((1 - g_mu) a_j)_{s,a} conj(b_i)_{s,b}, v = U_i w, p_i += -s Im Tr(lambda v)
//   backward: w_ab = sum_s ((1 + g_mu) a_i)_{s,a} conj(b_j)_{s,b}, v = w U_i^+, p_i += s Im Tr(lambda v)
__global__ __launch_bounds__(128) void k_dsf_dir(double *__restrict__ p, const double2 *__restrict__ b,
                                                 const double2 *__restrict__ a, const double2 *__restrict__ u,
                                                 const int *__restrict__ nn, int vh, int mu, int fwd, double s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const int j = nn[i] - 1;
  const int ia = fwd ? j : i, ib = fwd ? i : j;
  double2 w[3][3];
  for (int ca = 0; ca < 3; ca++) {
    double2 ga[4];
    proj_full(ga, a, ia, ca, mu, fwd ? 1.0 : -1.0);
    for (int cb = 0; cb < 3; cb++) {
      double2 acc = make_double2(0.0, 0.0);
      for (int sp = 0; sp < 4; sp++) cfma(acc, ga[sp], cconj(ld_aos(b, ib, sp, cb)));
      w[ca][cb] = acc;
    }
  }
  // U_i: u(row, col, i) at i*9 + col*3 + row
  const double2 *U = u + (size_t)i * 9;
  double2 v[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      double2 acc = make_double2(0.0, 0.0);
      for (int k = 0; k < 3; k++) {
        if (fwd) cfma(acc, U[k * 3 + r], w[k][c]);              // (U w)_rc
        else cfma(acc, w[r][k], cconj(U[k * 3 + c]));         // (w U^+)_rc
      }
      v[r][c] = acc;
    }
  // p_j += t Im Tr(lambda_j v) (su3.F90:196-213)
  const double t = fwd ? -s : s;
  double *q = p + (size_t)i * 8;
  q[0] += t * (v[0][1].y + v[1][0].y);
  q[1] += t * (v[0][1].x - v[1][0].x);
  q[2] += t * (v[0][0].y - v[1][1].y);
  q[3] += t * (v[0][2].y + v[2][0].y);
  q[4] += t * (v[0][2].x - v[2][0].x);
  q[5] += t * (v[1][2].y + v[2][1].y);
  q[6] += t * (v[1][2].x - v[2][1].x);
  q[7] += t * (v[0][0].y + v[1][1].y - 2.0 * v[2][2].y) / 1.7320508075688772;
}

}  // namespace

}  // namespace hb

using namespace hb;

extern "C" {

int hb_dsf_dir(hb_context *ctx, int mu, int forward, double *p, const double *b, const double *a, double s,
               const double *u, const int *nn, int n_tot) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c && p && b && a && u && nn, HB_E_ARG, "hb_dsf_dir: null argument");
  HB_CHECK(mu >= 0 && mu < 4, HB_E_ARG, "hb_dsf_dir: direction %d", mu);
  const int vh = c->g.vh;
  HB_CHECK(n_tot >= vh, HB_E_ARG, "hb_dsf_dir: n_tot %d < volh %d", n_tot, vh);
  // every neighbour index must address a site of the spinor fields
  for (int i = 0; i < vh; i++)
    HB_CHECK(nn[i] >= 1 && nn[i] <= n_tot, HB_E_ARG, "hb_dsf_dir: nn(%d) = %d outside 1..%d", i + 1, nn[i], n_tot);
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  const size_t sp_bytes = sizeof(double2) * 12 * (size_t)n_tot;
  double2 *bd = static_cast<double2 *>(scratch(c, SLOT_PLUGIN + 0, sp_bytes, &rc)); HB_TRY(rc);
  double2 *ad = static_cast<double2 *>(scratch(c, SLOT_PLUGIN + 1, sp_bytes, &rc)); HB_TRY(rc);
  double2 *ud = static_cast<double2 *>(scratch(c, SLOT_PLUGIN + 2, sizeof(double2) * 9 * (size_t)vh, &rc)); HB_TRY(rc);
  int *nnd = static_cast<int *>(scratch(c, SLOT_PLUGIN + 3, sizeof(int) * (size_t)vh, &rc)); HB_TRY(rc);
  double *pd = static_cast<double *>(scratch(c, SLOT_PLUGIN + 4, sizeof(double) * 8 * (size_t)vh, &rc)); HB_TRY(rc);
  HB_HIP(hipMemcpyAsync(bd, b, sp_bytes, hipMemcpyHostToDevice, c->stream));
  HB_HIP(hipMemcpyAsync(ad, a, sp_bytes, hipMemcpyHostToDevice, c->stream));
  HB_HIP(hipMemcpyAsync(ud, u, sizeof(double2) * 9 * (size_t)vh, hipMemcpyHostToDevice, c->stream));
  HB_HIP(hipMemcpyAsync(nnd, nn, sizeof(int) * (size_t)vh, hipMemcpyHostToDevice, c->stream));
  HB_HIP(hipMemcpyAsync(pd, p, sizeof(double) * 8 * (size_t)vh, hipMemcpyHostToDevice, c->stream));
  k_dsf_dir<<<nblocks(vh, 128), 128, 0, c->stream>>>(pd, bd, ad, ud, nnd, vh, mu, forward != 0, s);
  HB_HIP(hipGetLastError());
  HB_HIP(hipMemcpyAsync(p, pd, sizeof(double) * 8 * (size_t)vh, hipMemcpyDeviceToHost, c->stream));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

int hb_clover_mult_field(hb_context *ctx, int mode, const double *clov, double *out, const double *in) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c && clov && out && in, HB_E_ARG, "hb_clover_mult_field: null argument");
  HB_CHECK(mode == 0 || mode == 1, HB_E_ARG, "hb_clover_mult_field: mode %d (0: A, 1: LDL^+ solve)", mode);
  HB_HIP(hipSetDevice(c->device));
  const int vh = c->g.vh;
  int rc = 0;
  double2 *C = static_cast<double2 *>(scratch(c, SLOT_PLUGIN + 5, sizeof(double2) * 36 * (size_t)vh, &rc)); HB_TRY(rc);
  double2 *di = spinor_scratch(c, SLOT_IN, &rc); HB_TRY(rc);
  double2 *dout = spinor_scratch(c, SLOT_OUT, &rc); HB_TRY(rc);
  HB_TRY(ensure_dev_stage(c, sizeof(double2) * 36 * (size_t)vh));
  HB_HIP(hipMemcpyAsync(c->dev_stage, clov, sizeof(double2) * 36 * (size_t)vh, hipMemcpyHostToDevice, c->stream));
  k_clover_parity_to_soa<<<nblocks(vh), kBlock, 0, c->stream>>>(C, static_cast<const double2 *>(c->dev_stage), vh);
  HB_HIP(hipGetLastError());
  HB_TRY(spinor_upload(c, di, in));
  HB_TRY(clover_mult_parity_f64(c, mode, C, dout, di));
  return spinor_download(c, out, dout);
}

int hb_clover_dsf(hb_context *ctx, int eo, double *p, int site_stride, const double *b, const double *a,
                  double s) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p && b && a, HB_E_ARG, "hb_clover_dsf: null argument");
  HB_CHECK(eo == 0 || eo == 1, HB_E_ARG, "hb_clover_dsf: parity %d", eo);
  HB_CHECK(site_stride >= c->g.vh, HB_E_ARG, "hb_clover_dsf: site_stride < volh");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double *P = mom_scratch(c, &rc); HB_TRY(rc);
  double2 *bd = spinor_scratch(c, SLOT_PLUGIN + 0, &rc); HB_TRY(rc);
  double2 *ad = spinor_scratch(c, SLOT_PLUGIN + 1, &rc); HB_TRY(rc);
  HB_TRY(gen_upload(c, P, p, site_stride));
  HB_TRY(spinor_upload(c, bd, b));
  HB_TRY(spinor_upload(c, ad, a));
  HB_TRY(clover_dsf_dev(c, P, eo, bd, ad, s));
  return gen_download(c, p, P, site_stride);
}

int hb_clover_dsd(hb_context *ctx, int eo, double *p, int site_stride, double s) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p, HB_E_ARG, "hb_clover_dsd: null argument");
  HB_CHECK(eo == 0 || eo == 1, HB_E_ARG, "hb_clover_dsd: parity %d", eo);
  HB_CHECK(site_stride >= c->g.vh, HB_E_ARG, "hb_clover_dsd: site_stride < volh");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double *P = mom_scratch(c, &rc); HB_TRY(rc);
  HB_TRY(gen_upload(c, P, p, site_stride));
  HB_TRY(clover_dsd_dev(c, P, eo, s));
  return gen_download(c, p, P, site_stride);
}

int hb_clover_action_field(hb_context *ctx, const double *b, double *s_det) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c && b && s_det, HB_E_ARG, "hb_clover_action_field: null argument");
  HB_HIP(hipSetDevice(c->device));
  const int vh = c->g.vh;
  int rc = 0;
  double2 *C = static_cast<double2 *>(scratch(c, SLOT_PLUGIN + 5, sizeof(double2) * 36 * (size_t)vh, &rc)); HB_TRY(rc);
  double *res = static_cast<double *>(scratch(c, SLOT_PLUGIN + 4, sizeof(double) * 8 * (size_t)vh, &rc)); HB_TRY(rc);
  HB_TRY(ensure_dev_stage(c, sizeof(double2) * 36 * (size_t)vh));
  HB_HIP(hipMemcpyAsync(c->dev_stage, b, sizeof(double2) * 36 * (size_t)vh, hipMemcpyHostToDevice, c->stream));
  k_clover_parity_to_soa<<<nblocks(vh), kBlock, 0, c->stream>>>(C, static_cast<const double2 *>(c->dev_stage), vh);
  HB_HIP(hipGetLastError());
  HB_TRY(clover_action_parity_dev(c, C, 0, res));
  HB_HIP(hipMemcpyAsync(s_det, res, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

}  // extern "C"
