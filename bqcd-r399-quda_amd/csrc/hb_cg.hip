// hb_cg.hip - even-odd operators (m_tilde.F90, mtdagmt.F90, w_mult.F90) built
// from the fused hopping-matrix kernels, and the FP64 CG (cg.F90:51-137).
#include <chrono>
#include <cmath>
#include <cstring>

#include "hb_blas.h"
#include "hb_dslash.h"
#include "hb_ops.h"

namespace hb {

bool use_clover(const Context *c) { return c->have_clover[HB_CLOVER_A] && c->have_clover[HB_CLOVER_AINV]; }

DslashArgs<SF64> dslash_args64(Context *c, int out_parity, double2 *out, const double2 *in) {
  DslashArgs<SF64> a{};
  a.g = c->g;
  a.out_parity = out_parity;
  a.out = {out, nullptr};
  a.in = {const_cast<double2 *>(in), nullptr};
  a.U = c->gauge;
  a.partials = c->red;
  return a;
}

// half of M~ (or M~^+): t = A_oo^-1 D_oe in ; out = A_ee in - kappa^2 D_eo t + rho in
// (m_tilde.F90:61-71 / :98-108 with w_mult.F90's sc_axpy folded in).  With
// dot_out set, also Re<dotv, out> (sc_dot) in the same pass.  G = SF64 reads
// all 18 reals of each link, G = SF64R the two-row copy U (row 2 rebuilt).
template <class G>
int mtil_f64_g(Context *c, const typename G::ST *U, bool dag, double kappa, double rho, double2 *out,
               const double2 *in, double2 *t, double *dot_out, const double2 *dotv, const int *skip,
               double *red) {
  const bool clov = use_clover(c);
  DslashArgs<SF64, G> a1{};
  a1.g = c->g;
  a1.out_parity = 1;
  a1.out = {t, nullptr};
  a1.in = {const_cast<double2 *>(in), nullptr};
  a1.U = U;
  a1.partials = c->red;
  a1.C = clov ? c->clov[HB_CLOVER_AINV] : nullptr;
  a1.skip = skip;
  HB_TRY(launch_dslash(a1, dag, EPI_CLINV, c->stream));
  DslashArgs<SF64, G> a2 = a1;
  a2.out_parity = 0;
  a2.out = {out, nullptr};
  a2.in = {t, nullptr};
  a2.C = clov ? c->clov[HB_CLOVER_A] : nullptr;
  a2.x0 = {const_cast<double2 *>(in), nullptr};
  a2.mkappa2 = -kappa * kappa;
  a2.rho = rho;
  a2.skip = skip;
  if (dot_out) {
    a2.dotv = {const_cast<double2 *>(dotv), nullptr};
    a2.dot_out = dot_out;
    if (red) a2.partials = red;
    HB_TRY(launch_dslash(a2, dag, EPI_MTDOT, c->stream));
  } else {
    HB_TRY(launch_dslash(a2, dag, EPI_MTIL, c->stream));
  }
  return HB_OK;
}

int mtil_f64(Context *c, bool dag, double kappa, double rho, double2 *out, const double2 *in,
             double2 *t, double *dot_out, const double2 *dotv, const int *skip, double *red, bool recon) {
  if (recon)
    return mtil_f64_g<SF64R>(c, c->gauge64r, dag, kappa, rho, out, in, t, dot_out, dotv, skip, red);
  return mtil_f64_g<SF64>(c, c->gauge, dag, kappa, rho, out, in, t, dot_out, dotv, skip, red);
}

// rows 0 and 1 of the FP64 links; the largest |row 2 - conj(row0 x row1)| of
// any link goes to *dev (as float bits, non-negative: integer max = float max)
__global__ __launch_bounds__(kBlock) void k_links_two_rows(double2 *__restrict__ dst,
                                                           const double2 *__restrict__ src, int vh,
                                                           unsigned int *dev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int slab = blockIdx.y;  // mu*2 + eo
  float m = 0.f;
  if (i < vh) {
    const double2 *s = src + (size_t)slab * 9 * vh + i;
    double2 u[9];
#pragma unroll
    for (int k = 0; k < 9; k++) u[k] = s[(size_t)k * vh];
#pragma unroll
    for (int k = 0; k < 6; k++) dst[((size_t)slab * 6 + k) * vh + i] = u[k];
    auto cross = [](double2 a0, double2 b1, double2 a1, double2 b0) {  // conj(a0 b1 - a1 b0)
      double2 t = cmul(a0, b1), q = cmul(a1, b0);
      return make_double2(t.x - q.x, -(t.y - q.y));
    };
    const double2 r[3] = {cross(u[1], u[5], u[2], u[4]), cross(u[2], u[3], u[0], u[5]),
                          cross(u[0], u[4], u[1], u[3])};
#pragma unroll
    for (int k = 0; k < 3; k++)
      m = fmaxf(m, (float)fmax(fabs(r[k].x - u[6 + k].x), fabs(r[k].y - u[6 + k].y)));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(dev, __float_as_uint(m));
}

// Brings the two-row copy of the links up to date (allocating it on first
// use).  Returns in *use whether the FP64 CG may iterate with it: the context
// asks for 12-real links and every link is SU(3) to 1e-12, so that the
// rebuilt row 2 equals the stored one to rounding.
int ensure_links_two_rows(Context *c, bool *use) {
  *use = false;
  if (c->link_recon != 12) return HB_OK;
  if (c->r64_gauge_version != c->gauge_version) {
    const size_t vh = c->g.vh;
    if (!c->gauge64r) HB_HIP(hipMalloc(&c->gauge64r, sizeof(double2) * 48 * vh));
    unsigned int *dev = reinterpret_cast<unsigned int *>(c->red);
    HB_TRY(zero_f64(c, reinterpret_cast<double2 *>(c->red), 1));
    k_links_two_rows<<<dim3(nblocks((int)vh), 8), kBlock, 0, c->stream>>>(c->gauge64r, c->gauge, (int)vh, dev);
    HB_HIP(hipGetLastError());
    unsigned int bits = 0;
    HB_HIP(hipMemcpyAsync(&bits, dev, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HB_HIP(hipStreamSynchronize(c->stream));
    float d = 0.f;
    std::memcpy(&d, &bits, sizeof(d));
    c->r64_ok = d <= 1e-12f;
    c->r64_gauge_version = c->gauge_version;
  }
  *use = c->r64_ok;
  return HB_OK;
}

int apply_f64(Context *c, int op, double kappa, double rho, double2 *out, const double2 *in,
              double2 *t, double2 *y, double *dot_out, const int *skip, double *red, bool recon) {
  switch (op) {
    case HB_OP_MTIL: return mtil_f64(c, false, kappa, 0.0, out, in, t, nullptr, nullptr, skip, nullptr, recon);
    case HB_OP_MTIL_DAG: return mtil_f64(c, true, kappa, 0.0, out, in, t, nullptr, nullptr, skip, nullptr, recon);
    case HB_OP_W: return mtil_f64(c, false, kappa, rho, out, in, t, nullptr, nullptr, skip, nullptr, recon);
    case HB_OP_W_DAG: return mtil_f64(c, true, kappa, rho, out, in, t, nullptr, nullptr, skip, nullptr, recon);
    case HB_OP_MTDAGMT:
    case HB_OP_WDAGW: {
      double r = op == HB_OP_WDAGW ? rho : 0.0;
      HB_TRY(mtil_f64(c, false, kappa, r, y, in, t, nullptr, nullptr, skip, nullptr, recon));
      return mtil_f64(c, true, kappa, r, out, y, t, dot_out, in, skip, red, recon);
    }
  }
  set_error("unknown operator %d", op);
  return HB_E_ARG;
}

int check_ready(const Context *c) {
  HB_CHECK(c, HB_E_ARG, "null context");
  HB_CHECK(c->have_gauge, HB_E_STATE, "no gauge field loaded (hb_load_gauge)");
  return HB_OK;
}

size_t spinor_bytes(const Context *c) { return sizeof(double2) * 12 * (size_t)c->g.vh; }

double2 *spinor_scratch(Context *c, int slot, int *rc) {
  return static_cast<double2 *>(scratch(c, slot, spinor_bytes(c), rc));
}

// The FP64 CG iteration on device fields, with its work fields given: x is
// the initial guess in and the solution out.  `hs` is host memory the state
// is read back into between chunks of iterations.
int cg_f64_run(Context *c, const CgWork &w, int op, double kappa, double rho, double2 *x,
               const double2 *b, float rest, int maxiter, int log, int *iters, double *true_rtr,
               int *n_matvec) {
  HB_CHECK(op == HB_OP_MTDAGMT || op == HB_OP_WDAGW, HB_E_ARG,
           "cg: operator must be HB_OP_MTDAGMT or HB_OP_WDAGW (got %d)", op);
  HB_CHECK(maxiter >= 0, HB_E_ARG, "cg: maxiter < 0");
  CgState *st = w.st, *hs = w.hs;
  HB_TRY(cg_init(c, st, (double)rest, maxiter));
  // r = A x; r = b - r; p = r; rtr_old = |r|^2
  HB_TRY(apply_f64(c, op, kappa, rho, w.r, x, w.t, w.y, nullptr, nullptr, nullptr, w.recon));
  HB_TRY(cg_start(c, w.r, w.p, b, st, w.red));
  int *skip = &st->done;
  int nmv = 1;
  if (maxiter == 0) {
    // the reference's loop body never runs: niter = 0, x unchanged
    HB_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HB_HIP(hipStreamSynchronize(c->stream));
  } else {
    // iterations in chunks; every kernel returns immediately once done != 0, so
    // the chunk length only sets how often the host looks at the state
    int chunk = 8;
    for (;;) {
      for (int k = 0; k < chunk; k++) {
        HB_TRY(apply_f64(c, op, kappa, rho, w.q, w.p, w.t, w.y, &st->pap, skip, w.red, w.recon));
        HB_TRY(cg_update(c, w.r, w.q, st, w.red));
        HB_TRY(cg_pupdate(c, x, w.p, w.r, st));
      }
      HB_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
      HB_HIP(hipStreamSynchronize(c->stream));
      if (hs->done) break;
      if (chunk < 64) chunk *= 2;
    }
    nmv += hs->iter;
  }
  const int niter = hs->iter;
  const int done = hs->done;
  const double rtr = hs->rtr;
  // true residual |b - A x|^2 in FP64 (invert_quda.F90:155-168), with all
  // 18 reals of each link
  HB_TRY(apply_f64(c, op, kappa, rho, w.q, x, w.t, w.y, nullptr, nullptr));
  HB_TRY(residual(c, w.q, b, &st->res[0], w.red));
  HB_HIP(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
  HB_HIP(hipStreamSynchronize(c->stream));
  if (n_matvec) *n_matvec = nmv + 1;
  if (iters) *iters = niter;
  if (true_rtr) *true_rtr = hs->res[0];
  if (maxiter > 0 && done != 1 && log != 2) {
    set_error("cg(): no convergence; rtr = %.17g", rtr);
    return HB_E_NOCONV;
  }
  return HB_OK;
}

// FP64 CG on device fields with the context's scratch slots, timed.
int cg_f64(Context *c, int op, double kappa, double rho, double2 *x, const double2 *b, float rest,
           int maxiter, int log, int *iters, double *true_rtr) {
  int rc = 0;
  CgWork w;
  w.t = spinor_scratch(c, SLOT_T, &rc); HB_TRY(rc);
  w.y = spinor_scratch(c, SLOT_Y, &rc); HB_TRY(rc);
  w.r = spinor_scratch(c, SLOT_R, &rc); HB_TRY(rc);
  w.p = spinor_scratch(c, SLOT_P, &rc); HB_TRY(rc);
  w.q = spinor_scratch(c, SLOT_Q, &rc); HB_TRY(rc);
  w.st = static_cast<CgState *>(scratch(c, SLOT_STATE, sizeof(CgState), &rc)); HB_TRY(rc);
  w.red = c->red;
  HB_TRY(ensure_host_stage(c, sizeof(CgState)));
  w.hs = static_cast<CgState *>(c->host_stage);
  HB_TRY(ensure_links_two_rows(c, &w.recon));

  EventTimer tm;
  HB_TRY(tm.create());
  HB_TRY(tm.start(c->stream));
  int nmv = 0;
  int r = cg_f64_run(c, w, op, kappa, rho, x, b, rest, maxiter, log, iters, true_rtr, &nmv);
  if (r != HB_OK && r != HB_E_NOCONV) return r;
  HB_TRY(tm.stop(c->stream));
  float ms = 0.f;
  HB_TRY(tm.elapsed(&ms));
  c->last_solve_ms = ms;
  c->last_n_matvec = nmv;
  return r;
}

}  // namespace hb

using namespace hb;

extern "C" {

int hb_dslash_d(hb_context *ctx, int e, int dagger, double *out_d, const double *in_d) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(e == 0 || e == 1, HB_E_ARG, "hb_dslash: parity %d", e);
  HB_CHECK(out_d && in_d && out_d != in_d, HB_E_ARG, "hb_dslash: bad pointers");
  DslashArgs<SF64> a = dslash_args64(c, e, reinterpret_cast<double2 *>(out_d),
                                     reinterpret_cast<const double2 *>(in_d));
  return launch_dslash(a, dagger != 0, EPI_PLAIN, c->stream);
}

int hb_dslash(hb_context *ctx, int e, int dagger, double *out, const double *in) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(out && in, HB_E_ARG, "hb_dslash: null pointer");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double2 *di = spinor_scratch(c, SLOT_IN, &rc); HB_TRY(rc);
  double2 *dout = spinor_scratch(c, SLOT_OUT, &rc); HB_TRY(rc);
  HB_TRY(spinor_upload(c, di, in));
  HB_TRY(hb_dslash_d(ctx, e, dagger, reinterpret_cast<double *>(dout), reinterpret_cast<double *>(di)));
  return spinor_download(c, out, dout);
}

int hb_clover_mult(hb_context *ctx, int which, int eo, double *out, const double *in) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c && out && in, HB_E_ARG, "hb_clover_mult: null argument");
  HB_CHECK(which >= 0 && which <= 2, HB_E_ARG, "hb_clover_mult: which=%d", which);
  HB_CHECK(eo == 0 || eo == 1, HB_E_ARG, "hb_clover_mult: parity %d", eo);
  HB_CHECK(c->have_clover[which], HB_E_STATE, "hb_clover_mult: clover field %d not loaded", which);
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double2 *di = spinor_scratch(c, SLOT_IN, &rc); HB_TRY(rc);
  double2 *dout = spinor_scratch(c, SLOT_OUT, &rc); HB_TRY(rc);
  HB_TRY(spinor_upload(c, di, in));
  HB_TRY(clover_mult_f64(c, which, eo, dout, di));
  return spinor_download(c, out, dout);
}

int hb_apply_d(hb_context *ctx, int op, double kappa, double rho, double *out_d, const double *in_d) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(out_d && in_d && out_d != in_d, HB_E_ARG, "hb_apply: bad pointers");
  int rc = 0;
  double2 *t = spinor_scratch(c, SLOT_T, &rc); HB_TRY(rc);
  double2 *y = spinor_scratch(c, SLOT_Y, &rc); HB_TRY(rc);
  return apply_f64(c, op, kappa, rho, reinterpret_cast<double2 *>(out_d),
                   reinterpret_cast<const double2 *>(in_d), t, y, nullptr, nullptr);
}

int hb_apply(hb_context *ctx, int op, double kappa, double rho, double *out, const double *in) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(out && in, HB_E_ARG, "hb_apply: null pointer");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double2 *di = spinor_scratch(c, SLOT_IN, &rc); HB_TRY(rc);
  double2 *dout = spinor_scratch(c, SLOT_OUT, &rc); HB_TRY(rc);
  HB_TRY(spinor_upload(c, di, in));
  HB_TRY(hb_apply_d(ctx, op, kappa, rho, reinterpret_cast<double *>(dout), reinterpret_cast<double *>(di)));
  return spinor_download(c, out, dout);
}

int hb_cg_d(hb_context *ctx, int op, double kappa, double rho, double *x_d, const double *b_d,
            float rest, int maxiter, int log, int *iters, double *true_rtr) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(x_d && b_d && x_d != b_d, HB_E_ARG, "hb_cg: bad pointers");
  HB_HIP(hipSetDevice(c->device));
  return cg_f64(c, op, kappa, rho, reinterpret_cast<double2 *>(x_d),
                reinterpret_cast<const double2 *>(b_d), rest, maxiter, log, iters, true_rtr);
}

int hb_cg(hb_context *ctx, int op, double kappa, double rho, double *x, const double *b, float rest,
          int maxiter, int log, int *iters, double *true_rtr) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(x && b, HB_E_ARG, "hb_cg: null pointer");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double2 *dx = spinor_scratch(c, SLOT_X, &rc); HB_TRY(rc);
  double2 *db = spinor_scratch(c, SLOT_B, &rc); HB_TRY(rc);
  HB_TRY(spinor_upload(c, dx, x));
  HB_TRY(spinor_upload(c, db, b));
  int r = cg_f64(c, op, kappa, rho, dx, db, rest, maxiter, log, iters, true_rtr);
  if (r != HB_OK && r != HB_E_NOCONV) return r;
  HB_TRY(spinor_download(c, x, dx));
  return r;
}

}  // extern "C"
