// hb_blas.hip - spinor BLAS-1 with single-pass grid reductions, standalone
// clover multiplies, and precision conversions of spinors, links and clover.
#include "hb_blas.h"

namespace hb {

// ---------------------------------------------------------------------------
// standalone clover multiplies (clover/clover_mult_a.F90, _ao, _b)
// ---------------------------------------------------------------------------
template <class S, int MODE>
__global__ __launch_bounds__(kBlock) void k_clover_mult(SpinorPtr<S> out, SpinorPtr<S> in,
                                                        const typename S::ST *C, const float *Cn,
                                                        int eo, int vh) {
  using V = typename S::V;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  V x[12], y[12];
  load_spinor<S>(x, in, i, vh);
  clover_apply<S, MODE>(y, x, C, Cn, eo, i, vh);
  store_spinor<S>(out, i, vh, y);
}

int clover_mult_f64(Context *c, int which, int eo, double2 *out, const double2 *in) {
  SpinorPtr<SF64> o{out, nullptr}, s{const_cast<double2 *>(in), nullptr};
  if (which == HB_CLOVER_B)
    k_clover_mult<SF64, 1><<<nblocks(c->g.vh), kBlock, 0, c->stream>>>(o, s, c->clov[which], nullptr, eo, c->g.vh);
  else
    k_clover_mult<SF64, 0><<<nblocks(c->g.vh), kBlock, 0, c->stream>>>(o, s, c->clov[which], nullptr, eo, c->g.vh);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

// ---------------------------------------------------------------------------
// FP64 CG vector kernels (cg.F90:81-116); scalars live in CgState on device
// ---------------------------------------------------------------------------
// r = b - r, p = r, rtr_old = |r|^2  (cg.F90:84-89)
__global__ __launch_bounds__(kBlock) void k_cg_start(double2 *__restrict__ r, double2 *__restrict__ p,
                                                     const double2 *__restrict__ b, int n,
                                                     CgState *st, double *partials, unsigned *cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  for (int k = i; k < n; k += gridDim.x * blockDim.x) {
    double2 bb = b[k], rr = r[k];
    double2 v = make_double2(bb.x - rr.x, bb.y - rr.y);
    r[k] = v;
    p[k] = v;
    s += v.x * v.x + v.y * v.y;
  }
  double vals[1] = {s};
  grid_reduce<1>(vals, partials, cnt, [st](const double *t) {
    st->rtr_old = t[0];
    st->rtr = t[0];
    st->iter = 0;
    st->done = 0;
  });
}

// alpha = rtr_old / pAp; x += alpha p; r -= alpha q; rtr = |r|^2; convergence;
// beta = rtr / rtr_old (cg.F90:96-115)
__global__ __launch_bounds__(kBlock) void k_cg_update(double2 *__restrict__ x, double2 *__restrict__ r,
                                                      const double2 *__restrict__ p,
                                                      const double2 *__restrict__ q, int n,
                                                      CgState *st, double *partials, unsigned *cnt) {
  if (st->done) return;
  const double alpha = st->rtr_old / st->pap;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  for (int k = i; k < n; k += gridDim.x * blockDim.x) {
    double2 pp = p[k], qq = q[k], xx = x[k], rr = r[k];
    xx.x += alpha * pp.x; xx.y += alpha * pp.y;
    rr.x -= alpha * qq.x; rr.y -= alpha * qq.y;
    x[k] = xx;
    r[k] = rr;
    s += rr.x * rr.x + rr.y * rr.y;
  }
  double vals[1] = {s};
  grid_reduce<1>(vals, partials, cnt, [st](const double *t) {
    double rtr = t[0];
    st->iter += 1;
    st->rtr = rtr;
    if (rtr <= st->rest) {
      st->done = 1;
    } else if (st->iter >= st->maxiter) {
      st->done = 2;
    } else {
      st->beta = rtr / st->rtr_old;
      st->rtr_old = rtr;
    }
  });
}

// p = r + beta p  (sc_xpby, sc.F90:184-201)
__global__ __launch_bounds__(kBlock) void k_cg_pupdate(double2 *__restrict__ p, const double2 *__restrict__ r,
                                                       int n, const CgState *st) {
  if (st->done) return;
  const double beta = st->beta;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = i; k < n; k += gridDim.x * blockDim.x) {
    double2 pp = p[k], rr = r[k];
    p[k] = make_double2(rr.x + beta * pp.x, rr.y + beta * pp.y);
  }
}

// out = b - out, result |out|^2 -> *res
__global__ __launch_bounds__(kBlock) void k_residual(double2 *__restrict__ out, const double2 *__restrict__ b,
                                                     int n, double *res, double *partials, unsigned *cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  for (int k = i; k < n; k += gridDim.x * blockDim.x) {
    double2 bb = b[k], o = out[k];
    double2 v = make_double2(bb.x - o.x, bb.y - o.y);
    out[k] = v;
    s += v.x * v.x + v.y * v.y;
  }
  double vals[1] = {s};
  grid_reduce<1>(vals, partials, cnt, [res](const double *t) { res[0] = t[0]; });
}

// Re<x, y> and |x|^2 of FP64 fields
__global__ __launch_bounds__(kBlock) void k_dot_norm(const double2 *__restrict__ x, const double2 *__restrict__ y,
                                                     int n, double *res, double *partials, unsigned *cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  for (int k = i; k < n; k += gridDim.x * blockDim.x) {
    double2 a = x[k], b = y[k];
    s0 += a.x * b.x + a.y * b.y;
    s1 += a.x * a.x + a.y * a.y;
  }
  double vals[2] = {s0, s1};
  grid_reduce<2>(vals, partials, cnt, [res](const double *t) { res[0] = t[0]; res[1] = t[1]; });
}

__global__ void k_zero(double2 *p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t k = i; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = make_double2(0.0, 0.0);
}

__global__ void k_copy(double2 *__restrict__ d, const double2 *__restrict__ s, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t k = i; k < n; k += (size_t)gridDim.x * blockDim.x) d[k] = s[k];
}

int reduce_grid(const Context *c, int n) {
  int b = nblocks(n);
  int cap = c->red_blocks;
  return b < cap ? b : cap;
}

int cg_start(Context *c, double2 *r, double2 *p, const double2 *b, CgState *st) {
  int n = 12 * c->g.vh;
  k_cg_start<<<reduce_grid(c, n), kBlock, 0, c->stream>>>(r, p, b, n, st, c->red, c->red_count);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int cg_update(Context *c, double2 *x, double2 *r, const double2 *p, const double2 *q, CgState *st) {
  int n = 12 * c->g.vh;
  k_cg_update<<<reduce_grid(c, n), kBlock, 0, c->stream>>>(x, r, p, q, n, st, c->red, c->red_count + 1);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int cg_pupdate(Context *c, double2 *p, const double2 *r, const CgState *st) {
  int n = 12 * c->g.vh;
  k_cg_pupdate<<<nblocks(n), kBlock, 0, c->stream>>>(p, r, n, st);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int residual(Context *c, double2 *out, const double2 *b, double *res_d) {
  int n = 12 * c->g.vh;
  k_residual<<<reduce_grid(c, n), kBlock, 0, c->stream>>>(out, b, n, res_d, c->red, c->red_count + 2);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int dot_norm(Context *c, const double2 *x, const double2 *y, double *res_d) {
  int n = 12 * c->g.vh;
  k_dot_norm<<<reduce_grid(c, n), kBlock, 0, c->stream>>>(x, y, n, res_d, c->red, c->red_count + 3);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int zero_f64(Context *c, double2 *p, size_t n) {
  k_zero<<<nblocks((int)std::min<size_t>(n, (size_t)1 << 24)), kBlock, 0, c->stream>>>(p, n);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int copy_f64(Context *c, double2 *d, const double2 *s, size_t n) {
  k_copy<<<nblocks((int)std::min<size_t>(n, (size_t)1 << 24)), kBlock, 0, c->stream>>>(d, s, n);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

}  // namespace hb
