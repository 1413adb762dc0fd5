// hb_blas.hip - spinor BLAS-1 with deterministic grid reductions, standalone
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

// out = A in (mode 0) or the LDL^+ solve (mode 1) with parity 0 of the FP64
// clover field C ([2][36][vh] device layout)
int clover_mult_parity_f64(Context *c, int mode, const double2 *C, double2 *out, const double2 *in) {
  SpinorPtr<SF64> o{out, nullptr}, s{const_cast<double2 *>(in), nullptr};
  if (mode == 1)
    k_clover_mult<SF64, 1><<<nblocks(c->g.vh), kBlock, 0, c->stream>>>(o, s, C, nullptr, 0, c->g.vh);
  else
    k_clover_mult<SF64, 0><<<nblocks(c->g.vh), kBlock, 0, c->stream>>>(o, s, C, nullptr, 0, c->g.vh);
  HB_HIP(hipGetLastError());
  return HB_OK;
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
// Vector kernels: one thread per site, the 12 components of a site in 12
// coalesced planes (component k of site i at k*vh + i).
#define HB_SITE_LOOP for (int k = 0; k < 12; k++)
#define HB_AT(ptr, k) (ptr)[(size_t)(k) * vh + i]
// streaming (non-temporal) store of one complex number, see store_spinor
#ifndef HB_CG_NT_STORE
#define HB_CG_NT_STORE 1
#endif
#if HB_CG_NT_STORE
#define HB_ST(ptr, k, v) SF64::st_nt((ptr), (size_t)(k) * vh + i, (v))
#else
#define HB_ST(ptr, k, v) ((ptr)[(size_t)(k) * vh + i] = (v))
#endif

// r = b - r, p = r, rtr_old = |r|^2  (cg.F90:84-89)
__global__ __launch_bounds__(kBlock) void k_cg_start(double2 *__restrict__ r, double2 *__restrict__ p,
                                                     const double2 *__restrict__ b, int vh,
                                                     double *partials) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < vh) {
#pragma unroll
    HB_SITE_LOOP {
      double2 bb = HB_AT(b, k), rr = HB_AT(r, k);
      double2 v = make_double2(bb.x - rr.x, bb.y - rr.y);
      HB_AT(r, k) = v;
      HB_AT(p, k) = v;
      s += v.x * v.x + v.y * v.y;
    }
  }
  double vals[1] = {s};
  block_partials<1>(vals, partials);
}

// alpha = rtr_old / pAp; r -= alpha q; rtr = |r|^2 (cg.F90:96-109).  The
// x update of the same iteration is done by k_cg_pupdate, which reads p anyway.
__global__ __launch_bounds__(kBlock) void k_cg_update(double2 *__restrict__ r, const double2 *__restrict__ q,
                                                      int vh, const CgState *st, double *partials) {
  if (st->done) return;
  const double alpha = st->rtr_old / st->pap;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < vh) {
#pragma unroll
    HB_SITE_LOOP {
      double2 qq = HB_AT(q, k), rr = HB_AT(r, k);
      rr.x -= alpha * qq.x; rr.y -= alpha * qq.y;
      HB_ST(r, k, rr);
      s += rr.x * rr.x + rr.y * rr.y;
    }
  }
  double vals[1] = {s};
  block_partials<1>(vals, partials);
}

// x += alpha p of the iteration just finished (cg.F90:97-98), then, unless the
// solve has ended, p = r + beta p (sc_xpby, sc.F90:184-201); the same
// arithmetic on the same values as updating x beside r
__global__ __launch_bounds__(kBlock) void k_cg_pupdate(double2 *__restrict__ x, double2 *__restrict__ p,
                                                       const double2 *__restrict__ r, int vh,
                                                       const CgState *st) {
  if (!st->xpend) return;
  const double alpha = st->alpha, beta = st->beta;
  const bool next = st->done == 0;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
#pragma unroll
  HB_SITE_LOOP {
    double2 pp = HB_AT(p, k), xx = HB_AT(x, k);
    xx.x += alpha * pp.x; xx.y += alpha * pp.y;
    HB_ST(x, k, xx);
    if (next) {
      double2 rr = HB_AT(r, k);
      HB_ST(p, k, make_double2(rr.x + beta * pp.x, rr.y + beta * pp.y));
    }
  }
}

// out = b - out, result |out|^2
__global__ __launch_bounds__(kBlock) void k_residual(double2 *__restrict__ out, const double2 *__restrict__ b,
                                                     int vh, double *partials) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < vh) {
#pragma unroll
    HB_SITE_LOOP {
      double2 bb = HB_AT(b, k), o = HB_AT(out, k);
      double2 v = make_double2(bb.x - o.x, bb.y - o.y);
      HB_AT(out, k) = v;
      s += v.x * v.x + v.y * v.y;
    }
  }
  double vals[1] = {s};
  block_partials<1>(vals, partials);
}

// Re<x, y> and |x|^2 of FP64 fields
__global__ __launch_bounds__(kBlock) void k_dot_norm(const double2 *__restrict__ x, const double2 *__restrict__ y,
                                                     int vh, double *partials) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  if (i < vh) {
#pragma unroll
    HB_SITE_LOOP {
      double2 a = HB_AT(x, k), b = HB_AT(y, k);
      s0 += a.x * b.x + a.y * b.y;
      s1 += a.x * a.x + a.y * a.y;
    }
  }
  double vals[2] = {s0, s1};
  block_partials<2>(vals, partials);
}

#undef HB_SITE_LOOP
#undef HB_AT

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

// every field of the CG state cleared, then the stopping bound and maxiter
__global__ void k_cg_init(CgState *st, double rest, int maxiter) {
  if (threadIdx.x == 0) {
    CgState s{};
    s.rest = rest;
    s.maxiter = maxiter;
    *st = s;
  }
}

int cg_init(Context *c, CgState *st, double rest, int maxiter) {
  k_cg_init<<<1, 64, 0, c->stream>>>(st, rest, maxiter);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int cg_start(Context *c, double2 *r, double2 *p, const double2 *b, CgState *st, double *red) {
  const int vh = c->g.vh, grid = nblocks(vh);
  if (!red) red = c->red;
  k_cg_start<<<grid, kBlock, 0, c->stream>>>(r, p, b, vh, red);
  k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(red, grid, nullptr, FinCgStart{st});
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int cg_update(Context *c, double2 *r, const double2 *q, CgState *st, double *red) {
  const int vh = c->g.vh, grid = nblocks(vh);
  if (!red) red = c->red;
  k_cg_update<<<grid, kBlock, 0, c->stream>>>(r, q, vh, st, red);
  k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(red, grid, nullptr, FinCgUpdate{st});
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int cg_pupdate(Context *c, double2 *x, double2 *p, const double2 *r, const CgState *st) {
  const int vh = c->g.vh;
  k_cg_pupdate<<<nblocks(vh), kBlock, 0, c->stream>>>(x, p, r, vh, st);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int residual(Context *c, double2 *out, const double2 *b, double *res_d, double *red) {
  const int vh = c->g.vh, grid = nblocks(vh);
  if (!red) red = c->red;
  k_residual<<<grid, kBlock, 0, c->stream>>>(out, b, vh, red);
  k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(red, grid, nullptr, FinStore1{res_d});
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int dot_norm(Context *c, const double2 *x, const double2 *y, double *res_d) {
  const int vh = c->g.vh, grid = nblocks(vh);
  k_dot_norm<<<grid, kBlock, 0, c->stream>>>(x, y, vh, c->red);
  k_finalize<2><<<1, kFinThreads, 0, c->stream>>>(c->red, grid, nullptr, FinStore2{res_d});
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
