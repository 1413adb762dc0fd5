// hb_mixed.hip - mixed-precision CG with reliable updates.
//
// The inner iteration is the reference's CG (cg.F90:93-117) on the operator
// built from FP32 or 16-bit copies of the links and clover term, with FP32
// arithmetic and the search direction, residual and operator temporaries in
// the same storage precision (16-bit spinors carry a per-site FP32 scale).  The
// dot products accumulate in FP64.  A reliable update (defect correction)
// folds the low-precision solution increment into the FP64 solution and
// replaces the iterated residual by the FP64 true residual b - A x whenever the
// iterated residual has dropped by `delta` against its maximum since the last
// update, after `max_inner` iterations, or when it reaches the stopping bound.
// The solve stops only on the FP64 true residual (|r|^2 <= rest).
#include <cmath>

#include "hb_ops.h"

namespace hb {

struct MixState {
  double rtr, rtr_new, pap, beta, maxr;
  double alpha;  // of the last iteration, for the x_lo update
  int xpend;     // the last launched iteration ran: its x_lo update is due
  double rp_re, rp_im;  // <r, p> / |r|^2 after a reliable update
  double rest, delta;
  int iter, since, max_inner, maxiter;
  int flag;      // 0 iterate, 1 reliable update due, 2 converged, 3 maxiter reached
  int n_reliable;
  double res[4];
};

// ---- precision conversion ----------------------------------------------------
template <class S>
__global__ void k_cvt_field(typename S::ST *__restrict__ dst, const double2 *__restrict__ src, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t k = i; k < n; k += (size_t)gridDim.x * blockDim.x) {
    double2 v = src[k];
    S::st(dst, k, cmk<typename S::V>((typename S::A)v.x, (typename S::A)v.y));
  }
}

// links in reduced precision, rows 0 and 1 only (U[4][2][6][vh], see load_link)
template <class S>
__global__ void k_cvt_gauge_rows(typename S::ST *__restrict__ dst, const double2 *__restrict__ src, size_t vh) {
  const size_t n = 48 * vh;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const size_t slab = k / (6 * vh), rem = k - slab * 6 * vh;
    const double2 v = src[slab * 9 * vh + rem];
    S::st(dst, k, cmk<typename S::V>((typename S::A)v.x, (typename S::A)v.y));
  }
}

// links in 16 bits: rows 0 and 1, a 16-byte and an 8-byte plane per (mu, eo)
// slab (see load_link_f)
__global__ void k_cvt_gauge16(short2 *__restrict__ dst, const double2 *__restrict__ src, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int slab = blockIdx.y;  // mu*2 + eo
  if (i >= vh) return;
  link_to_rec16(dst + (size_t)slab * 6 * vh, i, vh, src + (size_t)slab * 9 * vh + i, vh);
}

// clover in 16-bit storage: per site and block, c 1 + m Q (see clover_block_scale)
// in five 16-byte planes per [eo][J] slab (see load_clover_block_raw)
__global__ void k_cvt_clover16(short2 *__restrict__ dst, const double2 *__restrict__ src, int vh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int eb = blockIdx.y;  // eo*2 + J
  if (i >= vh) return;
  clover_to_rec16(dst + (size_t)eb * kClov16Words * vh, i, vh, src + (size_t)eb * 18 * vh + i, vh);
}

// spinor FP64 -> S (thread per site: FP16 needs the site's norm)
template <class S>
__global__ void k_cvt_spinor(SpinorPtr<S> dst, const double2 *__restrict__ src, int vh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  double2 v[12];
#pragma unroll
  for (int k = 0; k < 12; k++) v[k] = src[(size_t)k * vh + i];
  store_spinor<S>(dst, i, vh, v);
}

// reduced-precision field back to an FP64 spinor (the stored values, decoded)
template <class S>
__global__ void k_spinor_to_f64(double2 *__restrict__ dst, SpinorPtr<S> src, int vh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typename S::V v[12];
  load_spinor<S>(v, src, i, vh);
#pragma unroll
  for (int k = 0; k < 12; k++) dst[(size_t)k * vh + i] = make_double2((double)v[k].x, (double)v[k].y);
}

// ---- inner iteration kernels -----------------------------------------------------
// The solution increment since the last reliable update, x_lo, is FP32 (in 16
// bits its rounding, amplified by the operator's high modes, stalled the FP32
// inner solver).  The update x_lo += alpha p happens in the p update, which
// reads p anyway.

// alpha = rtr / pAp; r -= alpha q; rtr_new = |r|^2
template <class S, class R>
__global__ __launch_bounds__(kBlock) void k_mix_update(SpinorPtr<R> r, SpinorPtr<S> q, int vh,
                                                       const MixState *st, double *partials) {
  using V = typename S::V;
  if (st->flag) return;
  const float alpha = (float)(st->rtr / st->pap);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < vh) {
    V qv[12], rv[12];
    load_spinor<S>(qv, q, i, vh);
    load_spinor<R>(rv, r, i, vh);
#pragma unroll
    for (int k = 0; k < 12; k++) {
      rv[k].x -= alpha * qv[k].x;
      rv[k].y -= alpha * qv[k].y;
      s += (double)rv[k].x * rv[k].x + (double)rv[k].y * rv[k].y;
    }
    store_spinor<R>(r, i, vh, rv);
  }
  double vals[1] = {s};
  block_partials<1>(vals, partials);
}

// iterated |r|^2 -> reliable-update decision or beta (cg.F90:109-115)
// (launched unconditionally: after a skipped update it only clears xpend, so
// that the p update of a skipped iteration adds nothing to x_lo)
struct FinMixUpdate {
  MixState *st;
  __device__ void operator()(const double (&t)[1]) const {
    if (st->flag) {
      st->xpend = 0;
      return;
    }
    st->alpha = st->rtr / st->pap;  // the alpha of k_mix_update
    st->xpend = 1;
    const double rn = t[0];
    st->iter += 1;
    st->since += 1;
    st->rtr_new = rn;
    const double rnorm = sqrt(rn);
    if (rnorm > st->maxr) st->maxr = rnorm;
    if (rn <= st->rest || rnorm < st->delta * st->maxr || st->since >= st->max_inner ||
        st->iter >= st->maxiter) {
      st->flag = 1;
    } else {
      st->beta = rn / st->rtr;
      st->rtr = rn;
    }
  }
};

// x_lo += alpha p (every iteration that ran), then p = r + beta p (unless a
// reliable update or the end is due)
template <class S, class R>
__global__ __launch_bounds__(kBlock) void k_mix_pupdate(SpinorPtr<S> p, SpinorPtr<R> r, float2 *__restrict__ xlo,
                                                        int vh, const MixState *st) {
  using V = typename S::V;
  if (!st->xpend) return;
  const float alpha = (float)st->alpha;
  const float beta = (float)st->beta;
  const bool next = st->flag == 0;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  V pv[12];
  load_spinor<S>(pv, p, i, vh);
#pragma unroll
  for (int k = 0; k < 12; k++) {
    float2 xv = xlo[(size_t)k * vh + i];
    xv.x += alpha * pv[k].x;
    xv.y += alpha * pv[k].y;
    xlo[(size_t)k * vh + i] = xv;
  }
  if (next) {
    V rv[12];
    load_spinor<R>(rv, r, i, vh);
#pragma unroll
    for (int k = 0; k < 12; k++) {
      pv[k].x = rv[k].x + beta * pv[k].x;
      pv[k].y = rv[k].y + beta * pv[k].y;
    }
    store_spinor<S>(p, i, vh, pv);
  }
}

// ---- reliable update kernels ------------------------------------------------------
// x += x_lo; x_lo = 0
__global__ void k_mix_accum(double2 *__restrict__ x, float2 *__restrict__ xlo, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t k = i; k < n; k += (size_t)gridDim.x * blockDim.x) {
    float2 d = xlo[k];
    double2 v = x[k];
    x[k] = make_double2(v.x + (double)d.x, v.y + (double)d.y);
    xlo[k] = make_float2(0.f, 0.f);
  }
}

// r = b - Ax (in place in ax), |r|^2; then r_lo = r and the scalar bookkeeping
template <class R>
__global__ __launch_bounds__(kBlock) void k_mix_true_residual(double2 *__restrict__ ax,
                                                              const double2 *__restrict__ b,
                                                              SpinorPtr<R> rlo, int vh,
                                                              double *partials) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < vh) {
    double2 v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
      double2 bb = b[(size_t)k * vh + i], a = ax[(size_t)k * vh + i];
      v[k] = make_double2(bb.x - a.x, bb.y - a.y);
      ax[(size_t)k * vh + i] = v[k];
      s += v[k].x * v[k].x + v[k].y * v[k].y;
    }
    store_spinor<R>(rlo, i, vh, v);
  }
  double vals[1] = {s};
  block_partials<1>(vals, partials);
}

// true |r|^2 -> convergence, or restart of the inner iteration
struct FinMixTrue {
  MixState *st;
  int first;
  __device__ void operator()(const double (&t)[1]) const {
    const double rt = t[0];
    st->res[0] = rt;
    if (!first) st->n_reliable += 1;
    st->since = 0;
    st->maxr = sqrt(rt);
    if (rt <= st->rest) {
      st->flag = 2;
    } else if (st->iter >= st->maxiter) {
      st->flag = 3;
    } else {
      // the true residual replaces the iterated one; the search direction is kept
      st->beta = first ? 0.0 : rt / st->rtr;
      st->rtr = rt;
      st->flag = 0;
    }
  }
};

// Re-orthogonalisation of the kept search direction at a reliable update:
// <r_lo, p> (complex) over the new residual, so that the restart can take out
// the component of p along r that the replaced residual no longer cancels
// (the reliable-update CG of Clark et al., "Solving lattice QCD systems of
// equations using mixed precision solvers on GPUs", 2010)
#ifndef HB_MIX_REORTH
#define HB_MIX_REORTH 1
#endif
template <class S, class R>
__global__ __launch_bounds__(kBlock) void k_mix_rdotp(SpinorPtr<S> p, SpinorPtr<R> r, int vh, const MixState *st,
                                                      double *partials) {
  using V = typename S::V;
  if (st->flag) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double re = 0.0, im = 0.0;
  if (i < vh) {
    V pv[12], rv[12];
    load_spinor<S>(pv, p, i, vh);
    load_spinor<R>(rv, r, i, vh);
#pragma unroll
    for (int k = 0; k < 12; k++) {  // conj(r) p
      re += (double)rv[k].x * pv[k].x + (double)rv[k].y * pv[k].y;
      im += (double)rv[k].x * pv[k].y - (double)rv[k].y * pv[k].x;
    }
  }
  double vals[2] = {re, im};
  block_partials<2>(vals, partials);
}

struct FinMixRdotp {
  MixState *st;
  __device__ void operator()(const double (&t)[2]) const {
    st->rp_re = t[0] / st->rtr;
    st->rp_im = t[1] / st->rtr;
  }
};

// p = r_lo + beta (p - rp r_lo) after a reliable update (rp = <r, p>/|r|^2, or
// 0; beta = 0 at the start: p = r)
template <class S, class R>
__global__ __launch_bounds__(kBlock) void k_mix_prestart(SpinorPtr<S> p, SpinorPtr<R> r, int vh,
                                                         const MixState *st) {
  using V = typename S::V;
  if (st->flag) return;
  const float beta = (float)st->beta;
  const float cr = (float)st->rp_re, ci = (float)st->rp_im;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  V pv[12], rv[12];
  load_spinor<R>(rv, r, i, vh);
  if (beta != 0.f) {
    load_spinor<S>(pv, p, i, vh);
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) pv[k] = cmk<V>(0.f, 0.f);
  }
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const float ux = pv[k].x - (cr * rv[k].x - ci * rv[k].y);
    const float uy = pv[k].y - (cr * rv[k].y + ci * rv[k].x);
    pv[k].x = rv[k].x + beta * ux;
    pv[k].y = rv[k].y + beta * uy;
  }
  store_spinor<S>(p, i, vh, pv);
}

// ---- low-precision copies of links and clover ------------------------------------
template <class S> struct LP;
template <> struct LP<SF32> {
  static const float2 *gauge(Context *c) { return c->gauge32; }
  static const float2 *clov(Context *c, int k) { return c->clov32[k]; }
  static const float *norm(Context *, int) { return nullptr; }
};
template <> struct LP<SF16> {
  static const short2 *gauge(Context *c) { return c->gauge16; }
  static const short2 *clov(Context *c, int k) { return c->clov16[k]; }
  static const float *norm(Context *, int) { return nullptr; }  // factors stored with the words
};

int ensure_lowprec(Context *c, int prec) {
  const size_t vh = c->g.vh;
  const bool clov = use_clover(c);
  if (c->lp_gauge_version == c->gauge_version * 1000 + prec &&
      c->lp_clover_version == c->clover_version * 1000 + prec)
    return HB_OK;
  const int grid = nblocks((int)std::min<size_t>(72 * vh, (size_t)1 << 22));
  if (prec == 32) {
    if (!c->gauge32) HB_HIP(hipMalloc(&c->gauge32, sizeof(float2) * 48 * vh));
    k_cvt_gauge_rows<SF32><<<grid, kBlock, 0, c->stream>>>(c->gauge32, c->gauge, vh);
    HB_HIP(hipGetLastError());
    if (clov) {
      for (int k = 0; k < 2; k++) {
        if (!c->clov32[k]) HB_HIP(hipMalloc(&c->clov32[k], sizeof(float2) * 72 * vh));
        k_cvt_field<SF32><<<grid, kBlock, 0, c->stream>>>(c->clov32[k], c->clov[k], 72 * vh);
        HB_HIP(hipGetLastError());
      }
    }
  } else {
    if (!c->gauge16) HB_HIP(hipMalloc(&c->gauge16, sizeof(short2) * 48 * vh));
    k_cvt_gauge16<<<dim3(nblocks((int)vh), 8), kBlock, 0, c->stream>>>(c->gauge16, c->gauge, (int)vh);
    HB_HIP(hipGetLastError());
    if (clov) {
      for (int k = 0; k < 2; k++) {
        if (!c->clov16[k]) HB_HIP(hipMalloc(&c->clov16[k], sizeof(short2) * 4 * kClov16Words * vh));
        k_cvt_clover16<<<dim3(nblocks((int)vh), 4), kBlock, 0, c->stream>>>(c->clov16[k], c->clov[k], (int)vh);
        HB_HIP(hipGetLastError());
      }
    }
  }
  c->lp_gauge_version = c->gauge_version * 1000 + prec;
  c->lp_clover_version = c->clover_version * 1000 + prec;
  return HB_OK;
}

// half of M~ (or M~^+) with spinors in S and links/clover in G; see mtil_f64
template <class S, class G>
int mtil_lp(Context *c, bool dag, double kappa, double rho, SpinorPtr<S> out, SpinorPtr<S> in,
            SpinorPtr<S> t, double *dot_out, SpinorPtr<S> dotv, const int *skip) {
  const bool clov = use_clover(c);
  DslashArgs<S, G> a1{};
  a1.g = c->g;
  a1.out_parity = 1;
  a1.out = t;
  a1.in = in;
  a1.U = LP<G>::gauge(c);
  a1.C = clov ? LP<G>::clov(c, HB_CLOVER_AINV) : nullptr;
  a1.Cn = clov ? LP<G>::norm(c, HB_CLOVER_AINV) : nullptr;
  a1.skip = skip;
  HB_TRY(launch_dslash(a1, dag, EPI_CLINV, c->stream));
  DslashArgs<S, G> a2{};
  a2.g = c->g;
  a2.out_parity = 0;
  a2.out = out;
  a2.in = t;
  a2.U = LP<G>::gauge(c);
  a2.C = clov ? LP<G>::clov(c, HB_CLOVER_A) : nullptr;
  a2.Cn = clov ? LP<G>::norm(c, HB_CLOVER_A) : nullptr;
  a2.x0 = in;
  a2.mkappa2 = -kappa * kappa;
  a2.rho = rho;
  a2.skip = skip;
  a2.partials = c->red;
  if (dot_out) {
    a2.dotv = dotv;
    a2.dot_out = dot_out;
    HB_TRY(launch_dslash(a2, dag, EPI_MTDOT, c->stream));
  } else {
    HB_TRY(launch_dslash(a2, dag, EPI_MTIL, c->stream));
  }
  return HB_OK;
}

template <class S> struct SpinorAlloc {
  static SpinorPtr<S> get(Context *c, int slot, int nslot, int *rc) {
    size_t vh = c->g.vh;
    size_t bytes = sizeof(typename S::ST) * 12 * vh;
    SpinorPtr<S> p{static_cast<typename S::ST *>(scratch(c, slot, bytes, rc)), nullptr};
    if (*rc) return p;
    if (S::kNorm) p.norm = static_cast<float *>(scratch(c, nslot, sizeof(float) * vh, rc));
    return p;
  }
};

template <class S, class G, class R = S>
int cg_mixed_impl(Context *c, int op, double kappa, double rho, double2 *x, const double2 *b,
                  float rest, int maxiter, int log, double delta, int max_inner, int *iters,
                  int *n_rel, double *true_rtr) {
  const int vh = c->g.vh;
  const double r_op = op == HB_OP_WDAGW ? rho : 0.0;
  int rc = 0;
  // FP64 work fields
  double2 *t64 = spinor_scratch(c, SLOT_T, &rc); HB_TRY(rc);
  double2 *y64 = spinor_scratch(c, SLOT_Y, &rc); HB_TRY(rc);
  double2 *ax = spinor_scratch(c, SLOT_MIX0, &rc); HB_TRY(rc);
  float2 *xlo = static_cast<float2 *>(scratch(c, SLOT_MIX1, sizeof(float2) * 12 * (size_t)vh, &rc)); HB_TRY(rc);
  // low-precision fields (norm slots follow SLOT_LPN)
  SpinorPtr<S> p = SpinorAlloc<S>::get(c, SLOT_LP0, SLOT_LPN, &rc); HB_TRY(rc);
  SpinorPtr<R> r = SpinorAlloc<R>::get(c, SLOT_LP1, SLOT_LPN1, &rc); HB_TRY(rc);
  SpinorPtr<S> q = SpinorAlloc<S>::get(c, SLOT_LP2, SLOT_LPN2, &rc); HB_TRY(rc);
  SpinorPtr<S> t = SpinorAlloc<S>::get(c, SLOT_LP3, SLOT_LPN3, &rc); HB_TRY(rc);
  SpinorPtr<S> y = SpinorAlloc<S>::get(c, SLOT_LP4, SLOT_LPN4, &rc); HB_TRY(rc);
  MixState *st = static_cast<MixState *>(scratch(c, SLOT_STATE, sizeof(MixState), &rc)); HB_TRY(rc);
  HB_TRY(ensure_host_stage(c, sizeof(MixState)));
  MixState *hs = static_cast<MixState *>(c->host_stage);
  HB_TRY(ensure_lowprec(c, G::kBytes == 8 ? 32 : 16));

  EventTimer tm;
  HB_TRY(tm.create());
  HB_TRY(tm.start(c->stream));
  MixState init{};
  init.rest = (double)rest;
  init.delta = delta;
  init.max_inner = max_inner;
  init.maxiter = maxiter;
  *hs = init;
  HB_HIP(hipMemcpyAsync(st, hs, sizeof(MixState), hipMemcpyHostToDevice, c->stream));
  HB_TRY(zero_f64(c, reinterpret_cast<double2 *>(xlo), (size_t)6 * vh));

  int nmv = 0;
  auto reliable = [&](int first) -> int {
    if (!first) {
      k_mix_accum<<<nblocks(std::min(12 * vh, 1 << 22)), kBlock, 0, c->stream>>>(x, xlo, (size_t)12 * vh);
      HB_HIP(hipGetLastError());
    }
    HB_TRY(apply_f64(c, op, kappa, rho, ax, x, t64, y64, nullptr, nullptr));
    k_mix_true_residual<R><<<nblocks(vh), kBlock, 0, c->stream>>>(ax, b, r, vh, c->red);
    k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(c->red, nblocks(vh), nullptr, FinMixTrue{st, first});
    HB_HIP(hipGetLastError());
    if (!first && HB_MIX_REORTH) {
      k_mix_rdotp<S, R><<<nblocks(vh), kBlock, 0, c->stream>>>(p, r, vh, st, c->red);
      k_finalize<2><<<1, kFinThreads, 0, c->stream>>>(c->red, nblocks(vh), &st->flag, FinMixRdotp{st});
      HB_HIP(hipGetLastError());
    }
    k_mix_prestart<S, R><<<nblocks(vh), kBlock, 0, c->stream>>>(p, r, vh, st);
    HB_HIP(hipGetLastError());
    nmv++;
    return HB_OK;
  };
  HB_TRY(reliable(1));
  const int *skip = &st->flag;
  int chunk = 8;
  for (;;) {
    for (int k = 0; k < chunk; k++) {
      HB_TRY((mtil_lp<S, G>(c, false, kappa, r_op, y, p, t, nullptr, SpinorPtr<S>{}, skip)));
      HB_TRY((mtil_lp<S, G>(c, true, kappa, r_op, q, y, t, &st->pap, p, skip)));
      k_mix_update<S, R><<<nblocks(vh), kBlock, 0, c->stream>>>(r, q, vh, st, c->red);
      k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(c->red, nblocks(vh), nullptr, FinMixUpdate{st});
      HB_HIP(hipGetLastError());
      k_mix_pupdate<S, R><<<nblocks(vh), kBlock, 0, c->stream>>>(p, r, xlo, vh, st);
      HB_HIP(hipGetLastError());
    }
    HB_HIP(hipMemcpyAsync(hs, st, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
    HB_HIP(hipStreamSynchronize(c->stream));
    if (hs->flag == 2 || hs->flag == 3) break;
    if (hs->flag == 1) {
      HB_TRY(reliable(0));
      HB_HIP(hipMemcpyAsync(hs, st, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
      HB_HIP(hipStreamSynchronize(c->stream));
      if (hs->flag == 2 || hs->flag == 3) break;
      // next chunk length: about the iterations one reliable interval took
      chunk = 8;
    } else if (chunk < 32) {
      chunk *= 2;
    }
  }
  HB_TRY(tm.stop(c->stream));
  float ms = 0.f;
  HB_TRY(tm.elapsed(&ms));
  c->last_solve_ms = ms;
  c->last_n_matvec = nmv + hs->iter;
  if (iters) *iters = hs->iter;
  if (n_rel) *n_rel = hs->n_reliable;
  if (true_rtr) *true_rtr = hs->res[0];
  if (hs->flag != 2 && log != 2) {
    set_error("cg_mixed(): no convergence; true rtr = %.17g", hs->res[0]);
    return HB_E_NOCONV;
  }
  return HB_OK;
}

int cg_mixed(Context *c, int op, double kappa, double rho, double2 *x, const double2 *b, float rest,
             int maxiter, int log, int prec, double delta, int max_inner, int *iters, int *n_rel,
             double *true_rtr) {
  HB_CHECK(op == HB_OP_MTDAGMT || op == HB_OP_WDAGW, HB_E_ARG,
           "cg_mixed: operator must be HB_OP_MTDAGMT or HB_OP_WDAGW (got %d)", op);
  HB_CHECK(delta > 0.0 && delta < 1.0, HB_E_ARG, "cg_mixed: delta must be in (0, 1)");
  HB_CHECK(max_inner >= 1, HB_E_ARG, "cg_mixed: max_inner must be >= 1");
  HB_CHECK(maxiter >= 1, HB_E_ARG, "cg_mixed: maxiter must be >= 1");
  if (prec == HB_PREC_F64)
    return cg_f64(c, op, kappa, rho, x, b, rest, maxiter, log, iters, true_rtr);
  if (prec == HB_PREC_F32)
    return cg_mixed_impl<SF32, SF32>(c, op, kappa, rho, x, b, rest, maxiter, log, delta, max_inner, iters, n_rel, true_rtr);
  if (prec == HB_PREC_F16)
    return cg_mixed_impl<SF16, SF16>(c, op, kappa, rho, x, b, rest, maxiter, log, delta, max_inner, iters, n_rel, true_rtr);
  if (prec == HB_PREC_F16_GAUGE)
    return cg_mixed_impl<SF32, SF16>(c, op, kappa, rho, x, b, rest, maxiter, log, delta, max_inner, iters, n_rel, true_rtr);
  set_error("cg_mixed: precision %d (use 64, 32, 16 or 18)", prec);
  return HB_E_ARG;
}

// average time of the first half of the reduced-precision operator, the
// hopping kernel with the A_oo^-1 epilogue (out = A^-1 D in, odd output), on the
// context stream
template <class S>
int time_dslash_lp(Context *c, int reps, const double2 *in64, double *ms_out) {
  int rc = 0;
  SpinorPtr<S> in = SpinorAlloc<S>::get(c, SLOT_LP0, SLOT_LPN, &rc); HB_TRY(rc);
  SpinorPtr<S> out = SpinorAlloc<S>::get(c, SLOT_LP3, SLOT_LPN3, &rc); HB_TRY(rc);
  k_cvt_spinor<S><<<nblocks(c->g.vh), kBlock, 0, c->stream>>>(in, in64, c->g.vh);
  HB_HIP(hipGetLastError());
  const bool clov = use_clover(c);
  DslashArgs<S, S> a{};
  a.g = c->g;
  a.out_parity = 1;
  a.out = out;
  a.in = in;
  a.U = LP<S>::gauge(c);
  a.C = clov ? LP<S>::clov(c, HB_CLOVER_AINV) : nullptr;
  a.Cn = clov ? LP<S>::norm(c, HB_CLOVER_AINV) : nullptr;
  for (int k = 0; k < 2; k++) HB_TRY(launch_dslash(a, false, EPI_CLINV, c->stream));
  EventTimer tm;
  HB_TRY(tm.create());
  HB_TRY(tm.start(c->stream));
  for (int k = 0; k < reps; k++) HB_TRY(launch_dslash(a, false, EPI_CLINV, c->stream));
  HB_TRY(tm.stop(c->stream));
  float ms = 0.f;
  HB_TRY(tm.elapsed(&ms));
  *ms_out = ms / reps;
  return HB_OK;
}

// the mixed solver's operator at reduced precision on an FP64 device spinor:
// in is stored at precision S (links and clover at G), the operator kernels of
// the inner iteration are applied (M~ or M~^+, with rho for W; M~^+ M~ or W^+ W
// as the solver chains them), and the stored result is decoded to FP64
template <class S, class G>
int apply_lp(Context *c, int op, double kappa, double rho, double2 *out64, const double2 *in64) {
  int rc = 0;
  SpinorPtr<S> in = SpinorAlloc<S>::get(c, SLOT_LP0, SLOT_LPN, &rc); HB_TRY(rc);
  SpinorPtr<S> out = SpinorAlloc<S>::get(c, SLOT_LP2, SLOT_LPN2, &rc); HB_TRY(rc);
  SpinorPtr<S> t = SpinorAlloc<S>::get(c, SLOT_LP3, SLOT_LPN3, &rc); HB_TRY(rc);
  SpinorPtr<S> y = SpinorAlloc<S>::get(c, SLOT_LP4, SLOT_LPN4, &rc); HB_TRY(rc);
  HB_TRY(ensure_lowprec(c, G::kBytes == 8 ? 32 : 16));
  const int vh = c->g.vh;
  k_cvt_spinor<S><<<nblocks(vh), kBlock, 0, c->stream>>>(in, in64, vh);
  HB_HIP(hipGetLastError());
  const bool w = op == HB_OP_W || op == HB_OP_W_DAG || op == HB_OP_WDAGW;
  const double r = w ? rho : 0.0;
  if (op == HB_OP_MTDAGMT || op == HB_OP_WDAGW) {
    HB_TRY((mtil_lp<S, G>(c, false, kappa, r, y, in, t, nullptr, SpinorPtr<S>{}, nullptr)));
    HB_TRY((mtil_lp<S, G>(c, true, kappa, r, out, y, t, nullptr, SpinorPtr<S>{}, nullptr)));
  } else {
    const bool dag = op == HB_OP_MTIL_DAG || op == HB_OP_W_DAG;
    HB_TRY((mtil_lp<S, G>(c, dag, kappa, r, out, in, t, nullptr, SpinorPtr<S>{}, nullptr)));
  }
  k_spinor_to_f64<S><<<nblocks(vh), kBlock, 0, c->stream>>>(out64, out, vh);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

}  // namespace hb

using namespace hb;

extern "C" {

int hb_time_dslash_lp(hb_context *ctx, int precision, int reps, const double *in_d, double *ms) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(in_d && ms, HB_E_ARG, "hb_time_dslash_lp: null pointer");
  HB_CHECK(reps >= 1, HB_E_ARG, "hb_time_dslash_lp: reps must be >= 1");
  HB_CHECK(precision == HB_PREC_F32 || precision == HB_PREC_F16, HB_E_ARG,
           "hb_time_dslash_lp: precision %d (use 32 or 16)", precision);
  HB_HIP(hipSetDevice(c->device));
  HB_TRY(ensure_lowprec(c, precision));
  const double2 *in = reinterpret_cast<const double2 *>(in_d);
  return precision == HB_PREC_F16 ? time_dslash_lp<SF16>(c, reps, in, ms)
                                  : time_dslash_lp<SF32>(c, reps, in, ms);
}

int hb_apply_lp_d(hb_context *ctx, int op, double kappa, double rho, int precision, double *out_d,
                  const double *in_d) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(out_d && in_d && out_d != in_d, HB_E_ARG, "hb_apply_lp_d: bad pointers");
  HB_CHECK(op >= HB_OP_MTIL && op <= HB_OP_WDAGW, HB_E_ARG, "hb_apply_lp_d: operator %d", op);
  HB_HIP(hipSetDevice(c->device));
  double2 *out = reinterpret_cast<double2 *>(out_d);
  const double2 *in = reinterpret_cast<const double2 *>(in_d);
  if (precision == HB_PREC_F32) return apply_lp<SF32, SF32>(c, op, kappa, rho, out, in);
  if (precision == HB_PREC_F16) return apply_lp<SF16, SF16>(c, op, kappa, rho, out, in);
  if (precision == HB_PREC_F16_GAUGE) return apply_lp<SF32, SF16>(c, op, kappa, rho, out, in);
  set_error("hb_apply_lp_d: precision %d (use 32, 16 or 18)", precision);
  return HB_E_ARG;
}

int hb_cg_mixed_d(hb_context *ctx, int op, double kappa, double rho, double *x_d, const double *b_d,
                  float rest, int maxiter, int log, int precision, double delta, int max_inner,
                  int *iters, int *n_reliable, double *true_rtr) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(x_d && b_d && x_d != b_d, HB_E_ARG, "hb_cg_mixed: bad pointers");
  HB_HIP(hipSetDevice(c->device));
  if (n_reliable) *n_reliable = 0;
  return cg_mixed(c, op, kappa, rho, reinterpret_cast<double2 *>(x_d), reinterpret_cast<const double2 *>(b_d),
                  rest, maxiter, log, precision, delta, max_inner, iters, n_reliable, true_rtr);
}

int hb_cg_mixed(hb_context *ctx, int op, double kappa, double rho, double *x, const double *b,
                float rest, int maxiter, int log, int precision, double delta, int max_inner,
                int *iters, int *n_reliable, double *true_rtr) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(x && b, HB_E_ARG, "hb_cg_mixed: null pointer");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double2 *dx = spinor_scratch(c, SLOT_X, &rc); HB_TRY(rc);
  double2 *db = spinor_scratch(c, SLOT_B, &rc); HB_TRY(rc);
  HB_TRY(spinor_upload(c, dx, x));
  HB_TRY(spinor_upload(c, db, b));
  if (n_reliable) *n_reliable = 0;
  int r = cg_mixed(c, op, kappa, rho, dx, db, rest, maxiter, log, precision, delta, max_inner,
                   iters, n_reliable, true_rtr);
  if (r != HB_OK && r != HB_E_NOCONV) return r;
  HB_TRY(spinor_download(c, x, dx));
  return r;
}

}  // extern "C"
