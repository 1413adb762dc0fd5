// hb_multi.hip - multi right-hand-side CG on a block hopping matrix.
//
// NR right-hand sides are stored RHS-fastest (component k of site i of RHS r
// at (k*vh + i)*NR + r).  One thread handles one (site, RHS) pair; the NR
// threads of a site are adjacent, so their spinor accesses are contiguous and
// their link and clover loads are the same addresses (one fetch serves the
// block).  Right-hand sides may carry different kappa (Hasenbusch shifts):
// the hopping term is kappa independent, the clover term is taken from the
// RHS's kappa group.  Each RHS runs its own CG (cg.F90:93-117) with its own
// scalars; a converged RHS is masked out of every kernel.  Inner iterations
// use FP32 or 16-bit storage; the FP64 true residual of each RHS is
// recomputed at its reliable updates, which also end its solve.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "hb_hop.h"
#include "hb_ops.h"

namespace hb {

constexpr int kMaxNR = 12;

double *multi_partials(Context *c, int NR, int *rc);

struct MultiState {
  double rtr[16], pap[16], beta[16], maxr[16], rest[16], res[16];
  double alpha[16];  // of each RHS's last iteration, for its x_lo update
  int xpend[16];     // that iteration ran: its x_lo update is due
  int flag[16];   // 0 iterate, 1 reliable update due, 2 converged, 3 maxiter, 4 unused lane
  int iter[16], since[16], nrel[16];
  double delta;
  int max_inner, maxiter;
};

template <class S, class G> struct MultiArgs {
  Geom g;
  int out_parity;
  SpinorPtr<S> out, in, x0, dotv;   // block layout, NR RHS per site
  const typename G::ST *U;
  const typename G::ST *C;          // A^-1 for CLINV, A for MTIL: FP32 planes [ngroups][2][36][vh],
                                    // 16-bit [2][2] slabs of ngroups * vh elements (load_clover_block_g_raw)
  const float *Cn;                  // unused (16-bit records carry their factors)
  int ngroups;
  int grp[16];
  double mk2[16], rho[16];
  const int *flag;                  // MultiState::flag
  double *dot_out;                  // MultiState::pap
  double *partials;
};

// clover block J of parity eo at site i for kappa group gidx (FP32 planes)
template <class SG>
__device__ __forceinline__ void load_clover_block_g(typename SG::V (&w)[18], const typename SG::ST *C,
                                                    const float *Cn, int ngroups, int gidx, int eo,
                                                    int J, int i, int vh) {
  static_assert(!SG::kNorm, "16-bit clover blocks are read with load_clover_block_g_raw");
  const uint32_t plane = (uint32_t)vh * SG::kBytes;
  rsrc_t r = make_rsrc(C, plane * 72u * (uint32_t)ngroups);
  const uint32_t vo = ((uint32_t)gidx * 72u * (uint32_t)vh + (uint32_t)i) * SG::kBytes;
  const uint32_t base = (uint32_t)(eo * 36 + J * 18);
#pragma unroll
  for (int k = 0; k < 18; k++) w[k] = SG::template bld<kNT>(r, vo, plane * (base + k));
}

// the same block of a 16-bit field: [eo][J] slabs of 16-byte planes whose
// elements are gidx * vh + i (one buffer per slab spans all groups; see
// load_clover_rec16); returns (m / 32767, c) -- the block is
// c 1 + (m / 32767) Q_int.  Cn is unused.
__device__ __forceinline__ float2 load_clover_block_g_raw(float2 (&w)[18], const SF16::ST *C,
                                                          const float *Cn, int ngroups, int gidx,
                                                          int eo, int J, int i, int vh) {
  const size_t slab = (size_t)ngroups * vh;
  rsrc_t r = make_rsrc(C + (size_t)(eo * 2 + J) * slab * kClov16Words, (uint32_t)(slab * kClov16));
  return load_clover_rec16(w, r, (uint32_t)gidx * (uint32_t)vh + (uint32_t)i, (uint32_t)slab * 16u);
}

template <class SG, class V>
__device__ __forceinline__ void clover_apply_g(V (&out)[12], const V (&in)[12], const typename SG::ST *C,
                                               const float *Cn, int ngroups, int gidx, int eo, int i,
                                               int vh) {
  constexpr bool kRaw = std::is_same<SG, SF16>::value;
  V xp[6], y1[6], y2[6];
#pragma unroll
  for (int m = 0; m < 6; m++) xp[m] = cadd(in[m], in[m + 6]);
  {
    typename SG::V w[18];
    float2 mc;
    if constexpr (kRaw) mc = load_clover_block_g_raw(w, C, Cn, ngroups, gidx, eo, 0, i, vh);
    else load_clover_block_g<SG>(w, C, Cn, ngroups, gidx, eo, 0, i, vh);
    V ww[18];
#pragma unroll
    for (int k = 0; k < 18; k++) ww[k] = cmk<V>(w[k].x, w[k].y);
    herm6(y1, ww, xp);
    if constexpr (kRaw) clover_raw_finish(y1, xp, mc);
  }
#pragma unroll
  for (int m = 0; m < 6; m++) asm volatile("" : "+v"(y1[m].x), "+v"(y1[m].y));
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < 6; m++) xp[m] = csub(in[m], in[m + 6]);
  {
    typename SG::V w[18];
    float2 mc;
    if constexpr (kRaw) mc = load_clover_block_g_raw(w, C, Cn, ngroups, gidx, eo, 1, i, vh);
    else load_clover_block_g<SG>(w, C, Cn, ngroups, gidx, eo, 1, i, vh);
    V ww[18];
#pragma unroll
    for (int k = 0; k < 18; k++) ww[k] = cmk<V>(w[k].x, w[k].y);
    herm6(y2, ww, xp);
    if constexpr (kRaw) clover_raw_finish(y2, xp, mc);
  }
#pragma unroll
  for (int m = 0; m < 6; m++) {
    out[m] = cadd(y1[m], y2[m]);
    out[m + 6] = csub(y1[m], y2[m]);
  }
}

template <class S, class G, int NR, bool DAG, int EPI>
__global__ __launch_bounds__(kBlock) void k_dslash_multi(MultiArgs<S, G> a) {
  using V = typename S::V;
  using A = typename S::A;
  const Geom g = a.g;
  const int vh = g.vh, vhe = vh * NR;
  const int tid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int site = tid / NR, r = tid - site * NR;
  const bool active = site < vh && a.flag[r] == 0;
  V acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = cmk<V>((A)0, (A)0);
  const int e = a.out_parity;
  if (active) {
    int xc[4];
    site_coords(g, e, site, xc[0], xc[1], xc[2], xc[3]);
    const SpinorView<S> in(a.in, vhe);
    if constexpr (HB_WAVE_SHARE && std::is_same<S, SF16>::value && NR <= 2) {
      // the lanes of a wave hold consecutive (site, RHS) elements; the x and y
      // neighbours (same RHS) come from other lanes where they are in the wave.
      // From NR = 4 on a wave spans 16 sites or fewer, and the extra registers of
      // the own-index spinor cost more than the shared x hops save (measured)
      uint32_t w[13];
      in.load_words(w, tid);
      const WaveView16 wv{in, w, tid - (int)(threadIdx.x & 63), g.L[0] / 2 * NR <= 32};
      hop_all<S, G, DAG, NR, WaveView16>(acc, a.U, wv, g, e, site, xc, r);
    } else {
      hop_all<S, G, DAG, NR>(acc, a.U, in, g, e, site, xc, r);
    }
  }
  const int ie = site * NR + r;
  if constexpr (EPI == EPI_CLINV) {
    if (active) {
      if (a.C) {
        V o[12];
        clover_apply_g<G>(o, acc, a.C, a.Cn, a.ngroups, a.grp[r], e, site, vh);
        store_spinor<S>(a.out, ie, vhe, o);
      } else {
        store_spinor<S>(a.out, ie, vhe, acc);
      }
    }
  } else {
    double vals[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) vals[k] = 0.0;
    if (active) {
      V x0[12], o[12];
      load_spinor<S>(x0, a.x0, ie, vhe);
      HB_SCHED_FENCE();
      if (a.C) clover_apply_g<G>(o, x0, a.C, a.Cn, a.ngroups, a.grp[r], e, site, vh);
      else {
#pragma unroll
        for (int k = 0; k < 12; k++) o[k] = x0[k];
      }
      const A mk2 = (A)a.mk2[r], rho = (A)a.rho[r];
#pragma unroll
      for (int k = 0; k < 12; k++) {
        o[k].x += mk2 * acc[k].x + rho * x0[k].x;
        o[k].y += mk2 * acc[k].y + rho * x0[k].y;
      }
      store_spinor<S>(a.out, ie, vhe, o);
      if constexpr (EPI == EPI_MTDOT) {
        V v[12];
        load_spinor<S>(v, a.dotv, ie, vhe);
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < 12; k++) d += (double)v[k].x * (double)o[k].x + (double)v[k].y * (double)o[k].y;
#pragma unroll
        for (int k = 0; k < NR; k++)
          if (k == r) vals[k] = d;
      }
    }
    if constexpr (EPI == EPI_MTDOT) block_partials<NR>(vals, a.partials);
  }
}

// per-RHS finalize functors ------------------------------------------------------
template <int NR> struct FinMultiStore {
  double *out;
  const int *flag;
  __device__ void operator()(const double (&t)[NR]) const {
#pragma unroll
    for (int k = 0; k < NR; k++)
      if (flag[k] == 0) out[k] = t[k];
  }
};

template <int NR> struct FinMultiUpdate {
  MultiState *st;
  __device__ void operator()(const double (&t)[NR]) const {
#pragma unroll
    for (int k = 0; k < NR; k++) {
      if (st->flag[k] != 0) {
        st->xpend[k] = 0;  // the p update of a skipped iteration adds nothing to x_lo
        continue;
      }
      st->alpha[k] = st->rtr[k] / st->pap[k];  // the alpha of k_multi_update
      st->xpend[k] = 1;
      const double rn = t[k];
      st->iter[k] += 1;
      st->since[k] += 1;
      const double rnorm = sqrt(rn);
      if (rnorm > st->maxr[k]) st->maxr[k] = rnorm;
      if (rn <= st->rest[k] || rnorm < st->delta * st->maxr[k] || st->since[k] >= st->max_inner ||
          st->iter[k] >= st->maxiter) {
        st->flag[k] = 1;
      } else {
        st->beta[k] = rn / st->rtr[k];
        st->rtr[k] = rn;
      }
    }
  }
};

// vector kernels (thread per (site, RHS)) ------------------------------------------
// As in the single-RHS mixed solver, the FP32 increments x_lo are updated in
// the p update.
template <class S, int NR>
__global__ __launch_bounds__(kBlock) void k_multi_update(SpinorPtr<S> r, SpinorPtr<S> q, int vh,
                                                         const MultiState *st, double *partials) {
  using V = typename S::V;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int site = tid / NR, k0 = tid - site * NR;
  const int vhe = vh * NR;
  double vals[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) vals[k] = 0.0;
  if (site < vh && st->flag[k0] == 0) {
    const float alpha = (float)(st->rtr[k0] / st->pap[k0]);
    V qv[12], rv[12];
    load_spinor<S>(qv, q, tid, vhe);
    load_spinor<S>(rv, r, tid, vhe);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      rv[k].x -= alpha * qv[k].x;
      rv[k].y -= alpha * qv[k].y;
      s += (double)rv[k].x * rv[k].x + (double)rv[k].y * rv[k].y;
    }
    store_spinor<S>(r, tid, vhe, rv);
#pragma unroll
    for (int k = 0; k < NR; k++)
      if (k == k0) vals[k] = s;
  }
  block_partials<NR>(vals, partials);
}

// x_lo += alpha p for every RHS whose iteration ran, then p = r + beta p for
// those that go on
template <class S, int NR>
__global__ __launch_bounds__(kBlock) void k_multi_pupdate(SpinorPtr<S> p, SpinorPtr<S> r, float2 *__restrict__ xlo,
                                                          int vh, const MultiState *st) {
  using V = typename S::V;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int site = tid / NR, k0 = tid - site * NR;
  if (site >= vh || !st->xpend[k0]) return;
  const int vhe = vh * NR;
  const float alpha = (float)st->alpha[k0], beta = (float)st->beta[k0];
  V pv[12];
  load_spinor<S>(pv, p, tid, vhe);
#pragma unroll
  for (int k = 0; k < 12; k++) {
    float2 xv = xlo[(size_t)k * vhe + tid];
    xv.x += alpha * pv[k].x;
    xv.y += alpha * pv[k].y;
    xlo[(size_t)k * vhe + tid] = xv;
  }
  if (st->flag[k0] != 0) return;
  V rv[12];
  load_spinor<S>(rv, r, tid, vhe);
#pragma unroll
  for (int k = 0; k < 12; k++) {
    pv[k].x = rv[k].x + beta * pv[k].x;
    pv[k].y = rv[k].y + beta * pv[k].y;
  }
  store_spinor<S>(p, tid, vhe, pv);
}

// reliable update of RHS k0: x += x_lo[k0]; x_lo[k0] = 0
template <int NR>
__global__ void k_multi_accum(double2 *__restrict__ x, float2 *__restrict__ xlo, int vh, int k0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const int vhe = vh * NR;
#pragma unroll
  for (int k = 0; k < 12; k++) {
    float2 d = xlo[(size_t)k * vhe + (size_t)i * NR + k0];
    double2 v = x[(size_t)k * vh + i];
    x[(size_t)k * vh + i] = make_double2(v.x + d.x, v.y + d.y);
    xlo[(size_t)k * vhe + (size_t)i * NR + k0] = make_float2(0.f, 0.f);
  }
}

// r = b - Ax for RHS k0 (FP64, plain layout), |r|^2, r_lo[k0] = r, p[k0] = r + beta p
template <class S, int NR>
__global__ __launch_bounds__(kBlock) void k_multi_true_residual(double2 *__restrict__ ax,
                                                                const double2 *__restrict__ b,
                                                                SpinorPtr<S> rlo, int vh, int k0,
                                                                double *partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < vh) {
    double2 v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
      double2 bb = b[(size_t)k * vh + i], aa = ax[(size_t)k * vh + i];
      v[k] = make_double2(bb.x - aa.x, bb.y - aa.y);
      s += v[k].x * v[k].x + v[k].y * v[k].y;
    }
    store_spinor<S>(rlo, i * NR + k0, vh * NR, v);
  }
  double vals[1] = {s};
  block_partials<1>(vals, partials);
}

struct FinMultiTrue {
  MultiState *st;
  int k0, first;
  __device__ void operator()(const double (&t)[1]) const {
    const double rt = t[0];
    st->res[k0] = rt;
    if (!first) st->nrel[k0] += 1;
    st->since[k0] = 0;
    st->maxr[k0] = sqrt(rt);
    if (rt <= st->rest[k0]) {
      st->flag[k0] = 2;
    } else if (st->iter[k0] >= st->maxiter) {
      st->flag[k0] = 3;
    } else {
      st->beta[k0] = first ? 0.0 : rt / st->rtr[k0];
      st->rtr[k0] = rt;
      st->flag[k0] = 0;
    }
  }
};

template <class S, int NR>
__global__ __launch_bounds__(kBlock) void k_multi_prestart(SpinorPtr<S> p, SpinorPtr<S> r, int vh,
                                                           int k0, const MultiState *st) {
  using V = typename S::V;
  if (st->flag[k0] != 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  const int vhe = vh * NR, ie = i * NR + k0;
  const float beta = (float)st->beta[k0];
  V pv[12], rv[12];
  load_spinor<S>(rv, r, ie, vhe);
  if (beta != 0.f) load_spinor<S>(pv, p, ie, vhe);
  else {
#pragma unroll
    for (int k = 0; k < 12; k++) pv[k] = cmk<V>(0.f, 0.f);
  }
#pragma unroll
  for (int k = 0; k < 12; k++) {
    pv[k].x = rv[k].x + beta * pv[k].x;
    pv[k].y = rv[k].y + beta * pv[k].y;
  }
  store_spinor<S>(p, ie, vhe, pv);
}

// conversions of the per-group clover fields --------------------------------------
// group gi of ngroups: FP32 planes [group][2][36][vh]; 16-bit [eo][J] slabs of
// 16-byte planes, element gi * vh + i (see load_clover_block_g_raw)
template <class S>
__global__ void k_cvt_clover_lp(typename S::ST *__restrict__ dst, const double2 *__restrict__ src, int vh,
                                int ngroups, int gi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int eb = blockIdx.y;  // eo*2 + J
  if (i >= vh) return;
  const double2 *s = src + (size_t)eb * 18 * vh + i;
  if constexpr (S::kNorm) {
    const size_t n = (size_t)ngroups * vh;
    clover_to_rec16(dst + (size_t)eb * kClov16Words * n, (size_t)gi * vh + i, n, s, vh);
  } else {
    typename S::ST *d = dst + (size_t)gi * 72 * vh;
    for (int k = 0; k < 18; k++) {
      double2 v = s[(size_t)k * vh];
      S::st(d, (size_t)eb * 18 * vh + (size_t)k * vh + i, make_float2((float)v.x, (float)v.y));
    }
  }
}

// links in reduced precision, rows 0 and 1 only: FP32 planes U[4][2][6][vh]
// (see load_link), 16-bit a 16-byte and an 8-byte plane per (mu, eo) slab
template <class S>
__global__ void k_cvt_gauge_lp(typename S::ST *__restrict__ dst, const double2 *__restrict__ src, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int slab = blockIdx.y;  // mu*2 + eo
  if (i >= vh) return;
  const double2 *s = src + (size_t)slab * 9 * vh + i;
  if constexpr (S::kNorm) {
    link_to_rec16(dst + (size_t)slab * 6 * vh, i, vh, s, vh);
  } else {
    for (int k = 0; k < 6; k++) {
      const double2 v = s[(size_t)k * vh];
      S::st(dst, ((size_t)slab * 6 + k) * vh + i, make_float2((float)v.x, (float)v.y));
    }
  }
}

// kernel building the clover term into given arrays (hb_gauge.hip)
int build_clover_into(Context *c, double csw_kappa, double2 *Ca, double2 *Ci, double2 *Cb);

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
struct MultiPlan {
  int ngroups = 0;
  std::vector<double> gkappa;         // kappa of each group
  std::vector<int> group_of;          // group of each RHS
  std::vector<double2 *> clovA, clovI;  // FP64 per group (null: context's / none)
  bool clover = false;
};

template <class S, class G>
struct MultiFields {
  const typename G::ST *U = nullptr;
  typename G::ST *C[2] = {nullptr, nullptr};  // A (even used), A^-1 (odd used), all groups
  float *Cn[2] = {nullptr, nullptr};
};

template <class G> size_t st_bytes() { return sizeof(typename G::ST); }

template <class S, class G, int NR>
int launch_multi_dslash(Context *c, const MultiArgs<S, G> &a, bool dag, int epi) {
  const int n = c->g.vh * NR;
  dim3 grid(nblocks(n, c->multi_tile)), block(c->multi_tile);
#define HB_DM(D, E) hipLaunchKernelGGL((k_dslash_multi<S, G, NR, D, E>), grid, block, 0, c->stream, a)
  if (!dag) {
    if (epi == EPI_CLINV) HB_DM(false, EPI_CLINV);
    else if (epi == EPI_MTIL) HB_DM(false, EPI_MTIL);
    else HB_DM(false, EPI_MTDOT);
  } else {
    if (epi == EPI_CLINV) HB_DM(true, EPI_CLINV);
    else if (epi == EPI_MTIL) HB_DM(true, EPI_MTIL);
    else HB_DM(true, EPI_MTDOT);
  }
#undef HB_DM
  if (epi == EPI_MTDOT)
    k_finalize<NR><<<1, kFinThreads, 0, c->stream>>>(a.partials, (int)grid.x, nullptr,
                                            FinMultiStore<NR>{a.dot_out, a.flag});
  HB_HIP(hipGetLastError());
  return HB_OK;
}

// one batched application of op (M~^+M~ or W^+W) to the block p -> q, with
// Re<p, q> per RHS; t, y are block temporaries
template <class S, class G, int NR>
int multi_apply(Context *c, const MultiFields<S, G> &f, int ngroups, const int *grp,
                const double *kappa, const double *rho, SpinorPtr<S> q, SpinorPtr<S> p,
                SpinorPtr<S> t, SpinorPtr<S> y, MultiState *st, bool clover) {
  MultiArgs<S, G> a{};
  a.g = c->g;
  a.U = f.U;
  a.ngroups = ngroups;
  for (int k = 0; k < NR; k++) {
    a.grp[k] = grp[k];
    a.mk2[k] = -kappa[k] * kappa[k];
    a.rho[k] = rho[k];
  }
  a.flag = st->flag;
  int rc = 0;
  a.partials = multi_partials(c, NR, &rc);
  HB_TRY(rc);
  a.dot_out = st->pap;
  for (int dag = 0; dag < 2; dag++) {
    SpinorPtr<S> src = dag ? y : p, dst = dag ? q : y;
    MultiArgs<S, G> a1 = a;
    a1.out_parity = 1;
    a1.out = t;
    a1.in = src;
    a1.C = clover ? f.C[1] : nullptr;
    a1.Cn = clover ? f.Cn[1] : nullptr;
    HB_TRY((launch_multi_dslash<S, G, NR>(c, a1, dag == 1, EPI_CLINV)));
    MultiArgs<S, G> a2 = a;
    a2.out_parity = 0;
    a2.out = dst;
    a2.in = t;
    a2.x0 = src;
    a2.dotv = p;
    a2.C = clover ? f.C[0] : nullptr;
    a2.Cn = clover ? f.Cn[0] : nullptr;
    HB_TRY((launch_multi_dslash<S, G, NR>(c, a2, dag == 1, dag ? EPI_MTDOT : EPI_MTIL)));
  }
  return HB_OK;
}

// per-block partial sums of the block kernels: NR values per block, sized for
// the smallest workgroup the autotuner tries
constexpr int kMinTile = 128;
double *multi_partials(Context *c, int NR, int *rc) {
  const size_t n = (size_t)NR * nblocks(c->g.vh * NR, kMinTile);
  return static_cast<double *>(scratch(c, SLOT_MULTI + 7, sizeof(double) * n, rc));
}

template <class S> SpinorPtr<S> block_field(Context *c, int slot, int nslot, int NR, int *rc) {
  const size_t n = (size_t)c->g.vh * NR;
  SpinorPtr<S> f{static_cast<typename S::ST *>(scratch(c, slot, sizeof(typename S::ST) * 12 * n, rc)), nullptr};
  if (*rc) return f;
  if (S::kNorm) f.norm = static_cast<float *>(scratch(c, nslot, sizeof(float) * n, rc));
  return f;
}

// low-precision links and per-group clover for the multi solver
template <class S, class G>
int multi_fields(Context *c, const MultiPlan &plan, MultiFields<S, G> &f) {
  const size_t vh = c->g.vh;
  int rc = 0;
  static_assert(G::kRecon, "reduced-precision links are stored as two rows");
  typename G::ST *U = static_cast<typename G::ST *>(scratch(c, SLOT_MULTI + 20, st_bytes<G>() * 48 * vh, &rc));
  HB_TRY(rc);
  k_cvt_gauge_lp<G><<<dim3(nblocks((int)vh), 8), kBlock, 0, c->stream>>>(U, c->gauge, (int)vh);
  HB_HIP(hipGetLastError());
  f.U = U;
  if (!plan.clover) return HB_OK;
  // bytes of one buffer: FP32 planes of all groups; 16-bit one (eo, J) slab of
  // all groups
  const unsigned long long span = G::kNorm ? (unsigned long long)kClov16 * vh * plan.ngroups
                                           : 72ULL * vh * plan.ngroups * st_bytes<G>();
  HB_CHECK(fits_buffer(span), HB_E_ARG,
           "cg_multi: the clover terms of %d kappa values at %zu sites per parity exceed the 4 GiB "
           "span of one buffer", plan.ngroups, vh);
  const size_t field = G::kNorm ? (size_t)4 * kClov16 * vh * plan.ngroups : st_bytes<G>() * 72 * vh * plan.ngroups;
  for (int w = 0; w < 2; w++) {
    f.C[w] = static_cast<typename G::ST *>(scratch(c, SLOT_MULTI + 21 + w, field, &rc));
    HB_TRY(rc);
    f.Cn[w] = nullptr;
    for (int gi = 0; gi < plan.ngroups; gi++) {
      const double2 *src = w == 0 ? plan.clovA[gi] : plan.clovI[gi];
      k_cvt_clover_lp<G><<<dim3(nblocks((int)vh), 4), kBlock, 0, c->stream>>>(f.C[w], src, (int)vh,
                                                                           plan.ngroups, gi);
      HB_HIP(hipGetLastError());
    }
  }
  return HB_OK;
}

// solves the RHS list `idx` (at most NR) as one block
template <class S, class G, int NR>
int cg_multi_block(Context *c, const MultiPlan &plan, const MultiFields<S, G> &f, int op,
                   const std::vector<int> &idx, const double *kappa, const double *rho,
                   double2 *const *x, const double2 *const *b, const float *rest, int maxiter,
                   double delta, int max_inner, int *iters, double *true_rtr, int *nrel_out) {
  const int vh = c->g.vh, nb = (int)idx.size();
  int rc = 0;
  SpinorPtr<S> p = block_field<S>(c, SLOT_MULTI + 0, SLOT_MULTI + 10, NR, &rc); HB_TRY(rc);
  SpinorPtr<S> r = block_field<S>(c, SLOT_MULTI + 1, SLOT_MULTI + 11, NR, &rc); HB_TRY(rc);
  SpinorPtr<S> q = block_field<S>(c, SLOT_MULTI + 2, SLOT_MULTI + 12, NR, &rc); HB_TRY(rc);
  SpinorPtr<S> t = block_field<S>(c, SLOT_MULTI + 3, SLOT_MULTI + 13, NR, &rc); HB_TRY(rc);
  SpinorPtr<S> y = block_field<S>(c, SLOT_MULTI + 4, SLOT_MULTI + 14, NR, &rc); HB_TRY(rc);
  float2 *xlo = static_cast<float2 *>(scratch(c, SLOT_MULTI + 5, sizeof(float2) * 12 * (size_t)vh * NR, &rc)); HB_TRY(rc);
  double2 *ax = spinor_scratch(c, SLOT_MIX0, &rc); HB_TRY(rc);
  double2 *t64 = spinor_scratch(c, SLOT_T, &rc); HB_TRY(rc);
  double2 *y64 = spinor_scratch(c, SLOT_Y, &rc); HB_TRY(rc);
  MultiState *st = static_cast<MultiState *>(scratch(c, SLOT_MULTI + 6, sizeof(MultiState), &rc)); HB_TRY(rc);
  HB_TRY(ensure_host_stage(c, sizeof(MultiState)));
  MultiState *hs = static_cast<MultiState *>(c->host_stage);

  int grp[16] = {0};
  double kap[16] = {0}, rh[16] = {0};
  MultiState init{};
  init.delta = delta;
  init.max_inner = max_inner;
  init.maxiter = maxiter;
  const bool wdw = op == HB_OP_WDAGW;
  for (int k = 0; k < NR; k++) {
    if (k < nb) {
      grp[k] = plan.group_of[idx[k]];
      kap[k] = kappa[idx[k]];
      rh[k] = wdw ? rho[idx[k]] : 0.0;
      init.rest[k] = (double)rest[idx[k]];
    } else {
      init.flag[k] = 4;
    }
  }
  *hs = init;
  HB_HIP(hipMemcpyAsync(st, hs, sizeof(MultiState), hipMemcpyHostToDevice, c->stream));
  HB_TRY(zero_f64(c, reinterpret_cast<double2 *>(xlo), (size_t)6 * vh * NR));
  HB_TRY(zero_f64(c, reinterpret_cast<double2 *>(p.p), (size_t)12 * vh * NR * sizeof(typename S::ST) / 16));

  double *part = multi_partials(c, NR, &rc);
  HB_TRY(rc);
  // FP64 true residual of RHS k -> r_lo[k], state; uses the RHS's kappa-group
  // clover term (or none), swapped into the context for the FP64 operator
  auto reliable = [&](int k, int first) -> int {
    double2 *sa = c->clov[HB_CLOVER_A], *si = c->clov[HB_CLOVER_AINV];
    const bool ha = c->have_clover[HB_CLOVER_A], hi = c->have_clover[HB_CLOVER_AINV];
    c->have_clover[HB_CLOVER_A] = c->have_clover[HB_CLOVER_AINV] = plan.clover;
    if (plan.clover) {
      c->clov[HB_CLOVER_A] = plan.clovA[grp[k]];
      c->clov[HB_CLOVER_AINV] = plan.clovI[grp[k]];
    }
    if (!first) {
      k_multi_accum<NR><<<nblocks(vh), kBlock, 0, c->stream>>>(x[idx[k]], xlo, vh, k);
      HB_HIP(hipGetLastError());
    }
    int rr = apply_f64(c, op, kap[k], rh[k], ax, x[idx[k]], t64, y64, nullptr, nullptr);
    c->clov[HB_CLOVER_A] = sa;
    c->clov[HB_CLOVER_AINV] = si;
    c->have_clover[HB_CLOVER_A] = ha;
    c->have_clover[HB_CLOVER_AINV] = hi;
    HB_TRY(rr);
    k_multi_true_residual<S, NR><<<nblocks(vh), kBlock, 0, c->stream>>>(ax, b[idx[k]], r, vh, k, c->red);
    k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(c->red, nblocks(vh), nullptr, FinMultiTrue{st, k, first});
    k_multi_prestart<S, NR><<<nblocks(vh), kBlock, 0, c->stream>>>(p, r, vh, k, st);
    HB_HIP(hipGetLastError());
    return HB_OK;
  };
  for (int k = 0; k < nb; k++) HB_TRY(reliable(k, 1));

  // one batched iteration; captured once into a graph and replayed
  auto iteration = [&]() -> int {
    HB_TRY((multi_apply<S, G, NR>(c, f, plan.ngroups, grp, kap, rh, q, p, t, y, st, plan.clover)));
    const int tile = c->multi_tile;
    const int grid = nblocks(vh * NR, tile);
    k_multi_update<S, NR><<<grid, tile, 0, c->stream>>>(r, q, vh, st, part);
    k_finalize<NR><<<1, kFinThreads, 0, c->stream>>>(part, grid, nullptr, FinMultiUpdate<NR>{st});
    k_multi_pupdate<S, NR><<<grid, tile, 0, c->stream>>>(p, r, xlo, vh, st);
    HB_HIP(hipGetLastError());
    return HB_OK;
  };
  const int chunk = 8;
  // the captured chunk and its executable, destroyed on every return path
  struct Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~Graph() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
    }
  } gr;
  HB_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int crc = HB_OK;
  for (int k = 0; k < chunk && crc == HB_OK; k++) crc = iteration();
  hipError_t ce = hipStreamEndCapture(c->stream, &gr.graph);
  HB_TRY(crc);
  if (ce != hipSuccess) {
    set_error("hipStreamEndCapture: %s", hipGetErrorString(ce));
    return HB_E_HIP;
  }
  HB_HIP(hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0));
  int result = HB_OK;
  for (;;) {
    hipError_t le = hipGraphLaunch(gr.exec, c->stream);
    if (le == hipSuccess) le = hipMemcpyAsync(hs, st, sizeof(MultiState), hipMemcpyDeviceToHost, c->stream);
    if (le == hipSuccess) le = hipStreamSynchronize(c->stream);
    if (le != hipSuccess) {
      set_error("multi-RHS iteration: %s", hipGetErrorString(le));
      result = HB_E_HIP;
      break;
    }
    bool any_active = false, any_rel = false;
    for (int k = 0; k < nb; k++) {
      if (hs->flag[k] == 0) any_active = true;
      if (hs->flag[k] == 1) any_rel = true;
    }
    if (any_rel) {
      for (int k = 0; k < nb; k++)
        if (hs->flag[k] == 1) {
          result = reliable(k, 0);
          if (result) break;
        }
      if (result) break;
      hipError_t le2 = hipMemcpyAsync(hs, st, sizeof(MultiState), hipMemcpyDeviceToHost, c->stream);
      if (le2 == hipSuccess) le2 = hipStreamSynchronize(c->stream);
      if (le2 != hipSuccess) {
        set_error("multi-RHS reliable update: %s", hipGetErrorString(le2));
        result = HB_E_HIP;
        break;
      }
      any_active = false;
      for (int k = 0; k < nb; k++)
        if (hs->flag[k] == 0) any_active = true;
    }
    if (!any_active) break;
  }
  HB_TRY(result);
  for (int k = 0; k < nb; k++) {
    if (iters) iters[idx[k]] = hs->iter[k];
    if (true_rtr) true_rtr[idx[k]] = hs->res[k];
    if (nrel_out) nrel_out[idx[k]] = hs->nrel[k];
  }
  return HB_OK;
}

// autotuner: seconds per RHS of one batched matvec for each block size and
// workgroup size of the block kernels
struct TuneKey {
  int L0, L1, L2, L3, prec, nrhs, ngroups;
  std::vector<int> group_sizes;  // RHS per kappa group, in group order
  bool operator<(const TuneKey &o) const {
    return std::tie(L0, L1, L2, L3, prec, nrhs, ngroups, group_sizes) <
           std::tie(o.L0, o.L1, o.L2, o.L3, o.prec, o.nrhs, o.ngroups, o.group_sizes);
  }
};

// The RHS of each kappa group in blocks of `block`, one group at a time: a
// block never mixes kappa values.  Systems of one kappa converge at about the
// same iteration count, while a lighter mass takes several times the
// iterations of a heavier one (24^3x48: ~830 / 257 / 107 at kappa 0.1355 /
// 0.13 / 0.12), and a block runs until its slowest member is done with the
// converged ones masked out but still occupying their lanes.
std::vector<std::vector<int>> plan_blocks(const MultiPlan &plan, int nrhs, int block) {
  std::vector<std::vector<int>> blocks;
  for (int gi = 0; gi < std::max(plan.ngroups, 1); gi++) {
    std::vector<int> cur;
    for (int k = 0; k < nrhs; k++) {
      if (plan.ngroups > 0 && plan.group_of[k] != gi) continue;
      cur.push_back(k);
      if ((int)cur.size() == block) {
        blocks.push_back(cur);
        cur.clear();
      }
    }
    if (!cur.empty()) blocks.push_back(cur);
  }
  return blocks;
}
static std::map<TuneKey, std::pair<int, int>> g_tuned;  // (block, tile)
static std::mutex g_tuned_mutex;                        // contexts may be used from several threads

template <class S, class G, int NR>
int time_block(Context *c, const MultiFields<S, G> &f, const MultiPlan &plan, double *ms_per_rhs) {
  int rc = 0;
  SpinorPtr<S> p = block_field<S>(c, SLOT_MULTI + 0, SLOT_MULTI + 10, NR, &rc); HB_TRY(rc);
  SpinorPtr<S> q = block_field<S>(c, SLOT_MULTI + 2, SLOT_MULTI + 12, NR, &rc); HB_TRY(rc);
  SpinorPtr<S> t = block_field<S>(c, SLOT_MULTI + 3, SLOT_MULTI + 13, NR, &rc); HB_TRY(rc);
  SpinorPtr<S> y = block_field<S>(c, SLOT_MULTI + 4, SLOT_MULTI + 14, NR, &rc); HB_TRY(rc);
  MultiState *st = static_cast<MultiState *>(scratch(c, SLOT_MULTI + 6, sizeof(MultiState), &rc)); HB_TRY(rc);
  HB_TRY(ensure_host_stage(c, sizeof(MultiState)));
  MultiState *hs = static_cast<MultiState *>(c->host_stage);
  *hs = MultiState{};
  HB_HIP(hipMemcpyAsync(st, hs, sizeof(MultiState), hipMemcpyHostToDevice, c->stream));
  // zero input so the timed kernels read finite data (not captured in a graph)
  HB_HIP(hipMemsetAsync(p.p, 0, sizeof(typename S::ST) * 12 * (size_t)c->g.vh * NR, c->stream));
  if (S::kNorm) HB_HIP(hipMemsetAsync(p.norm, 0, sizeof(float) * (size_t)c->g.vh * NR, c->stream));
  int grp[16] = {0};
  double kap[16], rh[16] = {0};
  for (int k = 0; k < 16; k++) kap[k] = 0.13;
  for (int k = 0; k < NR; k++) grp[k] = k % std::max(plan.ngroups, 1);
  EventTimer tm;
  HB_TRY(tm.create());
  HB_TRY((multi_apply<S, G, NR>(c, f, plan.ngroups, grp, kap, rh, q, p, t, y, st, plan.clover)));
  HB_TRY(tm.start(c->stream));
  const int reps = 3;
  for (int k = 0; k < reps; k++)
    HB_TRY((multi_apply<S, G, NR>(c, f, plan.ngroups, grp, kap, rh, q, p, t, y, st, plan.clover)));
  HB_TRY(tm.stop(c->stream));
  float ms = 0.f;
  HB_TRY(tm.elapsed(&ms));
  *ms_per_rhs = ms / reps / NR;
  return HB_OK;
}

template <class S, class G>
int cg_multi_prec(Context *c, MultiPlan &plan, int op, int nrhs, const double *kappa, const double *rho,
                  double2 *const *x, const double2 *const *b, const float *rest, int maxiter,
                  int block, double delta, int max_inner, int *iters, double *true_rtr, int prec) {
  MultiFields<S, G> f;
  HB_TRY((multi_fields<S, G>(c, plan, f)));
  static const int cand[5] = {1, 2, 4, 8, 12};
  static const int tiles[2] = {kMinTile, 256};
  int tile = c->user_tile;
  if (block == 0) {
    TuneKey key{c->g.L[0], c->g.L[1], c->g.L[2], c->g.L[3], prec, nrhs, plan.ngroups, {}};
    for (int gi = 0; gi < plan.ngroups; gi++)
      key.group_sizes.push_back((int)std::count(plan.group_of.begin(), plan.group_of.end(), gi));
    bool cached = false;
    {
      std::lock_guard<std::mutex> lock(g_tuned_mutex);
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) {
        block = it->second.first;
        tile = it->second.second;
        cached = true;
      }
    }
    if (!cached) {
      double best = 1e30;
      for (int nr : cand) {
        if (nr > nrhs && nr != 1) break;
        if (!multi_block_fits(c->g.vh, nr, prec)) break;  // larger blocks do not fit either
        for (int tl : tiles) {
          c->multi_tile = tl;
          double ms = 0.0;
          int r;
          switch (nr) {
            case 1: r = time_block<S, G, 1>(c, f, plan, &ms); break;
            case 2: r = time_block<S, G, 2>(c, f, plan, &ms); break;
            case 4: r = time_block<S, G, 4>(c, f, plan, &ms); break;
            case 8: r = time_block<S, G, 8>(c, f, plan, &ms); break;
            default: r = time_block<S, G, 12>(c, f, plan, &ms); break;
          }
          c->multi_tile = 256;
          HB_TRY(r);
          // a partial block (the last of a kappa group) costs as much as a
          // full one; the smaller workgroup must win by 3 % (a single matvec
          // favours it slightly over what whole solves show)
          const int nblk = (int)plan_blocks(plan, nrhs, nr).size();
          const double total = ms * nr * nblk * (tl < 256 ? 1.03 : 1.0);
          if (total < best) {
            best = total;
            block = nr;
            tile = tl;
          }
        }
      }
      std::lock_guard<std::mutex> lock(g_tuned_mutex);
      g_tuned[key] = {block, tile};
    }
  }
  HB_CHECK(block == 1 || block == 2 || block == 4 || block == 8 || block == 12, HB_E_ARG,
           "cg_multi: block size %d (use 0, 1, 2, 4, 8 or 12)", block);
  HB_CHECK(multi_block_fits(c->g.vh, block, prec), HB_E_ARG,
           "cg_multi: a block of %d right-hand sides at %d sites per parity exceeds the 4 GiB span "
           "of one buffer in precision %d (largest block: %d)", block, c->g.vh, prec,
           hb_multi_block_limit(c->g.vh, prec));
  c->last_block = block;
  c->last_tile = tile;
  c->multi_tile = tile;
  int result = HB_OK;
  for (const std::vector<int> &idx : plan_blocks(plan, nrhs, block)) {
    if (result != HB_OK) break;
    int r;
    switch (block) {
      case 1: r = cg_multi_block<S, G, 1>(c, plan, f, op, idx, kappa, rho, x, b, rest, maxiter, delta, max_inner, iters, true_rtr, nullptr); break;
      case 2: r = cg_multi_block<S, G, 2>(c, plan, f, op, idx, kappa, rho, x, b, rest, maxiter, delta, max_inner, iters, true_rtr, nullptr); break;
      case 4: r = cg_multi_block<S, G, 4>(c, plan, f, op, idx, kappa, rho, x, b, rest, maxiter, delta, max_inner, iters, true_rtr, nullptr); break;
      case 8: r = cg_multi_block<S, G, 8>(c, plan, f, op, idx, kappa, rho, x, b, rest, maxiter, delta, max_inner, iters, true_rtr, nullptr); break;
      default: r = cg_multi_block<S, G, 12>(c, plan, f, op, idx, kappa, rho, x, b, rest, maxiter, delta, max_inner, iters, true_rtr, nullptr); break;
    }
    result = r;
  }
  c->multi_tile = 256;
  return result;
}

int cg_multi(Context *c, int op, int nrhs, const double *kappa, const double *rho, double csw,
             double2 *const *x, const double2 *const *b, const float *rest, int maxiter, int log,
             int prec, int block, int *iters, double *true_rtr) {
  HB_CHECK(op == HB_OP_MTDAGMT || op == HB_OP_WDAGW, HB_E_ARG, "cg_multi: operator %d", op);
  HB_CHECK(nrhs >= 1 && nrhs <= 4096, HB_E_ARG, "cg_multi: nrhs %d", nrhs);
  HB_CHECK(maxiter >= 1, HB_E_ARG, "cg_multi: maxiter must be >= 1");
  MultiPlan plan;
  for (int k = 0; k < nrhs; k++) {
    int gsel = -1;
    for (int gi = 0; gi < plan.ngroups; gi++)
      if (plan.gkappa[gi] == kappa[k]) gsel = gi;
    if (gsel < 0) {
      gsel = plan.ngroups++;
      plan.gkappa.push_back(kappa[k]);
    }
    plan.group_of.push_back(gsel);
  }
  HB_CHECK(plan.ngroups <= kMaxKappaGroups, HB_E_ARG, "cg_multi: at most %d distinct kappa values",
           kMaxKappaGroups);
  int rc = 0;
  if (csw > 0.0) {
    plan.clover = true;
    for (int gi = 0; gi < plan.ngroups; gi++) {
      double2 *A = static_cast<double2 *>(scratch(c, SLOT_MULTI_CLOVER + 2 * gi, sizeof(double2) * 72 * (size_t)c->g.vh, &rc)); HB_TRY(rc);
      double2 *Ai = static_cast<double2 *>(scratch(c, SLOT_MULTI_CLOVER + 2 * gi + 1, sizeof(double2) * 72 * (size_t)c->g.vh, &rc)); HB_TRY(rc);
      double2 *B = static_cast<double2 *>(scratch(c, SLOT_MULTI_CLOVER_B, sizeof(double2) * 72 * (size_t)c->g.vh, &rc)); HB_TRY(rc);
      HB_TRY(build_clover_into(c, csw * plan.gkappa[gi], A, Ai, B));
      plan.clovA.push_back(A);
      plan.clovI.push_back(Ai);
    }
  } else if (csw < 0.0) {
    HB_CHECK(plan.ngroups == 1, HB_E_ARG, "cg_multi: the loaded clover term serves one kappa only");
    HB_CHECK(use_clover(c), HB_E_STATE, "cg_multi: no clover field loaded");
    plan.clover = true;
    plan.clovA.push_back(c->clov[HB_CLOVER_A]);
    plan.clovI.push_back(c->clov[HB_CLOVER_AINV]);
  }
  const double delta = 0.1;
  const int max_inner = 1000;
  int r;
  if (prec == HB_PREC_F32)
    r = cg_multi_prec<SF32, SF32>(c, plan, op, nrhs, kappa, rho, x, b, rest, maxiter, block, delta, max_inner, iters, true_rtr, prec);
  else if (prec == HB_PREC_F16)
    r = cg_multi_prec<SF16, SF16>(c, plan, op, nrhs, kappa, rho, x, b, rest, maxiter, block, delta, max_inner, iters, true_rtr, prec);
  else {
    set_error("cg_multi: precision %d (use 32 or 16)", prec);
    return HB_E_ARG;
  }
  HB_TRY(r);
  HB_HIP(hipStreamSynchronize(c->stream));
  if (log != 2 && true_rtr) {
    for (int k = 0; k < nrhs; k++)
      if (!(true_rtr[k] <= (double)rest[k])) {
        set_error("cg_multi(): RHS %d did not converge; true rtr = %.17g", k, true_rtr[k]);
        return HB_E_NOCONV;
      }
  }
  return HB_OK;
}

}  // namespace hb

using namespace hb;

extern "C" {

int hb_cg_multi_d(hb_context *ctx, int op, int nrhs, const double *kappa, const double *rho,
                  double csw, double *const *x_d, const double *const *b_d, const float *rest,
                  int maxiter, int log, int precision, int block, int *iters, double *true_rtr) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(kappa && rho && x_d && b_d && rest && iters && true_rtr, HB_E_ARG, "hb_cg_multi: null argument");
  HB_HIP(hipSetDevice(c->device));
  return cg_multi(c, op, nrhs, kappa, rho, csw, reinterpret_cast<double2 *const *>(x_d),
                  reinterpret_cast<const double2 *const *>(b_d), rest, maxiter, log, precision, block,
                  iters, true_rtr);
}

int hb_cg_multi(hb_context *ctx, int op, int nrhs, const double *kappa, const double *rho, double csw,
                double *const *x, const double *const *b, const float *rest, int maxiter, int log,
                int precision, int block, int *iters, double *true_rtr) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(kappa && rho && x && b && rest && iters && true_rtr, HB_E_ARG, "hb_cg_multi: null argument");
  HB_CHECK(nrhs >= 1 && nrhs <= 64, HB_E_ARG, "hb_cg_multi: nrhs %d (1..64 for host arrays)", nrhs);
  HB_HIP(hipSetDevice(c->device));
  const size_t bytes = spinor_bytes(c);
  std::vector<double2 *> dx(nrhs), db(nrhs);
  int rc = HB_OK;
  for (int k = 0; k < nrhs && rc == HB_OK; k++) {
    hipError_t e = hipMalloc(&dx[k], bytes);
    if (e == hipSuccess) e = hipMalloc(&db[k], bytes);
    if (e != hipSuccess) {
      set_error("hb_cg_multi: %s", hipGetErrorString(e));
      rc = HB_E_NOMEM;
    }
  }
  for (int k = 0; k < nrhs && rc == HB_OK; k++) {
    rc = spinor_upload(c, dx[k], x[k]);
    if (rc == HB_OK) rc = spinor_upload(c, db[k], b[k]);
  }
  int r = rc == HB_OK ? cg_multi(c, op, nrhs, kappa, rho, csw, dx.data(), db.data(), rest, maxiter, log,
                                  precision, block, iters, true_rtr)
                      : rc;
  if (r == HB_OK || r == HB_E_NOCONV)
    for (int k = 0; k < nrhs; k++) {
      int r2 = spinor_download(c, x[k], dx[k]);
      if (r2 != HB_OK) r = r2;
    }
  for (int k = 0; k < nrhs; k++) {
    (void)hipFree(dx[k]);
    (void)hipFree(db[k]);
  }
  return r;
}

}  // extern "C"
