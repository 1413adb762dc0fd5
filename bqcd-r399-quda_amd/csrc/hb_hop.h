// hb_hop.h - one hop of the hopping matrix and the scheduling fence between
// hops; shared by the single- and multi-RHS kernels.
#pragma once

#include "hb_kernels.h"

namespace hb {

// One hop: acc += (1 + s gamma_mu) U psi(x +- mu), U = U_mu(x) forward or
// U_mu(x - mu)^+ backward, with the boundary sign when the hop wraps.
// For a block of NR right-hand sides stored RHS-fastest, the spinor of site j
// for RHS r sits at index j*NR + r of the view (NR = 1: plain layout).
template <class S, class G, int MU, bool FWD, int SG, int NR = 1, class View = SpinorView<S>>
__device__ __forceinline__ void hop(typename S::V (&acc)[12], const typename G::ST *U,
                                    const View &in, const Geom &g, int e, int i,
                                    const int (&xc)[4], int r = 0) {
  using V = typename S::V;
  using A = typename S::A;
  const int vh = g.vh;
  int c[4] = {xc[0], xc[1], xc[2], xc[3]};
  A sgn = (A)1;
  if (FWD) {
    c[MU] += 1;
    if (c[MU] == g.L[MU]) { c[MU] = 0; sgn = (A)g.bc[MU]; }
  } else {
    c[MU] -= 1;
    if (c[MU] < 0) { c[MU] = g.L[MU] - 1; sgn = (A)g.bc[MU]; }
  }
  const int j = site_index(g, c[0], c[1], c[2], c[3]);
  V h[6], w[6];
  if constexpr (S::kNorm || link_raw<G>()) {
    // 16-bit words are used as stored: the spinor's site factor, the link's
    // 1/32767 and the boundary sign make one factor f of the hop, applied by
    // the multiply-adds that spread U h into acc
    A f;
    if constexpr (MU == 3) {
      V p[6];
      f = (A)2 * sgn * in.load_pair_raw(p, j * NR + r, SG < 0 ? 2 : 0);
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        h[2 * cc] = p[cc];
        h[2 * cc + 1] = p[3 + cc];
      }
    } else {
      V p[12];
      f = sgn * in.template load_raw_dir<MU>(p, j * NR + r);
      project<MU, SG>(h, p);
    }
    V u[9];
    if (FWD) f *= load_link_f<G>(u, U, MU, e, i, vh);
    else f *= load_link_f<G>(u, U, MU, 1 - e, j, vh);
    if (FWD) su3_mul_half(w, u, h);
    else su3_adj_mul_half(w, u, h);
    reconstruct<MU, SG, true>(acc, w, f);
  } else {
    if constexpr (MU == 3) {
      V p[6];
      in.load_pair_raw(p, j * NR + r, SG < 0 ? 2 : 0);  // values as stored: factor 1
      const A f = (A)2 * sgn;
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        h[2 * cc] = cscale<V>(f, p[cc]);
        h[2 * cc + 1] = cscale<V>(f, p[3 + cc]);
      }
    } else {
      V p[12];
      in.load(p, j * NR + r);
      project<MU, SG>(h, p);
      if (sgn < (A)0) {
#pragma unroll
        for (int k = 0; k < 6; k++) h[k] = cmk<V>(-h[k].x, -h[k].y);
      }
    }
    V u[9];
    if (FWD) {
      load_link<G>(u, U, MU, e, i, vh);
      su3_mul_half(w, u, h);
    } else {
      load_link<G>(u, U, MU, 1 - e, j, vh);
      su3_adj_mul_half(w, u, h);
    }
    reconstruct<MU, SG>(acc, w);
  }
}

// Completes one hop before the next starts: the accumulator is pinned in
// registers (so its arithmetic cannot be sunk to the final store) and the
// scheduler may not hoist the next hop's loads above this point.
template <class V> __device__ __forceinline__ void pin(V (&acc)[12]) {
#pragma unroll
  for (int k = 0; k < 12; k++) asm volatile("" : "+v"(acc[k].x), "+v"(acc[k].y));
}
#define HB_SCHED_FENCE() do { pin(acc); __builtin_amdgcn_sched_barrier(0); } while (0)

// Hops issued between fences: the loads of the hops of one group are in
// flight together.  FP64 keeps one (its 16-byte loads already hold enough
// bytes in flight per wave, and two FP64 hops spill); the reduced precisions
// issue two (the +mu and -mu hops), which overlaps one hop's arithmetic with
// the other's loads (2% on the 16-bit solver; four or eight hops gain no more
// and eight costs occupancy).
// the 16-bit hopping kernels take the x and y neighbours from other lanes of
// the wave where they are there (WaveView16)
#ifndef HB_WAVE_SHARE
#define HB_WAVE_SHARE 1
#endif
#ifndef HB_HOPS_PER_FENCE_REDUCED
#define HB_HOPS_PER_FENCE_REDUCED 2
#endif
template <class S> constexpr int hops_per_fence() {
  return sizeof(typename S::ST) == 16 ? 1 : HB_HOPS_PER_FENCE_REDUCED;
}

// ---- 16-bit hops with the loads of the next hop in flight ------------------
// A 16-bit hop is split into its loads (the stored words, no conversion) and
// its arithmetic; the hopping matrix fetches hop k+1 before it computes hop k,
// so that every hop's arithmetic overlaps the next hop's memory latency.
#ifndef HB_HOP_PIPELINE
#define HB_HOP_PIPELINE 1
#endif

struct HopWords16 {
  uint32_t s[13];  // neighbour spinor words (a t hop uses s[0..5]) and its scale word s[12]
  uint32_t l[6];   // link rows 0 and 1
  float sgn;       // boundary sign of the hop
};

template <class S, class G, int MU, bool FWD, int SG, int NR, class View>
__device__ __forceinline__ void hop_fetch16(HopWords16 &d, const typename G::ST *U, const View &in,
                                            const Geom &g, int e, int i, const int (&xc)[4], int r) {
  int c[4] = {xc[0], xc[1], xc[2], xc[3]};
  d.sgn = 1.0f;
  if (FWD) {
    c[MU] += 1;
    if (c[MU] == g.L[MU]) { c[MU] = 0; d.sgn = (float)g.bc[MU]; }
  } else {
    c[MU] -= 1;
    if (c[MU] < 0) { c[MU] = g.L[MU] - 1; d.sgn = (float)g.bc[MU]; }
  }
  const int j = site_index(g, c[0], c[1], c[2], c[3]);
  if constexpr (MU == 3) in.fetch_pair(d.s, j * NR + r, SG < 0 ? 2 : 0);
  else in.template fetch_dir<MU>(d.s, j * NR + r);
  if (FWD) fetch_link16(d.l, U, MU, e, i, g.vh);
  else fetch_link16(d.l, U, MU, 1 - e, j, g.vh);
}

template <int MU, bool FWD, int SG>
__device__ __forceinline__ void hop_finish16(float2 (&acc)[12], const HopWords16 &d) {
  float2 h[6], w[6];
  float f;
  if constexpr (MU == 3) {
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      h[2 * cc] = SF16::raw(as_short2(d.s[cc]));
      h[2 * cc + 1] = SF16::raw(as_short2(d.s[3 + cc]));
    }
    f = 2.0f * d.sgn * as_float(d.s[12]) * SF16::kInv;
  } else {
    float2 p[12];
#pragma unroll
    for (int k = 0; k < 12; k++) p[k] = SF16::raw(as_short2(d.s[k]));
    project<MU, SG>(h, p);
    f = d.sgn * as_float(d.s[12]) * SF16::kInv;
  }
  float2 u[9];
  decode_link16(u, d.l);
  f *= SF16::kInv;
  if (FWD) {
    su3_mul_half(w, u, h);  // colour 3 of w from row 3 alone
    reconstruct<MU, SG, true, true>(acc, w, f, f * SF16::kInv);
  } else {
    h[4] = cscale(SF16::kInv, h[4]);  // colour 3 of h meets row 3 alone
    h[5] = cscale(SF16::kInv, h[5]);
    su3_adj_mul_half(w, u, h);
    reconstruct<MU, SG, true>(acc, w, f);
  }
}

template <class S, class G, bool DAG, int NR, class View>
__device__ __forceinline__ void hop_all_pipelined16(float2 (&acc)[12], const typename G::ST *U,
                                                    const View &in, const Geom &g, int e, int i,
                                                    const int (&xc)[4], int r) {
  constexpr int SF = DAG ? 1 : -1;
  constexpr int SB = -SF;
  HopWords16 a, b;
#define HB_F(MU, FWD, SG, D) hop_fetch16<S, G, MU, FWD, SG, NR, View>(D, U, in, g, e, i, xc, r); \
  __builtin_amdgcn_sched_barrier(0);
#define HB_C(MU, FWD, SG, D) hop_finish16<MU, FWD, SG>(acc, D); HB_SCHED_FENCE();
  HB_F(3, true, SF, a)
  HB_F(3, false, SB, b) HB_C(3, true, SF, a)
  HB_F(2, true, SF, a) HB_C(3, false, SB, b)
  HB_F(2, false, SB, b) HB_C(2, true, SF, a)
  HB_F(1, true, SF, a) HB_C(2, false, SB, b)
  HB_F(1, false, SB, b) HB_C(1, true, SF, a)
  HB_F(0, true, SF, a) HB_C(1, false, SB, b)
  HB_F(0, false, SB, b) HB_C(0, true, SF, a)
  HB_C(0, false, SB, b)
#undef HB_F
#undef HB_C
}

// all 8 hops of the hopping matrix into acc, t first (the reference's
// direction order does not matter for the sum); the forward projector sign is
// (1 - gamma) for D and (1 + gamma) for D^+
template <class S, class G, bool DAG, int NR = 1, class View = SpinorView<S>>
__device__ __forceinline__ void hop_all(typename S::V (&acc)[12], const typename G::ST *U,
                                        const View &in, const Geom &g, int e, int i,
                                        const int (&xc)[4], int r = 0) {
  if constexpr (HB_HOP_PIPELINE && std::is_same<S, SF16>::value && std::is_same<G, SF16>::value) {
    hop_all_pipelined16<S, G, DAG, NR, View>(acc, U, in, g, e, i, xc, r);
    return;
  }
  constexpr int F = hops_per_fence<S>();
  constexpr int SF = DAG ? 1 : -1;
  constexpr int SB = -SF;
#define HB_HOP(MU, FWD, SG, N)                                    \
  hop<S, G, MU, FWD, SG, NR, View>(acc, U, in, g, e, i, xc, r);   \
  if constexpr ((N) % F == 0) HB_SCHED_FENCE();
  HB_HOP(3, true, SF, 1) HB_HOP(3, false, SB, 2)
  HB_HOP(2, true, SF, 3) HB_HOP(2, false, SB, 4)
  HB_HOP(1, true, SF, 5) HB_HOP(1, false, SB, 6)
  HB_HOP(0, true, SF, 7) HB_HOP(0, false, SB, 8)
#undef HB_HOP
}

}  // namespace hb
