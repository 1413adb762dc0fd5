// hb_hop.h - one hop of the hopping matrix and the scheduling fence between
// hops; shared by the single- and multi-RHS kernels.
#pragma once

#include "hb_kernels.h"

namespace hb {

// One hop: acc += (1 + s gamma_mu) U psi(x +- mu), U = U_mu(x) forward or
// U_mu(x - mu)^+ backward, with the boundary sign when the hop wraps.
// For a block of NR right-hand sides stored RHS-fastest, the spinor of site j
// for RHS r sits at index j*NR + r of the view (NR = 1: plain layout).
template <class S, class G, int MU, bool FWD, int SG, int NR = 1>
__device__ __forceinline__ void hop(typename S::V (&acc)[12], const typename G::ST *U,
                                    const SpinorView<S> &in, const Geom &g, int e, int i,
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
  if constexpr (MU == 3) {
    V p[6];
    in.load_pair(p, j * NR + r, SG < 0 ? 2 : 0);
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

// Completes one hop before the next starts: the accumulator is pinned in
// registers (so its arithmetic cannot be sunk to the final store) and the
// scheduler may not hoist the next hop's loads above this point.
template <class V> __device__ __forceinline__ void pin(V (&acc)[12]) {
#pragma unroll
  for (int k = 0; k < 12; k++) asm volatile("" : "+v"(acc[k].x), "+v"(acc[k].y));
}
#define HB_SCHED_FENCE() do { pin(acc); __builtin_amdgcn_sched_barrier(0); } while (0)

}  // namespace hb
