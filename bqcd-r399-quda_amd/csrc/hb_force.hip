// hb_force.hip - fermion force (dsf.F90 + d/DSFxyzt.F90 + clover/clover_d.F90),
// clover-determinant force (dsd.F90, clover_dsd), Wilson gauge force
// (dsg.F90), link update (hmc_u.F90) and the clover action
// (clover/clover_action.F90), all in FP64 on device-resident fields.
//
// Every force kernel is a gather over links: one thread owns one link
// (mu, eo, i) and adds all terms of that link to its 8 momentum coordinates,
// so no atomics are needed and the result is deterministic.  Momenta are
// device SoA P[4 mu][2 eo][8][vh].
#include "hb_ops.h"
#include "hb_su3.h"

namespace hb {

// ---------------------------------------------------------------------------
// spinor access helpers (FP64 SoA, component k = 3*spin + colour)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ld_spinor(double2 (&v)[12], const double2 *__restrict__ f, int i, int vh) {
#pragma unroll
  for (int k = 0; k < 12; k++) v[k] = f[(size_t)k * vh + i];
}

// (1 - sg*gamma_mu) a, all four spin components (d/DSFxyzt.F90:26-56):
// sg = +1 for the forward term, -1 for the backward term
template <int MU>
__device__ __forceinline__ void gamma_a(double2 (&ga)[12], const double2 (&a)[12], double sg) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double2 a0 = a[c], a1 = a[3 + c], a2 = a[6 + c], a3 = a[9 + c];
    if constexpr (MU == 3) {
      const double2 z = make_double2(0.0, 0.0);
      if (sg > 0) {
        ga[c] = z; ga[3 + c] = z;
        ga[6 + c] = cscale<double2>(2.0, a2); ga[9 + c] = cscale<double2>(2.0, a3);
      } else {
        ga[c] = cscale<double2>(2.0, a0); ga[3 + c] = cscale<double2>(2.0, a1);
        ga[6 + c] = z; ga[9 + c] = z;
      }
    } else {
      // h = a_up - sg*G a_lo ; lower = -sg*G^+ h
      double2 h0, h1, l0, l1;
      if constexpr (MU == 0) {  // G = [[0,i],[i,0]]
        h0 = csub(a0, cscale<double2>(sg, ci(a3)));
        h1 = csub(a1, cscale<double2>(sg, ci(a2)));
        l0 = cscale<double2>(sg, ci(h1));   // -sg * (-i h1)
        l1 = cscale<double2>(sg, ci(h0));
      } else if constexpr (MU == 1) {  // G = [[0,1],[-1,0]]
        h0 = csub(a0, cscale<double2>(sg, a3));
        h1 = cadd(a1, cscale<double2>(sg, a2));
        l0 = cscale<double2>(sg, h1);       // -sg * (-h1)
        l1 = cscale<double2>(-sg, h0);      // -sg * h0
      } else {  // G = [[i,0],[0,-i]]
        h0 = csub(a0, cscale<double2>(sg, ci(a2)));
        h1 = cadd(a1, cscale<double2>(sg, ci(a3)));
        l0 = cscale<double2>(sg, ci(h0));   // -sg * (-i h0)
        l1 = cscale<double2>(-sg, ci(h1));  // -sg * (i h1)
      }
      ga[c] = h0; ga[3 + c] = h1; ga[6 + c] = l0; ga[9 + c] = l1;
    }
  }
}

// w(ca, cb) = sum_s ga_s(ca) conj(b_s(cb))
__device__ __forceinline__ M3 outer(const double2 (&ga)[12], const double2 (&b)[12]) {
  M3 w;
#pragma unroll
  for (int ca = 0; ca < 3; ca++)
#pragma unroll
    for (int cb = 0; cb < 3; cb++) {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int s = 0; s < 4; s++) cfma(acc, ga[3 * s + ca], cconj(b[3 * s + cb]));
      w.m[3 * ca + cb] = acc;
    }
  return w;
}

__device__ __forceinline__ double *mom(double *P, int mu, int eo, int vh) {
  return P + (size_t)((mu * 2 + eo) * 8) * vh;
}
__device__ __forceinline__ void mom_add(double *P, int mu, int eo, int i, int vh, const double (&q)[8]) {
  double *b = mom(P, mu, eo, vh) + i;
#pragma unroll
  for (int j = 0; j < 8; j++) b[(size_t)j * vh] += q[j];
}

// ---------------------------------------------------------------------------
// hopping part of the fermion force (dsf.F90:77-98 with DSFxyzt.F90:77-117)
// ---------------------------------------------------------------------------
// Bvec/Avec of parity E (index 0) and of the other parity (index 1):
// E = EVEN: forward term (b_i, at_j), backward (a_i, bt_j);
// E = ODD : forward term (bt_i, a_j), backward (at_i, b_j).
template <int MU>
__device__ void hop_force_link(double (&q)[8], const double2 *__restrict__ U, const Geom &g, int e, int i,
                               const int (&x)[4], const double2 *Bi, const double2 *Aj,
                               const double2 *Ai, const double2 *Bj, double s) {
  const int vh = g.vh;
  int c[4] = {x[0], x[1], x[2], x[3]};
  double sgn = 1.0;
  c[MU] += 1;
  if (c[MU] == g.L[MU]) { c[MU] = 0; sgn = (double)g.bc[MU]; }
  const int j = site_index(g, c[0], c[1], c[2], c[3]);
  M3 u = link(U, MU, e, i, vh);
  m3_scale(u, sgn);  // the reference applies the boundary sign to the links (flip_bc)
  double2 av[12], ga[12], bv[12];
  {  // forward: v = U w, w = ((1 - g) A_j) (x) conj(B_i); p -= s Im Tr(lambda v)
    ld_spinor(av, Aj, j, vh);
    ld_spinor(bv, Bi, i, vh);
    gamma_a<MU>(ga, av, 1.0);
    M3 v = m3_mul(u, outer(ga, bv));
    im_tr_j(q, v, -s);
  }
  {  // backward: v = w U^+, w = ((1 + g) A_i) (x) conj(B_j); p += s Im Tr(lambda v)
    ld_spinor(av, Ai, i, vh);
    ld_spinor(bv, Bj, j, vh);
    gamma_a<MU>(ga, av, -1.0);
    M3 v = m3_muld(outer(ga, bv), u);
    im_tr_j(q, v, s);
  }
}

__global__ __launch_bounds__(128) void k_hop_force(double *P, const double2 *__restrict__ U, Geom g,
                                                   const double2 *a, const double2 *b, const double2 *at,
                                                   const double2 *bt, double s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int me = blockIdx.y, mu = me >> 1, e = me & 1;
  if (i >= g.vh) return;
  int x[4];
  site_coords(g, e, i, x[0], x[1], x[2], x[3]);
  const double2 *Bi = e == 0 ? b : bt, *Aj = e == 0 ? at : a;
  const double2 *Ai = e == 0 ? a : at, *Bj = e == 0 ? bt : b;
  double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  switch (mu) {
    case 0: hop_force_link<0>(q, U, g, e, i, x, Bi, Aj, Ai, Bj, s); break;
    case 1: hop_force_link<1>(q, U, g, e, i, x, Bi, Aj, Ai, Bj, s); break;
    case 2: hop_force_link<2>(q, U, g, e, i, x, Bi, Aj, Ai, Bj, s); break;
    default: hop_force_link<3>(q, U, g, e, i, x, Bi, Aj, Ai, Bj, s); break;
  }
  mom_add(P, mu, e, i, g.vh, q);
}

// ---------------------------------------------------------------------------
// clover parts: the six sigma_mu_nu fields w (clover_bsa / clover_ts) and the
// derivative of the clover leaves (clover_d)
// ---------------------------------------------------------------------------
__constant__ int c_pairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
// spin matrices of clover_bsa_12..34: (S a)_s = sum S[s][s'] a_s'  (entries 0, +-1, +-i)
__constant__ double2 c_bsa[6][4][4] = {
    {{{-1, 0}, {0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {1, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {-1, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}, {1, 0}}},
    {{{0, 0}, {0, -1}, {0, 0}, {0, 0}}, {{0, 1}, {0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}, {0, -1}}, {{0, 0}, {0, 0}, {0, 1}, {0, 0}}},
    {{{0, 0}, {0, 0}, {0, 0}, {1, 0}}, {{0, 0}, {0, 0}, {1, 0}, {0, 0}}, {{0, 0}, {1, 0}, {0, 0}, {0, 0}}, {{1, 0}, {0, 0}, {0, 0}, {0, 0}}},
    {{{0, 0}, {-1, 0}, {0, 0}, {0, 0}}, {{-1, 0}, {0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}, {-1, 0}}, {{0, 0}, {0, 0}, {-1, 0}, {0, 0}}},
    {{{0, 0}, {0, 0}, {0, 0}, {0, -1}}, {{0, 0}, {0, 0}, {0, 1}, {0, 0}}, {{0, 0}, {0, -1}, {0, 0}, {0, 0}}, {{0, 1}, {0, 0}, {0, 0}, {0, 0}}},
    {{{0, 0}, {0, 0}, {1, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}, {-1, 0}}, {{1, 0}, {0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {-1, 0}, {0, 0}, {0, 0}}},
};
// coefficients of t(s1,c1,s2,c2) in clover_ts_12..34
__constant__ double2 c_ts[6][4][4] = {
    {{{-1, 0}, {0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {1, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {-1, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}, {1, 0}}},
    {{{0, 0}, {0, 1}, {0, 0}, {0, 0}}, {{0, -1}, {0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}, {0, 1}}, {{0, 0}, {0, 0}, {0, -1}, {0, 0}}},
    {{{0, 0}, {0, 0}, {0, 0}, {1, 0}}, {{0, 0}, {0, 0}, {1, 0}, {0, 0}}, {{0, 0}, {1, 0}, {0, 0}, {0, 0}}, {{1, 0}, {0, 0}, {0, 0}, {0, 0}}},
    {{{0, 0}, {-1, 0}, {0, 0}, {0, 0}}, {{-1, 0}, {0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}, {-1, 0}}, {{0, 0}, {0, 0}, {-1, 0}, {0, 0}}},
    {{{0, 0}, {0, 0}, {0, 0}, {0, 1}}, {{0, 0}, {0, 0}, {0, -1}, {0, 0}}, {{0, 0}, {0, 1}, {0, 0}, {0, 0}}, {{0, -1}, {0, 0}, {0, 0}, {0, 0}}},
    {{{0, 0}, {0, 0}, {1, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}, {-1, 0}}, {{1, 0}, {0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {-1, 0}, {0, 0}, {0, 0}}},
};

// W[6][9][vh]: w_k = transpose(conj(b) S_k a) at the sites of one parity (clover_bsa)
__global__ __launch_bounds__(kBlock) void k_w_bsa(double2 *__restrict__ W, const double2 *__restrict__ b, const double2 *__restrict__ a,
                        int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  double2 av[12], bv[12];
  ld_spinor(av, a, i, vh);
  ld_spinor(bv, b, i, vh);
  for (int k = 0; k < 6; k++) {
    double2 sa[12];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int s = 0; s < 4; s++) {
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++) cfma(acc, c_bsa[k][s][s2], av[3 * s2 + c]);
        sa[3 * s + c] = acc;
      }
    M3 w = outer(sa, bv);
#pragma unroll
    for (int m = 0; m < 9; m++) W[((size_t)k * 9 + m) * vh + i] = w.m[m];
  }
}

// W from T = A^-1 (12x12, by LDL^+ solves with the B field, clover_t_init) and
// the coefficients of clover_ts
__global__ __launch_bounds__(128) void k_w_ts(double2 *__restrict__ W, const double2 *__restrict__ Cb, int eo,
                                              int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  double2 w[6][9];
#pragma unroll
  for (int k = 0; k < 6; k++)
#pragma unroll
    for (int m = 0; m < 9; m++) w[k][m] = make_double2(0.0, 0.0);
  for (int s2 = 0; s2 < 4; s2++)
    for (int c2 = 0; c2 < 3; c2++) {
      double2 e[12], col[12];
#pragma unroll
      for (int k = 0; k < 12; k++) e[k] = make_double2(k == 3 * s2 + c2 ? 1.0 : 0.0, 0.0);
      clover_apply<SF64, 1>(col, e, Cb, nullptr, eo, i, vh);  // column (s2,c2) of A^-1
#pragma unroll
      for (int k = 0; k < 6; k++)
#pragma unroll
        for (int s1 = 0; s1 < 4; s1++) {
          const double2 cf = c_ts[k][s1][s2];
          if (cf.x == 0.0 && cf.y == 0.0) continue;
#pragma unroll
          for (int c1 = 0; c1 < 3; c1++) cfma(w[k][3 * c1 + c2], cf, col[3 * s1 + c1]);
        }
    }
#pragma unroll
  for (int k = 0; k < 6; k++)
#pragma unroll
    for (int m = 0; m < 9; m++) W[((size_t)k * 9 + m) * vh + i] = w[k][m];
}

__device__ __forceinline__ M3 wmat(const double2 *__restrict__ W, int k, int i, int vh) {
  M3 r;
#pragma unroll
  for (int m = 0; m < 9; m++) r.m[m] = W[((size_t)k * 9 + m) * vh + i];
  return r;
}
// the same through a buffer resource over matrix k's nine planes (k uniform)
__device__ __forceinline__ M3 wmat_b(const double2 *__restrict__ W, int k, int i, int vh) {
  const uint32_t plane = (uint32_t)vh * 16u;
  const rsrc_t r = make_rsrc(W + (size_t)k * 9 * vh, plane * 9u);
  const uint32_t vo = (uint32_t)i * 16u;
  M3 m;
#pragma unroll
  for (int q = 0; q < 9; q++) m.m[q] = bld_f64x2(r, vo, plane * (uint32_t)q);
  return m;
}

// clover_d_loop (clover/clover_d.F90:152-200) for one link (mu, e, i) and
// plane partner nu; the w fields live on parity weo.  The eight terms of
// clover_d_same_eo (:203-233) or clover_d_diff_eo (:236-266) are evaluated in
// groups that load their links and w matrices when needed (reloads hit the
// caches), with a scheduling barrier between groups, so that at most five
// 3x3 matrices are live: loading all seven links and three w matrices first
// needs ~360 registers and spills.
#define HB_GROUP_END() __builtin_amdgcn_sched_barrier(0)
// an index the compiler cannot prove equal to its argument: a matrix loaded
// again through it is a fresh load, not a register kept live since the first
__device__ __forceinline__ int fresh(int i) {
  asm volatile("" : "+v"(i));
  return i;
}
__device__ void d_loop(double (&q)[8], const double2 *__restrict__ U, const Geom &g, int e, int mu, int nu,
                       int i, const int (&x)[4], const double2 *__restrict__ W, int k, int weo, double s) {
  const int o = 1 - e, vh = g.vh;
  const int id = nb(g, x[0], x[1], x[2], x[3], mu, 1);
  const int ie = nb(g, x[0], x[1], x[2], x[3], nu, 1);
  const int iff = nb(g, x[0], x[1], x[2], x[3], nu, -1);
  if (e == weo) {
    // clover_d_same_eo
    const int ib = nb2(g, x[0], x[1], x[2], x[3], mu, 1, nu, 1);
    const int ic = nb2(g, x[0], x[1], x[2], x[3], mu, 1, nu, -1);
    {
      const M3 u = m3_mul(link_b(U, mu, e, i, vh), link_b(U, nu, o, id, vh));   // u0 u1
      HB_GROUP_END();
      const M3 v = m3_mul(link_b(U, nu, e, i, vh), link_b(U, mu, o, ie, vh));   // u3 u2
      HB_GROUP_END();
      {
        const M3 wa = wmat_b(W, k, i, vh);
        re_tr_j_prod<0>(q, m3_muld(u, v), wa, s);
        HB_GROUP_END();
        re_tr_j_prod<2>(q, m3_mul(wa, v), u, s);
      }
      HB_GROUP_END();
      {
        const M3 wb = wmat_b(W, k, ib, vh);
        re_tr_j_prod<2>(q, m3_mul(u, wb), v, s);
        HB_GROUP_END();
        re_tr_j_prod<2>(q, m3_mul(v, wb), u, s);
      }
    }
    HB_GROUP_END();
    {
      const M3 u = m3_muld(link_b(U, mu, e, fresh(i), vh), link_b(U, nu, e, ic, vh));  // u0 u6^+
      HB_GROUP_END();
      const M3 v = m3_dmul(link_b(U, nu, o, iff, vh), link_b(U, mu, o, iff, vh));  // u4^+ u5
      HB_GROUP_END();
      {
        const M3 wa = wmat_b(W, k, fresh(i), vh);
        re_tr_j_prod<2>(q, m3_mul(wa, v), u, -s);
        HB_GROUP_END();
        re_tr_j_prod<0>(q, m3_muld(u, v), wa, -s);
      }
      HB_GROUP_END();
      {
        const M3 wc = wmat_b(W, k, ic, vh);
        re_tr_j_prod<2>(q, m3_mul(v, wc), u, -s);
        HB_GROUP_END();
        re_tr_j_prod<2>(q, m3_mul(u, wc), v, -s);
      }
    }
  } else {
    // clover_d_diff_eo
    {
      const M3 t = m3_muld(link_b(U, nu, o, id, vh), link_b(U, mu, o, ie, vh));  // u1 u2^+
      HB_GROUP_END();
      const M3 u = m3_muld(t, link_b(U, nu, e, i, vh));                          // u1 u2^+ u3^+
      HB_GROUP_END();
      {
        const M3 wd = wmat_b(W, k, id, vh);
        const M3 u0 = link_b(U, mu, e, fresh(i), vh);
        re_tr_j_prod<0>(q, m3_mul(u0, wd), u, s);
        HB_GROUP_END();
        re_tr_j_prod<2>(q, m3_dmul(u, wd), u0, s);
      }
      HB_GROUP_END();
      const M3 v = m3_dmul(t, m3_dag(link_b(U, mu, e, fresh(i), vh)));                  // u2 u1^+ u0^+
      HB_GROUP_END();
      {
        const M3 we = wmat_b(W, k, ie, vh);
        const M3 u3 = link_b(U, nu, e, fresh(i), vh);
        re_tr_j_prod<2>(q, m3_dmul(v, we), u3, s);
        HB_GROUP_END();
        re_tr_j_prod<0>(q, m3_mul(u3, we), v, s);
      }
    }
    HB_GROUP_END();
    {
      const M3 t = m3_mul(link_b(U, mu, o, iff, vh), link_b(U, nu, e, nb2(g, x[0], x[1], x[2], x[3], mu, 1, nu, -1), vh));  // u5 u6
      HB_GROUP_END();
      const M3 u = m3_dmul(t, link_b(U, nu, o, iff, vh));                        // u6^+ u5^+ u4
      HB_GROUP_END();
      {
        const M3 wd = wmat_b(W, k, fresh(id), vh);
        const M3 u0 = link_b(U, mu, e, fresh(i), vh);
        re_tr_j_prod<2>(q, m3_dmul(u, wd), u0, -s);
        HB_GROUP_END();
        re_tr_j_prod<0>(q, m3_mul(u0, wd), u, -s);
      }
      HB_GROUP_END();
      const M3 v = m3_muld(link_b(U, mu, e, fresh(i), vh), t);                          // u0 u6^+ u5^+
      HB_GROUP_END();
      {
        const M3 wf = wmat_b(W, k, iff, vh);
        const M3 u4 = link_b(U, nu, o, fresh(iff), vh);
        re_tr_j_prod<2>(q, m3_dmul(u4, wf), v, -s);
        HB_GROUP_END();
        re_tr_j_prod<0>(q, m3_mul(v, wf), u4, -s);
      }
    }
  }
}
#undef HB_GROUP_END

// all clover_d contributions of one link: up to two (W, parity, scale) sets
__global__ __launch_bounds__(64) void k_clover_force(double *P, const double2 *__restrict__ U, Geom g,
                                                     const double2 *W0, int weo0, double s0,
                                                     const double2 *W1, int weo1, double s1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int me = blockIdx.y, mu = me >> 1, e = me & 1;
  if (i >= g.vh) return;
  int x[4];
  site_coords(g, e, i, x[0], x[1], x[2], x[3]);
  double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // kept as loops: unrolled, the compiler merges the link loads of all twelve
  // (set, plane) bodies and keeps them live across the kernel
#pragma unroll 1
  for (int set = 0; set < 2; set++) {
    const double2 *W = set == 0 ? W0 : W1;
    const double s = set == 0 ? s0 : s1;
    const int weo = set == 0 ? weo0 : weo1;
    if (!W || s == 0.0) continue;
#pragma unroll 1
    for (int k = 0; k < 6; k++) {
      const int a = c_pairs[k][0], b = c_pairs[k][1];
      if (a != mu && b != mu) continue;
      const int nu = a == mu ? b : a;
      d_loop(q, U, g, e, mu, nu, i, x, W, k, weo, a == mu ? s : -s);
    }
  }
  mom_add(P, mu, e, i, g.vh, q);
}

// ---------------------------------------------------------------------------
// gauge force (dsg.F90:28-63) with the staple of staple.F90:28-76
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_gauge_force(double *P, const double2 *__restrict__ U, Geom g, double s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int me = blockIdx.y, mu = me >> 1, e = me & 1, o = 1 - e;
  if (i >= g.vh) return;
  const int vh = g.vh;
  int x[4];
  site_coords(g, e, i, x[0], x[1], x[2], x[3]);
  M3 st = m3_zero();
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    const int j1 = nb(g, x[0], x[1], x[2], x[3], mu, 1), j2 = nb(g, x[0], x[1], x[2], x[3], nu, 1);
    const int j3 = nb(g, x[0], x[1], x[2], x[3], nu, -1);
    const int j4 = nb2(g, x[0], x[1], x[2], x[3], nu, -1, mu, 1);
    // U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+  +  U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
    m3_acc(st, m3_muld(m3_muld(link_b(U, nu, o, j1, vh), link_b(U, mu, o, j2, vh)), link_b(U, nu, e, i, vh)));
    __builtin_amdgcn_sched_barrier(0);
    m3_acc(st, m3_dmul(m3_mul(link_b(U, mu, o, j3, vh), link_b(U, nu, e, j4, vh)), link_b(U, nu, o, j3, vh)));
    __builtin_amdgcn_sched_barrier(0);
  }
  double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  im_tr_j(q, m3_mul(link_b(U, mu, e, i, vh), st), s);
  mom_add(P, mu, e, i, vh, q);
}

// ---------------------------------------------------------------------------
// link update (hmc_u.F90:46-56): U <- exp(i step p.lambda) U, reunitarised
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_hmc_u(double2 *__restrict__ U, const double *__restrict__ P, Geom g,
                                               double step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int me = blockIdx.y, mu = me >> 1, e = me & 1;
  if (i >= g.vh) return;
  double h[8];
  const double *p = P + (size_t)(me * 8) * g.vh + i;
#pragma unroll
  for (int j = 0; j < 8; j++) h[j] = p[(size_t)j * g.vh] * step;
  M3 u = m3_mul(m3_expi(gellmann(h)), link(U, mu, e, i, g.vh));
  m3_reunitarize(u);
  store_link(U, mu, e, i, g.vh, u);
}

// 2 * sum_i log(prod_k stored B_kk) over one parity (clover_action.F90:40-56)
__global__ __launch_bounds__(kBlock) void k_clover_action(const double2 *__restrict__ Cb, int eo, int vh,
                                                          double *partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < vh) {
    double d = 1.0;
    for (int J = 0; J < 2; J++)
      for (int w = 15; w < 18; w++) {
        double2 v = Cb[(size_t)(eo * 36 + J * 18 + w) * vh + i];
        d *= v.x * v.y;
      }
    s = log(d);
  }
  double vals[1] = {s};
  block_partials<1>(vals, partials);
}

struct FinTwice {
  double *out;
  __device__ void operator()(const double (&t)[1]) const { out[0] = 2.0 * t[0]; }
};

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
double *mom_scratch(Context *c, int *rc) {
  return static_cast<double *>(scratch(c, SLOT_MOM, sizeof(double) * 64 * (size_t)c->g.vh, rc));
}

// dsf.F90:28-110 on device spinors a, b (even parity); P accumulates
int dsf_dev(Context *c, double *P, double step, double kappa, double csw_kappa, const double2 *a,
            const double2 *b) {
  const int vh = c->g.vh;
  int rc = 0;
  double2 *at = spinor_scratch(c, SLOT_FORCE0, &rc); HB_TRY(rc);
  double2 *bt = spinor_scratch(c, SLOT_FORCE1, &rc); HB_TRY(rc);
  double2 *W = static_cast<double2 *>(scratch(c, SLOT_FORCE2, sizeof(double2) * 54 * (size_t)vh, &rc)); HB_TRY(rc);
  double2 *W2 = static_cast<double2 *>(scratch(c, SLOT_FORCE3, sizeof(double2) * 54 * (size_t)vh, &rc)); HB_TRY(rc);
  double2 *tmp = spinor_scratch(c, SLOT_FORCE4, &rc); HB_TRY(rc);
  const bool clov = csw_kappa != 0.0;
  if (clov) HB_CHECK(c->have_clover[HB_CLOVER_B], HB_E_STATE, "dsf: clover B field not present");
  if (kappa != 0.0) {
    // A~ = D_oe a, B~ = D_eo^+ b (dsf.F90:56-57), then A_oo^-1 by LDL^+ (clover_mult_b)
    DslashArgs<SF64> d1 = dslash_args64(c, 1, clov ? tmp : at, a);
    HB_TRY(launch_dslash(d1, false, EPI_PLAIN, c->stream));
    if (clov) HB_TRY(clover_mult_f64(c, HB_CLOVER_B, 1, at, tmp));
    DslashArgs<SF64> d2 = dslash_args64(c, 1, clov ? tmp : bt, b);
    HB_TRY(launch_dslash(d2, true, EPI_PLAIN, c->stream));
    if (clov) HB_TRY(clover_mult_f64(c, HB_CLOVER_B, 1, bt, tmp));
  } else {
    HB_TRY(zero_f64(c, at, 12 * (size_t)vh));
    HB_TRY(zero_f64(c, bt, 12 * (size_t)vh));
  }
  const double s = -step * 2.0 * kappa * kappa;
  if (s != 0.0) {
    k_hop_force<<<dim3(nblocks(vh, 128), 8), 128, 0, c->stream>>>(P, c->gauge, c->g, a, b, at, bt, s);
    HB_HIP(hipGetLastError());
  }
  // clover part (dsf.F90:103-109) on physical links
  const double s1 = -step * 2.0 * (csw_kappa / 8.0);
  const double s2 = s1 * kappa * kappa;
  if (s1 != 0.0 || s2 != 0.0) {
    k_w_bsa<<<nblocks(vh), kBlock, 0, c->stream>>>(W, b, a, vh);
    k_w_bsa<<<nblocks(vh), kBlock, 0, c->stream>>>(W2, bt, at, vh);
    k_clover_force<<<dim3(nblocks(vh, 64), 8), 64, 0, c->stream>>>(P, c->gauge, c->g, W, 0, s1, W2, 1, s2);
    HB_HIP(hipGetLastError());
  }
  return HB_OK;
}

// dsd.F90:28-49 (clover_dsd on the odd sites)
int dsd_dev(Context *c, double *P, double step, double csw_kappa) {
  const double s = -step * 2.0 * (csw_kappa / 8.0);
  if (s == 0.0) return HB_OK;
  HB_CHECK(c->have_clover[HB_CLOVER_B], HB_E_STATE, "dsd: clover B field not present");
  const int vh = c->g.vh;
  int rc = 0;
  double2 *W = static_cast<double2 *>(scratch(c, SLOT_FORCE2, sizeof(double2) * 54 * (size_t)vh, &rc)); HB_TRY(rc);
  k_w_ts<<<nblocks(vh, 128), 128, 0, c->stream>>>(W, c->clov[HB_CLOVER_B], 1, vh);
  k_clover_force<<<dim3(nblocks(vh, 64), 8), 64, 0, c->stream>>>(P, c->gauge, c->g, W, 1, s, nullptr, 0, 0.0);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

// clover_dsf (clover/clover_d.F90:78-106) of spinors b, a of parity eo with
// scale s, on the loaded (physical) links; P accumulates
int clover_dsf_dev(Context *c, double *P, int eo, const double2 *b, const double2 *a, double s) {
  if (s == 0.0) return HB_OK;
  const int vh = c->g.vh;
  int rc = 0;
  double2 *W = static_cast<double2 *>(scratch(c, SLOT_FORCE2, sizeof(double2) * 54 * (size_t)vh, &rc)); HB_TRY(rc);
  k_w_bsa<<<nblocks(vh), kBlock, 0, c->stream>>>(W, b, a, vh);
  k_clover_force<<<dim3(nblocks(vh, 64), 8), 64, 0, c->stream>>>(P, c->gauge, c->g, W, eo, s, nullptr, 0, 0.0);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

// clover_dsd (clover/clover_d.F90:40-75) with the loaded LDL^+ field of parity eo
int clover_dsd_dev(Context *c, double *P, int eo, double s) {
  if (s == 0.0) return HB_OK;
  HB_CHECK(c->have_clover[HB_CLOVER_B], HB_E_STATE, "clover_dsd: clover B field not present");
  const int vh = c->g.vh;
  int rc = 0;
  double2 *W = static_cast<double2 *>(scratch(c, SLOT_FORCE2, sizeof(double2) * 54 * (size_t)vh, &rc)); HB_TRY(rc);
  k_w_ts<<<nblocks(vh, 128), 128, 0, c->stream>>>(W, c->clov[HB_CLOVER_B], eo, vh);
  k_clover_force<<<dim3(nblocks(vh, 64), 8), 64, 0, c->stream>>>(P, c->gauge, c->g, W, eo, s, nullptr, 0, 0.0);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

// clover_action (clover/clover_action.F90:29-62) of parity eo of the LDL^+
// field Cb (device layout [2][36][vh])
int clover_action_parity_dev(Context *c, const double2 *Cb, int eo, double *res_d) {
  const int vh = c->g.vh;
  k_clover_action<<<nblocks(vh), kBlock, 0, c->stream>>>(Cb, eo, vh, c->red);
  k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(c->red, nblocks(vh), nullptr, FinTwice{res_d});
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int dsg_dev(Context *c, double *P, double step, double beta) {
  if (beta == 0.0) return HB_OK;
  k_gauge_force<<<dim3(nblocks(c->g.vh, 128), 8), 128, 0, c->stream>>>(P, c->gauge, c->g, -step * beta / 3.0);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int build_clover_dev(Context *c, double csw_kappa);

int hmc_u_dev(Context *c, const double *P, double step, double csw_kappa) {
  k_hmc_u<<<dim3(nblocks(c->g.vh, 128), 8), 128, 0, c->stream>>>(c->gauge, P, c->g, step);
  HB_HIP(hipGetLastError());
  c->gauge_version++;
  if (csw_kappa != 0.0) HB_TRY(build_clover_dev(c, csw_kappa));
  return HB_OK;
}

int clover_action_dev(Context *c, double *res_d) {
  HB_CHECK(c->have_clover[HB_CLOVER_B], HB_E_STATE, "clover_action: clover B field not present");
  const int vh = c->g.vh;
  k_clover_action<<<nblocks(vh), kBlock, 0, c->stream>>>(c->clov[HB_CLOVER_B], 1, vh, c->red);
  k_finalize<1><<<1, kFinThreads, 0, c->stream>>>(c->red, nblocks(vh), nullptr, FinTwice{res_d});
  HB_HIP(hipGetLastError());
  return HB_OK;
}

}  // namespace hb

using namespace hb;

extern "C" {

int hb_dsf_d(hb_context *ctx, double *p_d, double step, double kappa, double csw_kappa, const double *a_d,
             const double *b_d) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p_d && a_d && b_d, HB_E_ARG, "hb_dsf_d: null pointer");
  HB_HIP(hipSetDevice(c->device));
  return dsf_dev(c, p_d, step, kappa, csw_kappa, reinterpret_cast<const double2 *>(a_d),
                 reinterpret_cast<const double2 *>(b_d));
}

int hb_dsf(hb_context *ctx, double *p, int site_stride, double step, double kappa, double csw_kappa,
           const double *a, const double *b) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p && a && b, HB_E_ARG, "hb_dsf: null pointer");
  HB_CHECK(site_stride >= c->g.vh, HB_E_ARG, "hb_dsf: site_stride < volh");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double *P = mom_scratch(c, &rc); HB_TRY(rc);
  double2 *da = spinor_scratch(c, SLOT_X, &rc); HB_TRY(rc);
  double2 *db = spinor_scratch(c, SLOT_B, &rc); HB_TRY(rc);
  HB_TRY(gen_upload(c, P, p, site_stride));
  HB_TRY(spinor_upload(c, da, a));
  HB_TRY(spinor_upload(c, db, b));
  HB_TRY(dsf_dev(c, P, step, kappa, csw_kappa, da, db));
  return gen_download(c, p, P, site_stride);
}

int hb_dsd(hb_context *ctx, double *p, int site_stride, double step, double csw_kappa) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p && site_stride >= c->g.vh, HB_E_ARG, "hb_dsd: bad arguments");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double *P = mom_scratch(c, &rc); HB_TRY(rc);
  HB_TRY(gen_upload(c, P, p, site_stride));
  HB_TRY(dsd_dev(c, P, step, csw_kappa));
  return gen_download(c, p, P, site_stride);
}

int hb_dsg(hb_context *ctx, double *p, int site_stride, double step, double beta) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p && site_stride >= c->g.vh, HB_E_ARG, "hb_dsg: bad arguments");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double *P = mom_scratch(c, &rc); HB_TRY(rc);
  HB_TRY(gen_upload(c, P, p, site_stride));
  HB_TRY(dsg_dev(c, P, step, beta));
  return gen_download(c, p, P, site_stride);
}

int hb_hmc_u(hb_context *ctx, const double *p, int site_stride, double step, double csw_kappa) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p && site_stride >= c->g.vh, HB_E_ARG, "hb_hmc_u: bad arguments");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double *P = mom_scratch(c, &rc); HB_TRY(rc);
  HB_TRY(gen_upload(c, P, p, site_stride));
  HB_TRY(hmc_u_dev(c, P, step, csw_kappa));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

int hb_hmc_u_d(hb_context *ctx, const double *p_d, double step, double csw_kappa) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p_d, HB_E_ARG, "hb_hmc_u_d: null pointer");
  HB_HIP(hipSetDevice(c->device));
  return hmc_u_dev(c, p_d, step, csw_kappa);
}

int hb_dsd_d(hb_context *ctx, double *p_d, double step, double csw_kappa) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p_d, HB_E_ARG, "hb_dsd_d: null pointer");
  HB_HIP(hipSetDevice(c->device));
  return dsd_dev(c, p_d, step, csw_kappa);
}

int hb_dsg_d(hb_context *ctx, double *p_d, double step, double beta) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(p_d, HB_E_ARG, "hb_dsg_d: null pointer");
  HB_HIP(hipSetDevice(c->device));
  return dsg_dev(c, p_d, step, beta);
}

int hb_clover_action(hb_context *ctx, double *s_det) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(s_det, HB_E_ARG, "hb_clover_action: null pointer");
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  double *r = static_cast<double *>(scratch(c, SLOT_STATE, sizeof(CgState), &rc)); HB_TRY(rc);
  HB_TRY(clover_action_dev(c, r));
  HB_HIP(hipMemcpyAsync(s_det, r, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

int hb_gauge_backup(hb_context *ctx, int slot) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(slot >= 0 && slot < kGaugeBackups, HB_E_ARG, "hb_gauge_backup: slot %d (0..%d)", slot, kGaugeBackups - 1);
  HB_HIP(hipSetDevice(c->device));
  int rc = 0;
  void *dst = scratch(c, SLOT_BACKUP + slot, sizeof(double2) * 72 * (size_t)c->g.vh, &rc); HB_TRY(rc);
  HB_HIP(hipMemcpyAsync(dst, c->gauge, sizeof(double2) * 72 * (size_t)c->g.vh, hipMemcpyDeviceToDevice, c->stream));
  return HB_OK;
}

int hb_gauge_restore(hb_context *ctx, int slot) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_TRY(check_ready(c));
  HB_CHECK(slot >= 0 && slot < kGaugeBackups, HB_E_ARG, "hb_gauge_restore: slot %d (0..%d)", slot, kGaugeBackups - 1);
  HB_CHECK((int)c->scratch.size() > SLOT_BACKUP + slot && c->scratch[SLOT_BACKUP + slot].p, HB_E_STATE,
           "hb_gauge_restore: slot %d never backed up", slot);
  HB_HIP(hipSetDevice(c->device));
  HB_HIP(hipMemcpyAsync(c->gauge, c->scratch[SLOT_BACKUP + slot].p, sizeof(double2) * 72 * (size_t)c->g.vh,
                        hipMemcpyDeviceToDevice, c->stream));
  c->gauge_version++;
  return HB_OK;
}

}  // extern "C"
