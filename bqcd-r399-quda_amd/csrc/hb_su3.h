// hb_su3.h - 3x3 complex matrices in registers (FP64) for the clover build,
// forces, link update and gauge generation kernels.
#pragma once

#include "hb_kernels.h"

namespace hb {

struct M3 {
  double2 m[9];  // row-major: m[3*r + c]
};

__device__ __forceinline__ M3 m3_zero() {
  M3 r;
#pragma unroll
  for (int k = 0; k < 9; k++) r.m[k] = make_double2(0.0, 0.0);
  return r;
}
__device__ __forceinline__ M3 m3_one() {
  M3 r = m3_zero();
  r.m[0].x = r.m[4].x = r.m[8].x = 1.0;
  return r;
}
__device__ __forceinline__ M3 m3_dag(const M3 &a) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) r.m[3 * i + j] = cconj(a.m[3 * j + i]);
  return r;
}
// s += a' b' with a' = conj(a) if CA, b' = conj(b) if CB: four chained FMAs
// (the force, clover and link-update kernels are bound by FP64 issue; the
// sign flips are free operand modifiers)
template <bool CA, bool CB>
__device__ __forceinline__ void cfma3(double2 &s, double2 a, double2 b) {
  const double ay = CA ? -a.y : a.y, by = CB ? -b.y : b.y;
  s.x = fma(a.x, b.x, s.x);
  s.x = fma(-ay, by, s.x);
  s.y = fma(a.x, by, s.y);
  s.y = fma(ay, b.x, s.y);
}
// Re(a' b') only
template <bool CA, bool CB>
__device__ __forceinline__ double cfma3_re(double s, double2 a, double2 b) {
  const double ay = CA ? -a.y : a.y, by = CB ? -b.y : b.y;
  return fma(-ay, by, fma(a.x, b.x, s));
}

// element (i, j) of a b (MODE 0), a^+ b (MODE 1) or a b^+ (MODE 2)
template <int MODE>
__device__ __forceinline__ double2 m3_elem(const M3 &a, const M3 &b, int i, int j) {
  double2 s = make_double2(0.0, 0.0);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if constexpr (MODE == 0) cfma3<false, false>(s, a.m[3 * i + k], b.m[3 * k + j]);
    else if constexpr (MODE == 1) cfma3<true, false>(s, a.m[3 * k + i], b.m[3 * k + j]);
    else cfma3<false, true>(s, a.m[3 * i + k], b.m[3 * j + k]);
  }
  return s;
}
template <int MODE>
__device__ __forceinline__ double m3_elem_re(const M3 &a, const M3 &b, int i, int j) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if constexpr (MODE == 0) s = cfma3_re<false, false>(s, a.m[3 * i + k], b.m[3 * k + j]);
    else if constexpr (MODE == 1) s = cfma3_re<true, false>(s, a.m[3 * k + i], b.m[3 * k + j]);
    else s = cfma3_re<false, true>(s, a.m[3 * i + k], b.m[3 * j + k]);
  }
  return s;
}
template <int MODE> __device__ __forceinline__ M3 m3_prod(const M3 &a, const M3 &b) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) r.m[3 * i + j] = m3_elem<MODE>(a, b, i, j);
  return r;
}
__device__ __forceinline__ M3 m3_mul(const M3 &a, const M3 &b) { return m3_prod<0>(a, b); }
// a^+ b
__device__ __forceinline__ M3 m3_dmul(const M3 &a, const M3 &b) { return m3_prod<1>(a, b); }
// a b^+
__device__ __forceinline__ M3 m3_muld(const M3 &a, const M3 &b) { return m3_prod<2>(a, b); }
__device__ __forceinline__ void m3_acc(M3 &a, const M3 &b) {
#pragma unroll
  for (int k = 0; k < 9; k++) a.m[k] = cadd(a.m[k], b.m[k]);
}
__device__ __forceinline__ void m3_scale(M3 &a, double s) {
#pragma unroll
  for (int k = 0; k < 9; k++) a.m[k] = cscale<double2>(s, a.m[k]);
}
__device__ __forceinline__ double m3_retr(const M3 &a) { return a.m[0].x + a.m[4].x + a.m[8].x; }

// link U_mu(x) of parity eo at site i (physical links, SoA gauge)
__device__ __forceinline__ M3 link(const double2 *__restrict__ U, int mu, int eo, int i, int vh) {
  M3 r;
  const double2 *b = U + (size_t)((mu * 2 + eo) * 9) * vh + i;
#pragma unroll
  for (int k = 0; k < 9; k++) r.m[k] = b[(size_t)k * vh];
  return r;
}
// the same through a buffer resource over the (mu, eo) slab (mu, eo
// wave-uniform): one 32-bit offset register per matrix instead of nine 64-bit
// addresses; slab offsets stay below 2^32 bytes up to vh = 2^32 / 144
__device__ __forceinline__ M3 link_b(const double2 *__restrict__ U, int mu, int eo, int i, int vh) {
  const uint32_t plane = (uint32_t)vh * 16u;
  const rsrc_t r = make_rsrc(U + (size_t)((mu * 2 + eo) * 9) * vh, plane * 9u);
  const uint32_t vo = (uint32_t)i * 16u;
  M3 m;
#pragma unroll
  for (int k = 0; k < 9; k++) m.m[k] = bld_f64x2(r, vo, plane * (uint32_t)k);
  return m;
}
__device__ __forceinline__ void store_link(double2 *__restrict__ U, int mu, int eo, int i, int vh,
                                           const M3 &v) {
  double2 *b = U + (size_t)((mu * 2 + eo) * 9) * vh + i;
#pragma unroll
  for (int k = 0; k < 9; k++) b[(size_t)k * vh] = v.m[k];
}

// neighbour index (parity of the result is 1 - eo for a single step)
__device__ __forceinline__ int nb(const Geom &g, int x, int y, int z, int t, int mu, int d) {
  int c[4] = {x, y, z, t};
  c[mu] += d;
  if (c[mu] >= g.L[mu]) c[mu] -= g.L[mu];
  if (c[mu] < 0) c[mu] += g.L[mu];
  return site_index(g, c[0], c[1], c[2], c[3]);
}
__device__ __forceinline__ int nb2(const Geom &g, int x, int y, int z, int t, int mu, int dm,
                                   int nu, int dn) {
  int c[4] = {x, y, z, t};
  c[mu] += dm;
  c[nu] += dn;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (c[k] >= g.L[k]) c[k] -= g.L[k];
    if (c[k] < 0) c[k] += g.L[k];
  }
  return site_index(g, c[0], c[1], c[2], c[3]);
}

// Gram-Schmidt of the first two columns, third column = conj(cross product)
// (the reunitarisation of su3.F90:313-365)
__device__ __forceinline__ void m3_reunitarize(M3 &u) {
  double len = 0.0;
#pragma unroll
  for (int r = 0; r < 3; r++) len += u.m[3 * r].x * u.m[3 * r].x + u.m[3 * r].y * u.m[3 * r].y;
  len = sqrt(len);
#pragma unroll
  for (int r = 0; r < 3; r++) u.m[3 * r] = make_double2(u.m[3 * r].x / len, u.m[3 * r].y / len);
  double2 f = make_double2(0.0, 0.0);
#pragma unroll
  for (int r = 0; r < 3; r++) cfmac(f, u.m[3 * r], u.m[3 * r + 1]);  // sum conj(u_r0) u_r1
#pragma unroll
  for (int r = 0; r < 3; r++) u.m[3 * r + 1] = csub(u.m[3 * r + 1], cmul(f, u.m[3 * r]));
  len = 0.0;
#pragma unroll
  for (int r = 0; r < 3; r++)
    len += u.m[3 * r + 1].x * u.m[3 * r + 1].x + u.m[3 * r + 1].y * u.m[3 * r + 1].y;
  len = sqrt(len);
#pragma unroll
  for (int r = 0; r < 3; r++) u.m[3 * r + 1] = make_double2(u.m[3 * r + 1].x / len, u.m[3 * r + 1].y / len);
  u.m[2] = cconj(csub(cmul(u.m[3], u.m[7]), cmul(u.m[6], u.m[4])));
  u.m[5] = cconj(csub(cmul(u.m[6], u.m[1]), cmul(u.m[0], u.m[7])));
  u.m[8] = cconj(csub(cmul(u.m[0], u.m[4]), cmul(u.m[3], u.m[1])));
}

// H = sum_j h_j lambda_j (Gell-Mann basis of su3.F90:196-233)
__device__ __forceinline__ M3 gellmann(const double (&h)[8]) {
  const double s3 = 0.57735026918962576;  // 1/sqrt(3)
  M3 H;
  H.m[0] = make_double2(h[2] + s3 * h[7], 0.0);
  H.m[4] = make_double2(-h[2] + s3 * h[7], 0.0);
  H.m[8] = make_double2(-2.0 * s3 * h[7], 0.0);
  H.m[1] = make_double2(h[0], -h[1]);
  H.m[3] = make_double2(h[0], h[1]);
  H.m[2] = make_double2(h[3], -h[4]);
  H.m[6] = make_double2(h[3], h[4]);
  H.m[5] = make_double2(h[5], -h[6]);
  H.m[7] = make_double2(h[5], h[6]);
  return H;
}

// exp(i H) for Hermitian H: Taylor series with scaling and squaring (the
// exponential map of gen2u, su3.F90:28-193, by a different method)
__device__ __forceinline__ M3 m3_expi(const M3 &H) {
  double nrm = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) nrm += fabs(H.m[k].x) + fabs(H.m[k].y);
  int sq = 0;
  while (nrm > 0.25 && sq < 60) { nrm *= 0.5; sq++; }
  const double sc = ldexp(1.0, -sq);
  M3 A, T = m3_one(), R = m3_one();
#pragma unroll
  for (int k = 0; k < 9; k++) A.m[k] = make_double2(-sc * H.m[k].y, sc * H.m[k].x);  // i sc H
  for (int n = 1; n <= 18; n++) {
    T = m3_mul(T, A);
    m3_scale(T, 1.0 / n);
    m3_acc(R, T);
  }
  for (int s = 0; s < sq; s++) R = m3_mul(R, R);
  return R;
}

// p_j += s * Im Tr(lambda_j v)  (su3.F90:196-213)
__device__ __forceinline__ void im_tr_j(double (&p)[8], const M3 &v, double s) {
  const double is3 = 0.57735026918962576;
  p[0] += s * (v.m[1].y + v.m[3].y);
  p[1] += s * (v.m[1].x - v.m[3].x);
  p[2] += s * (v.m[0].y - v.m[4].y);
  p[3] += s * (v.m[2].y + v.m[6].y);
  p[4] += s * (v.m[2].x - v.m[6].x);
  p[5] += s * (v.m[5].y + v.m[7].y);
  p[6] += s * (v.m[5].x - v.m[7].x);
  p[7] += s * (v.m[0].y + v.m[4].y - 2.0 * v.m[8].y) * is3;
}

// p_j += s * Re Tr(lambda_j a b) (MODE 0) or Re Tr(lambda_j a b^+) (MODE 2),
// the elements of the product formed one at a time (the diagonal only in its
// real part) instead of as a matrix: the clover force evaluates 24 such terms
// per link and plane and is bound by registers and FP64 issue
template <int MODE>
__device__ __forceinline__ void re_tr_j_prod(double (&p)[8], const M3 &a, const M3 &b, double s) {
  const double is3 = 0.57735026918962576;
  {
    const double d0 = m3_elem_re<MODE>(a, b, 0, 0), d1 = m3_elem_re<MODE>(a, b, 1, 1);
    const double d2 = m3_elem_re<MODE>(a, b, 2, 2);
    p[2] += s * (d0 - d1);
    p[7] += s * (d0 + d1 - 2.0 * d2) * is3;
  }
  {
    const double2 x01 = m3_elem<MODE>(a, b, 0, 1), x10 = m3_elem<MODE>(a, b, 1, 0);
    p[0] += s * (x01.x + x10.x);
    p[1] += s * (x10.y - x01.y);
  }
  {
    const double2 x02 = m3_elem<MODE>(a, b, 0, 2), x20 = m3_elem<MODE>(a, b, 2, 0);
    p[3] += s * (x02.x + x20.x);
    p[4] += s * (x20.y - x02.y);
  }
  {
    const double2 x12 = m3_elem<MODE>(a, b, 1, 2), x21 = m3_elem<MODE>(a, b, 2, 1);
    p[5] += s * (x12.x + x21.x);
    p[6] += s * (x21.y - x12.y);
  }
}

// p_j += s * Re Tr(lambda_j v)  (su3.F90:216-233)
__device__ __forceinline__ void re_tr_j(double (&p)[8], const M3 &v, double s) {
  const double is3 = 0.57735026918962576;
  p[0] += s * (v.m[1].x + v.m[3].x);
  p[1] += s * (v.m[3].y - v.m[1].y);
  p[2] += s * (v.m[0].x - v.m[4].x);
  p[3] += s * (v.m[2].x + v.m[6].x);
  p[4] += s * (v.m[6].y - v.m[2].y);
  p[5] += s * (v.m[5].x + v.m[7].x);
  p[6] += s * (v.m[7].y - v.m[5].y);
  p[7] += s * (v.m[0].x + v.m[4].x - 2.0 * v.m[8].x) * is3;
}

// counter-based random numbers: splitmix64 of (seed, stream, index)
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t s;
  __device__ __forceinline__ Rng(uint64_t seed, uint64_t id) : s(splitmix64(seed ^ splitmix64(id))) {}
  __device__ __forceinline__ double uniform() {  // (0, 1]
    s = splitmix64(s);
    return ((double)(s >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  }
  __device__ __forceinline__ void gauss2(double &a, double &b) {
    double r = sqrt(-2.0 * log(uniform()));
    double ph = 6.2831853071795862 * uniform();
    a = r * cos(ph);
    b = r * sin(ph);
  }
};

}  // namespace hb
