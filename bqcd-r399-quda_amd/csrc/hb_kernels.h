// hb_kernels.h - device building blocks shared by the hopping-matrix, clover,
// solver and force kernels: storage policies (FP64 / FP32 / FP16 with
// per-site norms), spinor and link access, spin projection, SU(3) products,
// and the packed 6x6 clover block products.
#pragma once

#include <type_traits>

#include "hb_internal.h"

namespace hb {

// ----------------------------------------------------------------------------
// raw buffer access: a buffer resource (base, byte size) in SGPRs plus a
// per-thread 32-bit byte offset in one VGPR and a wave-uniform plane offset in
// an SGPR -- all planes of a site share one offset register.
// ----------------------------------------------------------------------------
using rsrc_t = __amdgpu_buffer_rsrc_t;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// a stored 32-bit word as a short2 / float.  Taken by value: hipcc reads only
// element 0 of a vector when __builtin_bit_cast is applied to an element of
// it directly (q[k] of a u32x4), so every word goes through a scalar first.
__device__ __forceinline__ short2 as_short2(uint32_t w) { return __builtin_bit_cast(short2, w); }
__device__ __forceinline__ float as_float(uint32_t w) { return __builtin_bit_cast(float, w); }

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
// cache policy of buffer loads: 0 = default, kNT = non-temporal (streamed
// data read once per kernel, e.g. links, so that it does not evict the
// neighbour spinors that other sites reuse)
constexpr int kNT = 2;
// load policy of the 16-bit links and clover words (kNT unless tuned otherwise)
#ifndef HB_REDUCED_OPERATOR_POLICY
#define HB_REDUCED_OPERATOR_POLICY kNT
#endif
template <int POL = 0>
__device__ __forceinline__ double2 bld_f64x2(rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, POL));
}
template <int POL = 0>
__device__ __forceinline__ float2 bld_f32x2(rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, POL));
}
__device__ __forceinline__ float bld_f32(rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}

// ----------------------------------------------------------------------------
// storage policies: A = arithmetic type, V = complex of A, ST = stored complex
// ----------------------------------------------------------------------------
struct SF64 {
  using A = double;
  using V = double2;
  using ST = double2;
  static constexpr bool kNorm = false;
  static constexpr bool kRecon = false;  // links stored with all three rows
  static constexpr int kBytes = 16;
  __device__ __forceinline__ static V ld(const ST *p, size_t k) { return p[k]; }
  __device__ __forceinline__ static void st(ST *p, size_t k, V v) { p[k] = v; }
  __device__ __forceinline__ static void st_nt(ST *p, size_t k, V v) {
    __builtin_nontemporal_store(v.x, &p[k].x);
    __builtin_nontemporal_store(v.y, &p[k].y);
  }
  template <int POL = 0>
  __device__ __forceinline__ static V bld(rsrc_t r, uint32_t vo, uint32_t so) { return bld_f64x2<POL>(r, vo, so); }
};

// FP64 links stored as rows 0 and 1 ([4][2][6][vh]); row 2 is rebuilt as
// conj(row0 x row1), exact for SU(3) links (the FP64 solvers' operator)
struct SF64R : SF64 {
  static constexpr bool kRecon = true;
};

struct SF32 {
  using A = float;
  using V = float2;
  using ST = float2;
  static constexpr bool kNorm = false;
  static constexpr bool kRecon = true;   // links stored as two rows, third rebuilt
  static constexpr int kBytes = 8;
  __device__ __forceinline__ static V ld(const ST *p, size_t k) { return p[k]; }
  __device__ __forceinline__ static void st(ST *p, size_t k, V v) { p[k] = v; }
  __device__ __forceinline__ static void st_nt(ST *p, size_t k, V v) {
    __builtin_nontemporal_store(__builtin_bit_cast(uint64_t, v), reinterpret_cast<uint64_t *>(p + k));
  }
  template <int POL = 0>
  __device__ __forceinline__ static V bld(rsrc_t r, uint32_t vo, uint32_t so) { return bld_f32x2<POL>(r, vo, so); }
};

// 16-bit storage, FP32 arithmetic: signed 16-bit fixed point over [-1, 1]
// (resolution 2^-15, uniformly -- IEEE half would give only 2^-11 relative and
// ~1e-3 absolute spacing near 1, too coarse for links of the operator).
// Links are stored as is (|U_ij| <= 1); spinors and clover blocks carry a
// per-site FP32 scale (the site's largest magnitude).
struct SF16 {
  using A = float;
  using V = float2;
  using ST = short2;
  static constexpr bool kNorm = true;
  static constexpr bool kRecon = true;
  static constexpr int kBytes = 4;
  static constexpr float kScale = 32767.0f;
  static constexpr float kInv = 1.0f / 32767.0f;
  __device__ __forceinline__ static V cvt(ST s) { return make_float2(s.x * kInv, s.y * kInv); }
  // the stored integers as floats (value * 32767): kernels fold the 1/32767
  // into one per-site or per-hop scale instead of one multiply per word
  __device__ __forceinline__ static V raw(ST s) { return make_float2((float)s.x, (float)s.y); }
  __device__ __forceinline__ static short q(float v) {
    float r = rintf(v * kScale);
    r = fminf(fmaxf(r, -kScale), kScale);
    return (short)r;
  }
  // the pair (x, y) in [-1, 1] to 16 bits in one instruction (round to nearest,
  // clamped to +-32767, as q)
  __device__ __forceinline__ static ST qpack(float x, float y) {
    return __builtin_bit_cast(ST, __builtin_amdgcn_cvt_pknorm_i16(x, y));
  }
  __device__ __forceinline__ static V ld(const ST *p, size_t k) { return cvt(p[k]); }
  __device__ __forceinline__ static void st(ST *p, size_t k, V v) { p[k] = make_short2(q(v.x), q(v.y)); }
  __device__ __forceinline__ static void st_nt(ST *p, size_t k, V v) {
    __builtin_nontemporal_store(__builtin_bit_cast(uint32_t, make_short2(q(v.x), q(v.y))),
                                reinterpret_cast<uint32_t *>(p + k));
  }
  template <int POL = 0>
  __device__ __forceinline__ static V bld(rsrc_t r, uint32_t vo, uint32_t so) {
    return cvt(__builtin_bit_cast(ST, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, POL)));
  }
  template <int POL = 0>
  __device__ __forceinline__ static V bld_raw(rsrc_t r, uint32_t vo, uint32_t so) {
    return raw(__builtin_bit_cast(ST, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, POL)));
  }
};

// a spinor field of one parity: 12 planes of vh complex numbers (+ norms)
template <class S> struct SpinorPtr {
  typename S::ST *p;
  float *norm;  // FP16 only: one scale per site
};

// buffer view of a spinor field, built once per kernel (wave-uniform)
template <class S> struct SpinorView {
  rsrc_t r, rn;
  uint32_t plane;  // bytes per plane
  __device__ __forceinline__ SpinorView(const SpinorPtr<S> &f, int vh) {
    plane = (uint32_t)vh * S::kBytes;
    r = make_rsrc(f.p, plane * 12u);
    if constexpr (S::kNorm) rn = make_rsrc(f.norm, (uint32_t)vh * 4u);
  }
  // the spinor of site i as stored: for 16-bit storage the raw integers, with
  // the returned factor (site norm / 32767) still to be applied -- the hopping
  // term applies it once to the projected half spinor (12 instead of 48
  // multiplies); other storages return the values and 1
  __device__ __forceinline__ typename S::A load_raw(typename S::V (&v)[12], int i) const {
    const uint32_t vo = (uint32_t)i * S::kBytes;
    if constexpr (S::kNorm) {
#pragma unroll
      for (int k = 0; k < 12; k++) v[k] = S::bld_raw(r, vo, plane * k);
      return bld_f32(rn, (uint32_t)i * 4u, 0) * S::kInv;
    } else {
#pragma unroll
      for (int k = 0; k < 12; k++) v[k] = S::bld(r, vo, plane * k);
      return (typename S::A)1;
    }
  }
  // spin components s0, s0+1 (k = 3*s0 .. 3*s0+5) -- the t projector; same
  // convention as load_raw
  __device__ __forceinline__ typename S::A load_pair_raw(typename S::V (&v)[6], int i, int s0) const {
    const uint32_t vo = (uint32_t)i * S::kBytes;
    if constexpr (S::kNorm) {
#pragma unroll
      for (int k = 0; k < 6; k++) v[k] = S::bld_raw(r, vo, plane * (3 * s0 + k));
      return bld_f32(rn, (uint32_t)i * 4u, 0) * S::kInv;
    } else {
#pragma unroll
      for (int k = 0; k < 6; k++) v[k] = S::bld(r, vo, plane * (3 * s0 + k));
      return (typename S::A)1;
    }
  }
  __device__ __forceinline__ void load(typename S::V (&v)[12], int i) const {
    const typename S::A f = load_raw(v, i);
    if constexpr (S::kNorm) {
#pragma unroll
      for (int k = 0; k < 12; k++) { v[k].x *= f; v[k].y *= f; }
    }
  }
  // the neighbour in direction MU (same as load_raw; see WaveView)
  template <int MU>
  __device__ __forceinline__ typename S::A load_raw_dir(typename S::V (&v)[12], int j) const {
    return load_raw(v, j);
  }
};

// 16-bit fields are stored as planes of 16-byte vectors: plane q holds 16
// bytes (four short2 words) per element, so a 16-byte load of a wave reads
// 1 KiB of contiguous memory -- the cache lines that one 4-byte plane per word
// would touch, in a quarter of the load instructions.
//   spinor (n elements: sites, or (site, RHS) pairs of a block field): three
//     planes, plane q = components 4q .. 4q+3 (component k = 3*spin + colour);
//     the FP32 element scales in a separate array.  A full spinor is three
//     16-byte loads, a half spinor (two spins, the t projector) a 16- and an
//     8-byte load.
//   link slab (mu, eo): a 16-byte plane (words 0-3 of rows 0 and 1) and an
//     8-byte plane (words 4, 5)
//   clover slab (eo, J): five 16-byte planes, words 4q .. 4q+3; the last holds
//     words 16, 17 and the block's factors (m, c)
template <int POL = 0>
__device__ __forceinline__ u32x4 bld_u32x4(rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, POL));
}
template <int POL = 0>
__device__ __forceinline__ u32x2 bld_u32x2(rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, POL));
}

template <> struct SpinorView<SF16> {
  rsrc_t r, rn;
  uint32_t plane;  // bytes per 16-byte plane
  __device__ __forceinline__ SpinorView(const SpinorPtr<SF16> &f, int n) {
    plane = (uint32_t)n * 16u;
    r = make_rsrc(f.p, plane * 3u);
    rn = make_rsrc(f.norm, (uint32_t)n * 4u);
  }
  // the 12 stored words of element j and its scale word
  __device__ __forceinline__ void load_words(uint32_t (&w)[13], int j) const {
    const uint32_t vo = (uint32_t)j * 16u;
    const u32x4 a = bld_u32x4(r, vo, 0), b = bld_u32x4(r, vo, plane), c = bld_u32x4(r, vo, 2 * plane);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      w[k] = a[k];
      w[4 + k] = b[k];
      w[8 + k] = c[k];
    }
    w[12] = __builtin_amdgcn_raw_buffer_load_b32(rn, (uint32_t)j * 4u, 0, 0);
  }
  // the stored integers as floats and the factor (scale / 32767) still to be
  // applied (the hopping term applies it once to the projected half spinor)
  __device__ __forceinline__ float load_raw(float2 (&v)[12], int j) const {
    uint32_t w[13];
    load_words(w, j);
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = SF16::raw(as_short2(w[k]));
    return as_float(w[12]) * SF16::kInv;
  }
  // spin components s0, s0+1 (k = 3*s0 .. 3*s0+5): words 0-5 (plane 0 and the
  // first half of plane 1) or 6-11 (the second half of plane 1 and plane 2)
  __device__ __forceinline__ float load_pair_raw(float2 (&v)[6], int j, int s0) const {
    const uint32_t vo = (uint32_t)j * 16u;
    uint32_t w[6];
    if (s0 == 0) {
      const u32x4 a = bld_u32x4(r, vo, 0);
      const u32x2 b = bld_u32x2(r, vo, plane);
      w[0] = a[0]; w[1] = a[1]; w[2] = a[2]; w[3] = a[3]; w[4] = b[0]; w[5] = b[1];
    } else {
      const u32x2 a = bld_u32x2(r, vo + 8u, plane);
      const u32x4 b = bld_u32x4(r, vo, 2 * plane);
      w[0] = a[0]; w[1] = a[1]; w[2] = b[0]; w[3] = b[1]; w[4] = b[2]; w[5] = b[3];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = SF16::raw(as_short2(w[k]));
    return bld_f32(rn, (uint32_t)j * 4u, 0) * SF16::kInv;
  }
  __device__ __forceinline__ void load(float2 (&v)[12], int j) const {
    const float f = load_raw(v, j);
#pragma unroll
    for (int k = 0; k < 12; k++) { v[k].x *= f; v[k].y *= f; }
  }
  template <int MU>
  __device__ __forceinline__ float load_raw_dir(float2 (&v)[12], int j) const {
    return load_raw(v, j);
  }
  // the same reads without the conversion, for a hop fetched ahead of its
  // arithmetic: w[0..11] (or w[0..5] for a spin pair) and the scale in w[12]
  template <int MU>
  __device__ __forceinline__ void fetch_dir(uint32_t (&w)[13], int j) const { load_words(w, j); }
  __device__ __forceinline__ void fetch_pair(uint32_t (&w)[13], int j, int s0) const {
    const uint32_t vo = (uint32_t)j * 16u;
    if (s0 == 0) {
      const u32x4 a = bld_u32x4(r, vo, 0);
      const u32x2 b = bld_u32x2(r, vo, plane);
      w[0] = a[0]; w[1] = a[1]; w[2] = a[2]; w[3] = a[3]; w[4] = b[0]; w[5] = b[1];
    } else {
      const u32x2 a = bld_u32x2(r, vo + 8u, plane);
      const u32x4 b = bld_u32x4(r, vo, 2 * plane);
      w[0] = a[0]; w[1] = a[1]; w[2] = b[0]; w[3] = b[1]; w[4] = b[2]; w[5] = b[3];
    }
    w[12] = __builtin_amdgcn_raw_buffer_load_b32(rn, (uint32_t)j * 4u, 0, 0);
  }
};

// 16-bit neighbour spinors shared through the wavefront: every lane holds the
// stored words of the input site with its own index (13 words), and a hop in x
// or y whose neighbour index lies in the same wave takes the words from that
// lane (ds_bpermute) instead of the vector memory path; lanes whose neighbour
// is outside the wave load it (the load then runs for those lanes only).  The
// y neighbours are shared only where at least half of them are in the wave
// (share_y, uniform: the y stride in elements is at most 32).
struct WaveView16 {
  SpinorView<SF16> g;
  const uint32_t (&w)[13];
  int wave_base;
  bool share_y;
  __device__ __forceinline__ float load_raw(float2 (&v)[12], int j) const { return g.load_raw(v, j); }
  __device__ __forceinline__ float load_pair_raw(float2 (&v)[6], int j, int s0) const {
    return g.load_pair_raw(v, j, s0);
  }
  __device__ __forceinline__ void fetch_pair(uint32_t (&x)[13], int j, int s0) const { g.fetch_pair(x, j, s0); }
  // raw words of the neighbour j: from the lane that holds them, or loaded by
  // the lanes whose neighbour is outside the wave (no conversion here, so the
  // wait for those loads falls at the words' first use)
  template <int MU>
  __device__ __forceinline__ void fetch_dir(uint32_t (&x)[13], int j) const {
    if constexpr (MU >= 2) {
      g.load_words(x, j);
    } else {
      if (MU == 1 && !share_y) {
        g.load_words(x, j);
        return;
      }
      const int src = j - wave_base;
      const bool inwave = (unsigned)src < 64u;
      const int addr = (inwave ? src : 0) * 4;
#pragma unroll
      for (int k = 0; k < 13; k++) x[k] = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)w[k]);
      if (!inwave) g.load_words(x, j);
    }
  }
  template <int MU>
  __device__ __forceinline__ float load_raw_dir(float2 (&v)[12], int j) const {
    if constexpr (MU >= 2) {
      return g.load_raw(v, j);
    } else {
      if (MU == 1 && !share_y) return g.load_raw(v, j);
      const int src = j - wave_base;
      const bool inwave = (unsigned)src < 64u;
      const int addr = (inwave ? src : 0) * 4;
      uint32_t x[13];
#pragma unroll
      for (int k = 0; k < 13; k++) x[k] = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)w[k]);
      float f;
      if (inwave) {
#pragma unroll
        for (int k = 0; k < 12; k++) v[k] = SF16::raw(as_short2(x[k]));
        f = as_float(x[12]) * SF16::kInv;
      } else {
        f = g.load_raw(v, j);
      }
      return f;
    }
  }
};

template <class S>
__device__ __forceinline__ void load_spinor(typename S::V (&v)[12], const SpinorPtr<S> &f, int i,
                                            int vh) {
  if constexpr (S::kNorm) {
    // three 16-byte planes (see SpinorView<SF16>)
    const float n = f.norm[i] * S::kInv;
    const u32x4 *P = reinterpret_cast<const u32x4 *>(f.p);
    const u32x4 q[3] = {P[i], P[(size_t)vh + i], P[2 * (size_t)vh + i]};
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const uint32_t w = q[k >> 2][k & 3];
      v[k] = S::raw(as_short2(w));
      v[k].x *= n;
      v[k].y *= n;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = S::ld(f.p, (size_t)k * vh + i);
  }
}

#ifndef HB_NT_STORE_REDUCED
#define HB_NT_STORE_REDUCED 0
#endif
// FP64 results are written with non-temporal stores: an output field is read
// again only by a later kernel, long after it would have left L2, so
// streaming it past L2 keeps the L2 for the neighbour spinors the current
// kernel re-reads (the FP64 hopping matrix gains ~9% from this).  The 4- and
// 8-byte stores of the reduced precisions measured slower non-temporal.
template <class S, class VIn>
__device__ __forceinline__ void store_spinor(const SpinorPtr<S> &f, int i, int vh,
                                             const VIn (&v)[12]) {
  using V = typename S::V;
  using A = typename S::A;
  if constexpr (S::kNorm) {
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 12; k++) m = fmaxf(m, fmaxf(fabsf((float)v[k].x), fabsf((float)v[k].y)));
    float inv = m > 0.f ? 1.0f / m : 0.f;
    // three 16-byte planes (see SpinorView<SF16>): three 16-byte stores
    uint32_t w[12];
#pragma unroll
    for (int k = 0; k < 12; k++)
      w[k] = __builtin_bit_cast(uint32_t, S::qpack((float)v[k].x * inv, (float)v[k].y * inv));
    u32x4 *P = reinterpret_cast<u32x4 *>(f.p);
#if HB_NT_STORE_REDUCED
    __builtin_nontemporal_store(m, &f.norm[i]);
#pragma unroll
    for (int q = 0; q < 3; q++)
      __builtin_nontemporal_store(u32x4{w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]},
                                  &P[(size_t)q * vh + i]);
#else
    f.norm[i] = m;
#pragma unroll
    for (int q = 0; q < 3; q++) P[(size_t)q * vh + i] = u32x4{w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]};
#endif
  } else if constexpr (sizeof(typename S::ST) == 16) {
#pragma unroll
    for (int k = 0; k < 12; k++) S::st_nt(f.p, (size_t)k * vh + i, cmk<V>((A)v[k].x, (A)v[k].y));
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) S::st(f.p, (size_t)k * vh + i, cmk<V>((A)v[k].x, (A)v[k].y));
  }
}

// ----------------------------------------------------------------------------
// links: U[4][2][9][vh], element 3*row + col; one buffer per (mu, eo) slab.
// Reduced-precision copies keep rows 0 and 1 only (U[4][2][6][vh]): for
// U in SU(3) the third row is conj(row0 x row1), rebuilt after the load, so a
// link costs 6 instead of 9 complex numbers of traffic.
// ----------------------------------------------------------------------------
//
// 16-bit links are stored in a 16-byte and an 8-byte plane per (mu, eo) slab
// instead (24 bytes per site, see SpinorView<SF16>), one 16- and one 8-byte load.
template <class S> constexpr int link_planes() { return S::kRecon ? 6 : 9; }

template <class S, int POL = kNT>
__device__ __forceinline__ void load_link(typename S::V (&u)[9], const typename S::ST *U, int mu,
                                          int eo, int i, int vh) {
  static_assert(!std::is_same<S, SF16>::value, "16-bit links are read with load_link_f");
  constexpr int NP = link_planes<S>();
  const uint32_t plane = (uint32_t)vh * S::kBytes;
  rsrc_t r = make_rsrc(U + (size_t)((mu * 2 + eo) * NP) * vh, plane * (uint32_t)NP);
  const uint32_t vo = (uint32_t)i * S::kBytes;
#pragma unroll
  for (int k = 0; k < NP; k++) u[k] = S::template bld<POL>(r, vo, plane * k);
  if constexpr (S::kRecon) {
    using V = typename S::V;
    auto cross = [](V a0, V b1, V a1, V b0) {  // conj(a0 b1 - a1 b0)
      V t = cmul(a0, b1);
      V s = cmul(a1, b0);
      return cmk<V>(t.x - s.x, -(t.y - s.y));
    };
    u[6] = cross(u[1], u[5], u[2], u[4]);
    u[7] = cross(u[2], u[3], u[0], u[5]);
    u[8] = cross(u[0], u[4], u[1], u[3]);
  }
}

// 16-bit links are used as stored (the integers as floats, value * 32767):
// load_link_f returns the factor 1/32767 that the caller folds into the half
// spinor, and rebuilds row 3 = conj(row0 x row1) / 32767 in the same units.
// Other storages load the link as load_link and return 1.
template <class G> constexpr bool link_raw() { return std::is_same<G, SF16>::value; }

// the six stored words of a 16-bit link (rows 0 and 1), fetched ahead of use
template <int POL = HB_REDUCED_OPERATOR_POLICY>
__device__ __forceinline__ void fetch_link16(uint32_t (&w)[6], const SF16::ST *U, int mu, int eo, int i, int vh) {
  rsrc_t r = make_rsrc(U + (size_t)((mu * 2 + eo) * 6) * vh, (uint32_t)vh * 24u);
  const u32x4 a = bld_u32x4<POL>(r, (uint32_t)i * 16u, 0);
  const u32x2 b = bld_u32x2<POL>(r, (uint32_t)i * 8u, (uint32_t)vh * 16u);
  w[0] = a[0]; w[1] = a[1]; w[2] = a[2]; w[3] = a[3]; w[4] = b[0]; w[5] = b[1];
}

// the link of fetch_link16 as stored (integers as floats, value * 32767) with
// row 3 = conj(row0 x row1) rebuilt from them and left in the units of the
// product (value * 32767^2): its factor 1/32767 is applied once per hop by
// the caller, to the two colour-3 outputs or inputs of the multiplication
__device__ __forceinline__ void decode_link16(float2 (&u)[9], const uint32_t (&w)[6]) {
#pragma unroll
  for (int k = 0; k < 6; k++) u[k] = SF16::raw(as_short2(w[k]));
  auto cross = [](float2 a0, float2 b1, float2 a1, float2 b0) {  // conj(a0 b1 - a1 b0)
    const float re = fmaf(a0.x, b1.x, fmaf(-a0.y, b1.y, fmaf(-a1.x, b0.x, a1.y * b0.y)));
    const float mim = fmaf(a1.x, b0.y, fmaf(a1.y, b0.x, fmaf(-a0.x, b1.y, -(a0.y * b1.x))));
    return make_float2(re, mim);
  };
  u[6] = cross(u[1], u[5], u[2], u[4]);
  u[7] = cross(u[2], u[3], u[0], u[5]);
  u[8] = cross(u[0], u[4], u[1], u[3]);
}

template <class G, int POL = HB_REDUCED_OPERATOR_POLICY>
__device__ __forceinline__ typename G::A load_link_f(typename G::V (&u)[9], const typename G::ST *U,
                                                     int mu, int eo, int i, int vh) {
  if constexpr (link_raw<G>()) {
    using V = typename G::V;
    rsrc_t r = make_rsrc(U + (size_t)((mu * 2 + eo) * 6) * vh, (uint32_t)vh * 24u);
    const u32x4 a = bld_u32x4<POL>(r, (uint32_t)i * 16u, 0);
    const u32x2 b = bld_u32x2<POL>(r, (uint32_t)i * 8u, (uint32_t)vh * 16u);
    const uint32_t w[6] = {a[0], a[1], a[2], a[3], b[0], b[1]};
#pragma unroll
    for (int k = 0; k < 6; k++) u[k] = G::raw(as_short2(w[k]));
    auto cross = [](V a0, V b1, V a1, V b0) {  // conj(a0 b1 - a1 b0) / 32767
      V t = cmul(a0, b1);
      V s = cmul(a1, b0);
      return cmk<V>((t.x - s.x) * G::kInv, -(t.y - s.y) * G::kInv);
    };
    u[6] = cross(u[1], u[5], u[2], u[4]);
    u[7] = cross(u[2], u[3], u[0], u[5]);
    u[8] = cross(u[0], u[4], u[1], u[3]);
    return G::kInv;
  } else {
    load_link<G, POL>(u, U, mu, eo, i, vh);
    return (typename G::A)1;
  }
}

// w = U h  (h: 3 colours x 2 spins, index 2*c + s)
template <class V>
__device__ __forceinline__ void su3_mul_half(V (&w)[6], const V (&u)[9], const V (&h)[6]) {
#pragma unroll
  for (int r = 0; r < 3; r++) {
    V a0 = cmul(u[3 * r], h[0]);
    V a1 = cmul(u[3 * r], h[1]);
    cfma(a0, u[3 * r + 1], h[2]);
    cfma(a1, u[3 * r + 1], h[3]);
    cfma(a0, u[3 * r + 2], h[4]);
    cfma(a1, u[3 * r + 2], h[5]);
    w[2 * r] = a0;
    w[2 * r + 1] = a1;
  }
}

// w = U^+ h
template <class V>
__device__ __forceinline__ void su3_adj_mul_half(V (&w)[6], const V (&u)[9], const V (&h)[6]) {
#pragma unroll
  for (int r = 0; r < 3; r++) {
    V a0 = cmulc(u[r], h[0]);
    V a1 = cmulc(u[r], h[1]);
    cfmac(a0, u[3 + r], h[2]);
    cfmac(a1, u[3 + r], h[3]);
    cfmac(a0, u[6 + r], h[4]);
    cfmac(a1, u[6 + r], h[5]);
    w[2 * r] = a0;
    w[2 * r + 1] = a1;
  }
}

// ----------------------------------------------------------------------------
// spin projection (1 + s*gamma_mu) and reconstruction, for the gamma basis of
// the reference (d/D21.F90:79-182, d/D21xyzt.F90:27-78):
//   gamma_k = [[0, G_k], [G_k^+, 0]], G_x = [[0,i],[i,0]], G_y = [[0,1],[-1,0]],
//   G_z = [[i,0],[0,-i]];  gamma_t = diag(1,1,-1,-1).
// Half spinor h = psi_up + s*G psi_lo; lower half of the result = s*G^+ (U h).
// Spinor index k = 3*spin + colour; half spinor index 2*colour + spin.
// ----------------------------------------------------------------------------
template <int MU, int SG, class V>
__device__ __forceinline__ void project(V (&h)[6], const V (&p)[12]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    V p0 = p[c], p1 = p[3 + c], p2 = p[6 + c], p3 = p[9 + c];
    if constexpr (MU == 0) {  // h0 = p0 + s i p3, h1 = p1 + s i p2
      h[2 * c] = SG > 0 ? cadd(p0, ci(p3)) : csub(p0, ci(p3));
      h[2 * c + 1] = SG > 0 ? cadd(p1, ci(p2)) : csub(p1, ci(p2));
    } else if constexpr (MU == 1) {  // h0 = p0 + s p3, h1 = p1 - s p2
      h[2 * c] = SG > 0 ? cadd(p0, p3) : csub(p0, p3);
      h[2 * c + 1] = SG > 0 ? csub(p1, p2) : cadd(p1, p2);
    } else {  // MU == 2: h0 = p0 + s i p2, h1 = p1 - s i p3
      h[2 * c] = SG > 0 ? cadd(p0, ci(p2)) : csub(p0, ci(p2));
      h[2 * c + 1] = SG > 0 ? csub(p1, ci(p3)) : cadd(p1, ci(p3));
    }
  }
}

// acc += s w (s real)
template <class V, class A>
__device__ __forceinline__ V caxpy(V acc, A s, V w) {
  return cmk<V>(s * w.x + acc.x, s * w.y + acc.y);
}

// acc += (1 + s gamma_mu) spread of the half spinor w; with SCALED the half
// spinor is first multiplied by f (folded into the adds as multiply-adds)
// (with F3, colour 3 by f3 in place of f)
template <int MU, int SG, bool SCALED = false, bool F3 = false, class V>
__device__ __forceinline__ void reconstruct(V (&acc)[12], const V (&w)[6], decltype(V{}.x) f = 1,
                                            decltype(V{}.x) f3 = 1) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const decltype(V{}.x) fc = (F3 && c == 2) ? f3 : f;
    // a + fc v (plus) or a - fc v
    auto add = [fc](V a, V v, bool plus) {
      if constexpr (SCALED) return caxpy(a, plus ? fc : -fc, v);
      else return plus ? cadd(a, v) : csub(a, v);
    };
    V w0 = w[2 * c], w1 = w[2 * c + 1];
    if constexpr (MU == 3) {
      if constexpr (SG < 0) {  // (1 - gamma_t): spins 3,4
        acc[6 + c] = add(acc[6 + c], w0, true);
        acc[9 + c] = add(acc[9 + c], w1, true);
      } else {
        acc[c] = add(acc[c], w0, true);
        acc[3 + c] = add(acc[3 + c], w1, true);
      }
    } else {
      acc[c] = add(acc[c], w0, true);
      acc[3 + c] = add(acc[3 + c], w1, true);
      if constexpr (MU == 0) {  // lo0 = -s i w1, lo1 = -s i w0
        acc[6 + c] = add(acc[6 + c], ci(w1), SG < 0);
        acc[9 + c] = add(acc[9 + c], ci(w0), SG < 0);
      } else if constexpr (MU == 1) {  // lo0 = -s w1, lo1 = s w0
        acc[6 + c] = add(acc[6 + c], w1, SG < 0);
        acc[9 + c] = add(acc[9 + c], w0, SG > 0);
      } else {  // MU == 2: lo0 = -s i w0, lo1 = s i w1
        acc[6 + c] = add(acc[6 + c], ci(w0), SG < 0);
        acc[9 + c] = add(acc[9 + c], ci(w1), SG > 0);
      }
    }
  }
}

// ----------------------------------------------------------------------------
// packed Hermitian 6x6 blocks (type_clover_a word order, clover/clover.h):
//   w0 = (A11, A22) real pair, w1..5 = A12..A16, w6..9 = A23..A26,
//   w10 = (A33, A44), w11..13 = A34..A36, w14,15 = A45, A46, w16 = (A55, A66),
//   w17 = A56.
// LDL^+ factors (type_clover_b): w[r(r-1)/2 + c] = L_rc (r > c),
//   w15..17 = (D1, D2), (D3, D4), (D5, D6) with D_k the stored 1/2 * 1/d_k.
// ----------------------------------------------------------------------------
__host__ __device__ constexpr int a_off_word(int r, int c) {  // r < c
  return r == 0 ? c : r == 1 ? 4 + c : r == 2 ? 8 + c : r == 3 ? 10 + c : 17;
}

// y = A x for one block; w = the block's 18 words
template <class V>
__device__ __forceinline__ void herm6(V (&y)[6], const V (&w)[18], const V (&x)[6]) {
  using A = decltype(V{}.x);
  A d[6] = {w[0].x, w[0].y, w[10].x, w[10].y, w[16].x, w[16].y};
#pragma unroll
  for (int r = 0; r < 6; r++) y[r] = cscale<V>(d[r], x[r]);
#pragma unroll
  for (int r = 0; r < 6; r++)
#pragma unroll
    for (int c = r + 1; c < 6; c++) {
      V a = w[a_off_word(r, c)];
      cfma(y[r], a, x[c]);   // A_rc x_c
      cfmac(y[c], a, x[r]);  // A_cr x_r = conj(A_rc) x_r
    }
}

// x = (LDL^+)^-1 y for one block (forward then back substitution)
template <class V>
__device__ __forceinline__ void ldl6(V (&x)[6], const V (&w)[18], const V (&yin)[6]) {
  using A = decltype(V{}.x);
  V y[6];
#pragma unroll
  for (int r = 0; r < 6; r++) {
    V acc = yin[r];
#pragma unroll
    for (int c = 0; c < r; c++) {
      V l = w[r * (r - 1) / 2 + c];
      acc = csub(acc, cmul(l, y[c]));
    }
    y[r] = acc;
  }
  A d[6] = {w[15].x, w[15].y, w[16].x, w[16].y, w[17].x, w[17].y};
#pragma unroll
  for (int r = 5; r >= 0; r--) {
    V acc = cscale<V>(d[r], y[r]);
#pragma unroll
    for (int c = r + 1; c < 6; c++) {
      V l = w[c * (c - 1) / 2 + r];
      acc = csub(acc, cmulc(l, x[c]));
    }
    x[r] = acc;
  }
}

// 16-bit clover blocks: a block (the packed Hermitian 6x6 of A/2 or A^-1/2,
// real diagonal in words 0, 10 and 16) is stored as c 1 + m Q, with c the mean
// of its diagonal, m the largest magnitude of the remainder and Q in [-1, 1]
// in 16 bits; (m, c) is one float2 per block.  The blocks are close to a
// multiple of the identity, so taking c out first gives the small entries the
// whole 16-bit range (with a plain per-block scale the quantised clover term
// cost the mixed solver ~30% more iterations).
__device__ __forceinline__ bool clover_diag_word(int k) { return k == 0 || k == 10 || k == 16; }

__device__ __forceinline__ float2 clover_block_scale(const double2 *s, size_t stride) {
  const double c = (s[0].x + s[0].y + s[10 * stride].x + s[10 * stride].y + s[16 * stride].x +
                    s[16 * stride].y) / 6.0;
  double m = 0.0;
  for (int k = 0; k < 18; k++) {
    double2 v = s[k * stride];
    if (clover_diag_word(k)) {
      v.x -= c;
      v.y -= c;
    }
    m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
  }
  return make_float2((float)m, (float)c);
}

// Q word k of a block with scale mc (see clover_block_scale)
__device__ __forceinline__ float2 clover_block_word(double2 v, int k, float2 mc) {
  const float inv = mc.x > 0.f ? 1.0f / mc.x : 0.f;
  if (clover_diag_word(k)) {
    v.x -= mc.y;
    v.y -= mc.y;
  }
  return make_float2((float)v.x * inv, (float)v.y * inv);
}

// loads block J of the FP64 / FP32 clover field of parity eo at site i (16-bit
// fields are site-packed, see load_clover_block_raw)
template <class S>
__device__ __forceinline__ void load_clover_block(typename S::V (&w)[18], const typename S::ST *C,
                                                  const float *Cn, int eo, int J, int i, int vh) {
  static_assert(!S::kNorm, "16-bit clover blocks are read with load_clover_block_raw");
  const uint32_t plane = (uint32_t)vh * S::kBytes;
  rsrc_t r = make_rsrc(C + (size_t)(eo * 36 + J * 18) * vh, plane * 18u);
  const uint32_t vo = (uint32_t)i * S::kBytes;
#pragma unroll
  for (int k = 0; k < 18; k++) w[k] = S::template bld<kNT>(r, vo, plane * k);
}

// 16-bit clover blocks: per block and element five 16-byte planes (see
// SpinorView<SF16>): the 18 words of Q as short2, then (m, c) as a float2.
// Field layout [eo][J] slabs of vh elements; the multi-RHS solver's slabs hold
// its kappa groups, element gidx * vh + i.
constexpr uint32_t kClov16 = 80;  // bytes of one block of one element
constexpr int kClov16Words = 20;  // the same in short2 units

// block of element e of a slab (buffer r, plane = 16 * elements per slab
// bytes): Q as stored (the integers as floats) and the block's factors: the
// block is c 1 + (m / 32767) Q_int; returns (m / 32767, c)
__device__ __forceinline__ float2 load_clover_rec16(float2 (&w)[18], rsrc_t r, uint32_t e, uint32_t plane) {
  u32x4 q[5];
#pragma unroll
  for (int k = 0; k < 5; k++) q[k] = bld_u32x4<HB_REDUCED_OPERATOR_POLICY>(r, e * 16u, plane * k);
#pragma unroll
  for (int k = 0; k < 18; k++) {
    const uint32_t x = q[k >> 2][k & 3];
    w[k] = SF16::raw(as_short2(x));
  }
  const uint32_t m = q[4][2], c = q[4][3];
  return make_float2(as_float(m) * SF16::kInv, as_float(c));
}

// block J of the 16-bit clover field of parity eo at site i (Cn is unused: the
// factors are stored with the words)
__device__ __forceinline__ float2 load_clover_block_raw(float2 (&w)[18], const SF16::ST *C,
                                                        const float *Cn, int eo, int J, int i,
                                                        int vh) {
  rsrc_t r = make_rsrc(C + (size_t)((eo * 2 + J) * kClov16Words) * vh, (uint32_t)vh * kClov16);
  return load_clover_rec16(w, r, (uint32_t)i, (uint32_t)vh * 16u);
}

// writes the block of element e of a slab of n elements: the words Q (in
// [-1, 1]) and the factors (m, c)
__device__ __forceinline__ void store_clover_rec16(short2 *slab, size_t e, size_t n, const float2 (&q)[18],
                                                   float2 mc) {
  uint32_t w[20];
#pragma unroll
  for (int k = 0; k < 18; k++) w[k] = __builtin_bit_cast(uint32_t, make_short2(SF16::q(q[k].x), SF16::q(q[k].y)));
  w[18] = __builtin_bit_cast(uint32_t, mc.x);
  w[19] = __builtin_bit_cast(uint32_t, mc.y);
  u32x4 *P = reinterpret_cast<u32x4 *>(slab);
#pragma unroll
  for (int k = 0; k < 5; k++) P[(size_t)k * n + e] = u32x4{w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]};
}

// 16-bit form of the FP64 block s (words s[k * stride]) as element e of a slab
// of n elements, see clover_block_scale
__device__ __forceinline__ void clover_to_rec16(short2 *slab, size_t e, size_t n, const double2 *s,
                                                size_t stride) {
  const float2 mc = clover_block_scale(s, stride);
  float2 q[18];
#pragma unroll
  for (int k = 0; k < 18; k++) q[k] = clover_block_word(s[(size_t)k * stride], k, mc);
  store_clover_rec16(slab, e, n, q, mc);
}

// links in 16 bits: rows 0 and 1 of U (FP64 words s[k * stride]) as site i of
// a (mu, eo) slab of vh sites (a 16-byte and an 8-byte plane, see load_link_f)
__device__ __forceinline__ void link_to_rec16(short2 *slab, size_t i, size_t vh, const double2 *s,
                                              size_t stride) {
  uint32_t w[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const double2 v = s[(size_t)k * stride];
    w[k] = __builtin_bit_cast(uint32_t, make_short2(SF16::q((float)v.x), SF16::q((float)v.y)));
  }
  reinterpret_cast<u32x4 *>(slab)[i] = u32x4{w[0], w[1], w[2], w[3]};
  reinterpret_cast<u32x2 *>(slab + 4 * vh)[i] = u32x2{w[4], w[5]};
}

// y = (c 1 + f Q) x for one block from y = Q x (mc = (f, c))
template <class V>
__device__ __forceinline__ void clover_raw_finish(V (&y)[6], const V (&x)[6], float2 mc) {
#pragma unroll
  for (int m = 0; m < 6; m++) {
    y[m].x = y[m].x * mc.x + x[m].x * mc.y;
    y[m].y = y[m].y * mc.x + x[m].y * mc.y;
  }
}

// out = A in (Dirac basis) with A stored as two chiral blocks of A/2
// (clover_mult_a.F90:44-95); mode 0 = Hermitian product, 1 = LDL^+ solve.
// SG is the clover storage policy, V the arithmetic complex type.
template <class SG, int MODE, class V>
__device__ __forceinline__ void clover_apply(V (&out)[12], const V (&in)[12],
                                             const typename SG::ST *C, const float *Cn, int eo,
                                             int i, int vh) {
  V xp[6], y1[6], y2[6];
  constexpr bool kRaw = std::is_same<SG, SF16>::value;
  static_assert(!kRaw || MODE == 0, "16-bit clover blocks hold A/2 or A^-1/2, not LDL factors");
#pragma unroll
  for (int m = 0; m < 6; m++) xp[m] = cadd(in[m], in[m + 6]);
  {
    typename SG::V w[18];
    float2 mc;
    if constexpr (kRaw) mc = load_clover_block_raw(w, C, Cn, eo, 0, i, vh);
    else load_clover_block<SG>(w, C, Cn, eo, 0, i, vh);
    V ww[18];
#pragma unroll
    for (int k = 0; k < 18; k++) ww[k] = cmk<V>(w[k].x, w[k].y);
    if constexpr (MODE == 0) herm6(y1, ww, xp); else ldl6(y1, ww, xp);
    if constexpr (kRaw) clover_raw_finish(y1, xp, mc);
  }
  // finish block 0 before loading block 1 (bounds live registers)
#pragma unroll
  for (int m = 0; m < 6; m++) asm volatile("" : "+v"(y1[m].x), "+v"(y1[m].y));
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < 6; m++) xp[m] = csub(in[m], in[m + 6]);
  {
    typename SG::V w[18];
    float2 mc;
    if constexpr (kRaw) mc = load_clover_block_raw(w, C, Cn, eo, 1, i, vh);
    else load_clover_block<SG>(w, C, Cn, eo, 1, i, vh);
    V ww[18];
#pragma unroll
    for (int k = 0; k < 18; k++) ww[k] = cmk<V>(w[k].x, w[k].y);
    if constexpr (MODE == 0) herm6(y2, ww, xp); else ldl6(y2, ww, xp);
    if constexpr (kRaw) clover_raw_finish(y2, xp, mc);
  }
#pragma unroll
  for (int m = 0; m < 6; m++) {
    out[m] = cadd(y1[m], y2[m]);
    out[m + 6] = csub(y1[m], y2[m]);
  }
}

}  // namespace hb
