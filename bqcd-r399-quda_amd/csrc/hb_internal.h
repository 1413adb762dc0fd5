// hb_internal.h - shared definitions of the gfx950 BQCD hot-path library.
//
// Device layouts (site-major structure of arrays; `vh` = sites per parity):
//   spinor  T2 s[12][vh]            component k = 3*spin + colour
//   gauge   T2 U[4 mu][2 eo][9][vh] element k = 3*row + col; physical links
//                                   (the fermion boundary sign is applied in
//                                   the kernels, never stored)
//   clover  T2 C[2 eo][36][vh]      per parity: block J words 18*J + w, in the
//                                   reference's packed word order
//                                   (type_clover_a / type_clover_b)
//   momenta double P[4][2][8][vh]
// Site index within a parity: i = (x + LX*(y + LY*(z + LZ*t))) / 2, parity
// (x+y+z+t) & 1, the reference's even-odd order (init_common.F90:110-205).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <type_traits>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bqcd_hip.h"

namespace hb {

// ----------------------------------------------------------------------------
// errors
// ----------------------------------------------------------------------------
void set_error(const char *fmt, ...);

#define HB_HIP(call)                                                             \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ::hb::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #call,                \
                      hipGetErrorString(e_));                                    \
      return HB_E_HIP;                                                           \
    }                                                                            \
  } while (0)

#define HB_CHECK(cond, code, ...)                                                \
  do {                                                                           \
    if (!(cond)) {                                                               \
      ::hb::set_error(__VA_ARGS__);                                              \
      return (code);                                                             \
    }                                                                            \
  } while (0)

#define HB_TRY(expr)                                                             \
  do {                                                                           \
    int r_ = (expr);                                                             \
    if (r_ != HB_OK) return r_;                                                  \
  } while (0)

// two timing events on one stream, destroyed on every return path (an error
// return between start() and the end of a solve must not leak them)
struct EventTimer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  EventTimer() = default;
  EventTimer(const EventTimer &) = delete;
  EventTimer &operator=(const EventTimer &) = delete;
  ~EventTimer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  int create() {
    HB_HIP(hipEventCreate(&e0));
    HB_HIP(hipEventCreate(&e1));
    return HB_OK;
  }
  int start(hipStream_t s) {
    HB_HIP(hipEventRecord(e0, s));
    return HB_OK;
  }
  int stop(hipStream_t s) {
    HB_HIP(hipEventRecord(e1, s));
    return HB_OK;
  }
  // waits for the stop event, then the time between the two in ms
  int elapsed(float *ms) {
    HB_HIP(hipEventSynchronize(e1));
    HB_HIP(hipEventElapsedTime(ms, e0, e1));
    return HB_OK;
  }
};

// ----------------------------------------------------------------------------
// geometry passed by value to every kernel
// ----------------------------------------------------------------------------
struct Geom {
  int L[4];     // LX LY LZ LT
  int nxh;      // LX/2
  int vh;       // sites per parity
  int bc[4];    // fermion boundary sign per direction (+1/-1)
};

// coordinates of site i of parity eo
__device__ __forceinline__ void site_coords(const Geom &g, int eo, int i, int &x, int &y, int &z,
                                            int &t) {
  int xh = i % g.nxh;
  int r = i / g.nxh;
  y = r % g.L[1];
  r /= g.L[1];
  z = r % g.L[2];
  t = r / g.L[2];
  x = 2 * xh + ((y + z + t + eo) & 1);
}

__device__ __forceinline__ int site_index(const Geom &g, int x, int y, int z, int t) {
  return (x >> 1) + g.nxh * (y + g.L[1] * (z + g.L[2] * t));
}

// ----------------------------------------------------------------------------
// context
// ----------------------------------------------------------------------------
struct Field {
  void *p = nullptr;
  size_t bytes = 0;
};

struct Scalars;  // device CG state, see hb_cg.hip

struct Context {
  Geom g;
  int device = 0;
  hipStream_t stream = nullptr;      // stream in use
  hipStream_t own_stream = nullptr;  // created by hb_init
  // FP64 fields
  double2 *gauge = nullptr;          // [4][2][9][vh]
  double2 *clov[3] = {nullptr, nullptr, nullptr};  // A/2, A^-1/2, LDL^+ : [2][36][vh]
  bool have_gauge = false, have_clover[3] = {false, false, false};
  // low-precision copies (built lazily from the FP64 fields)
  float2 *gauge32 = nullptr;
  float2 *clov32[2] = {nullptr, nullptr};          // A, A^-1
  short2 *gauge16 = nullptr;                       // 16-bit fixed point
  short2 *clov16[2] = {nullptr, nullptr};          // words and (scale, diagonal offset) per block
  // FP64 links as rows 0 and 1 for the FP64 CG iteration (built on demand)
  double2 *gauge64r = nullptr;                     // [4][2][6][vh]
  uint64_t r64_gauge_version = 0;                  // gauge_version it was built from
  bool r64_ok = false;                             // the links are SU(3) to 1e-12
  int link_recon = 12;                             // links of the FP64 CG iteration: 12 or 18 reals
  uint64_t gauge_version = 1, clover_version = 1;
  uint64_t lp_gauge_version = 0, lp_clover_version = 0;
  // scratch
  std::vector<Field> scratch;
  double *red = nullptr;             // per-block partial sums of the reductions
  int red_blocks = 0;
  void *host_stage = nullptr;        // pinned staging
  size_t host_stage_bytes = 0;
  void *dev_stage = nullptr;         // device staging (AoS copies)
  size_t dev_stage_bytes = 0;
  double last_solve_ms = 0.0;
  int last_n_matvec = 0;
  int last_block = 0;                // RHS block size of the last multi-RHS solve
  int multi_tile = 256;              // workgroup size of the block kernels
  int last_tile = 0;                 // that of the last multi-RHS solve
  int user_tile = 256;               // for solves with an explicit block size
  // chronological initial guesses (mre.F90) of the two pseudofermion solves
  // of the HMC: basis vectors live in scratch slots, newest first
  int mre_nvec = 0;
  int mre_rank[2] = {0, 0};
  int mre_slot[2][16] = {};
  // model of the trajectory whose pseudofermion fields phi are in the HMC's
  // scratch slots (0: none); a backward test trajectory integrates with them
  char hmc_phi_model = 0;
};

int ensure_dev_stage(Context *c, size_t bytes);
int ensure_host_stage(Context *c, size_t bytes);
void *scratch(Context *c, int slot, size_t bytes, int *rc);
// looks up the device address of the no-alloc CG's static work area (hb_noalloc.hip)
void noalloc_setup(int device);

constexpr int kBlock = 256;

// Kernels address every field through buffer resources: a 32-bit byte offset
// per lane plus a 32-bit plane offset, so one buffer spans less than 4 GiB.
constexpr unsigned long long kBufferLimit = 1ULL << 32;
inline bool fits_buffer(unsigned long long bytes) { return bytes < kBufferLimit; }
// bytes of one stored complex number of a multi-RHS block field
inline int block_complex_bytes(int precision) { return precision == 32 ? 8 : 4; }
// a multi-RHS block field holds 12 complex numbers per (site, RHS) in one buffer
inline bool multi_block_fits(long long vh, int nr, int precision) {
  return fits_buffer(12ULL * (unsigned long long)vh * (unsigned long long)nr *
                     (unsigned long long)block_complex_bytes(precision));
}

// The hardware hands workgroup b to XCD (b mod 8).  This bijection of
// [0, nblk) gives each of the 8 XCDs (each with its own L2) one contiguous
// range of sites instead of every 8th block, so that neighbour reuse between
// adjacent blocks stays inside one L2.
__device__ __forceinline__ int xcd_block(int b, int nblk) {
  const int q = nblk >> 3, r = nblk & 7;
  const int x = b & 7, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}
inline int nblocks(int n, int b = kBlock) { return (n + b - 1) / b; }

// ----------------------------------------------------------------------------
// complex helpers
// ----------------------------------------------------------------------------
template <typename T> struct C2;
template <> struct C2<double> { using t = double2; };
template <> struct C2<float> { using t = float2; };

template <typename V> __device__ __forceinline__ V cmk(decltype(V{}.x) a, decltype(V{}.x) b) {
  V r; r.x = a; r.y = b; return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return cmk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return cmk<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return cmk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a^* b
template <typename V> __device__ __forceinline__ V cmulc(V a, V b) {
  return cmk<V>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename V> __device__ __forceinline__ V cconj(V a) { return cmk<V>(a.x, -a.y); }
template <typename V> __device__ __forceinline__ V ci(V a) { return cmk<V>(-a.y, a.x); }       // i a
template <typename V> __device__ __forceinline__ V cmi(V a) { return cmk<V>(a.y, -a.x); }      // -i a
template <typename V, typename S> __device__ __forceinline__ V cscale(S s, V a) {
  return cmk<V>(s * a.x, s * a.y);
}
// acc += a * b.  FP32 (the reduced-precision kernels, which are bound by
// VALU issue) chains four FMAs into acc; FP64 keeps acc + (product), the
// rounding the FP64 operator has always had (it is bound by memory).
template <typename V> __device__ __forceinline__ void cfma(V &acc, V a, V b) {
  if constexpr (std::is_same<decltype(V{}.x), float>::value) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
  } else {
    acc.x += a.x * b.x - a.y * b.y;
    acc.y += a.x * b.y + a.y * b.x;
  }
}
// acc += conj(a) * b
template <typename V> __device__ __forceinline__ void cfmac(V &acc, V a, V b) {
  if constexpr (std::is_same<decltype(V{}.x), float>::value) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(-a.y, b.x, acc.y);
  } else {
    acc.x += a.x * b.x + a.y * b.y;
    acc.y += a.x * b.y - a.y * b.x;
  }
}

// ----------------------------------------------------------------------------
// wavefront / block reductions (64-wide wavefronts)
// ----------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sums v over the block; result valid in thread 0.  `lds` needs blockDim/64 doubles.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < nw; k++) s += lds[k];
  }
  return s;
}

// Grid reductions in two launches, no fences: every block of the producing
// kernel stores its NV partial sums (plain stores, visible to later launches
// on the stream), then a one-block finalize kernel sums them in a fixed order
// -- deterministic, and no agent-scope release fence (an L2 write-back on
// CDNA) per block.
template <int NV>
__device__ __forceinline__ void block_partials(const double (&v)[NV], double *partials) {
  __shared__ double lds[NV][16];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) lds[k][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0.0;
      for (int w = 0; w < nw; w++) s += lds[k][w];
      partials[(size_t)k * gridDim.x + blockIdx.x] = s;
    }
  }
}

// one block: sums nparts partials per value (fixed order) and calls f(sums)
// from thread 0; `skip` gates it like the producing kernel.  1024 threads
// with four independent loads in flight each: the partials come from L2, and
// a single workgroup is latency bound, not bandwidth bound.
constexpr int kFinThreads = 1024;

template <int NV, class F>
__global__ __launch_bounds__(kFinThreads) void k_finalize(const double *__restrict__ partials, int nparts,
                                                          const int *skip, F f) {
  if (skip && *skip) return;
  __shared__ double lds[NV][kFinThreads / 64];
  constexpr int U = 4;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double *p = partials + (size_t)k * nparts;
    double a[U] = {0.0, 0.0, 0.0, 0.0};
    int b = threadIdx.x;
    for (; b + (U - 1) * kFinThreads < nparts; b += U * kFinThreads) {
#pragma unroll
      for (int u = 0; u < U; u++) a[u] += p[b + u * kFinThreads];
    }
    for (; b < nparts; b += kFinThreads) a[0] += p[b];
    double s = wave_sum((a[0] + a[1]) + (a[2] + a[3]));
    if ((threadIdx.x & 63) == 0) lds[k][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0.0;
      for (int w = 0; w < kFinThreads / 64; w++) s += lds[k][w];
      acc[k] = s;
    }
    f(acc);
  }
}

// finalize functor: store the sums
struct FinStore {
  double *out;
  template <int NV> __device__ void operator()(const double (&s)[NV]) const {
#pragma unroll
    for (int k = 0; k < NV; k++) out[k] = s[k];
  }
  __device__ void operator()(const double *s) const { out[0] = s[0]; }
};

}  // namespace hb
