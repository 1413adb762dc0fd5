// hb_fields.hip - conversions between the reference's host layouts and the
// library's site-major SoA device layouts; field upload/download.
#include "hb_internal.h"

namespace hb {

// ---- spinor: BQCD complex(4,3,volh)  <->  T2 s[12][vh], k = 3*spin + colour
__global__ void k_spinor_aos_to_soa(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                    int vh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int s = 0; s < 4; s++) dst[(size_t)(3 * s + c) * vh + i] = src[((size_t)i * 3 + c) * 4 + s];
}

__global__ void k_spinor_soa_to_aos(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                    int vh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int s = 0; s < 4; s++) dst[((size_t)i * 3 + c) * 4 + s] = src[(size_t)(3 * s + c) * vh + i];
}

// ---- gauge: BQCD complex(3,3,stride,0:1,4) <-> double2 U[4][2][9][vh], 3*row+col.
// `unflip` removes (or `flip` applies) the fermion boundary sign that flip_bc
// folds into the links leaving the last slice of an antiperiodic direction.
__global__ void k_gauge_aos_to_soa(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                   Geom g, int stride, int unflip) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int me = blockIdx.y;  // mu*2 + eo
  if (i >= g.vh) return;
  int mu = me >> 1, eo = me & 1;
  double sgn = 1.0;
  if (unflip && g.bc[mu] < 0) {
    int x[4];
    site_coords(g, eo, i, x[0], x[1], x[2], x[3]);
    if (x[mu] == g.L[mu] - 1) sgn = -1.0;
  }
  const double2 *s = src + ((size_t)me * stride + i) * 9;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 v = s[c * 3 + r];
      dst[((size_t)me * 9 + 3 * r + c) * g.vh + i] = make_double2(sgn * v.x, sgn * v.y);
    }
}

__global__ void k_gauge_soa_to_aos(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                   Geom g, int stride, int flip) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int me = blockIdx.y;
  if (i >= g.vh) return;
  int mu = me >> 1, eo = me & 1;
  double sgn = 1.0;
  if (flip && g.bc[mu] < 0) {
    int x[4];
    site_coords(g, eo, i, x[0], x[1], x[2], x[3]);
    if (x[mu] == g.L[mu] - 1) sgn = -1.0;
  }
  double2 *d = dst + ((size_t)me * stride + i) * 9;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 v = src[((size_t)me * 9 + 3 * r + c) * g.vh + i];
      d[c * 3 + r] = make_double2(sgn * v.x, sgn * v.y);
    }
}

// ---- clover: BQCD (18 words, 2 blocks, volh, 0:1) <-> double2 C[2][36][vh]
__global__ void k_clover_aos_to_soa(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                    int vh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int eo = blockIdx.y;
  if (i >= vh) return;
  const double2 *s = src + ((size_t)eo * vh + i) * 36;
  for (int w = 0; w < 36; w++) dst[((size_t)eo * 36 + w) * vh + i] = s[w];
}

__global__ void k_clover_soa_to_aos(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                    int vh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int eo = blockIdx.y;
  if (i >= vh) return;
  double2 *d = dst + ((size_t)eo * vh + i) * 36;
  for (int w = 0; w < 36; w++) d[w] = src[((size_t)eo * 36 + w) * vh + i];
}

// ---- momenta: BQCD real(8,stride,0:1,4) <-> double P[4][2][8][vh]
__global__ void k_gen_aos_to_soa(double *__restrict__ dst, const double *__restrict__ src, int vh,
                                 int stride) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int me = blockIdx.y;
  if (i >= vh) return;
  for (int j = 0; j < 8; j++) dst[((size_t)me * 8 + j) * vh + i] = src[((size_t)me * stride + i) * 8 + j];
}

__global__ void k_gen_soa_to_aos(double *__restrict__ dst, const double *__restrict__ src, int vh,
                                 int stride) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int me = blockIdx.y;
  if (i >= vh) return;
  for (int j = 0; j < 8; j++) dst[((size_t)me * stride + i) * 8 + j] = src[((size_t)me * 8 + j) * vh + i];
}

// helpers used by the host entry points -------------------------------------
int spinor_upload(Context *c, double2 *dst_d, const double *host) {
  size_t bytes = sizeof(double2) * 12 * (size_t)c->g.vh;
  HB_TRY(ensure_dev_stage(c, bytes));
  HB_HIP(hipMemcpyAsync(c->dev_stage, host, bytes, hipMemcpyHostToDevice, c->stream));
  k_spinor_aos_to_soa<<<nblocks(c->g.vh), kBlock, 0, c->stream>>>(
      dst_d, static_cast<const double2 *>(c->dev_stage), c->g.vh);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int spinor_download(Context *c, double *host, const double2 *src_d) {
  size_t bytes = sizeof(double2) * 12 * (size_t)c->g.vh;
  HB_TRY(ensure_dev_stage(c, bytes));
  k_spinor_soa_to_aos<<<nblocks(c->g.vh), kBlock, 0, c->stream>>>(
      static_cast<double2 *>(c->dev_stage), src_d, c->g.vh);
  HB_HIP(hipGetLastError());
  HB_HIP(hipMemcpyAsync(host, c->dev_stage, bytes, hipMemcpyDeviceToHost, c->stream));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

int gen_upload(Context *c, double *dst_d, const double *host, int stride) {
  size_t bytes = sizeof(double) * 8 * 8 * (size_t)stride;
  HB_TRY(ensure_dev_stage(c, bytes));
  HB_HIP(hipMemcpyAsync(c->dev_stage, host, bytes, hipMemcpyHostToDevice, c->stream));
  k_gen_aos_to_soa<<<dim3(nblocks(c->g.vh), 8), kBlock, 0, c->stream>>>(
      dst_d, static_cast<const double *>(c->dev_stage), c->g.vh, stride);
  HB_HIP(hipGetLastError());
  return HB_OK;
}

int gen_download(Context *c, double *host, const double *src_d, int stride) {
  size_t bytes = sizeof(double) * 8 * 8 * (size_t)stride;
  HB_TRY(ensure_dev_stage(c, bytes));
  // keep the padding sites of the caller's array intact
  HB_HIP(hipMemcpyAsync(c->dev_stage, host, bytes, hipMemcpyHostToDevice, c->stream));
  k_gen_soa_to_aos<<<dim3(nblocks(c->g.vh), 8), kBlock, 0, c->stream>>>(
      static_cast<double *>(c->dev_stage), src_d, c->g.vh, stride);
  HB_HIP(hipGetLastError());
  HB_HIP(hipMemcpyAsync(host, c->dev_stage, bytes, hipMemcpyDeviceToHost, c->stream));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

int alloc_fp64_fields(Context *c) {
  size_t vh = c->g.vh;
  if (!c->gauge) HB_HIP(hipMalloc(&c->gauge, sizeof(double2) * 72 * vh));
  for (int k = 0; k < 3; k++)
    if (!c->clov[k]) HB_HIP(hipMalloc(&c->clov[k], sizeof(double2) * 72 * vh));
  return HB_OK;
}

}  // namespace hb

using namespace hb;

extern "C" {

int hb_load_gauge(hb_context *ctx, const double *u, int site_stride, int flipped) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c && u, HB_E_ARG, "hb_load_gauge: null argument");
  HB_CHECK(site_stride >= c->g.vh, HB_E_ARG, "hb_load_gauge: site_stride %d < volh %d",
           site_stride, c->g.vh);
  HB_HIP(hipSetDevice(c->device));
  HB_TRY(alloc_fp64_fields(c));
  size_t bytes = sizeof(double2) * 9 * 8 * (size_t)site_stride;
  HB_TRY(ensure_dev_stage(c, bytes));
  HB_HIP(hipMemcpyAsync(c->dev_stage, u, bytes, hipMemcpyHostToDevice, c->stream));
  k_gauge_aos_to_soa<<<dim3(nblocks(c->g.vh), 8), kBlock, 0, c->stream>>>(
      c->gauge, static_cast<const double2 *>(c->dev_stage), c->g, site_stride, flipped ? 1 : 0);
  HB_HIP(hipGetLastError());
  HB_HIP(hipStreamSynchronize(c->stream));
  c->have_gauge = true;
  c->gauge_version++;
  return HB_OK;
}

int hb_save_gauge(hb_context *ctx, double *u, int site_stride, int flipped) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c && u, HB_E_ARG, "hb_save_gauge: null argument");
  HB_CHECK(c->have_gauge, HB_E_STATE, "hb_save_gauge: no gauge field loaded");
  HB_CHECK(site_stride >= c->g.vh, HB_E_ARG, "hb_save_gauge: site_stride < volh");
  HB_HIP(hipSetDevice(c->device));
  size_t bytes = sizeof(double2) * 9 * 8 * (size_t)site_stride;
  HB_TRY(ensure_dev_stage(c, bytes));
  HB_HIP(hipMemcpyAsync(c->dev_stage, u, bytes, hipMemcpyHostToDevice, c->stream));
  k_gauge_soa_to_aos<<<dim3(nblocks(c->g.vh), 8), kBlock, 0, c->stream>>>(
      static_cast<double2 *>(c->dev_stage), c->gauge, c->g, site_stride, flipped ? 1 : 0);
  HB_HIP(hipGetLastError());
  HB_HIP(hipMemcpyAsync(u, c->dev_stage, bytes, hipMemcpyDeviceToHost, c->stream));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

int hb_load_clover(hb_context *ctx, const double *a, const double *ainv, const double *b) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c, HB_E_ARG, "hb_load_clover: null context");
  HB_HIP(hipSetDevice(c->device));
  HB_TRY(alloc_fp64_fields(c));
  const double *src[3] = {a, ainv, b};
  size_t bytes = sizeof(double2) * 72 * (size_t)c->g.vh;
  HB_TRY(ensure_dev_stage(c, bytes));
  for (int k = 0; k < 3; k++) {
    if (!src[k]) continue;
    HB_HIP(hipMemcpyAsync(c->dev_stage, src[k], bytes, hipMemcpyHostToDevice, c->stream));
    k_clover_aos_to_soa<<<dim3(nblocks(c->g.vh), 2), kBlock, 0, c->stream>>>(
        c->clov[k], static_cast<const double2 *>(c->dev_stage), c->g.vh);
    HB_HIP(hipGetLastError());
    HB_HIP(hipStreamSynchronize(c->stream));
    c->have_clover[k] = true;
  }
  c->clover_version++;
  return HB_OK;
}

int hb_save_clover(hb_context *ctx, double *a, double *ainv, double *b) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c, HB_E_ARG, "hb_save_clover: null context");
  HB_HIP(hipSetDevice(c->device));
  double *dst[3] = {a, ainv, b};
  size_t bytes = sizeof(double2) * 72 * (size_t)c->g.vh;
  HB_TRY(ensure_dev_stage(c, bytes));
  for (int k = 0; k < 3; k++) {
    if (!dst[k]) continue;
    HB_CHECK(c->have_clover[k], HB_E_STATE, "hb_save_clover: clover field %d not present", k);
    k_clover_soa_to_aos<<<dim3(nblocks(c->g.vh), 2), kBlock, 0, c->stream>>>(
        static_cast<double2 *>(c->dev_stage), c->clov[k], c->g.vh);
    HB_HIP(hipGetLastError());
    HB_HIP(hipMemcpyAsync(dst[k], c->dev_stage, bytes, hipMemcpyDeviceToHost, c->stream));
    HB_HIP(hipStreamSynchronize(c->stream));
  }
  return HB_OK;
}

int hb_spinor_to_device(hb_context *ctx, double *dst_d, const double *src_host) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c && dst_d && src_host, HB_E_ARG, "hb_spinor_to_device: null argument");
  HB_HIP(hipSetDevice(c->device));
  HB_TRY(spinor_upload(c, reinterpret_cast<double2 *>(dst_d), src_host));
  HB_HIP(hipStreamSynchronize(c->stream));
  return HB_OK;
}

int hb_spinor_to_host(hb_context *ctx, double *dst_host, const double *src_d) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c && dst_host && src_d, HB_E_ARG, "hb_spinor_to_host: null argument");
  HB_HIP(hipSetDevice(c->device));
  return spinor_download(c, dst_host, reinterpret_cast<const double2 *>(src_d));
}

}  // extern "C"
