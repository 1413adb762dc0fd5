// hb_context.cpp - context lifecycle, staging buffers, error reporting.
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "hb_internal.h"

namespace hb {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

int ensure_dev_stage(Context *c, size_t bytes) {
  if (c->dev_stage_bytes >= bytes) return HB_OK;
  if (c->dev_stage) HB_HIP(hipFree(c->dev_stage));
  c->dev_stage = nullptr;
  c->dev_stage_bytes = 0;
  HB_HIP(hipMalloc(&c->dev_stage, bytes));
  c->dev_stage_bytes = bytes;
  return HB_OK;
}

int ensure_host_stage(Context *c, size_t bytes) {
  if (c->host_stage_bytes >= bytes) return HB_OK;
  if (c->host_stage) HB_HIP(hipHostFree(c->host_stage));
  c->host_stage = nullptr;
  c->host_stage_bytes = 0;
  HB_HIP(hipHostMalloc(&c->host_stage, bytes, hipHostMallocDefault));
  c->host_stage_bytes = bytes;
  return HB_OK;
}

// Scratch slots grow on demand and are kept for the life of the context.
void *scratch(Context *c, int slot, size_t bytes, int *rc) {
  *rc = HB_OK;
  if ((int)c->scratch.size() <= slot) c->scratch.resize(slot + 1);
  Field &f = c->scratch[slot];
  if (f.bytes >= bytes) return f.p;
  if (f.p) {
    hipError_t e = hipFree(f.p);
    if (e != hipSuccess) {
      set_error("hipFree(scratch): %s", hipGetErrorString(e));
      *rc = HB_E_HIP;
      return nullptr;
    }
  }
  f.p = nullptr;
  f.bytes = 0;
  hipError_t e = hipMalloc(&f.p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(scratch %zu bytes): %s", bytes, hipGetErrorString(e));
    *rc = HB_E_NOMEM;
    return nullptr;
  }
  f.bytes = bytes;
  return f.p;
}

}  // namespace hb

using namespace hb;

extern "C" {

const char *hb_last_error(void) { return g_err; }

int hb_init(const int L[4], const int bc[4], int device, hb_context **out) {
  HB_CHECK(L && bc && out, HB_E_ARG, "hb_init: null argument");
  for (int m = 0; m < 4; m++) {
    HB_CHECK(L[m] >= 2 && (L[m] % 2) == 0, HB_E_ARG, "hb_init: L[%d]=%d must be even and >= 2",
             m, L[m]);
    HB_CHECK(bc[m] == 1 || bc[m] == -1, HB_E_ARG, "hb_init: bc[%d]=%d must be +1 or -1", m,
             bc[m]);
  }
  long long vol = (long long)L[0] * L[1] * L[2] * L[3];
  // every field is addressed through buffer resources with 32-bit byte
  // offsets; the largest single-buffer field is an FP64 spinor (192 B per site)
  HB_CHECK(192LL * (vol / 2) < (1LL << 32), HB_E_ARG,
           "hb_init: lattice too large (%lld sites per parity; an FP64 spinor field must stay below 4 GiB)",
           (long long)(vol / 2));
  int ndev = 0;
  HB_HIP(hipGetDeviceCount(&ndev));
  HB_CHECK(device >= 0 && device < ndev, HB_E_ARG, "hb_init: device %d of %d", device, ndev);
  HB_HIP(hipSetDevice(device));
  Context *c = new Context();
  for (int m = 0; m < 4; m++) {
    c->g.L[m] = L[m];
    c->g.bc[m] = bc[m];
  }
  c->g.nxh = L[0] / 2;
  c->g.vh = (int)(vol / 2);
  c->device = device;
  // a blocking stream: ordered with the null stream, so a caller that works on
  // the null stream (and passes no stream of its own) needs no extra waits
  hipError_t e = hipStreamCreate(&c->own_stream);
  if (e != hipSuccess) {
    delete c;
    set_error("hipStreamCreate: %s", hipGetErrorString(e));
    return HB_E_HIP;
  }
  c->stream = c->own_stream;
  noalloc_setup(device);
  // reduction partials: up to 16 values per block of a site-parallel grid
  c->red_blocks = nblocks(2 * c->g.vh) + 64;
  e = hipMalloc(&c->red, sizeof(double) * 16 * (size_t)c->red_blocks);
  if (e != hipSuccess) {
    set_error("hb_init: %s", hipGetErrorString(e));
    hb_finalize(reinterpret_cast<hb_context *>(c));
    return HB_E_HIP;
  }
  *out = reinterpret_cast<hb_context *>(c);
  return HB_OK;
}

int hb_finalize(hb_context *ctx) {
  Context *c = reinterpret_cast<Context *>(ctx);
  if (!c) return HB_OK;
  // teardown: errors are ignored, every resource is released regardless
  (void)hipSetDevice(c->device);
  if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
  (void)hipFree(c->gauge);
  for (auto *p : c->clov) (void)hipFree(p);
  (void)hipFree(c->gauge64r);
  (void)hipFree(c->gauge32);
  for (auto *p : c->clov32) (void)hipFree(p);
  (void)hipFree(c->gauge16);
  for (auto *p : c->clov16) (void)hipFree(p);
  for (auto &f : c->scratch) (void)hipFree(f.p);
  (void)hipFree(c->red);
  (void)hipFree(c->dev_stage);
  if (c->host_stage) (void)hipHostFree(c->host_stage);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return HB_OK;
}

int hb_volh(const hb_context *ctx) {
  const Context *c = reinterpret_cast<const Context *>(ctx);
  return c ? c->g.vh : 0;
}

int hb_set_stream(hb_context *ctx, void *stream) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c, HB_E_ARG, "hb_set_stream: null context");
  c->stream = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
  return HB_OK;
}

int hb_set_link_recon(hb_context *ctx, int recon) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c, HB_E_ARG, "hb_set_link_recon: null context");
  HB_CHECK(recon == 12 || recon == 18, HB_E_ARG, "hb_set_link_recon: %d (use 12 or 18)", recon);
  c->link_recon = recon;
  return HB_OK;
}

int hb_multi_block_size(hb_context *ctx) {
  Context *c = reinterpret_cast<Context *>(ctx);
  return c ? c->last_block : 0;
}

int hb_multi_block_limit(int volh, int precision) {
  if (volh < 1 || (precision != HB_PREC_F32 && precision != HB_PREC_F16)) return 0;
  static const int cand[5] = {12, 8, 4, 2, 1};
  for (int nr : cand)
    if (hb::multi_block_fits(volh, nr, precision)) return nr;
  return 0;
}

int hb_multi_tile(hb_context *ctx) {
  Context *c = reinterpret_cast<Context *>(ctx);
  return c ? c->last_tile : 0;
}

int hb_multi_set_tile(hb_context *ctx, int tile) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c, HB_E_ARG, "hb_multi_set_tile: null context");
  HB_CHECK(tile == 128 || tile == 256, HB_E_ARG, "hb_multi_set_tile: tile %d (use 128 or 256)", tile);
  c->user_tile = tile;
  return HB_OK;
}

int hb_last_timing(hb_context *ctx, double *solve_ms, int *n_matvec) {
  Context *c = reinterpret_cast<Context *>(ctx);
  HB_CHECK(c, HB_E_ARG, "hb_last_timing: null context");
  if (solve_ms) *solve_ms = c->last_solve_ms;
  if (n_matvec) *n_matvec = c->last_n_matvec;
  return HB_OK;
}

}  // extern "C"
