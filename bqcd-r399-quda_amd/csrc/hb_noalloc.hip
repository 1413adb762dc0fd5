// hb_noalloc.hip - the FP64 CG of hb_cg_d with every work field in one
// statically allocated device array, for callers that cannot allocate device
// memory (code under stream capture, tight solver loops, harnesses that
// forbid hipMalloc).  The kernels and the iteration are those of hb_cg_d with
// all 18 reals of each link read (the two-row copy hb_cg_d may iterate with is
// allocated on demand, so this entry point does not use it).
//
// The array holds, at fixed offsets, the five work spinors of the CG (the
// operator temporaries t and y, and r, p, q), the device CG state and the
// reduction partials, sized for lattices of up to kNoAllocMaxVolh sites per
// parity.  Each GPU the library runs on holds its own copy until the process
// ends.  Every call on one GPU shares that GPU's copy: calls that overlap on
// one GPU (two streams, two threads) race and must be serialised by the
// caller.  The array keeps the last call's contents; a call writes every word
// it reads (the state is cleared by a kernel, the work fields are produced
// before they are read).
#include <mutex>

#include "hb_ops.h"

namespace {

constexpr int kNoAllocMaxVolh = HB_NOALLOC_MAX_VOLH;         // e.g. a 16^3 x 32 lattice
constexpr size_t kSpinorMax = 192ull * kNoAllocMaxVolh;      // one FP64 spinor field
constexpr size_t kStateOff = 5 * kSpinorMax;                 // CgState
constexpr size_t kRedOff = kStateOff + 256;                  // reduction partials
constexpr size_t kRedDoubles = 4096;                         // >= nblocks(kNoAllocMaxVolh)
constexpr size_t kScratchBytes = kRedOff + sizeof(double) * kRedDoubles;
static_assert(sizeof(hb::CgState) <= 256, "CG state does not fit its slot");
static_assert(kRedDoubles >= (kNoAllocMaxVolh + hb::kBlock - 1) / hb::kBlock, "partials slot too small");
static_assert(kStateOff % 256 == 0 && kRedOff % 256 == 0, "work fields must stay 256-byte aligned");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_base[kMaxDevices] = {};               // device address of `scratch` per device
hipError_t g_setup_err[kMaxDevices] = {};     // result of the address lookup

}  // namespace

// The work area.  In host code this name is only a handle for the address
// lookup below, never a device address.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace hb {

// Called by hb_init with `device` current: looks up this device's address of
// the work area once per device.  A failure is recorded and reported by the
// no-alloc entry point; it does not fail hb_init.
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_base[device]) return;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
  g_setup_err[device] = e;
  if (e == hipSuccess) g_base[device] = p;
}

}  // namespace hb

extern "C" {

int hb_cg_d_noalloc(hb_context *ctx, int op, double kappa, double rho, double *x_d, const double *b_d,
                    float rest, int maxiter, int log, int *iters, double *true_rtr) {
  hb::Context *c = reinterpret_cast<hb::Context *>(ctx);
  HB_TRY(hb::check_ready(c));
  HB_CHECK(x_d && b_d && x_d != b_d, HB_E_ARG, "hb_cg_d_noalloc: bad pointers");
  HB_CHECK(c->g.vh <= kNoAllocMaxVolh, HB_E_ARG,
           "hb_cg_d_noalloc: %d sites per parity exceed the static work area (at most %d)", c->g.vh,
           kNoAllocMaxVolh);
  HB_CHECK(c->device >= 0 && c->device < kMaxDevices, HB_E_ARG, "hb_cg_d_noalloc: device %d", c->device);
  void *base = nullptr;
  hipError_t setup_err = hipSuccess;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base = g_base[c->device];
    setup_err = g_setup_err[c->device];
  }
  HB_CHECK(base, HB_E_STATE, "hb_cg_d_noalloc: no work area on device %d (%s)", c->device,
           setup_err != hipSuccess ? hipGetErrorString(setup_err) : "hb_init has not run on it");
  HB_HIP(hipSetDevice(c->device));
  unsigned char *b = static_cast<unsigned char *>(base);
  hb::CgWork w;
  w.t = reinterpret_cast<double2 *>(b + 0 * kSpinorMax);
  w.y = reinterpret_cast<double2 *>(b + 1 * kSpinorMax);
  w.r = reinterpret_cast<double2 *>(b + 2 * kSpinorMax);
  w.p = reinterpret_cast<double2 *>(b + 3 * kSpinorMax);
  w.q = reinterpret_cast<double2 *>(b + 4 * kSpinorMax);
  w.st = reinterpret_cast<hb::CgState *>(b + kStateOff);
  w.red = reinterpret_cast<double *>(b + kRedOff);
  hb::CgState hs{};  // host copy of the state (pageable: nothing is allocated)
  w.hs = &hs;
  int nmv = 0;
  int r = hb::cg_f64_run(c, w, op, kappa, rho, reinterpret_cast<double2 *>(x_d),
                         reinterpret_cast<const double2 *>(b_d), rest, maxiter, log, iters, true_rtr, &nmv);
  if (r == HB_OK || r == HB_E_NOCONV) c->last_n_matvec = nmv;
  return r;
}

}  // extern "C"
