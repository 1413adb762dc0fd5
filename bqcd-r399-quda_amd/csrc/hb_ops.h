// hb_ops.h - operator and solver internals shared between translation units.
#pragma once

#include "hb_blas.h"
#include "hb_dslash.h"

namespace hb {

// scratch slots of the context (each grows on demand, kept for its lifetime)
enum {
  SLOT_T = 0,      // odd temporary of the operator
  SLOT_Y,          // even temporary (M~ p inside M~^+ M~)
  SLOT_R,
  SLOT_P,
  SLOT_Q,
  SLOT_X,          // host-API staging of x
  SLOT_B,          // host-API staging of b
  SLOT_IN,
  SLOT_OUT,
  SLOT_STATE,      // CgState
  SLOT_LP0,        // low-precision fields of the mixed solver
  SLOT_LP1,
  SLOT_LP2,
  SLOT_LP3,
  SLOT_LP4,
  SLOT_LPN,        // FP16 norms of SLOT_LP0..LP4
  SLOT_LPN1,
  SLOT_LPN2,
  SLOT_LPN3,
  SLOT_LPN4,
  SLOT_MIX0,       // FP64 fields of the mixed solver
  SLOT_MIX1,
  SLOT_FORCE0,     // force temporaries
  SLOT_FORCE1,
  SLOT_FORCE2,
  SLOT_FORCE3,
  SLOT_FORCE4,
  SLOT_MOM,
  SLOT_SINGLE_END
};

// limits that size the slot ranges below
constexpr int kMaxKappaGroups = 16;  // distinct kappa values of one multi-RHS solve
constexpr int kGaugeBackups = 4;     // hb_gauge_backup / hb_gauge_restore slots
constexpr int kMaxMre = 16;          // chronological guess vectors per pseudofermion

// Slot ranges after the single-RHS slots.  Every range is disjoint from the
// others (checked below), so no solver can overwrite another's fields or a
// caller's gauge backup.
enum {
  SLOT_MULTI = 64,                                  // multi-RHS block fields: +0 .. +24
  SLOT_MULTI_LAST = SLOT_MULTI + 24,
  SLOT_MULTI_CLOVER_B = SLOT_MULTI + 29,            // LDL^+ scratch of the per-kappa clover build
  SLOT_MULTI_CLOVER = SLOT_MULTI + 30,              // per kappa group gi: A/2 at +2 gi, A^-1/2 at +2 gi + 1
  SLOT_BACKUP = SLOT_MULTI_CLOVER + 2 * kMaxKappaGroups,  // caller-visible gauge backups
  SLOT_HMC = SLOT_BACKUP + kGaugeBackups,           // HMC fields (hb_hmc.hip), 24 slots
  SLOT_HMC_MV = SLOT_HMC + 24,                      // kMaxMre products M v_i
  SLOT_HMC_MRE = SLOT_HMC_MV + kMaxMre,             // 2 bases of kMaxMre vectors
  SLOT_PLUGIN = SLOT_HMC_MRE + 2 * kMaxMre,         // staging of the plugin entry points (hb_plugin.hip), 6 slots
  SLOT_END = SLOT_PLUGIN + 6
};
static_assert(SLOT_SINGLE_END <= SLOT_MULTI, "single-RHS slots overlap the multi-RHS range");
static_assert(SLOT_MULTI_LAST < SLOT_MULTI_CLOVER_B, "multi-RHS fields overlap the clover scratch");
static_assert(SLOT_MULTI_CLOVER_B < SLOT_MULTI_CLOVER, "clover scratch overlaps the per-kappa fields");
static_assert(SLOT_MULTI_CLOVER + 2 * kMaxKappaGroups <= SLOT_BACKUP, "per-kappa clover overlaps the backups");
static_assert(SLOT_BACKUP + kGaugeBackups <= SLOT_HMC, "gauge backups overlap the HMC fields");

// one-parity clover and force pieces behind the plugin entry points (hb_force.hip)
int clover_dsf_dev(Context *c, double *P, int eo, const double2 *b, const double2 *a, double s);
int clover_dsd_dev(Context *c, double *P, int eo, double s);
int clover_action_parity_dev(Context *c, const double2 *Cb, int eo, double *res_d);
int clover_mult_parity_f64(Context *c, int mode, const double2 *C, double2 *out, const double2 *in);
double *mom_scratch(Context *c, int *rc);

bool use_clover(const Context *c);
int check_ready(const Context *c);
size_t spinor_bytes(const Context *c);
double2 *spinor_scratch(Context *c, int slot, int *rc);
DslashArgs<SF64> dslash_args64(Context *c, int out_parity, double2 *out, const double2 *in);
// `red` (optional): partials buffer of the fused dot product, else the context's;
// `recon`: iterate with the two-row copy of the links (ensure_links_two_rows)
int mtil_f64(Context *c, bool dag, double kappa, double rho, double2 *out, const double2 *in,
             double2 *t, double *dot_out, const double2 *dotv, const int *skip, double *red = nullptr,
             bool recon = false);
int apply_f64(Context *c, int op, double kappa, double rho, double2 *out, const double2 *in,
              double2 *t, double2 *y, double *dot_out, const int *skip, double *red = nullptr,
              bool recon = false);
int ensure_links_two_rows(Context *c, bool *use);

// work fields of the FP64 CG: operator temporaries t (odd) and y (even), the
// CG vectors r, p, q = A p, the device state, the reduction partials (at
// least nblocks(vh) doubles) and host memory for reading the state back
struct CgWork {
  double2 *t = nullptr, *y = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
  CgState *st = nullptr;
  double *red = nullptr;
  CgState *hs = nullptr;
  bool recon = false;  // iterate with the two-row links (the true residual uses all 18 reals)
};
int cg_f64_run(Context *c, const CgWork &w, int op, double kappa, double rho, double2 *x,
               const double2 *b, float rest, int maxiter, int log, int *iters, double *true_rtr,
               int *n_matvec);
int cg_f64(Context *c, int op, double kappa, double rho, double2 *x, const double2 *b, float rest,
           int maxiter, int log, int *iters, double *true_rtr);

}  // namespace hb
