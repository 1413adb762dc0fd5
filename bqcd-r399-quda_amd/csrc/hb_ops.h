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
  SLOT_MULTI = 64  // multi-RHS fields start here
};

bool use_clover(const Context *c);
int check_ready(const Context *c);
size_t spinor_bytes(const Context *c);
double2 *spinor_scratch(Context *c, int slot, int *rc);
DslashArgs<SF64> dslash_args64(Context *c, int out_parity, double2 *out, const double2 *in);
int mtil_f64(Context *c, bool dag, double kappa, double rho, double2 *out, const double2 *in,
             double2 *t, double *dot_out, const double2 *dotv, const int *skip);
int apply_f64(Context *c, int op, double kappa, double rho, double2 *out, const double2 *in,
              double2 *t, double2 *y, double *dot_out, const int *skip);
int cg_f64(Context *c, int op, double kappa, double rho, double2 *x, const double2 *b, float rest,
           int maxiter, int log, int *iters, double *true_rtr);

}  // namespace hb
