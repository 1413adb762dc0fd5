// hb_blas.h - spinor vector kernels, reductions and CG device state.
#pragma once

#include "hb_kernels.h"

namespace hb {

// device-resident CG scalars (cg.F90:61-116); kernels read and update them so
// that an iteration needs no host round trip
struct CgState {
  double rtr_old, rtr, pap, beta;
  double alpha;    // step of the iteration whose x update is pending
  double rest;     // stopping bound on |r|^2 (cg_para%rest, promoted to FP64)
  int iter, maxiter;
  int done;        // 0 running, 1 converged, 2 maxiter reached
  int xpend;       // 1: x += alpha p of the last iteration is still to be done
  double res[4];   // scratch results (true residual etc.)
};

// finalize functors of the CG reductions (cg.F90:84-115)
struct FinCgStart {  // rtr_old = |r|^2
  CgState *st;
  __device__ void operator()(const double (&t)[1]) const {
    st->rtr_old = t[0];
    st->rtr = t[0];
    st->iter = 0;
    st->done = 0;
  }
};
// rtr = |r|^2; stop test (cg.F90:111); beta = rtr / rtr_old.  Launched
// without a skip flag: once the solve is done it only clears xpend, so that
// the x update of the last iteration (in k_cg_pupdate) happens exactly once.
struct FinCgUpdate {
  CgState *st;
  __device__ void operator()(const double (&t)[1]) const {
    if (st->done) {
      st->xpend = 0;
      return;
    }
    const double rtr = t[0];
    st->alpha = st->rtr_old / st->pap;
    st->xpend = 1;
    st->iter += 1;
    st->rtr = rtr;
    if (rtr <= st->rest) {
      st->done = 1;
    } else if (st->iter >= st->maxiter) {
      st->done = 2;
    } else {
      st->beta = rtr / st->rtr_old;
      st->rtr_old = rtr;
    }
  }
};
struct FinStore1 {
  double *out;
  __device__ void operator()(const double (&t)[1]) const { out[0] = t[0]; }
};
struct FinStore2 {
  double *out;
  __device__ void operator()(const double (&t)[2]) const { out[0] = t[0]; out[1] = t[1]; }
};

int reduce_grid(const Context *c, int n);
int clover_mult_f64(Context *c, int which, int eo, double2 *out, const double2 *in);
// `red`: per-block partial sums of the reductions (at least nblocks(vh)
// doubles); null selects the context's own buffer
int cg_init(Context *c, CgState *st, double rest, int maxiter);
int cg_start(Context *c, double2 *r, double2 *p, const double2 *b, CgState *st, double *red = nullptr);
int cg_update(Context *c, double2 *r, const double2 *q, CgState *st, double *red = nullptr);
int cg_pupdate(Context *c, double2 *x, double2 *p, const double2 *r, const CgState *st);
int residual(Context *c, double2 *out, const double2 *b, double *res_d, double *red = nullptr);
int dot_norm(Context *c, const double2 *x, const double2 *y, double *res_d);
int zero_f64(Context *c, double2 *p, size_t n);
int copy_f64(Context *c, double2 *d, const double2 *s, size_t n);

// host helpers from hb_fields.hip
int spinor_upload(Context *c, double2 *dst_d, const double *host);
int spinor_download(Context *c, double *host, const double2 *src_d);
int gen_upload(Context *c, double *dst_d, const double *host, int stride);
int gen_download(Context *c, double *host, const double *src_d, int stride);
int alloc_fp64_fields(Context *c);

}  // namespace hb
