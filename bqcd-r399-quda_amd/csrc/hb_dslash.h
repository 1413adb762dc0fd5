// hb_dslash.h - launch interface of the fused hopping-matrix kernels.
#pragma once

#include "hb_blas.h"

namespace hb {

enum { EPI_PLAIN = 0, EPI_CLINV = 1, EPI_MTIL = 2, EPI_MTDOT = 3 };

// S: spinor storage; G: link and clover storage (same arithmetic type as S)
template <class S, class G = S> struct DslashArgs {
  Geom g;
  int out_parity;
  SpinorPtr<S> out;
  SpinorPtr<S> in;
  const typename G::ST *U;   // links [4][2][9][vh]
  const typename G::ST *C;   // clover field for the epilogue (A^-1 for CLINV, A for MTIL), or null
  const float *Cn;           // 16-bit clover: (scale, diagonal offset) per site and block
  SpinorPtr<S> x0;           // MTIL: the vector the clover term acts on
  SpinorPtr<S> dotv;         // MTDOT: Re<dotv, out>
  double mkappa2;            // -kappa^2
  double rho;
  double *dot_out;           // MTDOT result (device)
  double *partials;          // MTDOT per-block partial sums (gridDim doubles)
  const int *skip;           // if non-null and *skip != 0 the launch does nothing
};

template <class S, class G> int launch_dslash(const DslashArgs<S, G> &a, bool dag, int epi, hipStream_t st);

}  // namespace hb
