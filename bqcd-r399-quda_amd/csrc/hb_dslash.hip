// hb_dslash.hip - fused even-odd clover-Wilson hopping matrix.
//
// One thread per output site.  Each thread gathers the 8 neighbour spinors
// (coalesced: neighbours of consecutive sites are consecutive in the SoA
// planes), projects them to half spinors in registers, multiplies by the 8
// links and reconstructs -- the whole of d/D21.F90:29-76 (projection, the
// four direction passes and the out read-modify-write) in a single pass that
// reads each operand once and writes the result once.  Periodic wrap and the
// antiperiodic sign are computed on the device from the site coordinates.
//
// Epilogues fuse the operator algebra of m_tilde.F90:43-114 into the write:
//   EPI_PLAIN  out = D in
//   EPI_CLINV  out = A^-1 (D in)                      (clover_mult_ao on conf%i)
//   EPI_MTIL   out = A x0 - kappa^2 (D in) + rho x0   (clover_mult_a + sc_xpby + sc_axpy)
//   EPI_MTDOT  EPI_MTIL and a grid reduction of Re<v, out> (sc_dot), single pass
#include "hb_dslash.h"
#include "hb_hop.h"

namespace hb {

template <class S, class G, bool DAG, int EPI>
__global__ __launch_bounds__(kBlock) void k_dslash(DslashArgs<S, G> a) {
  using V = typename S::V;
  using A = typename S::A;
  const Geom g = a.g;
  const int vh = g.vh;
  if (a.skip && *a.skip) return;
  const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  V acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = cmk<V>((A)0, (A)0);
  const int e = a.out_parity;
  const bool active = i < vh;
  if (active) {
    int xc[4];
    site_coords(g, e, i, xc[0], xc[1], xc[2], xc[3]);
    const SpinorView<S> in(a.in, vh);
    if constexpr (HB_WAVE_SHARE && std::is_same<S, SF16>::value) {
      uint32_t w[13];
      in.load_words(w, i);
      const WaveView16 wv{in, w, i - (int)(threadIdx.x & 63), g.L[0] / 2 <= 32};
      hop_all<S, G, DAG, 1, WaveView16>(acc, a.U, wv, g, e, i, xc);
    } else {
      hop_all<S, G, DAG>(acc, a.U, in, g, e, i, xc);
    }
  }

  if constexpr (EPI == EPI_PLAIN) {
    if (active) store_spinor<S>(a.out, i, vh, acc);
  } else if constexpr (EPI == EPI_CLINV) {
    if (active) {
      if (a.C) {
        V r[12];
        clover_apply<G, 0>(r, acc, a.C, a.Cn, e, i, vh);
        store_spinor<S>(a.out, i, vh, r);
      } else {
        store_spinor<S>(a.out, i, vh, acc);
      }
    }
  } else {
    double dot = 0.0;
    if (active) {
      V x0[12], r[12];
      load_spinor<S>(x0, a.x0, i, vh);
      // the hop sum is complete; keep it from being interleaved with the
      // clover loads
      HB_SCHED_FENCE();
      if (a.C) clover_apply<G, 0>(r, x0, a.C, a.Cn, e, i, vh);
      else {
#pragma unroll
        for (int k = 0; k < 12; k++) r[k] = x0[k];
      }
      const A mk2 = (A)a.mkappa2, rho = (A)a.rho;
#pragma unroll
      for (int k = 0; k < 12; k++) {
        r[k].x += mk2 * acc[k].x + rho * x0[k].x;
        r[k].y += mk2 * acc[k].y + rho * x0[k].y;
      }
      store_spinor<S>(a.out, i, vh, r);
      if constexpr (EPI == EPI_MTDOT) {
        V v[12];
        load_spinor<S>(v, a.dotv, i, vh);
#pragma unroll
        for (int k = 0; k < 12; k++) dot += (double)v[k].x * (double)r[k].x + (double)v[k].y * (double)r[k].y;
      }
    }
    if constexpr (EPI == EPI_MTDOT) {
      double vals[1] = {dot};
      block_partials<1>(vals, a.partials);
    }
  }
}

template <class S, class G>
int launch_dslash(const DslashArgs<S, G> &a, bool dag, int epi, hipStream_t st) {
  dim3 grid(nblocks(a.g.vh)), block(kBlock);
#define HB_DS(D, E) hipLaunchKernelGGL((k_dslash<S, G, D, E>), grid, block, 0, st, a)
  if (!dag) {
    switch (epi) {
      case EPI_PLAIN: HB_DS(false, EPI_PLAIN); break;
      case EPI_CLINV: HB_DS(false, EPI_CLINV); break;
      case EPI_MTIL: HB_DS(false, EPI_MTIL); break;
      default: HB_DS(false, EPI_MTDOT); break;
    }
  } else {
    switch (epi) {
      case EPI_PLAIN: HB_DS(true, EPI_PLAIN); break;
      case EPI_CLINV: HB_DS(true, EPI_CLINV); break;
      case EPI_MTIL: HB_DS(true, EPI_MTIL); break;
      default: HB_DS(true, EPI_MTDOT); break;
    }
  }
#undef HB_DS
  if (epi == EPI_MTDOT)
    k_finalize<1><<<1, kFinThreads, 0, st>>>(a.partials, (int)grid.x, a.skip, FinStore1{a.dot_out});
  HB_HIP(hipGetLastError());
  return HB_OK;
}

template int launch_dslash<SF64, SF64>(const DslashArgs<SF64, SF64> &, bool, int, hipStream_t);
// FP64 spinors with two-row FP64 links (the FP64 CG iteration)
template int launch_dslash<SF64, SF64R>(const DslashArgs<SF64, SF64R> &, bool, int, hipStream_t);
template int launch_dslash<SF32, SF32>(const DslashArgs<SF32, SF32> &, bool, int, hipStream_t);
template int launch_dslash<SF16, SF16>(const DslashArgs<SF16, SF16> &, bool, int, hipStream_t);
// FP32 spinors with FP16 links and clover
template int launch_dslash<SF32, SF16>(const DslashArgs<SF32, SF16> &, bool, int, hipStream_t);

}  // namespace hb
