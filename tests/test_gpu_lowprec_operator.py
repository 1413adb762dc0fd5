"""The mixed solver's reduced-precision operator (hb_apply_lp_d: the inner
iteration's kernels, with 16-bit or FP32 storage) against the FP64 operator
(hb_apply_d) on the same input.

Storing spinors, links and clover words in 16 bits (1/32767 of each site's or
entry's scale) and computing in FP32 bounds the difference to a few 1e-5 of the
result (measured on these lattices: 3.6e-5 norm-wise, 7.7e-5 of the field's rms
at the worst site; FP32 storage 1.5e-7 / 3.1e-7; FP32 spinors with 16-bit links
1.1e-5 / 1.9e-5).  A wrong hop, projector sign, boundary sign, neighbour index
or shared-neighbour lane changes the sites it touches by O(1): the 16-bit M~
against the FP64 M~^+ differs by 0.38 norm-wise.  The lattices cover the cases
of the 16-bit kernels' neighbour sharing: y neighbours taken from other lanes
(LX/2 <= 32), not taken (LX/2 > 32), site counts that are not a multiple of the
wave or the workgroup."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

# (norm-wise relative, largest site difference / rms of the result)
TOL = {"f16": (1e-4, 3e-4), "f16g": (5e-5, 1e-4), "f32": (1e-6, 2e-6)}
OPS = [("mtil", 0.0), ("mtil_dag", 0.0), ("mtdagmt", 0.0), ("w_mult", 0.1), ("w_mult_dag", 0.1),
       ("w_dagger_w", 0.1)]


@pytest.mark.parametrize("L", [(8, 8, 8, 8), (6, 4, 6, 6), (72, 4, 4, 4), (32, 4, 4, 6)])
@pytest.mark.parametrize("prec", ["f16", "f32", "f16g"])
def test_lowprec_operator_matches_fp64(L, prec):
    import torch
    import bqcd_hip
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(901, 0.3)
    kappa, csw = 0.125, 1.0
    a, ai, b = lat.clover_init(u, csw * kappa)
    g = bqcd_hip.Lattice(L)
    try:
        g.load_gauge(u)
        g.load_clover(a, ai, b)
        dev = torch.device("cuda:0")
        gen = torch.Generator().manual_seed(903)
        x = torch.randn((12, g.volh, 2), dtype=torch.float64, generator=gen).to(dev)
        ref = torch.empty_like(x)
        out = torch.empty_like(x)
        tol_norm, tol_site = TOL[prec]
        for op, rho in OPS:
            g.apply_d(op, kappa, rho, ref, x)
            out.fill_(np.nan)
            g.apply_lp_d(op, kappa, rho, out, x, precision=prec)
            torch.cuda.synchronize()
            d = (out - ref).cpu().numpy()
            r = ref.cpu().numpy()
            assert np.all(np.isfinite(d)), (op, prec)
            rel = np.linalg.norm(d) / np.linalg.norm(r)
            # per site: the 24 reals of a site of one parity
            site = np.sqrt((d ** 2).sum(axis=(0, 2)))
            rms = np.linalg.norm(r) / np.sqrt(g.volh)
            assert rel < tol_norm, (op, prec, L, rel)
            assert site.max() / rms < tol_site, (op, prec, L, site.max() / rms, int(site.argmax()))
            # not a copy of the FP64 path: the storage rounding is visible
            if prec != "f32":
                assert rel > 1e-7, (op, prec, rel)
    finally:
        g.close()


def test_lowprec_operator_rejects_bad_arguments():
    import torch
    import bqcd_hip
    L = (4, 4, 4, 4)
    lat = ol.Lattice(L)
    g = bqcd_hip.Lattice(L)
    try:
        g.load_gauge(lat.gauge_exp_random(5, 0.3))
        x = torch.zeros((12, g.volh, 2), dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        with pytest.raises(bqcd_hip.HipError, match="precision"):
            g.apply_lp_d("mtil", 0.12, 0.0, y, x, precision="f64")
        with pytest.raises(bqcd_hip.HipError, match="bad pointers"):
            g.apply_lp_d("mtil", 0.12, 0.0, x, x, precision="f16")
    finally:
        g.close()
