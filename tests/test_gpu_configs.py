"""The benchmark configurations at their own sizes on the GPU.

 - 8^4, kappa 0.13, csw 1.0 (the CPU-runnable configuration): the FP64 CG
   through the C ABI against the reference's own solve of the same system
   (tests/golden/l8888.npz, written by the reference cg() built from source):
   65 iterations on M~^+M~ and 53 on W^+W, the solution at 64 sampled sites
   to 1e-12 and its norm; the device clover build against the reference's
   clover_init at the same sites.  The mixed solvers on the same system stop
   on the FP64 true residual and agree with the FP64 solution to 1e-8.
 - 32^3x64, kappa 0.1355, csw 1.9, APE-smoothed Haar links (plaquette ~0.55):
   the FP64 CG and the mixed FP32 / 16-bit solvers (reliable update every 100
   iterations or on a 1e-2 drop, and the default 0.1 / 1000), each with
   |b - A x|^2 <= rest recomputed independently through hb_apply_d, and the
   mixed solutions within 1e-8 of the FP64 one.
 - 24^3x48, 12 right-hand sides over kappa 0.1355 / 0.13 / 0.12 (4 Gaussian
   sources each, csw 1.9): the autotuned block solve and fixed blocks 4 and
   12, every system against its own FP64 single-RHS solve (true residual
   <= rest recomputed with that kappa's clover term, solution within 1e-8).
 - 48^3x96, beta 5.29, csw 1.9, two Hasenbusch pseudofermions (rho 0.1),
   model B: a short Omelyan trajectory (tau 0.02 in 2 outer steps of 2 inner
   steps) from the same start and draws with the 16-bit batched solver and
   with the FP64 solver; the Hamiltonian terms agree to the solver tolerance.
   kappa is 0.13, the reference's own 48^3x96 benchmark value (work/input.mpp):
   on APE-smoothed random links kappa 0.1365 is past the critical value and a
   single solve runs for minutes, so the 0.1365 trajectories run from a
   thermalised start in bench.py --hmc instead.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def hip():
    import bqcd_hip
    return bqcd_hip


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    return torch, torch.device("cuda:0")


# ---------------------------------------------------------------------------
# 8^4 against the reference's own solve
# ---------------------------------------------------------------------------
def test_config_8888_fp64_cg_vs_reference(hip, golden):
    f = golden("l8888")
    L = tuple(int(x) for x in f["L"])
    lat = ol.Lattice(L)
    seed = int(f["seed"])
    u = lat.gauge_exp_random(seed, float(f["scale"]))
    assert np.allclose(f["u_checksum"], [np.sum(u), np.sum(u * u)], rtol=1e-14)
    eta_e = lat.spinor_gauss(seed + 1, 0)
    eta_o = lat.spinor_gauss(seed + 2, 1)
    kappa, csw, rho = float(f["kappa"]), float(f["csw"]), float(f["rho"])
    rest, maxiter = float(f["rest"]), int(f["maxiter"])
    sites = f["sites"]
    g = hip.Lattice(L)
    g.load_gauge(u)
    g.build_clover(csw * kappa)          # device clover_init
    a, ai, b = g.save_clover()
    for nm, fld in (("clover_a", a), ("clover_i", ai), ("clover_b", b)):
        assert rel(fld[:, sites], f[nm + "_sample"]) < 1e-13, nm
    print("8^4 FP64 CG", flush=True)
    x, it, tr = g.cg("mtdagmt", eta_e, kappa, rest=rest, maxiter=maxiter)
    assert it == int(f["info_cg_iters"]) == 65
    assert rel(x[sites], f["cg_x_sample"]) < 1e-12
    assert abs(np.sum(x * x) - float(f["cg_x_norm2"])) < 1e-12 * float(f["cg_x_norm2"])
    assert tr <= rest
    x, it, tr = g.cg("w_dagger_w", eta_e, kappa, rho, rest=rest, maxiter=maxiter, x0=0.01 * eta_o)
    assert it == int(f["info_cgw_iters"]) == 53
    assert rel(x[sites], f["cgw_x_sample"]) < 1e-12
    g.close()


@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_config_8888_mixed_vs_fp64(hip, golden, precision):
    """|r|/|b| <= 1e-10 (rest = 1e-20 |b|^2, the configuration's stopping rule)"""
    f = golden("l8888")
    L = tuple(int(x) for x in f["L"])
    lat = ol.Lattice(L)
    seed = int(f["seed"])
    u = lat.gauge_exp_random(seed, float(f["scale"]))
    eta_e = lat.spinor_gauss(seed + 1, 0)
    kappa, csw = float(f["kappa"]), float(f["csw"])
    rest = 1e-20 * float(np.sum(eta_e * eta_e))
    g = hip.Lattice(L)
    g.load_gauge(u)
    g.build_clover(csw * kappa)
    x64, it64, tr64 = g.cg("mtdagmt", eta_e, kappa, rest=rest, maxiter=4000)
    x, it, nrel, tr = g.cg_mixed("mtdagmt", eta_e, kappa, rest=rest, maxiter=8000, precision=precision)
    # the true residual recomputed from the returned solution
    r = eta_e - g.mtdagmt(x, kappa)
    print("8^4", precision, it64, it, nrel, tr / rest, flush=True)
    assert float(np.sum(r * r)) <= rest * (1 + 1e-6)
    assert tr <= rest
    assert np.sqrt(np.sum((x - x64) ** 2) / np.sum(x64 * x64)) < 1e-8
    g.close()


# ---------------------------------------------------------------------------
# 32^3x64: FP64 against mixed precision
# ---------------------------------------------------------------------------
def smeared(hip, L, seed, stream=None):
    g = hip.Lattice(L)
    g.gauge_random(seed, 0.0)
    g.gauge_smear(5, 0.36)
    return g


def true_rtr(g, torch, op, x, b, kappa, rho=0.0):
    ax = torch.empty_like(b)
    g.apply_d(op, kappa, rho, ax, x)
    return float(((b - ax) ** 2).sum())


def test_config_32x64_precision(hip, torch_dev):
    torch, dev = torch_dev
    L = (32, 32, 32, 64)
    kappa, csw = 0.1355, 1.9
    g = smeared(hip, L, 20261202)
    plaq = g.plaquette()
    assert 0.53 < plaq < 0.57
    g.build_clover(csw * kappa)
    gen = torch.Generator(device=dev)
    gen.manual_seed(20261203)
    b = torch.randn((12, g.volh, 2), dtype=torch.float64, device=dev, generator=gen) * 0.5 ** 0.5
    rest = 1e-20 * float((b * b).sum())
    x64 = torch.zeros_like(b)
    print("32^3x64 FP64 CG", flush=True)
    it64, tr64 = g.cg_d("mtdagmt", x64, b, kappa, rest=rest, maxiter=20000)
    print("FP64", it64, tr64 / rest, flush=True)
    assert tr64 <= rest
    assert true_rtr(g, torch, "mtdagmt", x64, b, kappa) <= rest * (1 + 1e-6)
    assert 900 < it64 < 1500
    nx64 = float(torch.linalg.vector_norm(x64))
    x = torch.zeros_like(b)
    for precision in ("f32", "f16"):
        for delta, max_inner in ((1e-2, 100), (0.1, 1000)):
            x.zero_()
            print("32^3x64", precision, delta, max_inner, flush=True)
            it, nrel, tr = g.cg_mixed_d("mtdagmt", x, b, kappa, rest=rest, maxiter=40000,
                                        precision=precision, delta=delta, max_inner=max_inner)
            err = float(torch.linalg.vector_norm(x - x64)) / nx64
            print(precision, delta, max_inner, it, nrel, tr / rest, err, flush=True)
            assert tr <= rest
            assert true_rtr(g, torch, "mtdagmt", x, b, kappa) <= rest * (1 + 1e-6)
            assert err < 1e-8
            assert it < 1.6 * it64
    g.close()


# ---------------------------------------------------------------------------
# 24^3x48: 12 right-hand sides over 3 kappa values
# ---------------------------------------------------------------------------
def test_config_24x48_twelve_rhs(hip, torch_dev):
    torch, dev = torch_dev
    L = (24, 24, 24, 48)
    csw = 1.9
    kappas = [0.1355] * 4 + [0.13] * 4 + [0.12] * 4
    g = smeared(hip, L, 20261209)
    gen = torch.Generator(device=dev)
    gen.manual_seed(20261210)
    bs = [torch.randn((12, g.volh, 2), dtype=torch.float64, device=dev, generator=gen) * 0.5 ** 0.5
          for _ in kappas]
    rests = [1e-20 * float((b * b).sum()) for b in bs]
    # every system's own FP64 single-RHS solve
    x64 = [torch.zeros_like(b) for b in bs]
    its64 = []
    for k, kap in enumerate(kappas):
        if k == 0 or kap != kappas[k - 1]:
            g.build_clover(csw * kap)
        it, tr = g.cg_d("mtdagmt", x64[k], bs[k], kap, rest=rests[k], maxiter=20000)
        assert tr <= rests[k]
        its64.append(it)
    print("FP64 iterations", its64, flush=True)
    # uneven convergence across the kappa groups
    assert min(its64[:4]) > 2 * max(its64[4:8]) and min(its64[4:8]) > 2 * max(its64[8:])
    for block in (0, 4, 12):
        xs = [torch.zeros_like(b) for b in bs]
        print("12-RHS block", block, flush=True)
        its, trs, blk = g.cg_multi_d("mtdagmt", xs, bs, kappas, csw=csw, rests=rests, maxiter=40000,
                                     precision="f16", block=block)
        print("block", block, "->", blk, its, flush=True)
        assert blk in (1, 2, 4, 8, 12) and (block == 0 or blk == block)
        for k, kap in enumerate(kappas):
            assert trs[k] <= rests[k]
            err = float(torch.linalg.vector_norm(xs[k] - x64[k]) / torch.linalg.vector_norm(x64[k]))
            assert err < 1e-8, (block, k, err)
        # the true residuals recomputed with each kappa's own clover term
        for k, kap in enumerate(kappas):
            if k == 0 or kap != kappas[k - 1]:
                g.build_clover(csw * kap)
            assert true_rtr(g, torch, "mtdagmt", xs[k], bs[k], kap) <= rests[k] * (1 + 1e-6)
    g.close()


# ---------------------------------------------------------------------------
# 48^3x96: the trajectory machinery at full size
# ---------------------------------------------------------------------------
def test_config_48x96_short_trajectory_mixed_vs_fp64(hip, torch_dev):
    torch, dev = torch_dev
    L = (48, 48, 48, 96)
    g = smeared(hip, L, 20261211)
    vh = g.volh
    p = torch.empty((4, 2, 8, vh), dtype=torch.float64, device=dev)
    eta1 = torch.empty((12, vh, 2), dtype=torch.float64, device=dev)
    eta2 = torch.empty_like(eta1)
    g.gauge_backup(0)
    res = {}
    for name, prec, batch in (("mixed16_batched", "f16", True), ("fp64", "f64", False)):
        g.gauge_restore(0)
        g.gaussian_d(p, 31, 1.0)
        g.gaussian_d(eta1, 32, 0.5)
        g.gaussian_d(eta2, 33, 0.5)
        prm = hip.hmc_params(beta=5.29, kappa=0.13, csw=1.9, rho=0.1, tau=0.02, ntau=2, m_scale=2,
                             model="B", integrator="omelyan", rest=1e-20 * 12 * vh, maxiter=50000,
                             precision=prec, mre_vectors=2, batch=batch)
        print("48^3x96 trajectory", name, flush=True)
        res[name] = g.hmc_trajectory_d(prm, p, eta1, eta2, accept_random=-1.0)
        print(name, res[name], flush=True)
    a, b = res["mixed16_batched"], res["fp64"]
    assert a["cg_ncall"] == b["cg_ncall"]
    # each action term to the solver tolerance (relative 1e-10 residuals)
    for t in ("sg_new", "sp_new", "sd_new", "sf1_new", "sf2_new"):
        assert abs(a[t] - b[t]) < 1e-8 * abs(b[t]) + 1e-6, t
    dh_a, dh_b = a["h_new"] - a["h_old"], b["h_new"] - b["h_old"]
    assert abs(dh_a - dh_b) < 1e-3
    g.close()
