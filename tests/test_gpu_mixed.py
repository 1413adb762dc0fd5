"""Mixed-precision CG (FP32 / FP16-storage inner solver, FP64 reliable updates):
the stopping rule is the FP64 true residual, so |b - A x|^2 <= rest must hold
exactly as for the FP64 solver; the solution must agree with the FP64 solve to
the accuracy the residual bound implies (condition number x 1e-10)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import bqcd_hip
    L = (8, 8, 8, 8)
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(321, 0.3)
    kappa, csw = 0.13, 1.0
    a, ai, b = lat.clover_init(u, csw * kappa)
    phi = lat.spinor_gauss(322, 0)
    g = bqcd_hip.Lattice(L)
    g.load_gauge(u)
    g.load_clover(a, ai, b)
    yield g, lat, phi, kappa
    g.close()


@pytest.mark.parametrize("prec", ["f32", "f16"])
@pytest.mark.parametrize("op,rho", [("mtdagmt", 0.0), ("w_dagger_w", 0.1)])
def test_mixed_true_residual(setup, prec, op, rho):
    g, lat, phi, kappa = setup
    rest = 1e-20 * float(np.sum(phi * phi))
    x64, it64, tr64 = g.cg(op, phi, kappa, rho, rest=rest, maxiter=5000)
    x, it, nrel, tr = g.cg_mixed(op, phi, kappa, rho, rest=rest, maxiter=5000, precision=prec,
                                 delta=0.01, max_inner=100)
    assert tr <= rest            # FP64 true residual, the stopping rule itself
    assert nrel >= 1
    # recompute the true residual independently with the FP64 operator
    r = phi - g.apply(op, x, kappa, rho)
    assert abs(float(np.sum(r * r)) - tr) <= 1e-3 * tr + 1e-30
    err = np.linalg.norm(x - x64) / np.linalg.norm(x64)
    assert err < 1e-7, err
    assert it <= 3 * it64


def test_mixed_from_guess_and_maxiter(setup):
    g, lat, phi, kappa = setup
    rest = 1e-20 * float(np.sum(phi * phi))
    x64, _, _ = g.cg("mtdagmt", phi, kappa, rest=rest, maxiter=5000)
    x, it, nrel, tr = g.cg_mixed("mtdagmt", phi, kappa, rest=rest, x0=x64, precision="f16")
    assert tr <= rest and it == 0 and nrel == 0
    import bqcd_hip
    with pytest.raises(bqcd_hip.NoConvergence):
        g.cg_mixed("mtdagmt", phi, kappa, rest=1e-40, maxiter=20, precision="f16")


def test_time_dslash_lp(setup):
    """the benchmarking entry point times the reduced-precision hopping kernel"""
    import bqcd_hip
    import torch
    g = setup[0] if isinstance(setup, tuple) else setup
    dev = torch.device("cuda:0")
    b = torch.randn((12, g.volh, 2), dtype=torch.float64, device=dev)
    for prec in ("f16", "f32"):
        ms = g.time_dslash_lp(b, prec, reps=3)
        assert 0.0 < ms < 1000.0
    with pytest.raises(bqcd_hip.HipError, match="precision"):
        g.time_dslash_lp(b, "f64", reps=1)


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_mixed_and_multi_without_clover(prec):
    """Wilson fermions (csw = 0: no clover term) through the reduced-precision
    single- and multi-RHS solvers, against the oracle's FP64 solve"""
    import bqcd_hip
    L = (6, 4, 4, 8)
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(91, 0.3)
    kappa = 0.12
    op = lat.op(lat.flip_bc(u), lat.clover(), lat.clover(), kappa, 0.0, 0.0)
    phis = [lat.spinor_gauss(400 + j, 0) for j in range(3)]
    rests = [1e-20 * float(np.sum(p * p)) for p in phis]
    g = bqcd_hip.Lattice(L)
    try:
        g.load_gauge(u)
        x, it, nrel, tr = g.cg_mixed("mtdagmt", phis[0], kappa, rest=rests[0], maxiter=5000,
                                     precision=prec)
        xr, _, _ = lat.cg(op, 2, phis[0], rests[0] * 1e-4, 5000)
        assert tr <= rests[0]
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-7
        xs, its, trs, used = g.cg_multi("mtdagmt", phis, [kappa] * 3, csw=0.0, rests=rests, block=4,
                                        precision=prec)
        for k in range(3):
            xr, _, _ = lat.cg(op, 2, phis[k], rests[k] * 1e-4, 5000)
            assert trs[k] <= rests[k]
            assert np.linalg.norm(xs[k] - xr) / np.linalg.norm(xr) < 1e-7
    finally:
        g.close()
