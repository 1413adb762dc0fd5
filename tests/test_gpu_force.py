"""Forces and link update on the GPU against the reference golden vectors:
dsf (hopping + clover fermion force, dsf.F90), dsd (clover determinant,
dsd.F90), dsg (gauge force, dsg.F90), clover_action and hmc_u (hmc_u.F90 with
clover_init after the update).  Bar: 1e-13 relative to the field maximum."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
TOL = 1e-13


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def hip():
    import bqcd_hip
    return bqcd_hip


@pytest.mark.parametrize("case", ["l4444", "l4846"])
def test_forces_vs_reference(hip, golden, case):
    f = golden(case)
    lat = ol.Lattice(tuple(int(x) for x in f["L"]))
    seed = int(f["seed"])
    u = lat.gauge_exp_random(seed, float(f["scale"]))
    eta_e = lat.spinor_gauss(seed + 1, 0)
    mom = lat.momenta_gauss(seed + 3)
    kappa, csw, step, beta = float(f["kappa"]), float(f["csw"]), float(f["step"]), float(f["beta"])
    g = hip.Lattice(lat.L)
    g.load_gauge(u)
    g.build_clover(csw * kappa)
    assert rel(g.dsf(eta_e, f["mtil"], step, kappa, csw * kappa), f["dsf"]) < TOL
    assert rel(g.dsd(step, csw * kappa), f["dsd"]) < TOL
    assert rel(g.dsg(step, beta), f["dsg"]) < TOL
    sa = g.clover_action()
    assert abs(sa - float(f["info_clover_action"])) < 1e-12 * abs(sa)
    if "u_new" in f:
        g.hmc_u(mom, step, csw * kappa)
        assert rel(g.save_gauge(), f["u_new"]) < TOL
        a, _, _ = g.save_clover()
        assert rel(a, f["clover_a_new"]) < TOL
    g.close()


def test_force_accumulates_and_padded_stride(hip):
    L = (4, 4, 4, 6)
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(3, 0.4)
    g = hip.Lattice(L)
    g.load_gauge(u)
    p0 = np.random.default_rng(0).standard_normal((4, 2, lat.volh + 5, 8))
    p = g.dsg(0.1, 5.29, p=p0)
    ref = lat.dsg(u, 0.1, 5.29)
    assert rel(p[:, :, :lat.volh] - p0[:, :, :lat.volh], ref) < 1e-12
    assert np.array_equal(p[:, :, lat.volh:], p0[:, :, lat.volh:])  # padding untouched
    g.close()


def test_wilson_force_no_clover(hip):
    """csw = 0: only the hopping part, checked against the oracle"""
    L = (4, 6, 4, 4)
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(8, 0.3)
    a = lat.spinor_gauss(9, 0)
    b = lat.spinor_gauss(10, 0)
    op = lat.op(lat.flip_bc(u), lat.clover(), lat.clover(), 0.12, 0.0)
    ref = lat.dsf(op, lat.flip_bc(u), 0.07, a, b, lat.clover())
    g = hip.Lattice(L)
    g.load_gauge(u)
    assert rel(g.dsf(a, b, 0.07, 0.12, 0.0), ref) < TOL
    g.close()
