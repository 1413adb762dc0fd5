"""Pins the oracle's forces by what the mathematics fixes: each force routine
at step = 1 is minus the derivative of its action along a link perturbation
U -> exp(i eps v.lambda) U (the map hmc_u applies, hmc_u.F90:45-57), checked
by central finite differences along random su(3) directions v:

  dsf  (dsf.F90:28-110):  S_f   = Re phi^+ (M~^+ M~)^-1 phi
  dsd  (dsd.F90:28-49):   S_det = clover_action(B_oo) (clover_action.F90)
  dsg  (dsg.F90:28-63):   S_g   = -beta/3 sum_p Re Tr U_p

A dropped term, a wrong sign or a transposed operand in the force changes
the directional derivative and fails these checks; the golden vectors then
pin the same routines against the reference itself."""
import numpy as np
import pytest

import oracle_lib as ol

EPS = 1e-5
L = (4, 4, 4, 4)
KAPPA, CSW = 0.13, 1.0


@pytest.fixture(scope="module")
def setup():
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(11, 0.3)
    phi = lat.spinor_gauss(12, 0)
    rng = np.random.default_rng(2024)
    dirs = [rng.standard_normal((4, 2, lat.volh, 8)) for _ in range(8)]
    return lat, u, phi, dirs


def s_f(lat, u, phi):
    ck = CSW * KAPPA
    a, ai, b = lat.clover_init(u, ck)
    uf = lat.flip_bc(u)
    op = lat.op(uf, a, ai, KAPPA, ck)
    x, it, rc = lat.cg(op, ol.OPS["mtdagmt"], phi, 1e-28, 5000)
    assert rc == 0
    return float(np.sum(phi * x)), x, op, uf, b


def s_det(lat, u):
    _, _, b = lat.clover_init(u, CSW * KAPPA)
    return lat.clover_action(np.ascontiguousarray(b[1]))


def fd(f, lat, u, v):
    return (f(lat.hmc_u(u, v, EPS)) - f(lat.hmc_u(u, v, -EPS))) / (2 * EPS)


def test_dsf_is_minus_gradient_of_s_f(setup):
    lat, u, phi, dirs = setup
    _, x, op, uf, b = s_f(lat, u, phi)
    mx = lat.apply(op, ol.OPS["mtil"], x)
    force = lat.dsf(op, uf, 1.0, x, mx, b)
    for v in dirs:
        d = fd(lambda w: s_f(lat, w, phi)[0], lat, u, v)
        assert abs(d + np.sum(force * v)) < 1e-6 * abs(d)


def test_dsd_is_minus_gradient_of_clover_action(setup):
    lat, u, phi, dirs = setup
    _, _, b = lat.clover_init(u, CSW * KAPPA)
    force = lat.dsd(b, 1.0, CSW * KAPPA, u)
    for v in dirs[:4]:
        d = fd(lambda w: s_det(lat, w), lat, u, v)
        assert abs(d + np.sum(force * v)) < 1e-6 * abs(d)


def test_dsg_is_minus_gradient_of_gauge_action(setup):
    lat, u, phi, dirs = setup
    beta = 5.29
    force = lat.dsg(u, 1.0, beta)
    for v in dirs[:4]:
        d = fd(lambda w: -beta / 3 * lat.plaquette_sum(w), lat, u, v)
        assert abs(d + np.sum(force * v)) < 1e-6 * abs(d)
