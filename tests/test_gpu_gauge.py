"""Device-side gauge kernels: clover_init on the GPU against the reference
golden clover fields and the oracle, the plaquette sum against the oracle, and
the synthetic-field generators against SU(3) invariants."""
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


def to_mats(u):
    """(4,2,vh,3col,3row,2) -> complex (..., row, col)"""
    c = u[..., 0] + 1j * u[..., 1]
    return np.swapaxes(c, -1, -2)


@pytest.mark.parametrize("case", ["l4444", "l4846"])
def test_build_clover_vs_reference(hip, golden, case):
    f = golden(case)
    lat = ol.Lattice(tuple(int(x) for x in f["L"]))
    u = lat.gauge_exp_random(int(f["seed"]), float(f["scale"]))
    g = hip.Lattice(lat.L)
    g.load_gauge(u)
    g.build_clover(float(f["csw"]) * float(f["kappa"]))
    a, ai, b = g.save_clover()
    assert rel(a, f["clover_a"]) < 1e-13
    assert rel(ai, f["clover_i"]) < 1e-13
    assert rel(b, f["clover_b"]) < 1e-13
    g.close()


def test_build_clover_rough_field_and_plaquette(hip):
    L = (6, 4, 4, 8)
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(77, 0.9)
    a0, ai0, b0 = lat.clover_init(u, 1.9 * 0.1355)
    g = hip.Lattice(L)
    g.load_gauge(u)
    g.build_clover(1.9 * 0.1355)
    a, ai, b = g.save_clover()
    assert rel(a, a0) < 1e-13 and rel(ai, ai0) < 1e-13 and rel(b, b0) < 1e-13
    ps = lat.plaquette_sum(u)
    assert abs(g.plaquette() - ps / (3 * 6 * 2 * lat.volh)) < 1e-14
    g.close()


@pytest.mark.parametrize("scale", [0.0, 0.3])
def test_gauge_random_is_su3(hip, scale):
    L = (4, 4, 4, 4)
    g = hip.Lattice(L)
    g.gauge_random(12345, scale)
    m = to_mats(g.save_gauge())
    eye = np.eye(3)
    uu = m @ np.conj(np.swapaxes(m, -1, -2))
    assert np.max(np.abs(uu - eye)) < 1e-13
    assert np.max(np.abs(np.linalg.det(m) - 1)) < 1e-13
    g.gauge_random(12345, scale)
    assert np.array_equal(to_mats(g.save_gauge()), m)  # deterministic
    g.close()


def test_ape_smearing_raises_plaquette_and_stays_su3(hip):
    L = (4, 4, 4, 8)
    g = hip.Lattice(L)
    g.gauge_random(99, 0.0)
    p0 = g.plaquette()
    g.gauge_smear(2, 0.5)
    p1 = g.plaquette()
    assert abs(p0) < 0.05 and p1 > p0 + 0.1
    m = to_mats(g.save_gauge())
    assert np.max(np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3))) < 1e-13
    assert np.max(np.abs(np.linalg.det(m) - 1)) < 1e-13
    g.close()
