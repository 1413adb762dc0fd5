"""Pins the oracle (CPU restatement) against the reference BQCD Fortran.

Golden vectors come from tests/golden/make_golden.py, which runs the reference
CPU path compiled from source (oracle/ref/Makefile).  The RNG known-answer
file is the reference's own ranf_test.reference, restated by the sequence of
calls in ranf_test.F90:25-80.
"""
import os
import re

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-13  # relative to the field's max magnitude: the parity bar of the FP64 operator


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def setup_case(f):
    lat = ol.Lattice(tuple(int(x) for x in f["L"]))
    seed = int(f["seed"])
    u = lat.gauge_exp_random(seed, float(f["scale"]))
    eta_e = lat.spinor_gauss(seed + 1, 0)
    eta_o = lat.spinor_gauss(seed + 2, 1)
    mom = lat.momenta_gauss(seed + 3)
    kappa, csw, rho = float(f["kappa"]), float(f["csw"]), float(f["rho"])
    return lat, u, eta_e, eta_o, mom, kappa, csw, rho


def test_ranf_known_answers():
    """ranf/ranset/ranget against ranf_test.reference (integers exact)."""
    text = open(os.path.join(HERE, "golden", "ranf_test.reference")).read()
    blocks = re.split(r"\n\s*\n\s*\n", text)
    rows = [tuple(line.split()) for line in text.splitlines() if re.match(r"^\s+-?\d+\s+\d+\s+[\d.]+$", line)]
    assert len(rows) == 10 + 10 + 31 + 31 + 48 + 48
    got = []
    ol.ranset(0, 0)
    for i in range(1, 11):
        s = ol.ranget(); got.append((i, s, ol.ranf()))
    ol.ranset(4711, 0)
    for i in range(1, 11):
        s = ol.ranget(); got.append((i, s, ol.ranf()))
    for i in range(-10, 21):
        ol.ranset(i, 0); s = ol.ranget(); got.append((i, s, ol.ranf()))
    for i in range(-10, 21):
        ol.ranset(0, i); s = ol.ranget(); got.append((i, s, ol.ranf()))
    for i in range(48):
        ol.ranset(2 ** i, 0); s = ol.ranget(); got.append((i, s, ol.ranf()))
    for i in range(48):
        ol.ranset(0, 2 ** i); s = ol.ranget(); got.append((i, s, ol.ranf()))
    for (ri, rs, rx), (gi, gs, gx) in zip(rows, got):
        assert int(ri) == gi and int(rs) == gs
        assert abs(float(rx) - gx) <= 1e-15
    assert len(blocks) > 1


def test_generator_inputs_bit_exact(golden):
    f = golden("l4444")
    lat, u, eta_e, eta_o, mom, *_ = setup_case(f)
    assert np.array_equal(u, f["u"])
    assert np.array_equal(eta_e, f["eta_e"])
    assert np.array_equal(eta_o, f["eta_o"])
    assert np.array_equal(mom, f["mom"])


@pytest.mark.parametrize("case", ["l4444", "l4846"])
def test_oracle_full_fields(golden, case):
    f = golden(case)
    lat, u, eta_e, eta_o, mom, kappa, csw, rho = setup_case(f)
    if "u_checksum" in f:
        assert np.allclose(f["u_checksum"], [np.sum(u), np.sum(u * u)], rtol=1e-14)
    a, ai, b = lat.clover_init(u, csw * kappa)
    assert rel(a, f["clover_a"]) < TOL
    assert rel(ai, f["clover_i"]) < TOL
    assert rel(b, f["clover_b"]) < TOL
    uf = lat.flip_bc(u)
    assert rel(lat.d(0, 0, eta_o, uf), f["d_eo"]) < TOL
    assert rel(lat.d(1, 0, eta_e, uf), f["d_oe"]) < TOL
    assert rel(lat.d(0, 1, eta_o, uf), f["ddag_eo"]) < TOL
    assert rel(lat.d(1, 1, eta_e, uf), f["ddag_oe"]) < TOL
    assert rel(lat.clover_mult_a(a[0], eta_e), f["cma_e"]) < TOL
    assert rel(lat.clover_mult_a(a[1], eta_o), f["cma_o"]) < TOL
    assert rel(lat.clover_mult_a(ai[1], eta_o), f["cmao_o"]) < TOL
    assert rel(lat.clover_mult_a(ai[0], eta_e), f["cmao_e"]) < TOL
    assert rel(lat.clover_mult_b(b[1], eta_o), f["cmb_o"]) < TOL
    op = lat.op(uf, a, ai, kappa, csw * kappa, rho)
    for name, which in (("mtil", 0), ("mtil_dag", 1), ("mtdagmt", 2), ("w_mult", 3), ("wdw", 5)):
        assert rel(lat.apply(op, which, eta_e), f[name]) < TOL, name
    x, it, rc = lat.cg(op, 2, eta_e, float(f["rest"]), int(f["maxiter"]))
    assert rc == 0 and it == int(f["info_cg_iters"])
    assert rel(x, f["cg_x"]) < 1e-12
    x, it, rc = lat.cg(op, 5, eta_e, float(f["rest"]), int(f["maxiter"]), x0=0.01 * eta_o)
    assert rc == 0 and it == int(f["info_cgw_iters"])
    assert rel(x, f["cgw_x"]) < 1e-12
    bvec = lat.apply(op, 0, eta_e)
    step = float(f["step"])
    assert rel(lat.dsf(op, uf, step, eta_e, bvec, b), f["dsf"]) < TOL
    assert rel(lat.dsd(b, step, csw * kappa, u), f["dsd"]) < TOL
    assert rel(lat.dsg(u, step, float(f["beta"])), f["dsg"]) < TOL
    assert abs(lat.clover_action(b[1]) - float(f["info_clover_action"])) < 1e-12 * abs(
        float(f["info_clover_action"]))
    if "u_new" in f:
        un = lat.hmc_u(u, mom, step)
        assert rel(un, f["u_new"]) < TOL
        a2, _, _ = lat.clover_init(un, csw * kappa)
        assert rel(a2, f["clover_a_new"]) < TOL


def test_oracle_sampled_8888(golden):
    f = golden("l8888")
    lat, u, eta_e, eta_o, mom, kappa, csw, rho = setup_case(f)
    assert np.allclose(f["u_checksum"], [np.sum(u), np.sum(u * u)], rtol=1e-14)
    sites = f["sites"]
    a, ai, b = lat.clover_init(u, csw * kappa)
    uf = lat.flip_bc(u)
    op = lat.op(uf, a, ai, kappa, csw * kappa, rho)
    got = {
        "d_eo": lat.d(0, 0, eta_o, uf),
        "ddag_oe": lat.d(1, 1, eta_e, uf),
        "mtdagmt": lat.apply(op, 2, eta_e),
        "wdw": lat.apply(op, 5, eta_e),
    }
    for k, v in got.items():
        assert rel(v[sites], f[k + "_sample"]) < TOL, k
        assert abs(np.sum(v * v) - f[k + "_norm2"]) < 1e-12 * f[k + "_norm2"]
    x, it, rc = lat.cg(op, 2, eta_e, float(f["rest"]), int(f["maxiter"]))
    assert it == int(f["info_cg_iters"])
    assert rel(x[sites], f["cg_x_sample"]) < 1e-12
    p = lat.dsf(op, uf, float(f["step"]), eta_e, lat.apply(op, 0, eta_e), b)
    assert rel(p[:, :, sites], f["dsf_sample"]) < TOL
    p = lat.dsd(b, float(f["step"]), csw * kappa, u)
    assert rel(p[:, :, sites], f["dsd_sample"]) < TOL


def test_golden_trajectory_is_the_reference_quick_test(golden):
    """The golden trajectory traj_quick (the reference's hmc() run by
    oracle/ref/traj_driver.F90 with the parameters of work/input.quick-test)
    reproduces the reference's own printed result for that input,
    work/out.quick-test.reference:44: plaquette energy 0.2460441746,
    exp(-Delta H) 0.4559411236, 43 CG calls, 133 iterations, at most 8."""
    f = golden("traj_quick")
    vol = 4 ** 4
    assert round(float(f["info_sg_new"]) / (6 * vol), 10) == 0.2460441746
    assert round(float(f["info_exp_dh"]), 10) == 0.4559411236
    assert int(f["info_cg_ncall"]) == 43
    assert int(f["info_cg_niter_tot"]) == 133
    assert int(f["info_cg_niter_max"]) == 8


def test_golden_trajectory_is_the_reference_smp_benchmark(golden):
    """traj_smp is the reference's hmc() (r399, libd21, OpenMP) run with the
    parameters of its own benchmark input work/input.smp (24^3x48, cold start,
    default seed, model A, fixed 100-iteration CGs); it reproduces the
    reference's printed result, work/out.smp.reference:40,47 (written by an
    earlier revision with D2 on 40 threads): PlaqEnergy 0.2512245607, 11 CG
    calls, 1100 iterations, at most 100, and the final ranf seed."""
    f = golden("traj_smp")
    vol = 24 ** 3 * 48
    assert round(float(f["info_sg_new"]) / (6 * vol), 10) == 0.2512245607
    assert int(f["info_cg_ncall"]) == 11
    assert int(f["info_cg_niter_tot"]) == 1100
    assert int(f["info_cg_niter_max"]) == 100
    assert int(f["info_seed_after"]) == 23583982622317


@pytest.mark.parametrize("L,ntraj,accept_draws,seed_after", [
    ((24, 24, 24, 48), 1, 0, 23583982622317),     # work/out.smp.reference:47
    ((48, 48, 48, 96), 2, 1, 166035151690961),    # work/out.mpp.reference:50
])
def test_benchmark_draw_counts(L, ntraj, accept_draws, seed_after):
    """The reference's benchmark runs end on the ranf seed they print.  A model
    A trajectory draws 8 x 4 complex Gaussians per site for the momenta
    (hmc_init_p) and 12 for the noise (hmc_init_phi), two uniforms each, and
    a Metropolis step one more uniform: skipping that many numbers from the
    default seed reproduces the printed seed, which pins the order and number
    of draws the GPU trajectories are fed with at these sizes."""
    vh = L[0] * L[1] * L[2] * L[3] // 2
    ol.ranset(0, ntraj * 88 * vh + accept_draws)
    assert ol.ranget() == seed_after


def test_draws_consume_the_stream_as_the_reference():
    """drawing the fields of one trajectory (hmc_draws) advances ranf by 88
    numbers per site, as the skip-ahead above assumes"""
    lat = ol.Lattice((4, 4, 4, 6))
    lat.hmc_draws(0, hasenbusch=False)
    s1 = ol.ranget()
    ol.ranset(0, 88 * lat.volh)
    assert ol.ranget() == s1
    lat.hmc_draws(None, hasenbusch=False)   # the next trajectory continues the stream
    s2 = ol.ranget()
    ol.ranset(0, 2 * 88 * lat.volh)
    assert ol.ranget() == s2
