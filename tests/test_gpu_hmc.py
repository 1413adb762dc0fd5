"""The HMC trajectory on the GPU (hb_hmc_trajectory) against the reference's
own hmc() on identical start links and identical random draws (the momenta
and pseudofermion noise that hmc() draws from ranf after ranset(seed), made
here by the oracle's ranf, which reproduces ranf_test.reference exactly):

 - traj_quick: the reference's work/input.quick-test (cold start, model C,
   MRE with 7 vectors, solver bound 1e-5); its golden value is the one the
   reference itself printed (work/out.quick-test.reference:44);
 - traj_b: model B on a rough start, tight solver, no MRE;
 - traj_a: model A with MRE 3;
 - traj_ragged: model B with MRE 4 on a 6x4x2x8 lattice (every extent
   different, partial workgroups).
Bars: exp(-Delta H) to 1e-9 relative, final plaquette action and links to
1e-10, identical CG call and iteration counts.

Then properties no reference pins: Omelyan and leapfrog Delta H fall as
step^2; batched (multi-RHS) and sequential mixed-precision model-B
trajectories agree; rejection restores the start links; the reversibility
test of hmc_test.F90 (forward, then backward with the step negated) returns
links, momenta and H to the start."""
import ast

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import bqcd_hip
    return bqcd_hip


def start_links(lat, f):
    start = str(f["start"])
    if start == "cold":
        u = lat.gauge()
        for k in range(3):
            u[:, :, :, k, k, 0] = 1.0
        return u
    seed, scale = ast.literal_eval(start)  # the (seed, scale) tuple of make_golden.py
    return lat.gauge_exp_random(int(seed), float(scale))


def params(hip, f, **kw):
    d = dict(beta=float(f["beta"]), kappa=float(f["kappa"]), csw=float(f["csw"]), rho=float(f["rho"]),
             tau=float(f["tau"]), ntau=int(f["ntau"]), m_scale=int(f["m_scale"]), model=str(f["model"]),
             rest=float(f["rest"]), maxiter=int(f["maxiter"]), mre_vectors=int(f["mre"]))
    d.update(kw)
    return hip.hmc_params(**d)


@pytest.mark.parametrize("case", ["traj_quick", "traj_b", "traj_a", "traj_ragged"])
def test_trajectory_vs_reference(hip, golden, case):
    f = golden(case)
    L = tuple(int(x) for x in f["L"])
    lat = ol.Lattice(L)
    u0 = start_links(lat, f)
    p, eta1, eta2 = lat.hmc_draws(int(f["seed"]), hasenbusch=str(f["model"]) != "A")
    g = hip.Lattice(L)
    g.load_gauge(u0)
    print(case, flush=True)
    out = g.hmc_trajectory(params(hip, f), p, eta1, eta2)
    print(out, flush=True)
    assert abs(out["exp_dh"] - float(f["info_exp_dh"])) < 1e-9 * float(f["info_exp_dh"])
    assert abs(out["sg_new"] - float(f["info_sg_new"])) < 1e-10 * float(f["info_sg_new"])
    assert out["cg_ncall"] == int(f["info_cg_ncall"])
    assert out["cg_niter_tot"] == int(f["info_cg_niter_tot"])
    assert out["cg_niter_max"] == int(f["info_cg_niter_max"])
    u = g.save_gauge()
    assert np.max(np.abs(u - f["u_new"])) < 1e-10
    g.close()


def benchmark_params(hip, ntau, tau):
    """work/input.smp and work/input.mpp: beta 5, kappa 0.13, csw 1.923076923076923,
    model A (rho 0), leapfrog with m_scale 1, solver_rest 1e-99 (0 as float32),
    solver_maxiter 100, solver_ignore_no_convergence 2, no MRE"""
    return hip.hmc_params(beta=5.0, kappa=0.13, csw=1.923076923076923, rho=0.0, tau=tau, ntau=ntau,
                          m_scale=1, model="A", integrator="leapfrog", rest=1e-99, maxiter=100, log=2,
                          mre_vectors=0)


def cold_links(lat):
    u = lat.gauge()
    for k in range(3):
        u[:, :, :, k, k, 0] = 1.0
    return u


def test_reference_benchmark_smp(hip, golden):
    """The reference's own benchmark input work/input.smp: 24^3x48, cold start,
    default seed, one force-accepted trajectory of 10 steps of 0.02 with every
    CG stopped after 100 iterations.  Against the reference hmc() run here
    (traj_smp): the final plaquette action to 1e-10, the links at 64 sites to
    1e-9 and the CG counts; and against the reference's printed result
    (work/out.smp.reference:40): PlaqEnergy 0.2512245607, 11 calls, 1100
    iterations, at most 100."""
    f = golden("traj_smp")
    L = (24, 24, 24, 48)
    lat = ol.Lattice(L)
    vol = 2 * lat.volh
    p, eta1, _ = lat.hmc_draws(0, hasenbusch=False)
    assert ol.ranget() == int(f["info_seed_after"])
    g = hip.Lattice(L)
    g.load_gauge(cold_links(lat))
    print("smp benchmark trajectory", flush=True)
    out = g.hmc_trajectory(benchmark_params(hip, 10, 0.02), p, eta1)
    print(out, out["sg_new"] / (6 * vol), flush=True)
    assert (out["cg_ncall"], out["cg_niter_tot"], out["cg_niter_max"]) == (11, 1100, 100)
    assert abs(out["sg_new"] - float(f["info_sg_new"])) < 1e-10 * float(f["info_sg_new"])
    assert round(out["sg_new"] / (6 * vol), 10) == 0.2512245607
    u = g.save_gauge()
    sites = f["sites"]
    assert np.max(np.abs(u[:, :, sites] - f["u_new_sample"])) < 1e-9
    g.close()


@pytest.mark.timeout(900)
def test_reference_benchmark_mpp(hip):
    """The reference's benchmark input work/input.mpp: 48^3x96, cold start,
    default seed, 20 steps of 0.01, fixed 100-iteration CGs, two trajectories:
    the first force-accepted, the second a Metropolis step with the next ranf
    number.  The reference printed (work/out.mpp.reference:41,45,50) PlaqEnergy
    0.2507903838 after both (the second was rejected with exp(-Delta H) = 0),
    21 CG calls with 2100 iterations and at most 100 in each, and the final
    seed 166035151690961."""
    L = (48, 48, 48, 96)
    lat = ol.Lattice(L)
    vol = 2 * lat.volh
    g = hip.Lattice(L)
    g.load_gauge(cold_links(lat))
    prm = benchmark_params(hip, 20, 0.01)
    print("mpp benchmark trajectory 1 (force accepted)", flush=True)
    p, eta1, _ = lat.hmc_draws(0, hasenbusch=False)
    out1 = g.hmc_trajectory(prm, p, eta1)
    print(out1, out1["sg_new"] / (6 * vol), flush=True)
    assert (out1["cg_ncall"], out1["cg_niter_tot"], out1["cg_niter_max"]) == (21, 2100, 100)
    assert round(out1["sg_new"] / (6 * vol), 10) == 0.2507903838
    print("mpp benchmark trajectory 2", flush=True)
    p, eta1, _ = lat.hmc_draws(None, hasenbusch=False)
    del _
    r = ol.ranf()
    assert ol.ranget() == 166035151690961
    out2 = g.hmc_trajectory(prm, p, eta1, accept_random=r)
    del p, eta1
    print(out2, flush=True)
    assert (out2["cg_ncall"], out2["cg_niter_tot"], out2["cg_niter_max"]) == (21, 2100, 100)
    assert out2["accepted"] == 0 and out2["exp_dh"] < r
    # rejected: the links (and plaquette) of the first trajectory's end
    assert round(out2["plaquette"], 12) == round(1.0 - out1["sg_new"] / (6 * vol), 12)
    assert round(1.0 - out2["plaquette"], 10) == 0.2507903838
    g.close()


def dh_at(hip, g, u0, prm, draws):
    g.load_gauge(u0)
    out = g.hmc_trajectory(prm, *draws)
    return out["h_new"] - out["h_old"]


@pytest.mark.parametrize("integrator", ["leapfrog", "omelyan"])
def test_delta_h_scales_as_step_squared(hip, integrator):
    L = (4, 4, 4, 4)
    lat = ol.Lattice(L)
    u0 = lat.gauge_exp_random(5, 0.2)
    draws = lat.hmc_draws(77)
    g = hip.Lattice(L)
    dh = []
    for n in (16, 32):
        prm = hip.hmc_params(beta=5.29, kappa=0.13, csw=1.9, rho=0.1, tau=0.2 / n, ntau=n, m_scale=1,
                             model="B", integrator=integrator, rest=1e-20, maxiter=5000)
        dh.append(dh_at(hip, g, u0, prm, draws))
    print(integrator, dh, dh[0] / dh[1], flush=True)
    assert 3.0 < dh[0] / dh[1] < 5.5
    g.close()


@pytest.mark.parametrize("model", ["B", "C"])
def test_batched_matches_sequential(hip, model):
    """Model B solves both pseudofermion systems of an outer step as one 2-RHS
    block; model C does so where an inner and an outer kick of the Omelyan
    integrator meet (the S_f1 and S_f2 systems at the same links).  Against
    the same trajectory with single-RHS solves.  Omelyan merges the kicks
    between two link updates: 4 ntau m + 1 inner and 2 ntau + 1 outer force
    evaluations."""
    L = (4, 4, 4, 6)
    lat = ol.Lattice(L)
    u0 = lat.gauge_exp_random(6, 0.2)
    draws = lat.hmc_draws(78)
    g = hip.Lattice(L)
    res = {}
    ntau, m = 4, 2
    for batch in (False, True):
        prm = hip.hmc_params(beta=5.29, kappa=0.13, csw=1.9, rho=0.1, tau=0.05, ntau=ntau, m_scale=m,
                             model=model, integrator="omelyan", rest=1e-20, maxiter=5000,
                             precision="f32", batch=batch)
        g.load_gauge(u0)
        res[batch] = (g.hmc_trajectory(prm, *draws), g.save_gauge())
    a, b = res[False], res[True]
    print(a[0], b[0], flush=True)
    assert abs(a[0]["h_new"] - a[0]["h_old"] - (b[0]["h_new"] - b[0]["h_old"])) < 1e-7
    assert np.max(np.abs(a[1] - b[1])) < 1e-9
    for out in (a[0], b[0]):
        assert (out["n_force_uv"], out["n_force_ir"]) == (4 * ntau * m + 1, 2 * ntau + 1)
        assert out["cg_ncall"] == a[0]["cg_ncall"]
    g.close()


def test_rejection_restores_links(hip):
    L = (4, 4, 4, 4)
    lat = ol.Lattice(L)
    u0 = lat.gauge_exp_random(8, 0.2)
    draws = lat.hmc_draws(79)
    g = hip.Lattice(L)
    g.load_gauge(u0)
    prm = hip.hmc_params(beta=5.29, kappa=0.13, csw=1.9, rho=0.1, tau=0.05, ntau=2, model="C",
                         rest=1e-22, maxiter=5000)
    out = g.hmc_trajectory(prm, *draws, accept_random=2.0e300)
    assert out["accepted"] == 0
    assert np.array_equal(g.save_gauge(), u0)
    g.close()


@pytest.mark.parametrize("model,integrator", [("B", "leapfrog"), ("C", "omelyan")])
def test_forward_backward_returns_to_start(hip, model, integrator):
    """hmc_test.F90:28-50: a force-accepted trajectory, then one with the step
    negated from its final momenta and the same pseudofermion fields.  The
    integrators are symmetric, so the links and momenta return to the start and
    H to its value there, up to the solver tolerance (|r|^2 <= 1e-24, no
    chronological guesses).  A backward call with no trajectory before it is an
    error."""
    import torch

    L = (4, 4, 4, 6)
    lat = ol.Lattice(L)
    u0 = lat.gauge_exp_random(9, 0.2)
    g = hip.Lattice(L)
    g.load_gauge(u0)
    vh = g.volh
    dev = torch.device("cuda:0")
    p = torch.empty((4, 2, 8, vh), dtype=torch.float64, device=dev)
    eta1 = torch.empty((12, vh, 2), dtype=torch.float64, device=dev)
    eta2 = torch.empty_like(eta1)
    g.gaussian_d(p, 101, 1.0)
    g.gaussian_d(eta1, 102, 0.5)
    g.gaussian_d(eta2, 103, 0.5)
    torch.cuda.synchronize()
    p0 = p.clone()
    prm = hip.hmc_params(beta=5.29, kappa=0.13, csw=1.9, rho=0.1, tau=0.1, ntau=5, m_scale=2, model=model,
                         integrator=integrator, rest=1e-24, maxiter=5000, mre_vectors=0)
    print("reversibility", model, integrator, flush=True)
    with pytest.raises(hip.HipError, match="backward"):
        g.hmc_test_d(prm, -1, p)
    fwd = g.hmc_test_d(prm, +1, p, eta1, eta2)
    u1 = g.save_gauge()
    bwd = g.hmc_test_d(prm, -1, p)
    u2 = g.save_gauge()
    torch.cuda.synchronize()
    dp = float((p - p0).abs().max())
    print(fwd["h_old"], fwd["h_new"], bwd["h_new"], np.max(np.abs(u2 - u0)), dp, flush=True)
    assert fwd["accepted"] == 1 and bwd["accepted"] == 1
    assert np.max(np.abs(u1 - u0)) > 1e-3   # the forward trajectory moved the links
    assert np.max(np.abs(u2 - u0)) < 1e-10
    assert dp < 1e-9
    assert abs(bwd["h_new"] - fwd["h_old"]) < 1e-9 * abs(fwd["h_old"])
    assert bwd["h_old"] == 0.0 and bwd["exp_dh"] == 0.0
    g.close()
