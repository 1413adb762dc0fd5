"""Force adjoints behind torch.autograd on the GPU, on the differentiation
workload: 16^3x32, Haar links smoothed by 5 APE-like sweeps (alpha 0.36),
kappa = 0.135, csw = 1.9, two-flavour pseudofermion phi = M~^+ eta.

 - backward of FermionAction (the dsf kernels) against central finite
   differences of S_f (eps = 1e-5) along 8 random su(3) directions; bound
   1e-6 relative (the FD truncation error is O(eps^2), the CG stops at
   |r|/|b| <= 1e-12);
 - the same gradient against the oracle's dsf (pinned to the reference dsf by
   the golden vectors) on identical fields and solution; bound 1e-12;
 - recompute mode equals stored mode; mixed-precision solves give the same
   gradient to the solve tolerance;
 - CloverDetAction and GaugeAction against finite differences."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

L = (16, 16, 16, 32)
KAPPA, CSW = 0.135, 1.9
EPS = 1e-5


@pytest.fixture(scope="module")
def env():
    import torch
    import bqcd_hip
    import actions

    g = bqcd_hip.Lattice(L)
    g.gauge_random(2718, 0.0)
    g.gauge_smear(5, 0.36)
    g.build_clover(CSW * KAPPA)
    lat = ol.Lattice(L)
    eta_h = lat.spinor_gauss(99, 0)
    eta = torch.empty((12, g.volh, 2), dtype=torch.float64, device="cuda")
    g.to_device(eta_h, eta)
    phi = actions.pseudofermion(g, eta, KAPPA)
    rest = float(1e-24 * torch.sum(phi * phi).item())
    gen = torch.Generator(device="cpu").manual_seed(7)
    dirs = [torch.randn((4, 2, 8, g.volh), generator=gen, dtype=torch.float64).cuda()
            for _ in range(8)]
    yield torch, g, actions, lat, phi, rest, dirs
    g.close()


def test_fermion_force_vs_finite_differences(env):
    torch, g, actions, lat, phi, rest, dirs = env

    def S(om):
        return actions.FermionAction.apply(om, g, phi, KAPPA, CSW * KAPPA, rest, 20000)

    omega = torch.zeros_like(dirs[0], requires_grad=True)
    s0 = S(omega)
    s0.backward()
    grad = omega.grad.detach()
    print("S_f", s0.item(), flush=True)
    for k, v in enumerate(dirs):
        with torch.no_grad():
            fd = (S(EPS * v).item() - S(-EPS * v).item()) / (2 * EPS)
        an = float(torch.sum(grad * v).item())
        print("direction", k, "fd", fd, "autograd", an, "rel", abs(fd - an) / abs(fd), flush=True)
        assert abs(fd - an) < 1e-6 * abs(fd)


def test_fermion_force_vs_oracle_dsf(env):
    torch, g, actions, lat, phi, rest, dirs = env
    x = torch.zeros_like(phi)
    g.set_stream(torch.cuda.current_stream())
    g.cg_d("mtdagmt", x, phi, KAPPA, 0.0, rest, 20000, log=1)
    omega = torch.zeros_like(dirs[0], requires_grad=True)
    actions.FermionAction.apply(omega, g, phi, KAPPA, CSW * KAPPA, rest, 20000).backward()
    p_gpu = -omega.grad.detach().cpu().numpy()             # dsf(step=1), (4,2,8,volh)
    # oracle dsf on the same links and solution (host layouts)
    u = g.save_gauge()
    ck = CSW * KAPPA
    a, ai, b = lat.clover_init(u, ck)
    uf = lat.flip_bc(u)
    op = lat.op(uf, a, ai, KAPPA, ck)
    xh = g.to_host(x)
    mx = lat.apply(op, ol.OPS["mtil"], xh)
    p_ref = lat.dsf(op, uf, 1.0, xh, mx, b)                 # (4,2,volh,8)
    p_ref = np.ascontiguousarray(np.transpose(p_ref, (0, 1, 3, 2)))
    rel = np.max(np.abs(p_gpu - p_ref)) / np.max(np.abs(p_ref))
    print("dsf gpu vs oracle", rel, flush=True)
    # the solve is deterministic: both forces contract the same solution
    assert rel < 1e-12


def test_recompute_and_mixed_match_stored(env):
    torch, g, actions, lat, phi, rest, dirs = env
    grads = {}
    for mode in ("stored", "recompute", "f32"):
        omega = torch.zeros_like(dirs[0], requires_grad=True)
        kw = dict(recompute=(mode == "recompute"), precision="f32" if mode == "f32" else "f64")
        actions.FermionAction.apply(omega, g, phi, KAPPA, CSW * KAPPA, rest, 20000,
                                    kw["precision"], kw["recompute"]).backward()
        grads[mode] = omega.grad.detach()
    n = torch.max(torch.abs(grads["stored"])).item()
    assert torch.max(torch.abs(grads["recompute"] - grads["stored"])).item() <= 1e-14 * n
    assert torch.max(torch.abs(grads["f32"] - grads["stored"])).item() < 1e-8 * n


def test_clover_det_and_gauge_forces_vs_finite_differences(env):
    torch, g, actions, lat, phi, rest, dirs = env
    for name, S in (("det", lambda om: actions.CloverDetAction.apply(om, g, CSW * KAPPA)),
                    ("gauge", lambda om: actions.GaugeAction.apply(om, g, 5.29))):
        omega = torch.zeros_like(dirs[0], requires_grad=True)
        S(omega).backward()
        grad = omega.grad.detach()
        for v in dirs[:3]:
            with torch.no_grad():
                fd = (S(EPS * v).item() - S(-EPS * v).item()) / (2 * EPS)
            an = float(torch.sum(grad * v).item())
            print(name, "fd", fd, "autograd", an, flush=True)
            assert abs(fd - an) < 1e-6 * abs(fd)


def test_links_restored_after_calls(env):
    torch, g, actions, lat, phi, rest, dirs = env
    u0 = g.save_gauge()
    with torch.no_grad():
        actions.FermionAction.apply(1e-3 * dirs[0], g, phi, KAPPA, CSW * KAPPA, rest, 20000)
        actions.GaugeAction.apply(1e-3 * dirs[1], g, 5.29)
    assert np.array_equal(g.save_gauge(), u0)
