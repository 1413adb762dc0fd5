"""The allocation-free FP64 CG (hb_cg_d_noalloc) against hb_cg_d, through the
C ABI.  Both run the same kernels with the same deterministic reductions (with
all 18 reals of each link read, hb_set_link_recon 18), so the iteration
counts, solutions and true residuals must be identical.

Two calls of hb_cg_d_noalloc in a row share one static work area: first a
16^3x32 lattice (the largest it takes, volh = HB_NOALLOC_MAX_VOLH) on
M~^+M~, then a smaller 8^4 lattice on W^+W from a nonzero initial guess,
which runs over what the first call left in the area.  A lattice above the
limit is refused with HB_E_ARG before any work is queued."""
import pytest

pytestmark = pytest.mark.gpu

MAX_VOLH = 65536  # HB_NOALLOC_MAX_VOLH in include/bqcd_hip.h


@pytest.fixture(scope="module")
def hip():
    import bqcd_hip
    return bqcd_hip


def setup(hip, torch, dev, L, seed, kappa, csw, guess):
    g = hip.Lattice(L)
    g.gauge_random(seed, 0.0)
    g.gauge_smear(5, 0.36)
    g.build_clover(csw * kappa)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 1)
    b = torch.randn((12, g.volh, 2), dtype=torch.float64, device=dev, generator=gen)
    x0 = 0.01 * torch.roll(b, 1, dims=1) if guess else torch.zeros_like(b)
    return g, b, x0


def test_noalloc_cg_matches_cg_d_on_two_calls():
    import torch
    import bqcd_hip as hip

    dev = torch.device("cuda:0")
    cases = [  # (lattice, seed, kappa, csw, operator, rho, initial guess)
        ((16, 16, 16, 32), 20261220, 0.1355, 1.9, "mtdagmt", 0.0, False),
        ((8, 8, 8, 8), 20261230, 0.13, 1.0, "w_dagger_w", 0.1, True),
    ]
    setups, refs = [], []
    for L, seed, kappa, csw, op, rho, guess in cases:
        g, b, x0 = setup(hip, torch, dev, L, seed, kappa, csw, guess)
        assert g.volh <= MAX_VOLH
        g.set_link_recon(18)
        rest = 1e-20 * float((b * b).sum())
        x = x0.clone()
        it, tr = g.cg_d(op, x, b, kappa, rho, rest=rest, maxiter=5000)
        refs.append((x, it, tr))
        setups.append((g, b, x0, kappa, rho, op, rest))
    # the two no-alloc calls back to back, the second on the smaller lattice
    outs = []
    for (g, b, x0, kappa, rho, op, rest) in setups:
        print("no-alloc CG on", g.L, op, flush=True)
        x = x0.clone()
        it, tr = g.cg_d_noalloc(op, x, b, kappa, rho, rest=rest, maxiter=5000)
        outs.append((x, it, tr))
    for (xr, itr, trr), (x, it, tr), (g, b, *_r) in zip(refs, outs, setups):
        print(g.L, "iterations", itr, it, "true |r|^2", trr, tr, flush=True)
        assert it == itr and it > 0
        assert tr == trr
        assert torch.equal(x, xr)
        g.close()


def test_noalloc_cg_refuses_lattice_above_limit(hip):
    import torch

    dev = torch.device("cuda:0")
    g = hip.Lattice((16, 16, 16, 34))
    assert g.volh > MAX_VOLH
    g.gauge_random(5, 0.0)
    b = torch.ones((12, g.volh, 2), dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    with pytest.raises(hip.HipError, match="static work area"):
        g.cg_d_noalloc("mtdagmt", x, b, 0.13)
    g.close()
