"""Multi-RHS solver: every block size and the autotuner, right-hand sides with
different kappa (Hasenbusch-style shifts) in one block, W^+W with rho.  Each
system must reach its FP64 true residual bound and agree with an independent
single-RHS FP64 solve (clover term built by the oracle for its kappa)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
CSW = 1.9


@pytest.fixture(scope="module")
def setup():
    import bqcd_hip
    L = (4, 4, 4, 8)
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(555, 0.35)
    kappas = [0.125, 0.12, 0.11]
    refs = {}
    for k in kappas:
        a, ai, b = lat.clover_init(u, CSW * k)
        refs[k] = lat.op(lat.flip_bc(u), a, ai, k, CSW * k, 0.05)
    phis = [lat.spinor_gauss(900 + j, 0) for j in range(12)]
    g = bqcd_hip.Lattice(L)
    g.load_gauge(u)
    yield g, lat, refs, phis, kappas
    g.close()


def check(lat, refs, xs, its, trs, phis, kap, rests, which, rhos=None):
    for k, x in enumerate(xs):
        assert trs[k] <= rests[k], (k, trs[k], rests[k])
        assert its[k] > 0
        xr, _, _ = lat.cg(refs[kap[k]], which, phis[k], rests[k] * 1e-4, 5000)
        err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        assert err < 1e-7, (k, err)


@pytest.mark.parametrize("block", [1, 2, 4, 8, 12])
def test_multi_block_sizes_mixed_kappa(setup, block):
    g, lat, refs, phis, kappas = setup
    kap = [kappas[j % 3] for j in range(12)]
    rests = [1e-20 * float(np.sum(p * p)) for p in phis]
    xs, its, trs, used = g.cg_multi("mtdagmt", phis, kap, csw=CSW, rests=rests, block=block,
                                    precision="f16")
    assert used == block
    check(lat, refs, xs, its, trs, phis, kap, rests, 2)
    # uneven convergence: lighter kappa needs more iterations
    assert max(its[k] for k in range(12) if kap[k] == 0.125) > max(its[k] for k in range(12) if kap[k] == 0.11)


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_multi_autotune_wdw(setup, prec):
    g, lat, refs, phis, kappas = setup
    kap = [kappas[0]] * 5
    rests = [1e-20 * float(np.sum(p * p)) for p in phis[:5]]
    xs, its, trs, used = g.cg_multi("w_dagger_w", phis[:5], kap, rhos=[0.05] * 5, csw=CSW,
                                    rests=rests, block=0, precision=prec)
    assert used in (1, 2, 4, 8, 12)
    check(lat, refs, xs, its, trs, phis, kap, rests, 5)


@pytest.mark.parametrize("block", [4, 12])
def test_multi_tile_128(setup, block):
    """the 128-thread workgroups the autotuner may choose give the same solutions"""
    g, lat, refs, phis, kappas = setup
    kap = [kappas[j % 3] for j in range(12)]
    rests = [1e-20 * float(np.sum(p * p)) for p in phis]
    g.set_multi_tile(128)
    try:
        xs, its, trs, used = g.cg_multi("mtdagmt", phis, kap, csw=CSW, rests=rests, block=block,
                                        precision="f16")
        assert g.multi_tile() == 128
    finally:
        g.set_multi_tile(256)
    assert used == block
    check(lat, refs, xs, its, trs, phis, kap, rests, 2)


def test_multi_tile_rejects_other_sizes(setup):
    import bqcd_hip
    g = setup[0]
    with pytest.raises(bqcd_hip.HipError, match="tile"):
        g.set_multi_tile(64)


def test_multi_with_loaded_clover_single_kappa(setup):
    g, lat, refs, phis, kappas = setup
    u = g.save_gauge()
    a, ai, b = lat.clover_init(u, CSW * kappas[1])
    g.load_clover(a, ai, b)
    kap = [kappas[1]] * 3
    rests = [1e-20 * float(np.sum(p * p)) for p in phis[:3]]
    xs, its, trs, used = g.cg_multi("mtdagmt", phis[:3], kap, csw=-1.0, rests=rests, block=2)
    check(lat, refs, xs, its, trs, phis, kap, rests, 2)


@pytest.mark.parametrize("L", [(6, 4, 6, 6), (10, 6, 6, 12)])
@pytest.mark.parametrize("tile", [128, 256])
def test_multi_ragged_lattice_partial_blocks(L, tile):
    """site counts that are not multiples of the workgroup (partial last blocks in
    every block kernel), three kappa groups in one block of 4 and a block of 12"""
    import bqcd_hip
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(77, 0.3)
    kappas = [0.125, 0.12, 0.11]
    refs = {}
    for k in kappas:
        a, ai, b = lat.clover_init(u, CSW * k)
        refs[k] = lat.op(lat.flip_bc(u), a, ai, k, CSW * k, 0.0)
    phis = [lat.spinor_gauss(300 + j, 0) for j in range(12)]
    kap = [kappas[j % 3] for j in range(12)]
    rests = [1e-20 * float(np.sum(p * p)) for p in phis]
    g = bqcd_hip.Lattice(L)
    try:
        g.load_gauge(u)
        g.set_multi_tile(tile)
        for block in (4, 12):
            xs, its, trs, used = g.cg_multi("mtdagmt", phis, kap, csw=CSW, rests=rests, block=block,
                                            precision="f16")
            assert used == block and g.multi_tile() == tile
            check(lat, refs, xs, its, trs, phis, kap, rests, 2)
    finally:
        g.close()


def test_sixteen_kappa_groups_leave_gauge_backups_alone():
    """Regression: the per-kappa clover fields of a multi-RHS solve with the
    largest number of distinct kappa values (16) and the gauge backups live in
    disjoint scratch ranges; a backup taken before the solve restores the
    links bit for bit, and an HMC trajectory (which keeps its own copy of the
    start links) leaves the caller's backup slots alone too."""
    import bqcd_hip

    L = (4, 4, 4, 4)
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(556, 0.3)
    u2 = lat.gauge_exp_random(557, 0.3)
    g = bqcd_hip.Lattice(L)
    g.load_gauge(u)
    for slot in range(4):
        g.gauge_backup(slot)
    kappas = [0.10 + 0.001 * j for j in range(16)]
    phis = [lat.spinor_gauss(950 + j, 0) for j in range(16)]
    rests = [1e-20 * float(np.sum(p * p)) for p in phis]
    print("16 kappa groups", flush=True)
    xs, its, trs, blk = g.cg_multi("mtdagmt", phis, kappas, csw=CSW, rests=rests, precision="f16", block=4)
    assert all(trs[k] <= rests[k] for k in range(16))
    g.load_gauge(u2)
    for slot in range(4):
        g.gauge_restore(slot)
        assert np.array_equal(g.save_gauge(), u), slot
    # a rejected trajectory restores its start from its own copy; slot 0 still holds u
    g.load_gauge(u2)
    draws = lat.hmc_draws(80)
    prm = bqcd_hip.hmc_params(beta=5.29, kappa=0.12, csw=CSW, rho=0.1, tau=0.05, ntau=2, model="C",
                              rest=1e-22, maxiter=5000)
    out = g.hmc_trajectory(prm, *draws, accept_random=2.0e300)
    assert out["accepted"] == 0
    assert np.array_equal(g.save_gauge(), u2)
    g.gauge_restore(0)
    assert np.array_equal(g.save_gauge(), u)
    g.close()
