"""The plugin libraries LIBD and LIBCLOVER replaced by the MI355X routines of
bqcd-r399-quda_amd/fortran/hip_plugins.F90, as BQCD's LIBS selection would
link them: oracle/_ref/bqcd_golden_plugin is the reference's own golden-vector
driver (the reference CPU path built from source) with the libd21 objects and
the LIBCLOVER objects that define d, d_dag, version_of_d, dsf_*, clover_init,
clover_mult_a/ao/b, clover_dsf, clover_dsd, clover_action and
allocate_clover_field_* left out.  Every hot-path routine the driver calls
then runs on the GPU: the clover term, D and D^+ on both parities, the clover
multiplies, M~, M~^+, M~^+M~, W, W^+W, the CG on both operators (BQCD's cg()
calling the plugins), dsf (its dsf_* pieces and clover_dsf), dsd
(clover_dsd), the link update with the clover rebuild, and clover_action.
Its outputs are compared with the golden vectors of the unmodified reference
(tests/golden/l4444.npz, l4846.npz): every field to 1e-12 relative and the
CG iteration counts exactly.

Also hb_dsf_dir through the C ABI with an explicit neighbour table that
points into appended halo copies of the spinors (as xbound_sc_field leaves
them on one process): the eight pieces, called as dsf.F90:80-98 calls them,
give the hopping part of the oracle's dsf.

The binary is built in the build container (make -C oracle/ref plugin);
without it the first test is skipped."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import bqcd_hip
    return bqcd_hip

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "bqcd_golden_plugin")

OUT_SC = ["d_eo", "d_oe", "ddag_eo", "ddag_oe", "cma_e", "cma_o", "cmao_o", "cmao_e", "cmb_o",
          "mtil", "mtil_dag", "mtdagmt", "w_mult", "wdw", "cg_x", "cgw_x"]
OUT_GEN = ["dsf", "dsd", "dsg"]
OUT_CLOVER = ["clover_a", "clover_i", "clover_b", "clover_a_new"]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/bqcd_golden_plugin not built")
@pytest.mark.parametrize("case", ["l4444", "l4846"])
def test_plugins_inside_reference_driver(golden, case):
    f = golden(case)
    L = tuple(int(x) for x in f["L"])
    lat = ol.Lattice(L)
    vh = lat.volh
    seed = int(f["seed"])
    fields = {"u": lat.gauge_exp_random(seed, float(f["scale"])), "eta_e": lat.spinor_gauss(seed + 1, 0),
              "eta_o": lat.spinor_gauss(seed + 2, 1), "mom": lat.momenta_gauss(seed + 3)}
    with tempfile.TemporaryDirectory() as d:
        for nm, arr in fields.items():
            arr.astype("<f8").tofile(os.path.join(d, nm + ".bin"))
        ctl = "%d %d %d %d\n%.17g %.17g %.17g\n%.9g %d\n%.17g %.17g\n%s\n%s\n0\n" % (
            *L, float(f["kappa"]), float(f["csw"]), float(f["rho"]), float(f["rest"]), int(f["maxiter"]),
            float(f["step"]), float(f["beta"]), d, d)
        print("bqcd_golden_plugin", case, flush=True)
        r = subprocess.run([BIN], input=ctl.encode(), capture_output=True, timeout=300, cwd=d)
        print(r.stdout.decode()[-2000:], r.stderr.decode()[-2000:])
        assert r.returncode == 0
        out = {}
        for nm in OUT_SC:
            out[nm] = np.fromfile(os.path.join(d, nm + ".bin"), "<f8").reshape(vh, 3, 4, 2)
        for nm in OUT_GEN:
            out[nm] = np.fromfile(os.path.join(d, nm + ".bin"), "<f8").reshape(4, 2, vh, 8)
        for nm in OUT_CLOVER:
            out[nm] = np.fromfile(os.path.join(d, nm + ".bin"), "<f8").reshape(2, vh, 2, 18, 2)
        out["u_new"] = np.fromfile(os.path.join(d, "u_new.bin"), "<f8").reshape(4, 2, vh, 3, 3, 2)
        info = {}
        for line in open(os.path.join(d, "info.txt")):
            k, v = line.split()
            info[k] = float(v)
    assert info["cg_iters"] == float(f["info_cg_iters"])
    assert info["cgw_iters"] == float(f["info_cgw_iters"])
    assert abs(info["clover_action"] - float(f["info_clover_action"])) <= 1e-12 * abs(float(f["info_clover_action"]))
    for nm, arr in out.items():
        if nm not in f:  # the link update's outputs are stored for one case only
            continue
        print(nm, rel(arr, f[nm]), flush=True)
        assert rel(arr, f[nm]) < 1e-12, nm


def fwd_neighbours(L):
    """nn(i, eo, mu, FWD) of init_common.F90:110-205, 0-based: [eo][mu] arrays
    of the index (in the other parity) of the forward neighbour of site i"""
    LX, LY, LZ, LT = L
    vh = LX * LY * LZ * LT // 2
    nn = np.zeros((2, 4, vh), dtype=np.int64)
    for t in range(LT):
        for z in range(LZ):
            for y in range(LY):
                for x in range(LX):
                    c = [x, y, z, t]
                    e = (x + y + z + t) & 1
                    i = (x + LX * (y + LY * (z + LZ * t))) // 2
                    for mu in range(4):
                        d = list(c)
                        d[mu] = (d[mu] + 1) % L[mu]
                        nn[e, mu, i] = (d[0] + LX * (d[1] + LY * (d[2] + LZ * d[3]))) // 2
    return nn


def test_dsf_dir_pieces_with_halo(hip):
    """dsf.F90:80-98 with the eight pieces of hb_dsf_dir, on spinors with
    appended halo copies addressed through a neighbour table, against the
    hopping part of the oracle's dsf (no clover term)"""
    L = (4, 4, 4, 6)
    lat = ol.Lattice(L)
    vh = lat.volh
    kappa, step = 0.13, 0.05
    uf = lat.flip_bc(lat.gauge_exp_random(55, 0.4))
    a = lat.spinor_gauss(56, 0)
    b = lat.spinor_gauss(57, 0)
    zero = lat.clover()
    p_ref = lat.dsf(lat.op(uf, zero, zero, kappa, 0.0), uf, step, a, b, zero)
    at = lat.d(1, 0, a, uf)                       # A~ = D_oe a
    bt = lat.d(1, 1, b, uf)                       # B~ = D_eo^+ b
    s = -step * 2.0 * kappa * kappa
    nbr = fwd_neighbours(L)
    g = hip.Lattice(L)
    p = np.zeros((4, 2, vh, 8))
    nhalo = 7
    for mu in range(4):
        for eo, bb, aa, fwd in ((0, b, at, 1), (1, bt, a, 1), (0, bt, a, 0), (1, b, at, 0)):
            # the first nhalo sites find their neighbour in an appended copy
            j = nbr[eo, mu]
            nn = (j + 1).astype(np.int32)
            nn[:nhalo] = vh + 1 + np.arange(nhalo)
            halo = lambda f: np.ascontiguousarray(np.concatenate([f, f[j[:nhalo]]], axis=0))
            g.dsf_dir(mu, fwd, p[mu, eo], halo(bb), halo(aa), s, uf[mu, eo], nn)
    print("dsf pieces against the oracle:", rel(p, p_ref), flush=True)
    assert rel(p, p_ref) < 1e-12
    g.close()
