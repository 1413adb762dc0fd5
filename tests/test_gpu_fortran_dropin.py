"""The Fortran drop-ins (bqcd-r399-quda_amd/fortran/hip_solver.F90) called
from BQCD itself: oracle/_ref/bqcd_dropin links the reference CPU path with
libbqcd_hip.so and runs, in one process on the same fields, each reference
routine next to the drop-in that replaces it (cg / hip_solver on M~+M~ and
W+W, dsf / hip_dsf, dsd / hip_dsd, dsg / hip_dsg, hmc_u / hip_hmc_u with the
clover rebuild), then whole trajectories: hmc() against hip_hmc() from the
same seed; the same with every CG stopped after 12 iterations (cg_para%log
= 2, BQCD's benchmark mode); and hmc_test's forward / backward pair
(HMC_TEST_FORWARDS, then HMC_TEST_BACKWARDS with tau negated).  The binary
reports every comparison and fails on: different CG iteration counts or
calls, solutions apart by more than 1e-9 of their maximum (both stop at the
same |r|^2 bound), forces and updated fields apart by more than 1e-12,
exp(-Delta H) apart by 1e-9 relative, links after a trajectory apart by 1e-10,
links left in a different boundary-flip state, or BQCD's running CG
iteration total (cg_iterations_total) advancing differently.

The binary is built in the build container (make -C oracle/ref dropin);
without it the test is skipped."""
import os
import subprocess
import tempfile

import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "bqcd_dropin")

CASES = {
    # lattice, kappa, csw, rho, rest, maxiter, step, beta, seed, gauge scale
    "l4444_rho": ((4, 4, 4, 4), 0.13, 1.0, 0.1, 1e-20, 2000, 0.05, 5.29, 4711, 0.3),
    "l4846": ((4, 8, 4, 6), 0.1355, 1.9, 0.0, 1e-18, 2000, 0.02, 5.29, 1234567, 0.5),
}


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/bqcd_dropin not built")
@pytest.mark.parametrize("case", sorted(CASES))
def test_fortran_dropin_against_reference(case):
    L, kappa, csw, rho, rest, maxiter, step, beta, seed, scale = CASES[case]
    lat = ol.Lattice(L)
    fields = {"u": lat.gauge_exp_random(seed, scale), "eta_e": lat.spinor_gauss(seed + 1, 0),
              "mom": lat.momenta_gauss(seed + 3)}
    with tempfile.TemporaryDirectory() as d:
        for nm, arr in fields.items():
            arr.astype("<f8").tofile(os.path.join(d, nm + ".bin"))
        ctl = "%d %d %d %d\n%.17g %.17g %.17g\n%.9g %d\n%.17g %.17g\n%s\n%s\n0\n" % (
            *L, kappa, csw, rho, rest, maxiter, step, beta, d, d)
        print("bqcd_dropin", case, flush=True)
        r = subprocess.run([BIN], input=ctl.encode(), capture_output=True, timeout=120, cwd=d)
    out = r.stdout.decode()
    print(out)
    print(r.stderr.decode()[-2000:])
    assert r.returncode == 0, out
    lines = dict(line.split()[1:3] for line in out.splitlines() if line.startswith("DROPIN"))
    assert lines["failures"] == "0"
    assert lines["mtdagmt_iters_ref"] == lines["mtdagmt_iters_hip"]
    if rho:
        assert lines["wdw_iters_ref"] == lines["wdw_iters_hip"]
    for name in ("hmc", "hmc_fixed12", "hmc_test_forward"):
        assert lines[name + "_cg_iters_ref"] == lines[name + "_cg_iters_hip"]
    assert lines["hmc_cg_total_ref"] == lines["hmc_cg_total_hip"]
    assert lines["hmc_test_backward_cg_iters_ref"] == lines["hmc_test_backward_cg_iters_hip"]
