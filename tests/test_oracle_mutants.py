"""The oracle's checks that stand without the reference -- the closed forms of
the hopping matrix, clover term, link update and CG (test_oracle_closed_form.py)
and the finite-difference gradients of the forces (test_oracle_fd.py) --
catch plausible mistakes: each case below plants one mistake in a copy of the
oracle's sources, builds it in a temporary directory, and runs those checks
against it, which must fail.  CPU only; the oracle itself is untouched."""
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
ORACLE = os.path.join(REPO, "oracle")
SOURCES = ("bqcd_oracle.c", "bqcd_oracle.h", "or_clover_force.inc")
C, INC = "bqcd_oracle.c", "or_clover_force.inc"

# name: (file, text, replacement)
MUTANTS = {
    # hopping matrix (d, d_dag)
    "x projector sign": (C, "s * (G[mu][k][0] * SPC(in, j, 2, c)",
                         "(mu == 0 ? -s : s) * (G[mu][k][0] * SPC(in, j, 2, c)"),
    "dropped backward y hop": (C, "int j = NN(g, i, e, mu, fb);",
                               "int j = NN(g, i, e, mu, fb);\n        if (mu == 1 && fb == 1) continue;"),
    "transposed link": (C, "cplx m = dag ? conj(UL(u, off, c, r)) : UL(u, off, r, c);",
                        "cplx m = dag ? conj(UL(u, off, c, r)) : UL(u, off, c, r);"),
    "backward link not conjugated": (C, "cplx m = dag ? conj(UL(u, off, c, r)) : UL(u, off, r, c);",
                                     "cplx m = dag ? UL(u, off, c, r) : UL(u, off, r, c);"),
    "z hop along y": (C, "y[mu] += fb == 0 ? 1 : -1;", "y[mu == 2 ? 1 : mu] += fb == 0 ? 1 : -1;"),
    "t projector factor 2 dropped": (C, "b[up][c] += 2.0 * w[c][0];", "b[up][c] += w[c][0];"),
    "t projector spins swapped": (C, "int up = (s < 0) ? 2 : 0; /* (1 - g_t)",
                                  "int up = (s < 0) ? 0 : 2; /* (1 - g_t)"),
    "lower half sign": (C, "b[2 + k][c] += s * (conj(", "b[2 + k][c] += -s * (conj("),
    # clover_init, clover_f_mu_nu, clover_inv
    "clover second block sign": (C, "cplx v2 = factor * (P[r][c] - Q[r][c]);",
                                 "cplx v2 = factor * (P[r][c] + Q[r][c]);"),
    "clover off-diagonal phase": (C, "P[r][c + 3] = f[r][c] + I * h[r][c];",
                                  "P[r][c + 3] = f[r][c] - I * h[r][c];"),
    "clover leaf link conjugated": (C, "load_u(c, u, volh, nu, e, xpm, 0); load_u(d, u, volh, mu, e, x, 1);",
                                    "load_u(c, u, volh, nu, e, xpm, 1); load_u(d, u, volh, mu, e, x, 1);"),
    "clover plane 2-4 for 3-4": (C, "clover_f_mu_nu(g, f, 2, 3, i, eo, u);", "clover_f_mu_nu(g, f, 1, 3, i, eo, u);"),
    "clover inverse row not conjugated": (C, "aset(ainv, col, c, 0.5 * conj(x[c]));",
                                          "aset(ainv, col, c, 0.5 * x[c]);"),
    "clover LDL conjugate dropped": (C, "acc -= L[r][k] * d[k] * conj(L[c][k]);",
                                     "acc -= L[r][k] * d[k] * L[c][k];"),
    # hmc_u: generator matrix, exponential, reunitarisation
    "lambda_2 sign": (C, "H[0][1] = h[0] - I * h[1];", "H[0][1] = h[0] + I * h[1];"),
    "lambda_8 normalisation": (C, "H[2][2] = -2.0 * s3 * h[7];", "H[2][2] = -s3 * h[7];"),
    "exponential series truncated": (C, "for (int k = 1; k <= 24; k++) {", "for (int k = 1; k <= 3; k++) {"),
    "half step": (C, "for (int j = 0; j < 8; j++) q[j] = pp[j] * step;",
                  "for (int j = 0; j < 8; j++) q[j] = pp[j] * 0.5 * step;"),
    "third column not conjugated": (C, "m[0][2] = conj(m[1][0] * m[2][1] - m[2][0] * m[1][1]);",
                                    "m[0][2] = (m[1][0] * m[2][1] - m[2][0] * m[1][1]);"),
    # operators and cg
    "operator kappa not squared": (C, "double bfac = -kappa * kappa;", "double bfac = -kappa;"),
    "cg beta inverted": (C, "double bk = rtr / rtrold;", "double bk = rtrold / rtr;"),
    "cg x updated along r": (C, "x[k] += ak * p[k];", "x[k] += ak * r[k];"),
    "W shift with the wrong sign": (C, "for (int k = 0; k < n; k++) out[k] += op->p.rho * in[k];",
                                    "for (int k = 0; k < n; k++) out[k] -= op->p.rho * in[k];"),
    # forces: im_tr_j / re_tr_j, dsg staples, dsf hopping and clover parts
    "im_tr_j component sign": (C, "p[1] += s * (creal(v[0][1]) - creal(v[1][0]));",
                               "p[1] += s * (creal(v[1][0]) - creal(v[0][1]));"),
    "re_tr_j component sign": (C, "p[6] += s * (cimag(v[2][1]) - cimag(v[1][2]));",
                               "p[6] += s * (cimag(v[1][2]) - cimag(v[2][1]));"),
    "staple link not conjugated": (C, "load_u(a, u, volh, nu, e, j4, 1); load_u(b, u, volh, mu, o, j3, 1);",
                                   "load_u(a, u, volh, nu, e, j4, 0); load_u(b, u, volh, mu, o, j3, 1);"),
    "dsf backward even term dropped": (C, "dsf_link(g, p, bt, a, s, u_mut, mu, 0, 0);", "/* dropped */"),
    "dsf odd clover factor": (C, "double s2 = s1 * kappa * kappa;", "double s2 = s1 * kappa;"),
    "dsf projector lower half sign": (C, "ga[2 + k][c] = -sg * (conj(G[mu][0][k])",
                                      "ga[2 + k][c] = sg * (conj(G[mu][0][k])"),
    "clover force leaf sign": (INC, "m3_mul3(r, v, wb, ud); re_tr_j(q, r, s);",
                               "m3_mul3(r, v, wb, ud); re_tr_j(q, r, -s);"),
    "clover force leaf dropped": (INC, "m3_mul3(r, vd, we, d3); re_tr_j(q, r, s);", "/* dropped */"),
}

RUN = """
import sys
sys.path.insert(0, {tests!r})
import oracle_lib
oracle_lib.LIB_PATH = {so!r}
import pytest
sys.exit(pytest.main(["-q", "-p", "no:cacheprovider"] + {tests_run!r}))
"""


@pytest.fixture(scope="module")
def tmp():
    d = tempfile.mkdtemp(prefix="oracle_mutants_")
    yield d
    shutil.rmtree(d, ignore_errors=True)


def build_and_check(d, mutation=None):
    """copy the oracle's sources into d, apply mutation (file, text,
    replacement), build, and run the checks against that build"""
    os.makedirs(d)
    for f in SOURCES:
        shutil.copy(os.path.join(ORACLE, f), d)
    if mutation:
        fname, a, b = mutation
        path = os.path.join(d, fname)
        src = open(path).read()
        assert src.count(a) == 1, "mutation site not found: %r" % a
        open(path, "w").write(src.replace(a, b))
    so = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=gnu99", "-fcx-limited-range", "-shared", "-o", so,
                    os.path.join(d, C), "-lm"], check=True)
    tests_run = [os.path.join(HERE, "test_oracle_closed_form.py"), os.path.join(HERE, "test_oracle_fd.py")]
    code = RUN.format(tests=HERE, so=so, tests_run=tests_run)
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=REPO, timeout=300)


needs_gcc = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")


@pytest.fixture(scope="module")
def results(tmp):
    """every build and check run, a few at a time (each is its own process)"""
    from concurrent.futures import ThreadPoolExecutor

    jobs = {"control": (os.path.join(tmp, "control"), None)}
    for k, name in enumerate(sorted(MUTANTS)):
        jobs[name] = (os.path.join(tmp, "m%d" % k), MUTANTS[name])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        futures = {name: pool.submit(build_and_check, *job) for name, job in jobs.items()}
        return {name: f.result() for name, f in futures.items()}


@needs_gcc
def test_unmodified_copy_passes(results):
    """the control: the same copy-and-build without a mistake passes every check"""
    r = results["control"]
    assert r.returncode == 0, r.stdout[-2000:]


@needs_gcc
@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_checks_fail_on_planted_mistake(results, name):
    r = results[name]
    assert r.returncode == 1, r.stdout[-2000:]   # 1: the checks ran and some failed
    assert " failed" in r.stdout
