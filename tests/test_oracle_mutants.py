"""The closed-form checks of the oracle's hopping matrix, clover term and
link update (test_oracle_closed_form.py) catch plausible mistakes: each case below plants
one mistake in a copy of oracle/bqcd_oracle.c, builds it in a temporary
directory, and runs those checks against it, which must fail.  CPU only; the
oracle itself is untouched."""
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = os.path.join(REPO, "oracle", "bqcd_oracle.c")

MUTANTS = {
    "x projector sign": ("s * (G[mu][k][0] * SPC(in, j, 2, c)",
                         "(mu == 0 ? -s : s) * (G[mu][k][0] * SPC(in, j, 2, c)"),
    "dropped backward y hop": ("int j = NN(g, i, e, mu, fb);",
                               "int j = NN(g, i, e, mu, fb);\n        if (mu == 1 && fb == 1) continue;"),
    "transposed link": ("cplx m = dag ? conj(UL(u, off, c, r)) : UL(u, off, r, c);",
                        "cplx m = dag ? conj(UL(u, off, c, r)) : UL(u, off, c, r);"),
    "backward link not conjugated": ("cplx m = dag ? conj(UL(u, off, c, r)) : UL(u, off, r, c);",
                                     "cplx m = dag ? UL(u, off, c, r) : UL(u, off, r, c);"),
    "z hop along y": ("y[mu] += fb == 0 ? 1 : -1;", "y[mu == 2 ? 1 : mu] += fb == 0 ? 1 : -1;"),
    "t projector factor 2 dropped": ("b[up][c] += 2.0 * w[c][0];", "b[up][c] += w[c][0];"),
    "t projector spins swapped": ("int up = (s < 0) ? 2 : 0; /* (1 - g_t)", "int up = (s < 0) ? 0 : 2; /* (1 - g_t)"),
    "lower half sign": ("b[2 + k][c] += s * (conj(", "b[2 + k][c] += -s * (conj("),
    # clover_init, clover_f_mu_nu, clover_inv
    "clover second block sign": ("cplx v2 = factor * (P[r][c] - Q[r][c]);",
                                 "cplx v2 = factor * (P[r][c] + Q[r][c]);"),
    "clover off-diagonal phase": ("P[r][c + 3] = f[r][c] + I * h[r][c];",
                                  "P[r][c + 3] = f[r][c] - I * h[r][c];"),
    "clover leaf link conjugated": ("load_u(c, u, volh, nu, e, xpm, 0); load_u(d, u, volh, mu, e, x, 1);",
                                    "load_u(c, u, volh, nu, e, xpm, 1); load_u(d, u, volh, mu, e, x, 1);"),
    "clover plane 2-4 for 3-4": ("clover_f_mu_nu(g, f, 2, 3, i, eo, u);", "clover_f_mu_nu(g, f, 1, 3, i, eo, u);"),
    "clover inverse row not conjugated": ("aset(ainv, col, c, 0.5 * conj(x[c]));", "aset(ainv, col, c, 0.5 * x[c]);"),
    "clover LDL conjugate dropped": ("acc -= L[r][k] * d[k] * conj(L[c][k]);", "acc -= L[r][k] * d[k] * L[c][k];"),
    # hmc_u: generator matrix, exponential, reunitarisation
    "lambda_2 sign": ("H[0][1] = h[0] - I * h[1];", "H[0][1] = h[0] + I * h[1];"),
    "lambda_8 normalisation": ("H[2][2] = -2.0 * s3 * h[7];", "H[2][2] = -s3 * h[7];"),
    "exponential series truncated": ("for (int k = 1; k <= 24; k++) {", "for (int k = 1; k <= 3; k++) {"),
    "half step": ("for (int j = 0; j < 8; j++) q[j] = pp[j] * step;",
                  "for (int j = 0; j < 8; j++) q[j] = pp[j] * 0.5 * step;"),
    "third column not conjugated": ("m[0][2] = conj(m[1][0] * m[2][1] - m[2][0] * m[1][1]);",
                                    "m[0][2] = (m[1][0] * m[2][1] - m[2][0] * m[1][1]);"),
}

RUN = """
import sys
sys.path.insert(0, {tests!r})
import oracle_lib
oracle_lib.LIB_PATH = {so!r}
import pytest
sys.exit(pytest.main(["-q", "-p", "no:cacheprovider", {test!r}]))
"""


@pytest.fixture(scope="module")
def tmp():
    d = tempfile.mkdtemp(prefix="oracle_mutants_")
    yield d
    shutil.rmtree(d, ignore_errors=True)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_closed_form_checks_fail_on_planted_mistake(tmp, name):
    a, b = MUTANTS[name]
    src = open(SRC).read()
    assert src.count(a) == 1, "mutation site not found: %s" % name
    k = sorted(MUTANTS).index(name)
    c, so = os.path.join(tmp, "m%d.c" % k), os.path.join(tmp, "m%d.so" % k)
    open(c, "w").write(src.replace(a, b))
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=gnu99", "-fcx-limited-range", "-shared",
                    "-I", os.path.join(REPO, "oracle"), "-o", so, c, "-lm"], check=True)
    code = RUN.format(tests=HERE, so=so, test=os.path.join(HERE, "test_oracle_closed_form.py"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=REPO, timeout=300)
    assert r.returncode == 1, r.stdout[-2000:]   # 1: tests ran and some failed
    assert " failed" in r.stdout
