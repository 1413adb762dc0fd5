"""Pins the oracle's hopping matrix d / d_dag (d/D21.F90:29-76) by what the
mathematics fixes, independently of the reference's golden vectors.

On unit links a plane wave in(x) = exp(i p.x) chi is an eigenfunction of the
hop: out(x) = exp(i p.x) K(p) chi with

    K(p) = sum_mu [(1 - g_mu) exp(i p_mu) + (1 + g_mu) exp(-i p_mu)]
         = 2 sum_mu cos p_mu - 2 i sum_mu sin p_mu g_mu,

and d_dag the same with g_mu -> -g_mu.  The g_mu read off the oracle must be
Hermitian, square to one, anticommute pairwise, be traceless, and act on spin
only (identity in colour); K must be linear in the sin p_mu.  With the links
of the last time slice negated (flip_bc) the field is antiperiodic in time,
and the momenta p_t = (2n + 1) pi / L_t are the eigenmodes.  A dropped hop,
a projector of the wrong sign, a direction that moves along another axis, a
transposed colour product or a missing boundary sign each break one of these.
"""
import numpy as np
import pytest

import oracle_lib as ol

L = (4, 6, 4, 8)


def coords(eo):
    """coordinates (x, y, z, t) of the sites of parity eo in the oracle's order
    i = (x + Lx (y + Ly (z + Lz t))) / 2"""
    lx, ly, lz, lt = L
    n = np.arange(lx * ly * lz * lt)
    x, y, z, t = n % lx, (n // lx) % ly, (n // (lx * ly)) % lz, n // (lx * ly * lz)
    m = (x + y + z + t) % 2 == eo
    c = np.stack([x[m], y[m], z[m], t[m]], 1)
    return c[np.argsort(n[m] // 2, kind="stable")]


def unit_links(lat):
    u = lat.gauge()
    for k in range(3):
        u[:, :, :, k, k, 0] = 1.0
    return u


def hop_matrix(lat, u, p, dagger, e=0):
    """K(p) (12 x 12, index 4 * colour + spin) from the oracle's d on 12 plane
    waves, checking that each output is exp(i p.x) times one constant spinor"""
    pe = np.exp(1j * coords(e) @ p)
    po = np.exp(1j * coords(1 - e) @ p)
    K = np.zeros((12, 12), complex)
    for k in range(12):
        chi = np.zeros(12, complex)
        chi[k] = 1.0
        inp = (po[:, None] * chi[None, :]).reshape(-1, 3, 4)
        out = lat.d(e, dagger, np.stack([inp.real, inp.imag], -1), u)
        v = (out[..., 0] + 1j * out[..., 1]).reshape(-1, 12) / pe[:, None]
        assert np.max(np.abs(v - v[0])) < 1e-12, "not an eigenfunction"
        K[:, k] = v[0]
    return K


def momentum(n, antiperiodic=False):
    p = 2 * np.pi * np.asarray(n, float) / np.asarray(L, float)
    if antiperiodic:
        p[3] += np.pi / L[3]
    return p


@pytest.fixture(scope="module")
def gammas():
    """g_mu read off the oracle: K(p) for p along mu alone"""
    lat = ol.Lattice(L)
    u = unit_links(lat)
    g = []
    for mu in range(4):
        n = [0, 0, 0, 0]
        n[mu] = 1
        p = momentum(n)
        K = hop_matrix(lat, u, p, 0)
        scalar = 2 * np.cos(p).sum()
        g.append((K - scalar * np.eye(12)) / (-2j * np.sin(p[mu])))
    return g


def test_zero_momentum_is_eight():
    lat = ol.Lattice(L)
    K = hop_matrix(lat, unit_links(lat), np.zeros(4), 0)
    assert np.max(np.abs(K - 8 * np.eye(12))) < 1e-13


def test_gammas_form_a_clifford_algebra(gammas):
    one = np.eye(12)
    for mu, gm in enumerate(gammas):
        g4 = gm.reshape(3, 4, 3, 4)
        # colour identity: only the (c, c) blocks are non-zero and they are equal
        for c in range(3):
            for c2 in range(3):
                want = g4[0, :, 0, :] if c == c2 else np.zeros((4, 4))
                assert np.max(np.abs(g4[c, :, c2, :] - want)) < 1e-12
        assert np.max(np.abs(gm - gm.conj().T)) < 1e-12      # Hermitian
        assert np.max(np.abs(gm @ gm - one)) < 1e-12         # squares to one
        assert abs(np.trace(gm)) < 1e-12                     # traceless
        for nu in range(mu):
            ac = gm @ gammas[nu] + gammas[nu] @ gm
            assert np.max(np.abs(ac)) < 1e-12, (mu, nu)      # anticommute


@pytest.mark.parametrize("antiperiodic", [False, True])
def test_general_momentum_and_dagger(gammas, antiperiodic):
    lat = ol.Lattice(L)
    u = unit_links(lat)
    if antiperiodic:
        u = lat.flip_bc(u)
    p = momentum([1, 2, 3, 1], antiperiodic)
    scalar = 2 * np.cos(p).sum() * np.eye(12)
    slash = sum(np.sin(p[mu]) * gammas[mu] for mu in range(4))
    for dagger, sign in ((0, -1), (1, 1)):
        for e in (0, 1):
            K = hop_matrix(lat, u, p, dagger, e)
            assert np.max(np.abs(K - (scalar + sign * 2j * slash))) < 1e-12, (dagger, e)


def constant_links(lat, V):
    """U_mu(x) = V[mu] everywhere; host layout [mu][eo][site][col][row][re, im]
    (BQCD's u(row, col, site, eo, mu))"""
    u = lat.gauge()
    for mu in range(4):
        u[mu, :, :, :, :, 0] = V[mu].T.real
        u[mu, :, :, :, :, 1] = V[mu].T.imag
    return u


@pytest.mark.parametrize("antiperiodic", [False, True])
def test_constant_su3_links(gammas, antiperiodic):
    """U_mu(x) = V_mu: K = sum_mu exp(i p_mu) V_mu (1 - g_mu) + exp(-i p_mu) V_mu^+ (1 + g_mu),
    colour (slow index) times spin; catches a transposed or un-daggered link"""
    from scipy.linalg import expm

    rng = np.random.default_rng(7)
    V = []
    for mu in range(4):
        h = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        h = 0.5 * (h + h.conj().T)
        h -= np.trace(h) / 3 * np.eye(3)
        V.append(expm(1j * 0.7 * h))
    lat = ol.Lattice(L)
    u = constant_links(lat, V)
    if antiperiodic:
        u = lat.flip_bc(u)
    p = momentum([1, 0, 2, 3], antiperiodic)
    one = np.eye(12)
    for dagger in (0, 1):
        s = 1 if dagger else -1   # d: (1 - g) forward; d_dag: (1 + g) forward
        want = sum(np.exp(1j * p[mu]) * np.kron(V[mu], np.eye(4)) @ (one + s * gammas[mu]) +
                   np.exp(-1j * p[mu]) * np.kron(V[mu].conj().T, np.eye(4)) @ (one - s * gammas[mu])
                   for mu in range(4))
        K = hop_matrix(lat, u, p, dagger)
        assert np.max(np.abs(K - want)) < 1e-12, dagger


def test_antiperiodic_needs_the_flipped_links():
    """the half-integer time momenta are eigenmodes only with the boundary sign"""
    lat = ol.Lattice(L)
    with pytest.raises(AssertionError, match="eigenfunction"):
        hop_matrix(lat, unit_links(lat), momentum([0, 0, 0, 0], True), 0)


def full_lattice_links(u):
    """U[mu][x, y, z, t] as 3 x 3 (row, col) matrices from the host layout"""
    U = np.zeros((4,) + L + (3, 3), complex)
    for eo in range(2):
        c = coords(eo)
        for mu in range(4):
            m = u[mu, eo, :, :, :, 0] + 1j * u[mu, eo, :, :, :, 1]   # [site][col][row]
            U[mu, c[:, 0], c[:, 1], c[:, 2], c[:, 3]] = m.transpose(0, 2, 1)
    return U


def field_strength(U, mu, nu):
    """F_mu nu = (Q - Q^+) / 8i with Q the sum of the four plaquette leaves at x
    in the mu-nu plane, all oriented mu before nu"""
    def at(X, *steps):  # X(x + sum of steps), a step being (direction, +-1)
        for d, s in steps:
            X = np.roll(X, -s, axis=d)
        return X

    def dag(X):
        return np.conj(np.swapaxes(X, -1, -2))

    Um, Un = U[mu], U[nu]
    q = (Um @ at(Un, (mu, 1)) @ dag(at(Um, (nu, 1))) @ dag(Un)
         + Un @ dag(at(Um, (mu, -1), (nu, 1))) @ dag(at(Un, (mu, -1))) @ at(Um, (mu, -1))
         + dag(at(Um, (mu, -1))) @ dag(at(Un, (mu, -1), (nu, -1))) @ at(Um, (mu, -1), (nu, -1)) @ at(Un, (nu, -1))
         + dag(at(Un, (nu, -1))) @ at(Um, (nu, -1)) @ at(Un, (mu, 1), (nu, -1)) @ dag(Um))
    return (q - dag(q)) / 8j


def site_matrices(lat, apply):
    """the 12 x 12 matrix (index 4 * colour + spin) of a site-local operator on every site"""
    n = lat.volh
    M = np.zeros((n, 12, 12), complex)
    for k in range(12):
        chi = np.zeros((n, 12), complex)
        chi[:, k] = 1.0
        out = apply(np.stack([chi.real, chi.imag], -1).reshape(n, 3, 4, 2))
        M[:, :, k] = (out[..., 0] + 1j * out[..., 1]).reshape(n, 12)
    return M


@pytest.mark.parametrize("e", [0, 1])
def test_clover_term_closed_form(gammas, e):
    """clover_init (clover/clover_init.F90:28-189): on every site
    A = 1 - csw kappa sum_{mu<nu} sigma_mu nu F_mu nu, sigma_mu nu = (i/2)[g_mu, g_nu],
    with the g_mu of the hopping matrix and F built here from the links; the
    stored inverse is A^-1, the LDL^+ substitution (clover_mult_b) applies the
    same A^-1, A is Hermitian with trace 12 and commutes with g_5"""
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(11, 0.4)
    ck = 0.25
    a, ai, b = lat.clover_init(u, ck)
    A = site_matrices(lat, lambda x: lat.clover_mult_a(a[e], x))
    Ai = site_matrices(lat, lambda x: lat.clover_mult_a(ai[e], x))
    U = full_lattice_links(u)
    c = coords(e)
    S = np.zeros_like(A)
    for mu in range(4):
        for nu in range(mu + 1, 4):
            F = field_strength(U, mu, nu)[c[:, 0], c[:, 1], c[:, 2], c[:, 3]]
            sigma = 0.5j * (gammas[mu] @ gammas[nu] - gammas[nu] @ gammas[mu])[:4, :4]
            S += np.einsum("nab,st->nasbt", F, sigma).reshape(lat.volh, 12, 12)
    one = np.eye(12)
    assert np.max(np.abs(A - one)) > 0.05                     # a non-trivial field
    assert np.max(np.abs(A - (one - ck * S))) < 1e-13
    assert np.max(np.abs(Ai @ A - one)) < 1e-13
    assert np.max(np.abs(A - np.conj(np.swapaxes(A, 1, 2)))) < 1e-14
    assert np.max(np.abs(np.trace(A, axis1=1, axis2=2) - 12)) < 1e-13
    g5 = gammas[0] @ gammas[1] @ gammas[2] @ gammas[3]
    assert np.max(np.abs(A @ g5 - g5 @ A)) < 1e-14
    chi = np.random.default_rng(3).normal(size=(lat.volh, 3, 4, 2))
    assert np.max(np.abs(lat.clover_mult_b(b[e], chi) - lat.clover_mult_a(ai[e], chi))) < 1e-13


def gell_mann():
    lam = np.zeros((8, 3, 3), complex)
    lam[0][0, 1] = lam[0][1, 0] = 1
    lam[1][0, 1], lam[1][1, 0] = -1j, 1j
    lam[2][0, 0], lam[2][1, 1] = 1, -1
    lam[3][0, 2] = lam[3][2, 0] = 1
    lam[4][0, 2], lam[4][2, 0] = -1j, 1j
    lam[5][1, 2] = lam[5][2, 1] = 1
    lam[6][1, 2], lam[6][2, 1] = -1j, 1j
    lam[7] = np.diag([1.0, 1.0, -2.0]) / np.sqrt(3.0)
    return lam


def test_hmc_u_is_the_exponential_map():
    """hmc_u (hmc_u.F90:28-69, gen2u su3.F90:28-193): U' = exp(i step sum_j p_j
    lambda_j) U for every link, with the standard Gell-Mann matrices (the
    generator coordinates of im_tr_j / re_tr_j, su3.F90:196-233), compared with
    scipy's matrix exponential; the updated links are SU(3)"""
    from scipy.linalg import expm

    lat = ol.Lattice((4, 4, 4, 4))
    u = lat.gauge_exp_random(5, 0.5)
    p = np.random.default_rng(1).normal(size=lat.gen().shape)
    step = 0.3
    un = lat.hmc_u(u, p, step)
    lam = gell_mann()
    m0 = (u[..., 0] + 1j * u[..., 1]).swapaxes(-1, -2)     # (mu, eo, site, row, col)
    m1 = (un[..., 0] + 1j * un[..., 1]).swapaxes(-1, -2)
    H = np.einsum("meij,jab->meiab", p, lam)
    want = np.array([[[expm(1j * step * H[m, e, i]) @ m0[m, e, i] for i in range(lat.volh)]
                      for e in range(2)] for m in range(4)])
    assert np.max(np.abs(m1 - want)) < 1e-13
    assert np.max(np.abs(m1.conj().swapaxes(-1, -2) @ m1 - np.eye(3))) < 1e-14
    assert np.max(np.abs(np.linalg.det(m1) - 1.0)) < 1e-14


@pytest.mark.parametrize("which,rho", [("mtdagmt", 0.0), ("w_dagger_w", 0.1)])
def test_cg_on_a_plane_wave_converges_in_two_iterations(which, rho):
    """cg (cg.F90:51-137) on M~^+ M~ (mtdagmt.F90, m_tilde.F90) and on W^+ W
    with W = M~ + rho (w_mult.F90), unit links, csw = 0: M~ = 1 - kappa^2 D_eo
    D_oe maps exp(i p.x) chi to exp(i p.x) (1 - kappa^2 K^2) chi, and M~^+
    uses K^+ (d_dag).  K = a - i b.g has two eigenvalues, so the normal
    operator has two on this invariant space and CG ends after two iterations
    with x = exp(i p.x) (M^+ M)^-1 chi, M = 1 - kappa^2 K^2 + rho."""
    lat = ol.Lattice(L)
    u = unit_links(lat)
    a, ai, b = lat.clover_init(u, 0.0)
    kappa = 0.12
    op = lat.op(u, a, ai, kappa, 0.0, rho)
    p = momentum([1, 2, 1, 3])
    K = hop_matrix(lat, u, p, 0)
    M = (1.0 + rho) * np.eye(12) - kappa ** 2 * K @ K
    chi = np.random.default_rng(5).normal(size=12) + 1j * np.random.default_rng(6).normal(size=12)
    pe = np.exp(1j * coords(0) @ p)
    src = (pe[:, None] * chi[None, :]).reshape(-1, 3, 4)
    x, iters, rc = lat.cg(op, ol.OPS[which], np.stack([src.real, src.imag], -1), 1e-24, 50)
    assert rc == 0 and iters <= 2, iters
    want = pe[:, None] * np.linalg.solve(M.conj().T @ M, chi)[None, :]
    got = (x[..., 0] + 1j * x[..., 1]).reshape(-1, 12)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
