/*
 * bqcd_oracle.c - CPU restatement of the BQCD fermion hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only the
 * tests, the smoke entry point and the benchmark's CPU-baseline leg may load
 * this library; the product path never does.
 *
 * Each function cites the reference file:line whose algorithm it restates.
 * Every function is pinned against the reference Fortran (built from source
 * under oracle/_ref, golden vectors in tests/golden) and/or against closed
 * forms and invariants in tests/test_oracle_*.py.
 */
#include "bqcd_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------------- */
/* layout helpers                                                            */
/* ------------------------------------------------------------------------- */
#define SP(ptr, i, s, c) (((cplx *)(ptr))[((size_t)(i) * 3 + (c)) * 4 + (s)])
#define SPC(ptr, i, s, c) (((const cplx *)(ptr))[((size_t)(i) * 3 + (c)) * 4 + (s)])
/* U(row r, col c) of link (mu, eo, i) */
static inline size_t u_off(int volh, int mu, int eo, int i) {
  return (((size_t)mu * 2 + eo) * volh + i) * 9;
}
#define UL(ptr, off, r, c) (((const cplx *)(ptr))[(off) + (c) * 3 + (r)])
#define ULM(ptr, off, r, c) (((cplx *)(ptr))[(off) + (c) * 3 + (r)])
/* clover block word k (0..17) of block J (0,1) at (eo, i) */
static inline size_t cl_off(int volh, int eo, int i, int J) {
  return (((size_t)eo * volh + i) * 2 + J) * 18;
}
static inline int NN(const or_geom *g, int i, int eo, int mu, int fb) {
  return g->nn[(((size_t)fb * 4 + mu) * 2 + eo) * g->volh + i];
}

/* ------------------------------------------------------------------------- */
/* geometry: init_common.F90:110-205, xyzt2i.F90:28-88, index.F90:83-102      */
/* ------------------------------------------------------------------------- */
void or_site_coords(const or_geom *g, int eo, int i, int x[4]) {
  int nxh = g->L[0] / 2;
  int r = i / nxh, xh = i % nxh;
  x[1] = r % g->L[1];
  x[2] = (r / g->L[1]) % g->L[2];
  x[3] = r / (g->L[1] * g->L[2]);
  x[0] = 2 * xh + ((x[1] + x[2] + x[3] + eo) & 1);
}

int or_site_index(const or_geom *g, const int x[4], int *eo) {
  int y[4];
  for (int m = 0; m < 4; m++) y[m] = ((x[m] % g->L[m]) + g->L[m]) % g->L[m];
  if (eo) *eo = (y[0] + y[1] + y[2] + y[3]) & 1;
  int ilex = y[0] + g->L[0] * (y[1] + g->L[1] * (y[2] + g->L[2] * y[3]));
  return ilex / 2;
}

int or_geom_init(or_geom *g, const int L[4]) {
  for (int m = 0; m < 4; m++) {
    if (L[m] <= 0 || (L[m] & 1)) return 1;
    g->L[m] = L[m];
  }
  g->volh = L[0] * L[1] * L[2] * L[3] / 2;
  g->nn = (int *)malloc(sizeof(int) * 16 * (size_t)g->volh);
  if (!g->nn) return 2;
  for (int eo = 0; eo < 2; eo++)
    for (int i = 0; i < g->volh; i++) {
      int x[4];
      or_site_coords(g, eo, i, x);
      for (int fb = 0; fb < 2; fb++)
        for (int mu = 0; mu < 4; mu++) {
          int y[4] = {x[0], x[1], x[2], x[3]};
          y[mu] += fb == 0 ? 1 : -1;
          g->nn[(((size_t)fb * 4 + mu) * 2 + eo) * g->volh + i] = or_site_index(g, y, NULL);
        }
    }
  return 0;
}

void or_geom_free(or_geom *g) {
  free(g->nn);
  g->nn = NULL;
}

/* flip_bc.F90:97-121: negate every t-link leaving the last time slice */
void or_flip_bc(const or_geom *g, double *u) {
  for (int eo = 0; eo < 2; eo++)
    for (int i = 0; i < g->volh; i++) {
      int x[4];
      or_site_coords(g, eo, i, x);
      if (x[3] != g->L[3] - 1) continue;
      double *p = u + 2 * u_off(g->volh, 3, eo, i);
      for (int k = 0; k < 18; k++) p[k] = -p[k];
    }
}

/* ------------------------------------------------------------------------- */
/* Wilson hopping matrix: d/D21.F90:29-182, d/D21xyzt.F90:81-197             */
/* out(x) = sum_mu (1 -+ g_mu) U_mu(x) in(x+mu) + (1 +- g_mu) U_mu(x-mu)^+ in(x-mu) */
/* (upper sign: d, lower sign: d_dag).  Per direction mu the half spinor is  */
/* h = psi_upper + s*G_mu psi_lower and the lower half of the result is      */
/* s*G_mu^+ (U h), with G_x = [[0,i],[i,0]], G_y = [[0,1],[-1,0]],           */
/* G_z = [[i,0],[0,-i]]; for t the projector keeps one spin pair (factor 2).  */
/* ------------------------------------------------------------------------- */
static const cplx G[3][2][2] = {
    {{0, I}, {I, 0}},
    {{0, 1}, {-1, 0}},
    {{I, 0}, {0, -I}},
};

static void su3_mult(cplx out[3][2], const double *u, size_t off, const cplx h[3][2], int dag) {
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 2; k++) {
      cplx acc = 0;
      for (int c = 0; c < 3; c++) {
        cplx m = dag ? conj(UL(u, off, c, r)) : UL(u, off, r, c);
        acc += m * h[c][k];
      }
      out[r][k] = acc;
    }
}

void or_d(const or_geom *g, int e, int dagger, double *out, const double *in, const double *u) {
  int o = 1 - e, volh = g->volh;
  for (int i = 0; i < volh; i++) {
    cplx b[4][3];
    memset(b, 0, sizeof b);
    for (int mu = 3; mu >= 0; mu--) { /* T initialises, then z, y, x (D21.F90:58-72) */
      for (int fb = 0; fb < 2; fb++) {
        int j = NN(g, i, e, mu, fb);
        /* projector sign: forward hop uses (1 - g), backward (1 + g); dagger flips */
        double s = (fb == 0 ? -1.0 : 1.0) * (dagger ? -1.0 : 1.0);
        cplx h[3][2], w[3][2];
        for (int c = 0; c < 3; c++) {
          if (mu == 3) {
            int up = (s < 0) ? 2 : 0; /* (1 - g_t) keeps spins 3,4; (1 + g_t) spins 1,2 */
            h[c][0] = SPC(in, j, up, c);
            h[c][1] = SPC(in, j, up + 1, c);
          } else {
            for (int k = 0; k < 2; k++)
              h[c][k] = SPC(in, j, k, c) +
                        s * (G[mu][k][0] * SPC(in, j, 2, c) + G[mu][k][1] * SPC(in, j, 3, c));
          }
        }
        size_t off = fb == 0 ? u_off(volh, mu, e, i) : u_off(volh, mu, o, j);
        su3_mult(w, u, off, h, fb == 1);
        for (int c = 0; c < 3; c++) {
          if (mu == 3) {
            int up = (s < 0) ? 2 : 0;
            b[up][c] += 2.0 * w[c][0];
            b[up + 1][c] += 2.0 * w[c][1];
          } else {
            b[0][c] += w[c][0];
            b[1][c] += w[c][1];
            /* lower = s * G^+ w */
            for (int k = 0; k < 2; k++)
              b[2 + k][c] += s * (conj(G[mu][0][k]) * w[c][0] + conj(G[mu][1][k]) * w[c][1]);
          }
        }
      }
    }
    for (int sp = 0; sp < 4; sp++)
      for (int c = 0; c < 3; c++) SP(out, i, sp, c) = b[sp][c];
  }
}

/* ------------------------------------------------------------------------- */
/* clover multiplies: clover/clover_mult_a.F90:30-96, clover_mult_ao.F90,     */
/* clover_mult_b.F90 + .h90; storage map clover/clover.h:28-82               */
/* ------------------------------------------------------------------------- */
/* packed Hermitian 6x6 (A) word positions, 0-based: diag (word, re/im) and   */
/* upper off-diagonal (r<c) words, in the type_clover_a sequence order        */
static const int A_DIAG_W[6] = {0, 0, 10, 10, 16, 16};
static const int A_DIAG_P[6] = {0, 1, 0, 1, 0, 1};
static int A_OFF_W[6][6]; /* filled by clover_tables() */
static int B_OFF_W[6][6]; /* strict lower (r>c) words of type_clover_b */
static const int B_DIAG_W[6] = {15, 15, 16, 16, 17, 17};
static int tables_done = 0;

static void clover_tables(void) {
  if (tables_done) return;
  /* type_clover_a: i11,i22 | i12..i16 | i23..i26 | i33,i44 | i34..i36 | i45,i46 | i55,i66 | i56 */
  int w = 1;
  for (int r = 0; r < 2; r++) {
    for (int c = r + 1; c < 6; c++) A_OFF_W[r][c] = w++;
  }
  w = 11;
  for (int r = 2; r < 4; r++)
    for (int c = r + 1; c < 6; c++) A_OFF_W[r][c] = w++;
  A_OFF_W[4][5] = 17;
  /* type_clover_b: i21 | i31 i32 | i41 i42 i43 | i51..i54 | i61..i65 | i11..i66 */
  w = 0;
  for (int r = 1; r < 6; r++)
    for (int c = 0; c < r; c++) B_OFF_W[r][c] = w++;
  tables_done = 1;
}

static inline cplx aget(const cplx *a, int r, int c) {
  if (r == c) {
    const double *d = (const double *)&a[A_DIAG_W[r]];
    return d[A_DIAG_P[r]];
  }
  if (r < c) return a[A_OFF_W[r][c]];
  return conj(a[A_OFF_W[c][r]]);
}
static inline void aset(cplx *a, int r, int c, cplx v) { /* r <= c */
  if (r == c) {
    double *d = (double *)&a[A_DIAG_W[r]];
    d[A_DIAG_P[r]] = creal(v);
  } else {
    a[A_OFF_W[r][c]] = v;
  }
}
static inline cplx bget_l(const cplx *b, int r, int c) { /* L(r,c), r>c */
  return b[B_OFF_W[r][c]];
}
static inline double bget_d(const cplx *b, int k) {
  const double *d = (const double *)&b[B_DIAG_W[k]];
  return d[k & 1];
}
static inline void bset_d(cplx *b, int k, double v) {
  double *d = (double *)&b[B_DIAG_W[k]];
  d[k & 1] = v;
}

static void herm6_mult(cplx y[6], const cplx *a, const cplx x[6]) {
  for (int r = 0; r < 6; r++) {
    cplx acc = 0;
    for (int c = 0; c < 6; c++) acc += aget(a, r, c) * x[c];
    y[r] = acc;
  }
}

/* 6-vector of block J: components k = 3*s + c, s in {0,1} (clover.h:182-193)
 * built from (upper +- lower) spins */
static void chiral_split(const void *in, int i, cplx xp[6], cplx xm[6]) {
  for (int s = 0; s < 2; s++)
    for (int c = 0; c < 3; c++) {
      cplx up = ((const cplx *)in)[((size_t)i * 3 + c) * 4 + s];
      cplx lo = ((const cplx *)in)[((size_t)i * 3 + c) * 4 + s + 2];
      xp[3 * s + c] = up + lo;
      xm[3 * s + c] = up - lo;
    }
}
static void chiral_join(void *out, int i, const cplx y1[6], const cplx y2[6]) {
  for (int s = 0; s < 2; s++)
    for (int c = 0; c < 3; c++) {
      ((cplx *)out)[((size_t)i * 3 + c) * 4 + s] = y1[3 * s + c] + y2[3 * s + c];
      ((cplx *)out)[((size_t)i * 3 + c) * 4 + s + 2] = y1[3 * s + c] - y2[3 * s + c];
    }
}

/* a points at the clover field of one parity: [volh][2][18] */
void or_clover_mult_a(int volh, double *out, const double *a, const double *in) {
  clover_tables();
  for (int i = 0; i < volh; i++) {
    cplx xp[6], xm[6], y1[6], y2[6];
    chiral_split(in, i, xp, xm);
    herm6_mult(y1, (const cplx *)a + cl_off(volh, 0, i, 0), xp);
    herm6_mult(y2, (const cplx *)a + cl_off(volh, 0, i, 1), xm);
    chiral_join(out, i, y1, y2);
  }
}

void or_clover_mult_ao(int volh, const double *a, double *x) {
  or_clover_mult_a(volh, x, a, x); /* per-site in-place is safe: inputs read first */
}

/* LDL^+ forward/back substitution, clover_mult_b.h90:1-15 */
static void ldl_solve(cplx x[6], const cplx *b, const cplx y_in[6]) {
  cplx y[6];
  for (int r = 0; r < 6; r++) {
    cplx acc = y_in[r];
    for (int c = 0; c < r; c++) acc -= bget_l(b, r, c) * y[c];
    y[r] = acc;
  }
  for (int r = 5; r >= 0; r--) {
    cplx acc = y[r] * bget_d(b, r);
    for (int c = r + 1; c < 6; c++) acc -= conj(bget_l(b, c, r)) * x[c];
    x[r] = acc;
  }
}

void or_clover_mult_b(int volh, const double *b, double *x) {
  clover_tables();
  for (int i = 0; i < volh; i++) {
    cplx xp[6], xm[6], y1[6], y2[6];
    chiral_split(x, i, xp, xm);
    ldl_solve(y1, (const cplx *)b + cl_off(volh, 0, i, 0), xp);
    ldl_solve(y2, (const cplx *)b + cl_off(volh, 0, i, 1), xm);
    chiral_join(x, i, y1, y2);
  }
}

/* clover_inv.F90:29-139 (LDL^+, halving) and :142-220 (explicit inverse) */
static void clover_inv(cplx *b, cplx *ainv, cplx *a) {
  double d[6];
  cplx L[6][6];
  for (int r = 0; r < 6; r++) {
    for (int c = 0; c < r; c++) {
      cplx acc = aget(a, r, c);
      for (int k = 0; k < c; k++) acc -= L[r][k] * d[k] * conj(L[c][k]);
      L[r][c] = acc / d[c];
    }
    double dr = creal(aget(a, r, r));
    for (int k = 0; k < r; k++) dr -= d[k] * (creal(L[r][k]) * creal(L[r][k]) + cimag(L[r][k]) * cimag(L[r][k]));
    d[r] = dr;
  }
  for (int r = 1; r < 6; r++)
    for (int c = 0; c < r; c++) b[B_OFF_W[r][c]] = L[r][c];
  for (int k = 0; k < 6; k++) bset_d(b, k, 1.0 / d[k]);
  /* explicit inverse, column by column; store 1/2 * A^-1 (upper triangle) */
  for (int col = 0; col < 6; col++) {
    cplx e[6] = {0}, x[6];
    e[col] = 1.0;
    ldl_solve(x, b, e);
    /* x = A^-1 e_col = column col; row col of A^-1 is conj(x) (Hermitian) */
    for (int c = col; c < 6; c++) aset(ainv, col, c, 0.5 * conj(x[c]));
  }
  for (int k = 0; k < 6; k++) bset_d(b, k, 0.5 * bget_d(b, k));
  for (int r = 0; r < 6; r++)
    for (int c = r; c < 6; c++) aset(a, r, c, 0.5 * aget(a, r, c));
}

/* clover_uuuu.F90:47-128: the four plaquette leaves around x in plane (mu,nu) */
static void m3_mul(cplx r[3][3], const cplx a[3][3], const cplx b[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cplx acc = 0;
      for (int k = 0; k < 3; k++) acc += a[i][k] * b[k][j];
      r[i][j] = acc;
    }
}
static void load_u(cplx m[3][3], const double *u, int volh, int mu, int eo, int i, int dag) {
  size_t off = u_off(volh, mu, eo, i);
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) m[r][c] = dag ? conj(UL(u, off, c, r)) : UL(u, off, r, c);
}

/* clover_f_mu_nu.F90:37-103: f = -i (Q - Q^+), Q = sum of the 4 leaves */
static void clover_f_mu_nu(const or_geom *g, cplx f[3][3], int mu, int nu, int x, int e,
                           const double *u) {
  int o = 1 - e, volh = g->volh;
  int xp_ = NN(g, x, e, mu, 0), xm_ = NN(g, x, e, mu, 1);
  int x_p = NN(g, x, e, nu, 0), x_m = NN(g, x, e, nu, 1);
  int xmp = NN(g, xm_, o, nu, 0), xmm = NN(g, xm_, o, nu, 1), xpm = NN(g, xp_, o, nu, 1);
  cplx q[3][3] = {{0}}, a[3][3], b[3][3], c[3][3], d[3][3], t1[3][3], t2[3][3], t3[3][3];
  /* leaf 1: U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ */
  load_u(a, u, volh, mu, e, x, 0); load_u(b, u, volh, nu, o, xp_, 0);
  load_u(c, u, volh, mu, o, x_p, 1); load_u(d, u, volh, nu, e, x, 1);
  m3_mul(t1, a, b); m3_mul(t2, t1, c); m3_mul(t3, t2, d);
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) q[i][j] += t3[i][j];
  /* leaf 2: U_nu(x) U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu) */
  load_u(a, u, volh, nu, e, x, 0); load_u(b, u, volh, mu, e, xmp, 1);
  load_u(c, u, volh, nu, o, xm_, 1); load_u(d, u, volh, mu, o, xm_, 0);
  m3_mul(t1, a, b); m3_mul(t2, t1, c); m3_mul(t3, t2, d);
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) q[i][j] += t3[i][j];
  /* leaf 3: U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu) */
  load_u(a, u, volh, mu, o, xm_, 1); load_u(b, u, volh, nu, e, xmm, 1);
  load_u(c, u, volh, mu, e, xmm, 0); load_u(d, u, volh, nu, o, x_m, 0);
  m3_mul(t1, a, b); m3_mul(t2, t1, c); m3_mul(t3, t2, d);
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) q[i][j] += t3[i][j];
  /* leaf 4: U_nu(x-nu)^+ U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^+ */
  load_u(a, u, volh, nu, o, x_m, 1); load_u(b, u, volh, mu, o, x_m, 0);
  load_u(c, u, volh, nu, e, xpm, 0); load_u(d, u, volh, mu, e, x, 1);
  m3_mul(t1, a, b); m3_mul(t2, t1, c); m3_mul(t3, t2, d);
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) q[i][j] += t3[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) f[i][j] = -I * (q[i][j] - conj(q[j][i]));
  for (int i = 0; i < 3; i++) f[i][i] = creal(f[i][i]);
}

/* clover_init.F90:28-189 */
void or_clover_init(const or_geom *g, double *a, double *ainv, double *b, const double *u,
                    double csw_kappa) {
  clover_tables();
  double factor = -csw_kappa / 8.0;
  int volh = g->volh;
  for (int eo = 0; eo < 2; eo++)
    for (int i = 0; i < volh; i++) {
      cplx f[3][3], h[3][3], P[6][6], Q[6][6];
      /* P: diag blocks from f_21, off-diagonal block f_32 + i f_31 (init1/init2) */
      memset(P, 0, sizeof P);
      memset(Q, 0, sizeof Q);
      clover_f_mu_nu(g, f, 1, 0, i, eo, u);
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          P[r][c] = f[r][c];
          P[r + 3][c + 3] = -f[r][c];
        }
      clover_f_mu_nu(g, f, 2, 1, i, eo, u);
      clover_f_mu_nu(g, h, 2, 0, i, eo, u);
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) P[r][c + 3] = f[r][c] + I * h[r][c];
      /* Q: f_34 diag blocks, off-diagonal f_14 + i f_42 */
      clover_f_mu_nu(g, f, 2, 3, i, eo, u);
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          Q[r][c] = f[r][c];
          Q[r + 3][c + 3] = -f[r][c];
        }
      clover_f_mu_nu(g, f, 0, 3, i, eo, u);
      clover_f_mu_nu(g, h, 3, 1, i, eo, u);
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) Q[r][c + 3] = f[r][c] + I * h[r][c];
      cplx *a1 = (cplx *)a + cl_off(volh, eo, i, 0);
      cplx *a2 = (cplx *)a + cl_off(volh, eo, i, 1);
      for (int r = 0; r < 6; r++)
        for (int c = r; c < 6; c++) {
          cplx v1 = factor * (P[r][c] + Q[r][c]);
          cplx v2 = factor * (P[r][c] - Q[r][c]);
          if (r == c) { v1 = creal(v1) + 1.0; v2 = creal(v2) + 1.0; }
          aset(a1, r, c, v1);
          aset(a2, r, c, v2);
        }
      clover_inv((cplx *)b + cl_off(volh, eo, i, 0), (cplx *)ainv + cl_off(volh, eo, i, 0), a1);
      clover_inv((cplx *)b + cl_off(volh, eo, i, 1), (cplx *)ainv + cl_off(volh, eo, i, 1), a2);
    }
}

/* clover/clover_action.F90:29-62: 2 * sum log(prod of stored B_kk) */
double or_clover_action(int volh, const double *b_odd) {
  clover_tables();
  double s = 0;
  for (int i = 0; i < volh; i++) {
    double d = 1.0;
    for (int J = 0; J < 2; J++) {
      const cplx *b = (const cplx *)b_odd + cl_off(volh, 0, i, J);
      for (int k = 0; k < 6; k++) d *= bget_d(b, k);
    }
    s += log(d);
  }
  return 2.0 * s;
}

/* ------------------------------------------------------------------------- */
/* BLAS-1 and operators: sc.F90, m_tilde.F90:43-114, mtdagmt.F90, w_mult.F90 */
/* ------------------------------------------------------------------------- */
double or_dot(int n, const double *x, const double *y) {
  double s = 0;
  for (int k = 0; k < n; k++) s += x[k] * y[k];
  return s;
}

static void mtil_impl(const or_op *op, int dag, double *out, const double *in) {
  const or_geom *g = op->g;
  int volh = g->volh, n = 24 * volh;
  double *tmp = (double *)malloc(sizeof(double) * n);
  double kappa = op->p.kappa;
  if (kappa != 0.0) {
    or_d(g, 1, dag, tmp, in, op->u);
    if (op->p.csw_kappa != 0.0)
      or_clover_mult_ao(volh, op->ainv + 2 * cl_off(volh, 1, 0, 0), tmp);
    or_d(g, 0, dag, out, tmp, op->u);
  } else {
    memset(out, 0, sizeof(double) * n);
  }
  double bfac = -kappa * kappa;
  if (op->p.csw_kappa != 0.0) {
    or_clover_mult_a(volh, tmp, op->a + 2 * cl_off(volh, 0, 0, 0), in);
    for (int k = 0; k < n; k++) out[k] = tmp[k] + bfac * out[k];
  } else {
    for (int k = 0; k < n; k++) out[k] = in[k] + bfac * out[k];
  }
  free(tmp);
}

void or_apply(const or_op *op, int which, double *out, const double *in) {
  int n = 24 * op->g->volh;
  double *tmp;
  switch (which) {
  case 0: mtil_impl(op, 0, out, in); break;
  case 1: mtil_impl(op, 1, out, in); break;
  case 2:
    tmp = (double *)malloc(sizeof(double) * n);
    mtil_impl(op, 0, tmp, in);
    mtil_impl(op, 1, out, tmp);
    free(tmp);
    break;
  case 3:
  case 4:
    mtil_impl(op, which == 4, out, in);
    for (int k = 0; k < n; k++) out[k] += op->p.rho * in[k];
    break;
  case 5:
    tmp = (double *)malloc(sizeof(double) * n);
    or_apply(op, 3, tmp, in);
    or_apply(op, 4, out, tmp);
    free(tmp);
    break;
  }
}

/* cg.F90:51-137 */
int or_cg(const or_op *op, int which, double *x, const double *b, float rest, int maxiter,
          int *iterations) {
  int n = 24 * op->g->volh;
  double *r = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n);
  double *q = (double *)malloc(sizeof(double) * n);
  or_apply(op, which, r, x);
  double rtrold = 0, rtr = 0;
  for (int k = 0; k < n; k++) {
    r[k] = b[k] - r[k];
    p[k] = r[k];
    rtrold += r[k] * r[k];
  }
  int niter, conv = 0;
  for (niter = 1; niter <= maxiter; niter++) {
    or_apply(op, which, q, p);
    double paap = or_dot(n, p, q);
    double ak = rtrold / paap;
    rtr = 0;
    for (int k = 0; k < n; k++) {
      x[k] += ak * p[k];
      r[k] -= ak * q[k];
      rtr += r[k] * r[k];
    }
    if (rtr <= (double)rest) { conv = 1; break; }
    double bk = rtr / rtrold;
    rtrold = rtr;
    for (int k = 0; k < n; k++) p[k] = r[k] + bk * p[k];
  }
  if (!conv) niter = maxiter;
  *iterations = niter;
  free(r); free(p); free(q);
  return conv ? 0 : 1;
}

/* ------------------------------------------------------------------------- */
/* SU(3): su3.F90:196-233 (im_tr_j / re_tr_j), :313-365 (u_normalize)         */
/* ------------------------------------------------------------------------- */
static void im_tr_j(double *p, const cplx v[3][3], double s) {
  const double sqrt3 = 1.7320508075688772;
  p[0] += s * (cimag(v[0][1]) + cimag(v[1][0]));
  p[1] += s * (creal(v[0][1]) - creal(v[1][0]));
  p[2] += s * (cimag(v[0][0]) - cimag(v[1][1]));
  p[3] += s * (cimag(v[0][2]) + cimag(v[2][0]));
  p[4] += s * (creal(v[0][2]) - creal(v[2][0]));
  p[5] += s * (cimag(v[1][2]) + cimag(v[2][1]));
  p[6] += s * (creal(v[1][2]) - creal(v[2][1]));
  p[7] += s * (cimag(v[0][0]) + cimag(v[1][1]) - 2.0 * cimag(v[2][2])) / sqrt3;
}

static void re_tr_j(double *p, const cplx v[3][3], double s) {
  const double sqrt3 = 1.7320508075688772;
  p[0] += s * (creal(v[0][1]) + creal(v[1][0]));
  p[1] += s * (cimag(v[1][0]) - cimag(v[0][1]));
  p[2] += s * (creal(v[0][0]) - creal(v[1][1]));
  p[3] += s * (creal(v[0][2]) + creal(v[2][0]));
  p[4] += s * (cimag(v[2][0]) - cimag(v[0][2]));
  p[5] += s * (creal(v[1][2]) + creal(v[2][1]));
  p[6] += s * (cimag(v[2][1]) - cimag(v[1][2]));
  p[7] += s * (creal(v[0][0]) + creal(v[1][1]) - 2.0 * creal(v[2][2])) / sqrt3;
}

static void u_normalize(cplx m[3][3]) {
  double len = 0;
  for (int r = 0; r < 3; r++) len += creal(m[r][0]) * creal(m[r][0]) + cimag(m[r][0]) * cimag(m[r][0]);
  len = sqrt(len);
  for (int r = 0; r < 3; r++) m[r][0] /= len;
  cplx f = 0;
  for (int r = 0; r < 3; r++) f += m[r][1] * conj(m[r][0]);
  for (int r = 0; r < 3; r++) m[r][1] -= f * m[r][0];
  len = 0;
  for (int r = 0; r < 3; r++) len += creal(m[r][1]) * creal(m[r][1]) + cimag(m[r][1]) * cimag(m[r][1]);
  len = sqrt(len);
  for (int r = 0; r < 3; r++) m[r][1] /= len;
  m[0][2] = conj(m[1][0] * m[2][1] - m[2][0] * m[1][1]);
  m[1][2] = conj(m[2][0] * m[0][1] - m[0][0] * m[2][1]);
  m[2][2] = conj(m[0][0] * m[1][1] - m[1][0] * m[0][1]);
}

/* u := exp(i sum_j h_j lambda_j), lambda_j the Gell-Mann matrices.  The
 * reference evaluates the exponential in closed form (su3.F90:28-193); here it
 * is the Taylor series of exp(i H / 2^s) to 24 terms, with H scaled until the
 * sum of its element magnitudes is at most 1/4, squared s times.  Both are the
 * exact exponential to rounding; the golden hmc_u vectors and a comparison
 * with a library matrix exponential check it. */
static void gellmann_h(cplx H[3][3], const double h[8]) {
  const double s3 = 1.0 / 1.7320508075688772;
  H[0][0] = h[2] + s3 * h[7];
  H[1][1] = -h[2] + s3 * h[7];
  H[2][2] = -2.0 * s3 * h[7];
  H[0][1] = h[0] - I * h[1]; H[1][0] = conj(H[0][1]);
  H[0][2] = h[3] - I * h[4]; H[2][0] = conj(H[0][2]);
  H[1][2] = h[5] - I * h[6]; H[2][1] = conj(H[1][2]);
}

static void exp_iH_taylor(cplx U[3][3], const cplx H[3][3]) {
  /* scaling and squaring of the Taylor series */
  double nrm = 0;
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) nrm += cabs(H[i][j]);
  int sq = 0;
  while (nrm > 0.25) { nrm *= 0.5; sq++; }
  double sc = ldexp(1.0, -sq);
  cplx A[3][3], T[3][3], R[3][3], tmp[3][3];
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
    A[i][j] = I * sc * H[i][j];
    R[i][j] = (i == j);
    T[i][j] = (i == j);
  }
  for (int k = 1; k <= 24; k++) {
    m3_mul(tmp, T, A);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
      T[i][j] = tmp[i][j] / k;
      R[i][j] += T[i][j];
    }
  }
  for (int s = 0; s < sq; s++) {
    m3_mul(tmp, R, R);
    memcpy(R, tmp, sizeof R);
  }
  memcpy(U, R, sizeof R);
}

void or_gen2u(double u[18], const double h[8]) {
  cplx H[3][3], U[3][3];
  gellmann_h(H, h);
  exp_iH_taylor(U, H);
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) ((cplx *)u)[c * 3 + r] = U[r][c];
}

/* hmc_u.F90:28-69: U <- exp(i step p.lambda) U, then reunitarise */
void or_hmc_u(const or_geom *g, double *u, const double *p, double step) {
  int volh = g->volh;
  for (int mu = 0; mu < 4; mu++)
    for (int eo = 0; eo < 2; eo++)
      for (int i = 0; i < volh; i++) {
        double q[8], v[18];
        const double *pp = p + ((((size_t)mu * 2 + eo) * volh) + i) * 8;
        for (int j = 0; j < 8; j++) q[j] = pp[j] * step;
        or_gen2u(v, q);
        cplx V[3][3], W[3][3], R[3][3];
        size_t off = u_off(volh, mu, eo, i);
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) {
            V[r][c] = ((cplx *)v)[c * 3 + r];
            W[r][c] = UL(u, off, r, c);
          }
        m3_mul(R, V, W);
        u_normalize(R);
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) ULM(u, off, r, c) = R[r][c];
      }
}

/* ------------------------------------------------------------------------- */
/* staples, gauge force and plaquette: staple.F90:28-76, uuu.F90:28-82,       */
/* dsg.F90:28-63, action.F90:70-123                                          */
/* ------------------------------------------------------------------------- */
static void staple(const or_geom *g, cplx S[3][3], const double *u, int i, int e, int mu) {
  int o = 1 - e, volh = g->volh;
  memset(S, 0, sizeof(cplx) * 9);
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    int j1 = NN(g, i, e, mu, 0), j2 = NN(g, i, e, nu, 0), j3 = NN(g, i, e, nu, 1);
    int j4 = NN(g, j3, o, mu, 0);
    cplx a[3][3], b[3][3], c[3][3], t1[3][3], t2[3][3];
    /* forward: U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ */
    load_u(a, u, volh, nu, o, j1, 0); load_u(b, u, volh, mu, o, j2, 1); load_u(c, u, volh, nu, e, i, 1);
    m3_mul(t1, a, b); m3_mul(t2, t1, c);
    for (int r = 0; r < 3; r++) for (int k = 0; k < 3; k++) S[r][k] += t2[r][k];
    /* backward: U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) */
    load_u(a, u, volh, nu, e, j4, 1); load_u(b, u, volh, mu, o, j3, 1); load_u(c, u, volh, nu, o, j3, 0);
    m3_mul(t1, a, b); m3_mul(t2, t1, c);
    for (int r = 0; r < 3; r++) for (int k = 0; k < 3; k++) S[r][k] += t2[r][k];
  }
}

void or_dsg(const or_geom *g, double *p, const double *u, double step, double beta) {
  if (beta == 0.0) return;
  double s = -step * beta / 3.0;
  int volh = g->volh;
  for (int mu = 0; mu < 4; mu++)
    for (int eo = 0; eo < 2; eo++)
      for (int i = 0; i < volh; i++) {
        cplx S[3][3], U[3][3], W[3][3];
        staple(g, S, u, i, eo, mu);
        load_u(U, u, volh, mu, eo, i, 0);
        m3_mul(W, U, S);
        im_tr_j(p + ((((size_t)mu * 2 + eo) * volh) + i) * 8, W, s);
      }
}

/* sum over all plaquettes of Re Tr U_p (each plaquette once) */
double or_plaquette_sum(const or_geom *g, const double *u) {
  int volh = g->volh;
  double s = 0;
  for (int eo = 0; eo < 2; eo++)
    for (int i = 0; i < volh; i++)
      for (int mu = 0; mu < 4; mu++)
        for (int nu = mu + 1; nu < 4; nu++) {
          int o = 1 - eo;
          int j1 = NN(g, i, eo, mu, 0), j2 = NN(g, i, eo, nu, 0);
          cplx a[3][3], b[3][3], c[3][3], d[3][3], t1[3][3], t2[3][3];
          load_u(a, u, volh, mu, eo, i, 0); load_u(b, u, volh, nu, o, j1, 0);
          load_u(c, u, volh, mu, o, j2, 1); load_u(d, u, volh, nu, eo, i, 1);
          m3_mul(t1, a, b); m3_mul(t2, t1, c);
          for (int r = 0; r < 3; r++)
            for (int k = 0; k < 3; k++) s += creal(t2[r][k] * d[k][r]);
        }
  return s;
}

/* ------------------------------------------------------------------------- */
/* hopping part of the fermion force: dsf.F90:28-110, d/DSFxyzt.F90:77-117    */
/* ------------------------------------------------------------------------- */
/* (1 - sg*g_mu) a at site j: the 4 spin components used in DSFxyzt           */
static void gamma_a(cplx ga[4][3], const double *a, int j, int mu, double sg) {
  for (int c = 0; c < 3; c++) {
    if (mu == 3) {
      int up = sg > 0 ? 2 : 0; /* forward (1 - g_t) keeps spins 3,4 */
      for (int s = 0; s < 4; s++) ga[s][c] = 0;
      ga[up][c] = 2.0 * SPC(a, j, up, c);
      ga[up + 1][c] = 2.0 * SPC(a, j, up + 1, c);
    } else {
      cplx h[2];
      for (int k = 0; k < 2; k++)
        h[k] = SPC(a, j, k, c) - sg * (G[mu][k][0] * SPC(a, j, 2, c) + G[mu][k][1] * SPC(a, j, 3, c));
      ga[0][c] = h[0];
      ga[1][c] = h[1];
      for (int k = 0; k < 2; k++)
        ga[2 + k][c] = -sg * (conj(G[mu][0][k]) * h[0] + conj(G[mu][1][k]) * h[1]);
    }
  }
}

static void dsf_link(const or_geom *g, double *p, const double *bv, const double *av, double s,
                     const double *u, int mu, int eo, int fwd) {
  int volh = g->volh;
  for (int i = 0; i < volh; i++) {
    int j = NN(g, i, eo, mu, 0);
    int ia = fwd ? j : i, ib = fwd ? i : j;
    cplx ga[4][3], w[3][3], v[3][3], U[3][3];
    gamma_a(ga, av, ia, mu, fwd ? 1.0 : -1.0);
    for (int ca = 0; ca < 3; ca++)
      for (int cb = 0; cb < 3; cb++) {
        cplx acc = 0;
        for (int sp = 0; sp < 4; sp++) acc += ga[sp][ca] * conj(SPC(bv, ib, sp, cb));
        w[ca][cb] = acc;
      }
    if (fwd) {
      load_u(U, u, volh, mu, eo, i, 0);
      m3_mul(v, U, w);
    } else {
      load_u(U, u, volh, mu, eo, i, 1);
      m3_mul(v, w, U);
    }
    im_tr_j(p + ((((size_t)mu * 2 + eo) * volh) + i) * 8, v, fwd ? -s : s);
  }
}

static void clover_dsf(const or_geom *g, int eo, double *p, const double *b, const double *a,
                       double s, const double *u);

/* dsf.F90:28-110.  u_mut must hold the flipped (antiperiodic) links on entry;
 * as in the reference it is un-flipped before the clover part and stays so. */
void or_dsf(const or_op *op, double *u_mut, double *p, double step, const double *a,
            const double *bvec, const double *b_clover) {
  const or_geom *g = op->g;
  int volh = g->volh, n = 24 * volh;
  double kappa = op->p.kappa;
  double *at = (double *)calloc(n, sizeof(double));
  double *bt = (double *)calloc(n, sizeof(double));
  if (kappa != 0.0) {
    or_d(g, 1, 0, at, a, u_mut);
    or_d(g, 1, 1, bt, bvec, u_mut);
    if (op->p.csw_kappa != 0.0) {
      or_clover_mult_b(volh, b_clover + 2 * cl_off(volh, 1, 0, 0), at);
      or_clover_mult_b(volh, b_clover + 2 * cl_off(volh, 1, 0, 0), bt);
    }
  }
  double s = -step * 2.0 * kappa * kappa;
  if (s != 0.0) {
    for (int mu = 0; mu < 4; mu++) {
      dsf_link(g, p, bvec, at, s, u_mut, mu, 0, 1);
      dsf_link(g, p, bt, a, s, u_mut, mu, 1, 1);
      dsf_link(g, p, bt, a, s, u_mut, mu, 0, 0);
      dsf_link(g, p, bvec, at, s, u_mut, mu, 1, 0);
    }
  }
  or_flip_bc(g, u_mut);
  double s1 = -step * 2.0 * (op->p.csw_kappa / 8.0);
  double s2 = s1 * kappa * kappa;
  if (s1 != 0.0) clover_dsf(g, 0, p, bvec, a, s1, u_mut);
  if (s2 != 0.0) clover_dsf(g, 1, p, bt, at, s2, u_mut);
  free(at);
  free(bt);
}

/* ------------------------------------------------------------------------- */
/* clover force: see or_clover_force.c                                        */
/* ------------------------------------------------------------------------- */
#include "or_clover_force.inc"

/* ------------------------------------------------------------------------- */
/* random numbers: ranf.F90:25-121 (48-bit LCG), ran.F90:28-110 (Box-Muller)  */
/* ------------------------------------------------------------------------- */
static const int64_t RAN_MH = 2651554, RAN_ML = 15184245;
static const int64_t MASK24 = 16777215, MASK48 = 281474976710655LL;
static const int64_t RAN_DEFAULT_SEED = 48131768981101LL;
static int64_t ran_seed = 48131768981101LL;

static inline int64_t lcg_mult(int64_t seed, int64_t mH, int64_t mL) {
  int64_t sH = (seed >> 24) & MASK24, sL = seed & MASK24;
  return (sL * mL + ((sL * mH + sH * mL) << 24)) & MASK48;
}

double or_ranf(void) {
  ran_seed = lcg_mult(ran_seed, RAN_MH, RAN_ML);
  return ldexp((double)ran_seed, -48);
}

int64_t or_ranget(void) { return ran_seed; }

void or_ranset(int64_t seed_in, int64_t n_skip) {
  ran_seed = seed_in == 0 ? RAN_DEFAULT_SEED : ((seed_in | 1) & MASK48);
  int64_t n = n_skip & MASK48, mmH = RAN_MH, mmL = RAN_ML;
  while (n > 0) {
    if (n & 1) ran_seed = lcg_mult(ran_seed, mmH, mmL);
    int64_t mm = (mmL * mmL + ((mmH * mmL) << 25)) & MASK48;
    mmH = (mm >> 24) & MASK24;
    mmL = mm & MASK24;
    n >>= 1;
  }
}

/* ran.F90:28-110, single process: sites in (t, z, y, x/2) order */
void or_ran_gauss_volh(const or_geom *g, int m, double *ran, double var, int eo) {
  double twovar = 2.0 * var;
  const double twopi = 6.2831853071795862;
  for (int i = 0; i < g->volh; i++)
    for (int j = 0; j < m; j++) {
      double r = sqrt(-twovar * log(or_ranf()));
      double phi = twopi * or_ranf();
      ran[2 * ((size_t)i * m + j)] = r * cos(phi);
      ran[2 * ((size_t)i * m + j) + 1] = r * sin(phi);
    }
  (void)eo;
}

/* ------------------------------------------------------------------------- */
/* synthetic inputs                                                          */
/* ------------------------------------------------------------------------- */
/* U = exp(i * scale * sum_j w_j lambda_j), w_j ~ N(0,1) from ranf Box-Muller */
void or_gauge_exp_random(const or_geom *g, double *u, int64_t seed, double scale) {
  or_ranset(seed, 0);
  int volh = g->volh;
  const double twopi = 6.2831853071795862;
  for (int mu = 0; mu < 4; mu++)
    for (int eo = 0; eo < 2; eo++)
      for (int i = 0; i < volh; i++) {
        double h[8];
        for (int j = 0; j < 8; j += 2) {
          double r = sqrt(-2.0 * log(or_ranf()));
          double phi = twopi * or_ranf();
          h[j] = scale * r * cos(phi);
          h[j + 1] = scale * r * sin(phi);
        }
        or_gen2u(u + 2 * u_off(volh, mu, eo, i), h);
      }
}

/* unit complex Gaussian per component (E|z|^2 = 1), as ran_gauss_volh(12,.,1/2) */
void or_spinor_gauss(const or_geom *g, double *s, int64_t seed, int eo) {
  or_ranset(seed, 0);
  or_ran_gauss_volh(g, 12, s, 0.5, eo);
}

/* unit-variance real Gaussians, 8 per link */
void or_momenta_gauss(const or_geom *g, double *p, int64_t seed) {
  or_ranset(seed, 0);
  size_t n = (size_t)32 * g->volh;
  or_geom tmp = *g;
  (void)tmp;
  for (size_t k = 0; k < n; k++) {
    const double twopi = 6.2831853071795862;
    double r = sqrt(-2.0 * log(or_ranf()));
    double phi = twopi * or_ranf();
    p[2 * k] = r * cos(phi);
    p[2 * k + 1] = r * sin(phi);
  }
}
