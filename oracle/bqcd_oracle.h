/*
 * bqcd_oracle.h - CPU restatement of the BQCD fermion hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only the tests, the smoke entry point and the CPU-baseline
 * leg of the benchmark may load it.  Nothing in the product path calls it.
 *
 * All fields use BQCD's own memory layouts (Fortran column-major), packed to
 * volh sites per parity:
 *   spinor    complex(4 spin, 3 colour, volh)          -> double[volh][3][4][2]
 *   gauge     complex(3 row, 3 col, volh, 0:1, 4 mu)   -> double[4][2][volh][3 col][3 row][2]
 *   clover    complex(18, 2 blocks, volh, 0:1)         -> double[2][volh][2][18][2]
 *   generator real(8, volh, 0:1, 4 mu)                 -> double[4][2][volh][8]
 * Site index within a parity: i = (x + Lx*(y + Ly*(z + Lz*t))) / 2 (0-based),
 * parity = (x+y+z+t) mod 2, as init_common.F90:110-205 / index.F90:83-102.
 */
#ifndef BQCD_ORACLE_H
#define BQCD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int L[4];
  int volh;
  int *nn; /* nn[((fb*4 + mu)*2 + eo)*volh + i], 0-based, fb 0=FWD 1=BWD */
} or_geom;

typedef struct {
  double kappa, csw_kappa, rho; /* h = 0 only (as every configuration here) */
} or_para;

/* geometry */
int  or_geom_init(or_geom *g, const int L[4]);
void or_geom_free(or_geom *g);
void or_site_coords(const or_geom *g, int eo, int i, int x[4]);
int  or_site_index(const or_geom *g, const int x[4], int *eo);
void or_flip_bc(const or_geom *g, double *u);

/* hopping matrix and clover */
void or_d(const or_geom *g, int e, int dagger, double *out, const double *in, const double *u);
void or_clover_mult_a(int volh, double *out, const double *a, const double *in);
void or_clover_mult_ao(int volh, const double *a, double *x);
void or_clover_mult_b(int volh, const double *b, double *x);
void or_clover_init(const or_geom *g, double *a, double *ainv, double *b, const double *u,
                    double csw_kappa);
double or_clover_action(int volh, const double *b_odd);

/* even-odd operators: op 0 = mtil, 1 = mtil_dag, 2 = mtdagmt, 3 = w_mult,
 * 4 = w_mult_dag, 5 = w_dagger_w */
typedef struct {
  const or_geom *g;
  const double *u, *a, *ainv;
  or_para p;
} or_op;
void or_apply(const or_op *op, int which, double *out, const double *in);

/* BLAS-1 on 24*volh reals */
double or_dot(int n, const double *x, const double *y);

/* CG, BQCD semantics (cg.F90:51-137). Returns 0, or 1 for no convergence. */
int or_cg(const or_op *op, int which, double *x, const double *b, float rest, int maxiter,
          int *iterations);

/* forces (p accumulates) */
void or_dsf(const or_op *op, double *u_mut, double *p, double step, const double *a,
            const double *bvec, const double *b_clover);
void or_dsd(const or_geom *g, double *p, const double *b_clover, double step, double csw_kappa,
            const double *u);
void or_dsg(const or_geom *g, double *p, const double *u, double step, double beta);
double or_plaquette_sum(const or_geom *g, const double *u);

/* SU(3) */
void or_gen2u(double u[18], const double h[8]);
void or_hmc_u(const or_geom *g, double *u, const double *p, double step);

/* BQCD random numbers (ranf.F90, ran.F90) */
void   or_ranset(int64_t seed, int64_t nskip);
int64_t or_ranget(void);
double or_ranf(void);
void   or_ran_gauss_volh(const or_geom *g, int m, double *ran, double var, int eo);

/* synthetic field generators (seeded, deterministic) */
void or_gauge_exp_random(const or_geom *g, double *u, int64_t seed, double scale);
void or_spinor_gauss(const or_geom *g, double *s, int64_t seed, int eo);
void or_momenta_gauss(const or_geom *g, double *p, int64_t seed);

#ifdef __cplusplus
}
#endif
#endif
