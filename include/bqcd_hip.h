/*
 * bqcd_hip.h - C ABI of the MI355X (gfx950) implementation of the BQCD
 * fermion hot path: even-odd preconditioned clover-Wilson hopping matrix,
 * clover multiplies/inverse, the CG solve on M~^+M~ (or W^+W), and the
 * fermion / clover-determinant / gauge forces.
 *
 * Plain C types only.  Every entry point returns 0 on success or a negative
 * HB_E_* code; hb_last_error() gives the calling thread's last message.  A
 * context is used by one host thread at a time (its calls are not re-entrant);
 * distinct contexts may be used from different threads.  A Fortran iso_c_binding
 * module (bqcd-r399-quda_amd/fortran/bqcd_hip.F90) declares every entry
 * point; bqcd-r399-quda_amd/fortran/hip_solver.F90 maps them onto BQCD's
 * cg() / quda_solver(), dsf, dsd, dsg and hmc_u interfaces.
 *
 * Host-side field layouts are BQCD's own (Fortran column-major):
 *   spinor   complex(4 spin, 3 colour, volh)                SPINCOL_FIELD
 *   gauge    complex(3, 3, site_stride, 0:1, 4)             GAUGE_FIELD
 *   clover   type_clover_a/b (2 blocks, volh, 0:1), 18 complex words/block
 *   momenta  real(8, site_stride, 0:1, 4)                   GENERATOR_FIELD
 * site_stride is volh_tot for arrays allocated by BQCD (only the first volh
 * sites of each (eo, mu) slab are read/written) or volh for packed arrays.
 *
 * Device-side ("_d") entry points take device pointers in the library's own
 * site-major SoA layout: a spinor is 12 planes of volh complex numbers,
 * component k = 3*spin + colour, i.e. double[12][volh][2]; momenta are
 * double[4][2][8][volh].  Parity 0 = EVEN, 1 = ODD, as BQCD.
 */
#ifndef BQCD_HIP_H
#define BQCD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hb_context hb_context;

enum {
  HB_OK = 0,
  HB_E_ARG = -1,       /* invalid argument (shape, parity, null pointer) */
  HB_E_HIP = -2,       /* HIP runtime error */
  HB_E_STATE = -3,     /* field not loaded / wrong order of calls */
  HB_E_NOCONV = -4,    /* CG did not converge within maxiter (log != 2) */
  HB_E_NOMEM = -5
};

/* operators, as the reference's operator routines */
enum {
  HB_OP_MTIL = 0,       /* m_tilde.F90:43      M~ = A_ee - kappa^2 D_eo A_oo^-1 D_oe */
  HB_OP_MTIL_DAG = 1,   /* m_tilde.F90:80      M~^+                                 */
  HB_OP_MTDAGMT = 2,    /* mtdagmt.F90:28      M~^+ M~                              */
  HB_OP_W = 3,          /* w_mult.F90:28       W = M~ + rho                          */
  HB_OP_W_DAG = 4,      /* w_mult.F90:46       W^+                                   */
  HB_OP_WDAGW = 5       /* w_mult.F90:64       W^+ W                                 */
};

/* clover fields */
enum { HB_CLOVER_A = 0, HB_CLOVER_AINV = 1, HB_CLOVER_B = 2 };

/* precisions of the inner solver: FP64; FP32; FP16 storage of spinors, links
 * and clover (per-site scales, FP32 arithmetic); FP16 links and clover with
 * FP32 spinors */
enum { HB_PREC_F64 = 64, HB_PREC_F32 = 32, HB_PREC_F16 = 16, HB_PREC_F16_GAUGE = 18 };

const char *hb_last_error(void);

/* ---- lifecycle: replaces init_quda / comm_set_gridsize (invert_quda hook,
 *      bqcd.F90:75-76) and end_quda (bqcd.F90:93) ------------------------- */
/* L = (LX, LY, LZ, LT) even extents; bc = fermion boundary signs (+1/-1) per
 * direction (bc_fermions, module_lattice.F90); device = HIP ordinal. */
int hb_init(const int L[4], const int bc[4], int device, hb_context **ctx);
int hb_finalize(hb_context *ctx);
int hb_volh(const hb_context *ctx);

/* ---- fields (host, BQCD layout) ---------------------------------------- */
/* Uploads conf%u.  links_are_bc_flipped != 0 means the caller has applied
 * flip_bc (flip_bc.F90:97-121) as BQCD does before every solve; the library
 * keeps physical links and applies the boundary sign in its kernels. */
int hb_load_gauge(hb_context *ctx, const double *u, int site_stride, int links_are_bc_flipped);
int hb_save_gauge(hb_context *ctx, double *u, int site_stride, int bc_flipped);
/* Uploads conf%a, conf%i, conf%b (any may be NULL) as clover_init left them
 * (A/2, A^-1/2, LDL^+ factors; replaces load_clover_quda,
 * invert_quda.F90:70-93).  The rho shift of W is not folded into A: it is the
 * rho argument of the operators, so one clover field serves W and M~. */
int hb_load_clover(hb_context *ctx, const double *a, const double *ainv, const double *b);
/* clover_init (clover/clover_init.F90:28-78) on the device from the loaded
 * physical links: builds A/2, A^-1/2 and the LDL^+ factors for both parities */
int hb_build_clover(hb_context *ctx, double csw_kappa);
int hb_save_clover(hb_context *ctx, double *a, double *ainv, double *b);
/* clover_action (clover/clover_action.F90:29-62) of the odd-parity B field */
int hb_clover_action(hb_context *ctx, double *s_det);

/* ---- operators (host spinors, BQCD layout) ------------------------------ */
/* d / d_dag (d/D21.F90:29-76): out (parity e) = D in (parity 1-e) */
int hb_dslash(hb_context *ctx, int out_parity, int dagger, double *out, const double *in);
/* clover_mult_a / clover_mult_ao / clover_mult_b (clover/clover_mult_*.F90)
 * on one parity: which = HB_CLOVER_A (A), AINV (A^-1 via conf%i),
 * B (A^-1 via the LDL^+ factors) */
int hb_clover_mult(hb_context *ctx, int which, int parity, double *out, const double *in);
/* mtil, mtil_dag, mtdagmt, w_mult, w_mult_dag, w_dagger_w (HB_OP_*) */
int hb_apply(hb_context *ctx, int op, double kappa, double rho, double *out, const double *in);

/* ---- solver: replaces cg() (cg.F90:51) and quda_solver()
 *      (invert_quda.F90:96-196), same contract -------------------------- */
/* Solves op x = b (op = HB_OP_MTDAGMT, or HB_OP_WDAGW with rho).  x is the
 * initial guess on entry and the solution on exit.  Stops when the iterated
 * |r|^2 <= rest (absolute, rest is the float32 cg_para%rest).  log == 2
 * accepts no convergence (returns 0 with iters = maxiter), otherwise
 * HB_E_NOCONV.  true_rtr (may be NULL) receives |b - op x|^2 recomputed in
 * FP64, the re-check quda_solver does (invert_quda.F90:155-168). */
int hb_cg(hb_context *ctx, int op, double kappa, double rho, double *x, const double *b,
          float rest, int maxiter, int log, int *iters, double *true_rtr);

/* Mixed precision CG with reliable updates (defect correction): inner
 * iterations in FP32 or FP16 storage (FP32 arithmetic); FP64 true residual,
 * solution accumulation and global sums.  A reliable update happens when the
 * iterated residual drops by delta against its maximum since the last update
 * or after max_inner iterations.  Same stopping rule and outputs as hb_cg;
 * iters counts inner iterations, n_reliable (may be NULL) the updates. */
int hb_cg_mixed(hb_context *ctx, int op, double kappa, double rho, double *x, const double *b,
                float rest, int maxiter, int log, int precision, double delta, int max_inner,
                int *iters, int *n_reliable, double *true_rtr);

/* Multi right-hand-side solve (no counterpart in the reference, which solves
 * one system per cg() call): nrhs systems op(kappa[k], rho[k]) x[k] = b[k]
 * with host spinors in BQCD layout.  A block of right-hand sides shares one
 * pass of the block hopping matrix (every link and clover word read once for
 * the block); the systems of each distinct kappa are blocked together (rho,
 * e.g. a Hasenbusch shift, may differ within a block) and converged systems
 * are masked out.  csw > 0 builds the clover term of
 * every distinct kappa on the device from the loaded links, csw == 0 means no
 * clover term, csw < 0 uses the loaded clover field (one kappa only).  Inner
 * iterations in precision (HB_PREC_F32 or HB_PREC_F16), FP64 reliable updates
 * per system; each stops on its FP64 true residual <= rest[k].  block = RHS
 * block size 1, 2, 4, 8 or 12, or 0 to let the autotuner time the candidates
 * (block size and workgroup size of the block kernels) and choose; an explicit
 * block size runs with 256-thread workgroups.  iters[k] and true_rtr[k] are
 * outputs. */
int hb_cg_multi(hb_context *ctx, int op, int nrhs, const double *kappa, const double *rho,
                double csw, double *const *x, const double *const *b, const float *rest,
                int maxiter, int log, int precision, int block, int *iters, double *true_rtr);
/* block size chosen by the last hb_cg_multi call */
int hb_multi_block_size(hb_context *ctx);
/* largest RHS block size (12, 8, 4, 2 or 1) whose block fields fit the 4 GiB
 * span of one buffer at volh sites per parity in precision HB_PREC_F32 or
 * HB_PREC_F16; 0 if none does or the arguments are invalid.  hb_cg_multi
 * rejects larger explicit blocks and its autotuner skips them.  Host-only (no
 * device or context needed); no counterpart in the reference. */
int hb_multi_block_limit(int volh, int precision);
/* workgroup size (threads) of the block kernels of the last hb_cg_multi call */
int hb_multi_tile(hb_context *ctx);
/* workgroup size for hb_cg_multi calls with an explicit block size: 128 or 256
 * (the default); the autotuner (block = 0) chooses its own */
int hb_multi_set_tile(hb_context *ctx, int tile);

/* ---- forces (host, BQCD layout; p accumulates) -------------------------- */
/* dsf (dsf.F90:28-110): hopping + clover part of dS_f/dU for spinors a, b
 * (even parity).  The links held by the context are physical. */
int hb_dsf(hb_context *ctx, double *p, int site_stride, double step, double kappa,
           double csw_kappa, const double *a, const double *b);
/* dsd (dsd.F90:28-49): clover-determinant force */
int hb_dsd(hb_context *ctx, double *p, int site_stride, double step, double csw_kappa);
/* dsg (dsg.F90:28-63): Wilson gauge force */
int hb_dsg(hb_context *ctx, double *p, int site_stride, double step, double beta);
/* hmc_u (hmc_u.F90:28-69): U <- exp(i step p.lambda) U, reunitarised; the
 * clover term is rebuilt when csw_kappa != 0 */
int hb_hmc_u(hb_context *ctx, const double *p, int site_stride, double step, double csw_kappa);
/* sum over plaquettes of Re Tr U_p (action.F90:70-123 uses it for S_g) */
int hb_plaquette(hb_context *ctx, double *re_tr_sum);
/* device copies of the links (up to 4 slots), e.g. the HMC backup before a
 * trajectory (hmc.F90 keeps conf_old) */
int hb_gauge_backup(hb_context *ctx, int slot);
int hb_gauge_restore(hb_context *ctx, int slot);

/* ---- device-pointer entry points (library SoA layout, device memory) --- */
/* stream: a hipStream_t, or NULL for the context's own stream, a blocking
 * stream (ordered with the null stream).  All later work of the context is
 * issued on it. */
int hb_set_stream(hb_context *ctx, void *stream);
int hb_dslash_d(hb_context *ctx, int out_parity, int dagger, double *out_d, const double *in_d);
int hb_apply_d(hb_context *ctx, int op, double kappa, double rho, double *out_d,
               const double *in_d);
int hb_cg_d(hb_context *ctx, int op, double kappa, double rho, double *x_d, const double *b_d,
            float rest, int maxiter, int log, int *iters, double *true_rtr);
/* Links the FP64 CG iterates with: 12 (default; rows 0 and 1 are read and row 2
 * is rebuilt as conj(row0 x row1), the reconstruction BQCD asked of QUDA,
 * invert_quda.F90:234) or 18 (every stored real).  The 12-real form is used
 * only while every loaded link is SU(3) to 1e-12, otherwise the CG reads all
 * 18.  The final true residual, hb_dslash, hb_apply and the forces always use
 * all 18 reals.  No counterpart in BQCD's own cg(). */
int hb_set_link_recon(hb_context *ctx, int recon);

/* hb_cg_d without any allocation.  It has no counterpart in the reference; it
 * is for callers that cannot allocate device memory (code under stream
 * capture, tight solver loops, harnesses that forbid hipMalloc).  Arguments,
 * kernels and results are those of hb_cg_d with 18-real links
 * (hb_set_link_recon(ctx, 18)).  Its work fields (five spinors,
 * the CG state, the reduction partials) live in one static device array per
 * GPU, sized for volh <= HB_NOALLOC_MAX_VOLH (60 MiB, held until the process
 * ends); a larger lattice returns HB_E_ARG.  Every call on one GPU uses that
 * GPU's array: calls that overlap on one GPU race.  hb_last_timing reports
 * its operator count, not its time. */
#define HB_NOALLOC_MAX_VOLH 65536
int hb_cg_d_noalloc(hb_context *ctx, int op, double kappa, double rho, double *x_d,
                    const double *b_d, float rest, int maxiter, int log, int *iters,
                    double *true_rtr);
int hb_cg_mixed_d(hb_context *ctx, int op, double kappa, double rho, double *x_d,
                  const double *b_d, float rest, int maxiter, int log, int precision,
                  double delta, int max_inner, int *iters, int *n_reliable, double *true_rtr);
int hb_cg_multi_d(hb_context *ctx, int op, int nrhs, const double *kappa, const double *rho,
                  double csw, double *const *x_d, const double *const *b_d, const float *rest,
                  int maxiter, int log, int precision, int block, int *iters, double *true_rtr);
/* BQCD layout <-> SoA conversions of one spinor (host <-> device) */
int hb_spinor_to_device(hb_context *ctx, double *dst_d, const double *src_host);
int hb_spinor_to_host(hb_context *ctx, double *dst_host, const double *src_d);
/* forces and link update on device fields; p_d is SoA momenta [4][2][8][volh] */
int hb_dsf_d(hb_context *ctx, double *p_d, double step, double kappa, double csw_kappa,
             const double *a_d, const double *b_d);
int hb_dsd_d(hb_context *ctx, double *p_d, double step, double csw_kappa);
int hb_dsg_d(hb_context *ctx, double *p_d, double step, double beta);
int hb_hmc_u_d(hb_context *ctx, const double *p_d, double step, double csw_kappa);
/* synthetic fields on the device (seeded, deterministic on a given GPU) */
int hb_gauge_random(hb_context *ctx, unsigned long long seed, double scale);
int hb_gauge_smear(hb_context *ctx, int sweeps, double alpha);

/* timing of the last solve: its time (ms) between HIP events recorded on the
 * context stream at its start and end, and its number of operator applications */
int hb_last_timing(hb_context *ctx, double *solve_ms, int *n_matvec);
/* average time (ms, HIP events on the context stream) of reps launches of the
 * reduced-precision hopping kernel of the mixed solver's operator (out = A^-1 D
 * in on the odd sites), precision HB_PREC_F16 or HB_PREC_F32, on the even
 * device spinor in_d (converted first); a benchmarking aid with no counterpart
 * in the reference */
int hb_time_dslash_lp(hb_context *ctx, int precision, int reps, const double *in_d, double *ms);
/* out_d = op in_d with the mixed solver's reduced-precision operator
 * (precision HB_PREC_F32, HB_PREC_F16, or HB_PREC_F16_GAUGE: FP32 spinors with
 * 16-bit links and clover): in_d is stored at that precision, the inner
 * iteration's operator kernels are applied (op as hb_apply_d; M~^+ M~ and
 * W^+ W as the solver chains M~ and M~^+), and the stored result is decoded to
 * FP64.  Device spinors of one parity.  The operator at QUDA's sloppy
 * precision (invert_quda.F90:313-316 sets cuda_prec_sloppy, to FP64 in the
 * reference); it exposes the inner operator for verification against
 * hb_apply_d */
int hb_apply_lp_d(hb_context *ctx, int op, double kappa, double rho, int precision, double *out_d,
                  const double *in_d);

/* ---- entry points of the link-time plugins LIBD (d/Makefile:113-129) and
 *      LIBCLOVER (clover/Makefile:44-71); host arrays in BQCD layouts, every
 *      call uploads its operands and downloads its result.  d / d_dag are
 *      hb_load_gauge + hb_dslash, clover_init is hb_load_gauge +
 *      hb_build_clover + hb_save_clover. --------------------------------- */
/* dsf_{x,y,z,t}{f,b} (d/DSFxyzt.F90:77-117): hopping-force piece of
 * direction mu (0..3) forward or backward.  p: real(8, volh) of one (parity,
 * mu), accumulates; b, a: spinors complex(4,3,n_tot) including any halo
 * sites; u: complex(3,3,volh) links of (parity, mu); nn(volh): 1-based
 * neighbour indices into 1..n_tot (checked on the host). */
int hb_dsf_dir(hb_context *ctx, int mu, int forward, double *p, const double *b, const double *a,
               double s, const double *u, const int *nn, int n_tot);
/* clover_mult_a (mode 0: out = A in) and clover_mult_b (mode 1: out = the
 * LDL^+ solve) with one parity of a clover field, type_clover_a/b(2, volh)
 * (clover/clover_mult_a.F90:30-96, clover_mult_b.F90:30-97); clover_mult_ao
 * is mode 0 with out == in */
int hb_clover_mult_field(hb_context *ctx, int mode, const double *clov, double *out, const double *in);
/* clover_dsf(eo, p, b, a, s, u) (clover/clover_d.F90:78-106) on the loaded
 * links, which must be physical; p real(8, site_stride, 0:1, 4) accumulates */
int hb_clover_dsf(hb_context *ctx, int eo, double *p, int site_stride, const double *b, const double *a,
                  double s);
/* clover_dsd(eo, p, b, s, u) (clover/clover_d.F90:40-75) with the loaded
 * LDL^+ field (hb_load_clover) and links */
int hb_clover_dsd(hb_context *ctx, int eo, double *p, int site_stride, double s);
/* clover_action(b) (clover/clover_action.F90:29-62) of one parity of an LDL^+
 * field, type_clover_b(2, volh) */
int hb_clover_action_field(hb_context *ctx, const double *b, double *s_det);

/* ---- HMC trajectory: replaces hmc() (hmc.F90:28-180) with the integrator
 *      of hmc_leap_frog.F90:28-79 / hmc_integrator.F90:28-103 ------------- */
enum { HB_INT_LEAPFROG = 0,   /* hmc_leap_frog.F90: multi-time-scale leapfrog */
       HB_INT_OMELYAN = 1 };  /* second-order minimum-norm (Omelyan) on both scales */

typedef struct {
  double beta, kappa, csw, rho;   /* hmc_para; csw = 0: no clover term */
  double tau;                     /* step of the outer (ir) scale: trajectory length / ntau */
  int ntau, m_scale;              /* outer steps; inner (uv) steps per outer step */
  char model;                     /* 'A', 'B' or 'C' (hmc_integrator.F90) */
  int integrator;                 /* HB_INT_* */
  double lambda;                  /* Omelyan parameter (0 = 0.1931833275037836) */
  float rest;                     /* cg_para%rest (absolute |r|^2) */
  int maxiter;                    /* cg_para%maxiter */
  int log;                        /* cg_para%log: 2 accepts a solve that reaches maxiter
                                     without convergence (cg.F90:119-124, BQCD's
                                     solver_ignore_no_convergence 2); otherwise such a
                                     solve ends the trajectory with HB_E_NOCONV */
  int precision;                  /* solver: HB_PREC_F64 (cg.F90), HB_PREC_F32, HB_PREC_F16 */
  double delta;                   /* reliable-update threshold of the mixed solver */
  int max_inner;
  int mre_vectors;                /* chronological guesses (mre.F90); 0: the source */
  int batch;                      /* model B: solve the two pseudofermions together
                                     (multi-RHS solver, precision F32/F16) */
  int verbose;                    /* 1: one progress line per outer step on stderr */
} hb_hmc_params;

typedef struct {
  double exp_dh, h_old, h_new;
  double sg_old, sg_new;          /* plaquette action without beta: 6V - sum Re Tr U_p / 3 */
  double sp_old, sp_new, sd_old, sd_new;
  double sf1_old, sf1_new, sf2_old, sf2_new;
  double plaquette;               /* <Re Tr U_p>/3 after the trajectory (accepted or not) */
  int accepted;
  int cg_ncall, cg_niter_tot, cg_niter_max;  /* get_cg_stat (cg.F90:166) */
  int n_force_ir, n_force_uv;     /* force evaluations per scale */
  double seconds;                 /* wall time of the trajectory */
} hb_hmc_out;

/* One trajectory from the loaded links (physical; the clover term is rebuilt
 * from them).  p: momenta, eta1/eta2: pseudofermion noise of hmc_init_phi
 * (eta2 is read only for models B and C), in BQCD layouts (p with
 * site_stride); they are the draws of hmc_init_p / hmc_init_phi, so a caller
 * that draws them with BQCD's ranf reproduces BQCD's trajectory.
 * accept_random < 0 forces acceptance; otherwise the trajectory is accepted
 * when accept_random < exp(-Delta H), and on rejection the links (and clover
 * term) of the start are restored. */
int hb_hmc_trajectory(hb_context *ctx, const hb_hmc_params *prm, const double *p, int site_stride,
                      const double *eta1, const double *eta2, double accept_random, hb_hmc_out *out);
/* the same on device fields: p_d [4][2][8][volh], eta*_d [12][volh][2] */
int hb_hmc_trajectory_d(hb_context *ctx, const hb_hmc_params *prm, const double *p_d,
                        const double *eta1_d, const double *eta2_d, double accept_random,
                        hb_hmc_out *out);
/* hmc() in its reversibility-test modes (hmc.F90:124-210, as hmc_test.F90:28-50
 * drives them), on device fields.  direction = +1 (HMC_TEST_FORWARDS): a
 * force-accepted trajectory from the momenta p_d and the noise eta*_d.
 * direction = -1 (HMC_TEST_BACKWARDS): a force-accepted trajectory with the
 * step negated, from the momenta p_d and the pseudofermion fields phi of the
 * trajectory run before it on this context (same model); no noise is read
 * (eta*_d may be NULL), and h_old, exp_dh and the _old terms of *out are 0, as
 * hmc() leaves them.  Both directions write the final momenta back to p_d.
 * Forward then backward returns the links and momenta to the start, and the
 * backward h_new to the forward h_old, up to the solver tolerance. */
int hb_hmc_test_d(hb_context *ctx, const hb_hmc_params *prm, int direction, double *p_d,
                  const double *eta1_d, const double *eta2_d, hb_hmc_out *out);
/* the same on host arrays in BQCD layouts (p with site_stride; eta1/eta2 are
 * read for direction +1 only): p holds the momenta on entry and the final
 * momenta on return, as hmc()'s saved momentum field between the forward and
 * the backward call of hmc_test (hmc_test.F90:41-47). */
int hb_hmc_test(hb_context *ctx, const hb_hmc_params *prm, int direction, double *p, int site_stride,
                const double *eta1, const double *eta2, hb_hmc_out *out);
/* ---- BQCD configuration files (host only, no device needed) -----------
 * conf_write / conf_read (conf.F90:305-470) with field_io (comm/field_io_mpi.F90)
 * and the ConfInfoHeader / CheckSum blocks (conf_info.F90): one big-endian file
 * per time slice holding the first two columns of every link, a .info file
 * with the lattice, parameters and the POSIX cksum of each slice file.  u is a
 * host gauge field in BQCD layout (physical links).  Files are named as
 * files.F90 does: bqcd.<run>.<ensemble>.<former>.<traj>.<t>.u and .info in dir;
 * info_path (may be NULL) receives the .info path. */
int hb_conf_write(const int L[4], const int bc[4], const double *u, int site_stride, const char *dir,
                  int run, int traj, int ensemble, int former, double beta, double kappa, double csw,
                  double h, double plaq_energy, char *info_path, int info_path_len);
/* reads the configuration an .info file describes (slice files relative to the
 * .info file's directory), checking the format, the lattice and every slice's
 * checksum; the third column of each link is rebuilt (u_complete,
 * su3.F90:301-310).  plaq_energy (may be NULL) receives the header's value. */
int hb_conf_read(const char *info_file, const int L[4], double *u, int site_stride, double *plaq_energy);

/* n Gaussian numbers of variance var on the device (seeded, counter-based):
 * momenta are var 1 (hmc_init_p), pseudofermion noise var 1/2 per real
 * component (ran_gauss_volh(12, ., 1/2, .)) */
int hb_gaussian_d(hb_context *ctx, unsigned long long seed, double var, double *x_d, long long n);

#ifdef __cplusplus
}
#endif
#endif
