!===============================================================================
!
! hip_solver.F90 - BQCD-side drop-ins for the MI355X library, compiled inside
! BQCD's own build (it uses BQCD's modules and defs.h).  Each routine has the
! signature of the reference routine it replaces, so a call site changes only
! its name:
!
!   init_hip_solver(device)    comm_set_gridsize + init_quda  (bqcd.F90:75-76)
!   end_hip_solver()           end_quda                       (bqcd.F90:93)
!   hip_solver(...)            quda_solver / cg               (invert_quda.F90:96-196,
!                                                              cg.F90:51-137)
!   hip_dsf(...)               dsf                            (dsf.F90:28-110)
!   hip_dsd(...)               dsd                            (dsd.F90:28-49)
!   hip_dsg(...)               dsg                            (dsg.F90:28-63)
!   hip_hmc_u(...)             hmc_u                          (hmc_u.F90:28-69)
!   hip_hmc(...)               hmc (one trajectory)           (hmc.F90:28-180)
!
! Conventions kept from the reference:
!  - hip_solver and hip_dsf are called with the links bc-flipped (flip_bc
!    applied by the caller, dsf1.F90:62, dsf2.F90:59, action.F90:50,
!    hmc_init_phi.F90:56); hip_dsf leaves them unflipped, as dsf does
!    (dsf.F90:103).  hip_dsd, hip_dsg and hip_hmc_u see physical links.
!  - cg_para%rest is the absolute bound on |r|^2, cg_para%maxiter and
!    cg_para%log (2 = no convergence tolerated) are honoured, and cg_stat /
!    cg_iterations_total are updated as cg() does.
!  - x is the initial guess on entry (cg.F90:81, use_init_guess in the QUDA
!    path).
! The fields are uploaded on every call, as quda_solver does with
! load_gauge_quda / load_clover_quda; device-resident callers use the _d
! entry points of the C ABI instead.
!
!-------------------------------------------------------------------------------
# include "defs.h"

module module_hip_solver
  use, intrinsic :: iso_c_binding
  use bqcd_hip
  implicit none

  type(c_ptr), save :: hip_ctx = c_null_ptr

  ! inner solver of hip_solver: HB_PREC_F64 runs the FP64 CG with the
  ! recursion of cg(); HB_PREC_F32 / HB_PREC_F16 run the mixed-precision CG
  ! with FP64 reliable updates (same stopping rule, on the FP64 true residual)
  integer(c_int), save :: hip_precision = HB_PREC_F64
  real(c_double), save :: hip_delta = 0.1d0
  integer(c_int), save :: hip_max_inner = 1000

contains

  subroutine hip_check(rc, where)
    integer(c_int), intent(in)   :: rc
    character(len=*), intent(in) :: where
    if (rc /= HB_OK) then
       call die(where // ": " // trim(hb_last_error()))
    endif
  end subroutine

end module module_hip_solver

!-------------------------------------------------------------------------------
subroutine init_hip_solver(device)

  use module_hip_solver
  use module_lattice
  implicit none
  integer, intent(in) :: device
  integer(c_int)      :: lat(4), bc(4)

  if (any(NPE /= 1)) call die("init_hip_solver(): one process per lattice only")
  if (any(gamma_index /= (/1, 2, 3, 4/))) call die("init_hip_solver(): gamma_index must be 1 2 3 4")

  lat = N
  bc = bc_fermions
  call hip_check(hb_init(lat, bc, int(device, c_int), hip_ctx), "init_hip_solver()")
end

!-------------------------------------------------------------------------------
subroutine end_hip_solver()

  use module_hip_solver
  implicit none

  if (c_associated(hip_ctx)) call hip_check(hb_finalize(hip_ctx), "end_hip_solver()")
  hip_ctx = c_null_ptr
end

!-------------------------------------------------------------------------------
subroutine hip_upload(conf, para, flipped, with_b)

  ! links (bc-flipped or physical) and, when the clover term is on, A/2 and
  ! A^-1/2 (plus the LDL^+ factors when with_b) as clover_init left them

  use module_hip_solver
  use module_vol
  use typedef_hmc
  implicit none
  type(hmc_conf), intent(in) :: conf
  type(hmc_para), intent(in) :: para
  integer, intent(in)        :: flipped
  logical, intent(in)        :: with_b

  if (.not. c_associated(hip_ctx)) call die("hip_upload(): init_hip_solver() not called")
  call hip_check(hb_load_gauge(hip_ctx, conf%u, int(volh_tot, c_int), int(flipped, c_int)), "hb_load_gauge")
  if (para%csw_kappa /= ZERO) then
     if (with_b) then
        call hip_check(hb_load_clover(hip_ctx, conf%a, conf%i, conf%b), "hb_load_clover")
     else
        call hip_check(hb_load_clover(hip_ctx, conf%a, conf%i), "hb_load_clover")
     endif
  endif
end

!-------------------------------------------------------------------------------
subroutine hip_solver(matrix_mult, x, b, para, conf, iterations, rho)

  ! solves (M~+rho)^+ (M~+rho) x = b; rho = 0 is mtdagmt, rho /= 0 is
  ! w_dagger_w, the two operators of the five call sites (action.F90:55,
  ! dsf1.F90:66, dsf2.F90:64, hmc_init_phi.F90:68, traces.F90:142)

  use module_hip_solver
  use module_cg
  use module_vol
  use typedef_hmc
  implicit none

  external                          :: matrix_mult
  SPINCOL_OVERINDEXED,  intent(inout) :: x
  SPINCOL_OVERINDEXED,  intent(in)  :: b
  type(hmc_para),       intent(in)  :: para
  type(hmc_conf),       intent(in)  :: conf
  integer,              intent(out) :: iterations
  REAL,                 intent(in)  :: rho

  integer(c_int) :: op, iters, nrel, rc
  real(c_double) :: true_rtr
  character(72)  :: msg

  TIMING_START(timing_bin_cg)

  if (para%h /= ZERO) call die("hip_solver(): h /= 0 is not supported")

  call hip_upload(conf, para, 1, .false.)

  op = HB_OP_MTDAGMT
  if (rho /= ZERO) op = HB_OP_WDAGW

  ! log = 2 here; the convergence decision below follows cg()
  if (hip_precision == HB_PREC_F64) then
     rc = hb_cg(hip_ctx, op, real(para%kappa, c_double), real(rho, c_double), x, b, &
                real(cg_para%rest, c_float), int(cg_para%maxiter, c_int), 2_c_int, iters, true_rtr)
  else
     rc = hb_cg_mixed(hip_ctx, op, real(para%kappa, c_double), real(rho, c_double), x, b, &
                      real(cg_para%rest, c_float), int(cg_para%maxiter, c_int), 2_c_int, &
                      hip_precision, hip_delta, hip_max_inner, iters, nrel, true_rtr)
  endif
  call hip_check(rc, "hip_solver()")

  if (true_rtr > cg_para%rest .and. cg_para%log /= 2) then
     write(msg, *) "cg(): no convergence; rtr = ", true_rtr
     call die(msg)
  endif

  cg_stat%ncall = cg_stat%ncall + 1
  cg_stat%niter = iters
  cg_stat%niter_max = max(cg_stat%niter_max, int(iters))
  cg_stat%niter_tot = cg_stat%niter_tot + iters
  cg_iterations_total = cg_iterations_total + iters

  iterations = iters

  TIMING_STOP(timing_bin_cg)
end

!-------------------------------------------------------------------------------
subroutine hip_dsf(p, conf, step, para, a, b)

  use module_hip_solver
  use module_vol
  use typedef_hmc
  implicit none

  type(hmc_para),  intent(in)    :: para
  type(hmc_conf),  intent(in)    :: conf
  GENERATOR_FIELD, intent(inout) :: p
  REAL,            intent(in)    :: step
  SPINCOL_FIELD,   intent(in)    :: a
  SPINCOL_FIELD,   intent(in)    :: b

  if (para%h /= ZERO) call die("hip_dsf(): h /= 0 is not supported")

  call hip_upload(conf, para, 1, .true.)
  call hip_check(hb_dsf(hip_ctx, p, int(volh_tot, c_int), real(step, c_double), &
                        real(para%kappa, c_double), real(para%csw_kappa, c_double), a, b), "hip_dsf()")

  ! dsf leaves the links physical (dsf.F90:103)
  call flip_bc(conf%u)
end

!-------------------------------------------------------------------------------
subroutine hip_dsd(p, conf, step, para)

  use module_hip_solver
  use module_hmc_forces
  use module_vol
  use typedef_hmc
  implicit none

  type(hmc_para),  intent(in)    :: para
  type(hmc_conf),  intent(in)    :: conf
  GENERATOR_FIELD, intent(inout) :: p
  REAL,            intent(in)    :: step

  if (para%csw_kappa == ZERO .or. step == ZERO) return

  call hmc_forces_old(p)
  call hip_upload(conf, para, 0, .true.)
  call hip_check(hb_dsd(hip_ctx, p, int(volh_tot, c_int), real(step, c_double), &
                        real(para%csw_kappa, c_double)), "hip_dsd()")
  call hmc_forces_new(p, step, i_sd)
end

!-------------------------------------------------------------------------------
subroutine hip_dsg(p, u, step, beta)

  use module_hip_solver
  use module_hmc_forces
  use module_vol
  implicit none

  GENERATOR_FIELD, intent(inout) :: p
  GAUGE_FIELD,     intent(in)    :: u
  REAL,            intent(in)    :: step, beta

  if (beta == ZERO) return
  if (.not. c_associated(hip_ctx)) call die("hip_dsg(): init_hip_solver() not called")

  TIMING_START(timing_bin_dsg)
  call hmc_forces_old(p)
  call hip_check(hb_load_gauge(hip_ctx, u, int(volh_tot, c_int), 0_c_int), "hb_load_gauge")
  call hip_check(hb_dsg(hip_ctx, p, int(volh_tot, c_int), real(step, c_double), &
                        real(beta, c_double)), "hip_dsg()")
  call hmc_forces_new(p, step, i_sg)
  TIMING_STOP(timing_bin_dsg)
end

!-------------------------------------------------------------------------------
subroutine hip_hmc_u(p, conf, step, para)

  ! U <- exp(i step p) U on the device, reunitarised; the clover term is
  ! rebuilt there and both are copied back into conf, which then holds what
  ! hmc_u + clover_init leave (hmc_u.F90:45-68)

  use module_hip_solver
  use module_switches
  use module_vol
  use typedef_hmc
  implicit none

  GENERATOR_FIELD, intent(in)    :: p
  type(hmc_conf),  intent(inout) :: conf
  REAL,            intent(in)    :: step
  type(hmc_para),  intent(in)    :: para

  TIMING_START(timing_bin_hmc_u)
  if (.not. c_associated(hip_ctx)) call die("hip_hmc_u(): init_hip_solver() not called")
  call hip_check(hb_load_gauge(hip_ctx, conf%u, int(volh_tot, c_int), 0_c_int), "hb_load_gauge")
  call hip_check(hb_hmc_u(hip_ctx, p, int(volh_tot, c_int), real(step, c_double), 0.0_c_double), &
                 "hip_hmc_u()")
  call hip_check(hb_save_gauge(hip_ctx, conf%u, int(volh_tot, c_int), 0_c_int), "hb_save_gauge")
  TIMING_STOP(timing_bin_hmc_u)

  TIMING_START(timing_bin_hmc_xbound_g)
  call xbound_g_field(conf%u)
  TIMING_STOP(timing_bin_hmc_xbound_g)

  ! clover_init on the updated links, also for csw_kappa = 0 (A = 1)
  if (switches%clover) then
     call hip_check(hb_build_clover(hip_ctx, real(para%csw_kappa, c_double)), "hb_build_clover")
     call hip_check(hb_save_clover(hip_ctx, conf%a, conf%i, conf%b), "hb_save_clover")
  endif
end

!-------------------------------------------------------------------------------
subroutine hip_hmc(para, conf, out, force_accept, test)

  ! one trajectory of hmc() (hmc.F90:28-180) on the device: the momenta and
  ! pseudofermion noise are drawn here with BQCD's own generator in hmc()'s
  ! order (hmc_init_p, then hmc_init_phi's ran_gauss_volh), the acceptance
  ! number likewise after them, so with the same seed this reproduces hmc();
  ! the leapfrog and force groups follow para%model.  conf%u and the clover
  ! term come back updated (or unchanged on rejection); conf%phi is not set.
  !
  ! test = HMC_TEST_FORWARDS / HMC_TEST_BACKWARDS are hmc()'s reversibility
  ! modes as hmc_test (hmc_test.F90:28-50) drives them: the forward call draws
  ! and integrates as usual; the backward call (para%tau negated by the
  ! caller) draws nothing and integrates from the final momenta of the forward
  ! call (kept in the saved momentum field, as hmc() keeps its p) with the
  ! pseudofermion fields of that call; both are force-accepted and both report
  ! through hmc_test_report, as hmc() does.
  !
  ! cg_para%log is honoured (2: solves that reach maxiter are accepted, as in
  ! BQCD's benchmark inputs), and the trajectory's solves are entered in
  ! cg_stat, cg_iterations_total and the CG timing bin as cg() would.

  use module_hip_solver
  use module_cg
  use module_mre
  use module_p_interface
  use module_switches
  use module_vol
  use typedef_hmc
#ifdef TIMING
  use module_timing_meas
#endif
  implicit none

  type(hmc_para), intent(in)    :: para
  type(hmc_conf), intent(inout) :: conf
  type(hmc_out),  intent(out)   :: out
  integer,        intent(in)    :: force_accept
  integer,        intent(in)    :: test

  P_GENERATOR_FIELD, save :: p, p_start
  P_GAUGE_FIELD, save     :: u_start
  P_SPINCOL_FIELD, save   :: eta1, eta2
  type(hb_hmc_params)     :: prm
  type(hb_hmc_out)        :: o
  real(c_double)          :: r
  integer(c_int)          :: rc
  REAL, external          :: ranf

  TIMING_START(timing_bin_hmc)

  if (para%h /= ZERO) call die("hip_hmc(): h /= 0 is not supported")
  if (.not. c_associated(hip_ctx)) call die("hip_hmc(): init_hip_solver() not called")
  if (test /= HMC_TEST_NONE .and. force_accept == 0) &
       call die("hip_hmc(): a reversibility test is force-accepted (hmc_test.F90:41-47)")

  if (.not. associated(p)) then
     call allocate_gen_field(p)
     call allocate_sc_field(eta1)
     call allocate_sc_field(eta2)
  endif

  if (test /= HMC_TEST_BACKWARDS) then
     call hmc_init_p(p)
     call ran_gauss_volh(NDIRAC * NCOL, eta1, HALF, EVEN)
     if (switches%hasenbusch) call ran_gauss_volh(NDIRAC * NCOL, eta2, HALF, EVEN)
  endif

  r = -1.0_c_double
  if (force_accept == 0) r = ranf()

  prm%beta = para%beta
  prm%kappa = para%kappa
  prm%csw = para%csw
  if (.not. switches%clover) prm%csw = 0.0_c_double
  prm%rho = para%rho
  prm%tau = para%tau
  prm%ntau = para%ntau
  prm%m_scale = para%m_scale
  prm%model = para%model
  prm%integrator = HB_INT_LEAPFROG
  prm%lambda = 0.0_c_double
  prm%rest = real(cg_para%rest, c_float)
  prm%maxiter = cg_para%maxiter
  prm%log = cg_para%log
  prm%precision = hip_precision
  prm%delta = hip_delta
  prm%max_inner = hip_max_inner
  prm%mre_vectors = mre_n_vec
  prm%batch = 0
  prm%verbose = 0

  call hip_check(hb_load_gauge(hip_ctx, conf%u, int(volh_tot, c_int), 0_c_int), "hb_load_gauge")
  if (test == HMC_TEST_NONE) then
     if (switches%hasenbusch) then
        rc = hb_hmc_trajectory(hip_ctx, prm, p, int(volh_tot, c_int), eta1, eta2, r, o)
     else
        rc = hb_hmc_trajectory(hip_ctx, prm, p, int(volh_tot, c_int), eta1, accept_random=r, out=o)
     endif
  else if (test == HMC_TEST_FORWARDS) then
     if (.not. associated(p_start)) then
        call allocate_gen_field(p_start)
        call allocate_g_field(u_start)
     endif
     p_start = p
     u_start = conf%u
     if (switches%hasenbusch) then
        rc = hb_hmc_test(hip_ctx, prm, 1_c_int, p, int(volh_tot, c_int), eta1, eta2, o)
     else
        rc = hb_hmc_test(hip_ctx, prm, 1_c_int, p, int(volh_tot, c_int), eta1, out=o)
     endif
  else
     ! hmc_test negated para%tau; the library negates the step of a backward call itself
     prm%tau = -para%tau
     rc = hb_hmc_test(hip_ctx, prm, -1_c_int, p, int(volh_tot, c_int), out=o)
  endif
  call hip_check(rc, "hip_hmc()")

  call hip_check(hb_save_gauge(hip_ctx, conf%u, int(volh_tot, c_int), 0_c_int), "hb_save_gauge")
  call xbound_g_field(conf%u)
  if (switches%clover) then
     call hip_check(hb_save_clover(hip_ctx, conf%a, conf%i, conf%b), "hb_save_clover")
  endif

  ! hmc.F90:145-147 (H with beta), hmc.F90:131-134; the report of hmc_test.F90
  if (test == HMC_TEST_FORWARDS) then
     call hmc_test_report(test, p_start, u_start, o%sp_old, o%sg_old * para%beta, &
                          o%sf1_old, o%sf2_old, o%sd_old)
  else if (test == HMC_TEST_BACKWARDS) then
     call hmc_test_report(test, p, conf%u, o%sp_new, o%sg_new * para%beta, &
                          o%sf1_new, o%sf2_new, o%sd_new)
  endif

  out%exp_dh = o%exp_dh
  out%accepted = o%accepted
  out%sg = o%sg_new
  out%sf = o%sf1_new
  out%cg_ncall = o%cg_ncall
  out%cg_niter_max = o%cg_niter_max
  out%cg_niter_tot = o%cg_niter_tot

  ! cg_stat as hmc() leaves it (init_cg_stat at its start, cg.F90:126-131 per
  ! solve), and the running totals of BQCD's timing report (timing.F90:278-285)
  cg_stat%ncall = o%cg_ncall
  cg_stat%niter_max = o%cg_niter_max
  cg_stat%niter_tot = o%cg_niter_tot
  cg_iterations_total = cg_iterations_total + o%cg_niter_tot
#ifdef TIMING
  meas(timing_bin_cg)%n_call = meas(timing_bin_cg)%n_call + o%cg_ncall
#endif

  TIMING_STOP(timing_bin_hmc)
end
