!===============================================================================
!
! dropin_driver.F90 - the MI355X drop-ins of hip_solver.F90 called from BQCD
! itself, next to the reference routines they replace, on the same fields in
! the same process.  Test infrastructure: links the reference CPU path (built
! from source by the Makefile next to this file) with libbqcd_hip.so and runs
! on a GPU host.
!
! stdin (as golden_driver): Lx Ly Lz Lt / kappa csw rho / rest maxiter /
! step beta / input directory (u.bin, eta_e.bin, mom.bin) / (unused) / 0
!
! The last comparisons are whole trajectories: hmc() against hip_hmc() after
! the same ranset (model C with rho, model A without), then the same with
! every CG stopped after a few iterations (cg_para%log = 2, the reference's
! benchmark mode), then hmc_test's forward / backward pair
! (HMC_TEST_FORWARDS, then HMC_TEST_BACKWARDS with tau negated) with each.
! BQCD's CG statistics (cg_stat, cg_iterations_total) must advance alike.
!
! Prints one "DROPIN name value" line per comparison (iteration counts, and
! max |hip - reference| / max |reference| of each result) and exits with
! status 1 if any comparison fails its bound.
!
!-------------------------------------------------------------------------------
# include "defs.h"

program dropin_driver

  use typedef_para
  use typedef_hmc
  use module_vol
  use module_cg
  use module_switches
  use module_p_interface
  use module_hip_solver
  use module_mre
  implicit none

  type(type_para) :: para
  type(hmc_para)  :: hp
  type(hmc_conf)  :: conf, conf2
  type(hmc_out)   :: out_ref, out_hip
  integer         :: lat(4), maxiter, bench_iters, it_ref, it_hip, mu, eo, nfail
  integer         :: tot_ref, tot_hip
  REAL            :: kappa, csw, rho, step, beta
  real            :: rest
  character(len=512) :: indir, outdir
  COMPLEX, allocatable :: eta_e(:,:,:), x_ref(:,:,:), x_hip(:,:,:), tmp(:,:,:)
  COMPLEX, allocatable :: ubuf(:,:,:,:,:)
  REAL, allocatable    :: pbuf(:,:,:,:)
  P_GENERATOR_FIELD    :: p_ref, p_hip, mom
  external             :: mtdagmt, w_dagger_w

  read(*,*) lat
  read(*,*) kappa, csw, rho
  read(*,*) rest, maxiter
  read(*,*) step, beta
  read(*,'(a)') indir
  read(*,'(a)') outdir
  read(*,*) bench_iters

  para%L = lat
  para%NPE = (/1, 1, 1, 1/)
  para%bc_fermions = (/1, 1, 1, -1/)
  para%gamma_index = (/1, 2, 3, 4/)

  call init_common(para)
  call init_modules()
  call init_flip_bc()
  call init_xbound()
  call init_cg_para(rest, maxiter, 1)
  call init_cg_stat()
  mre_n_vec = 0

  switches%clover = .true.
  switches%hasenbusch = .false.

  hp%beta = beta
  hp%kappa = kappa
  hp%csw = csw
  hp%csw_kappa = csw * kappa
  hp%h = ZERO
  hp%rho = rho

  nullify(conf%u, conf%a, conf%i, conf%b, conf2%u, conf2%a, conf2%i, conf2%b)
  nullify(conf%phi, conf%phi2, conf2%phi, conf2%phi2)
  nullify(p_ref, p_hip, mom)
  call allocate_g_field(conf%u)
  call allocate_clover_field_a(conf%a)
  call allocate_clover_field_a(conf%i)
  call allocate_clover_field_b(conf%b)
  call allocate_g_field(conf2%u)
  call allocate_clover_field_a(conf2%a)
  call allocate_clover_field_a(conf2%i)
  call allocate_clover_field_b(conf2%b)
  call allocate_gen_field(p_ref)
  call allocate_gen_field(p_hip)
  call allocate_gen_field(mom)

  allocate(ubuf(NCOL, NCOL, volh, EVEN:ODD, DIM))
  allocate(pbuf(NGEN, volh, EVEN:ODD, DIM))
  call rd(trim(indir)//'/u.bin', ubuf, 2 * size(ubuf))
  conf%u = ZERO
  do mu = 1, DIM
     do eo = EVEN, ODD
        conf%u(:, :, 1:volh, eo, mu) = ubuf(:, :, 1:volh, eo, mu)
     enddo
  enddo
  call xbound_g_field(conf%u)
  call rd(trim(indir)//'/mom.bin', pbuf, size(pbuf))
  mom = ZERO
  do mu = 1, DIM
     do eo = EVEN, ODD
        mom(:, 1:volh, eo, mu) = pbuf(:, 1:volh, eo, mu)
     enddo
  enddo

  allocate(eta_e(NDIRAC, NCOL, volh_tot), x_ref(NDIRAC, NCOL, volh_tot))
  allocate(x_hip(NDIRAC, NCOL, volh_tot), tmp(NDIRAC, NCOL, volh_tot))
  eta_e = ZERO ; x_ref = ZERO ; x_hip = ZERO ; tmp = ZERO
  call rd(trim(indir)//'/eta_e.bin', eta_e, 24 * volh)

  call clover_init(conf%a, conf%i, conf%b, conf%u, hp%csw_kappa)

  call init_hip_solver(0)
  nfail = 0

  ! --- solver: M~+M~ (rho = 0) and W+W (rho), links flipped as at the call sites
  call flip_bc(conf%u)

  x_ref = ZERO
  call cg(mtdagmt, x_ref, eta_e, hp, conf, it_ref)
  x_hip = ZERO
  call hip_solver(mtdagmt, x_hip, eta_e, hp, conf, it_hip, ZERO)
  call report_int('mtdagmt_iters_ref', it_ref)
  call report_int('mtdagmt_iters_hip', it_hip)
  if (it_ref /= it_hip) nfail = nfail + 1
  call report(x_hip, x_ref, 24 * volh_tot, 'mtdagmt_x', 1.0e-9_8)

  if (rho /= ZERO) then
     x_ref = ZERO ; x_ref(:, :, 1:volh) = 0.01_8 * eta_e(:, :, 1:volh)
     x_hip = x_ref
     call cg(w_dagger_w, x_ref, eta_e, hp, conf, it_ref)
     call hip_solver(w_dagger_w, x_hip, eta_e, hp, conf, it_hip, rho)
     call report_int('wdw_iters_ref', it_ref)
     call report_int('wdw_iters_hip', it_hip)
     if (it_ref /= it_hip) nfail = nfail + 1
     call report(x_hip, x_ref, 24 * volh_tot, 'wdw_x', 1.0e-9_8)
  endif

  ! --- fermion force, a = solution, b = M~ a; dsf un-flips the links --------
  call mtil(tmp, x_ref, hp, conf)
  call copy_conf(conf2, conf)
  p_ref = ZERO
  call dsf(p_ref, conf, step, hp, x_ref, tmp)
  p_hip = ZERO
  call hip_dsf(p_hip, conf2, step, hp, x_ref, tmp)
  call report(p_hip, p_ref, size(p_ref), 'dsf', 1.0e-12_8)
  call report(conf2%u, conf%u, 2 * size(conf%u), 'dsf_links_unflipped', 0.0_8)

  ! --- determinant and gauge forces (physical links) -----------------------
  p_ref = ZERO
  call dsd(p_ref, conf, step, hp)
  p_hip = ZERO
  call hip_dsd(p_hip, conf, step, hp)
  call report(p_hip, p_ref, size(p_ref), 'dsd', 1.0e-12_8)

  p_ref = ZERO
  call dsg(p_ref, conf%u, step, beta)
  p_hip = ZERO
  call hip_dsg(p_hip, conf%u, step, beta)
  call report(p_hip, p_ref, size(p_ref), 'dsg', 1.0e-12_8)

  ! --- link update + clover rebuild -----------------------------------------
  call copy_conf(conf2, conf)
  call hmc_u(mom, conf, step, hp)
  call hip_hmc_u(mom, conf2, step, hp)
  call report(conf2%u, conf%u, 2 * size(conf%u), 'hmc_u_links', 1.0e-12_8)
  call report(conf2%a, conf%a, 72 * 2 * volh, 'hmc_u_clover_a', 1.0e-12_8)
  call report(conf2%i, conf%i, 72 * 2 * volh, 'hmc_u_clover_i', 1.0e-12_8)
  call report(conf2%b, conf%b, 72 * 2 * volh, 'hmc_u_clover_b', 1.0e-12_8)

  ! --- a whole trajectory: hmc() against hip_hmc(), same seed ----------------
  hp%beta = beta
  hp%rho = rho
  hp%tau = 0.05_8
  hp%ntau = 4
  hp%m_scale = 2
  hp%model = 'C'
  if (rho == ZERO) hp%model = 'A'
  switches%dynamical = .true.
  switches%hasenbusch = (hp%model /= 'A')
  call allocate_sc_field(conf%phi)
  call allocate_sc_field(conf%phi2)
  call init_cg_para(rest, maxiter, 0)
  call copy_conf(conf2, conf)
  call init_cg_stat()
  call ranset(4711_8, 0_8)
  tot_ref = cg_iterations_total
  call hmc(hp, conf, out_ref, 1, 0)
  tot_ref = cg_iterations_total - tot_ref
  call ranset(4711_8, 0_8)
  tot_hip = cg_iterations_total
  call hip_hmc(hp, conf2, out_hip, 1, 0)
  tot_hip = cg_iterations_total - tot_hip
  call compare_traj('hmc', 1.0e-9_8, 1.0e-10_8)
  call report_int('hmc_cg_total_ref', tot_ref)
  call report_int('hmc_cg_total_hip', tot_hip)
  if (tot_ref /= tot_hip .or. tot_ref /= out_ref%cg_niter_tot) nfail = nfail + 1

  ! --- the benchmark mode: every CG stopped after 12 iterations (log = 2) ----
  call init_cg_para(0.0, 12, 2)
  call copy_conf(conf2, conf)
  call ranset(4712_8, 0_8)
  call hmc(hp, conf, out_ref, 1, 0)
  call ranset(4712_8, 0_8)
  call hip_hmc(hp, conf2, out_hip, 1, 0)
  call compare_traj('hmc_fixed12', 1.0e-9_8, 1.0e-10_8)
  if (out_hip%cg_niter_max /= 12 .or. out_hip%cg_niter_tot /= 12 * out_hip%cg_ncall) nfail = nfail + 1
  call init_cg_para(rest, maxiter, 0)

  ! --- hmc_test.F90:28-50: forward, then backward with tau negated -----------
  call copy_conf(conf2, conf)
  call ranset(4713_8, 0_8)
  call hmc(hp, conf, out_ref, 1, HMC_TEST_FORWARDS)
  call ranset(4713_8, 0_8)
  call hip_hmc(hp, conf2, out_hip, 1, HMC_TEST_FORWARDS)
  call compare_traj('hmc_test_forward', 1.0e-9_8, 1.0e-10_8)
  hp%tau = -hp%tau
  call hmc(hp, conf, out_ref, 1, HMC_TEST_BACKWARDS)
  call hip_hmc(hp, conf2, out_hip, 1, HMC_TEST_BACKWARDS)
  hp%tau = -hp%tau
  call report_real('hmc_test_backward_sg_ref', out_ref%sg)
  call report_real('hmc_test_backward_sg_hip', out_hip%sg)
  if (.not. abs(out_hip%sg - out_ref%sg) <= 1.0e-10_8 * abs(out_ref%sg)) nfail = nfail + 1
  call report_int('hmc_test_backward_cg_iters_ref', out_ref%cg_niter_tot)
  call report_int('hmc_test_backward_cg_iters_hip', out_hip%cg_niter_tot)
  if (out_ref%cg_niter_tot /= out_hip%cg_niter_tot) nfail = nfail + 1
  call report(conf2%u, conf%u, 2 * size(conf%u), 'hmc_test_backward_links', 1.0e-10_8)

  call end_hip_solver()

  write(*, '(a, i4)') 'DROPIN failures ', nfail
  if (nfail /= 0) stop 1

CONTAINS

  ! a trajectory of hip_hmc (conf2, out_hip) against hmc() (conf, out_ref)
  subroutine compare_traj(name, tol_dh, tol_u)
    character(len=*) :: name
    real(8) :: tol_dh, tol_u
    call report_int(name // '_cg_iters_ref', out_ref%cg_niter_tot)
    call report_int(name // '_cg_iters_hip', out_hip%cg_niter_tot)
    if (out_ref%cg_niter_tot /= out_hip%cg_niter_tot) nfail = nfail + 1
    if (out_ref%cg_ncall /= out_hip%cg_ncall) nfail = nfail + 1
    call report_real(name // '_exp_dh_ref', out_ref%exp_dh)
    call report_real(name // '_exp_dh_hip', out_hip%exp_dh)
    if (.not. abs(out_hip%exp_dh - out_ref%exp_dh) <= tol_dh * abs(out_ref%exp_dh)) nfail = nfail + 1
    call report(conf2%u, conf%u, 2 * size(conf%u), name // '_links', tol_u)
  end subroutine compare_traj

  subroutine copy_conf(dst, src)
    type(hmc_conf) :: dst, src
    dst%u = src%u
    dst%a = src%a
    dst%i = src%i
    dst%b = src%b
  end subroutine copy_conf

  subroutine report_real(name, v)
    character(len=*) :: name
    real(8) :: v
    write(*, '(a, a, a, es25.16)') 'DROPIN ', name, ' ', v
  end subroutine report_real

  subroutine report_int(name, v)
    character(len=*) :: name
    integer :: v
    write(*, '(a, a, a, i10)') 'DROPIN ', name, ' ', v
  end subroutine report_int

  ! max |h - r| / max |r| over n float64 words; fails above tol
  subroutine report(h, r, n, name, tol)
    integer :: n
    character(len=*) :: name
    real(8) :: tol
    type(*), target :: h(*), r(*)
    real(8), pointer :: hh(:), rr(:)
    real(8) :: d, m
    call c_f_pointer(c_loc(h), hh, [n])
    call c_f_pointer(c_loc(r), rr, [n])
    d = maxval(abs(hh - rr))
    m = maxval(abs(rr))
    if (m > 0) d = d / m
    write(*, '(a, a, a, es12.4)') 'DROPIN ', name, ' ', d
    if (.not. (d <= tol)) nfail = nfail + 1
  end subroutine report

end program dropin_driver

subroutine rd(fname, a, n)       ! reads n float64 words
  implicit none
  character(len=*) :: fname
  integer :: n
  REAL :: a(*)
  open(30, file=fname, access='stream', form='unformatted', status='old', &
       action='read')
  read(30) a(1:n)
  close(30)
end
