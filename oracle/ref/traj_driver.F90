!===============================================================================
!
! traj_driver.F90 - runs the reference's own HMC trajectory (hmc.F90: hmc_init_p,
! hmc_init_phi, hmc_leap_frog with the model A/B/C integrator, the Hamiltonian
! before and after) on a start configuration read from a raw file, and writes
! the pieces of H, exp(-Delta H), the final plaquette and the final links.
! Test infrastructure: golden trajectories for the GPU trajectory driver.
!
! stdin, one item per line:
!   Lx Ly Lz Lt
!   beta kappa csw rho
!   tau ntau m_scale model        (tau = trajectory length / ntau, as hmc_para)
!   rest maxiter mre_vectors log   (log = cg_para%log: 2 accepts solves that
!                                   reach maxiter, cg.F90:119-124)
!   seed                           (ranset; 0 = the default seed of ranf.F90)
!   input directory                (u.bin: physical links, complex(3,3,volh,0:1,4))
!   output directory               (info.txt, u_new.bin)
!
!-------------------------------------------------------------------------------
# include "defs.h"

program traj_driver

  use typedef_para
  use typedef_hmc
  use module_vol
  use module_cg
  use module_mre
  use module_switches
  use module_p_interface
  implicit none

  type(type_para) :: para
  type(hmc_para)  :: hp
  type(hmc_conf)  :: conf
  type(hmc_out)   :: out
  integer         :: lat(4), maxiter, ntau, m_scale, nmre, cglog, mu, eo
  integer(8)      :: seed
  REAL            :: kappa, csw, rho, tau, beta, plaq
  real            :: rest
  character(len=1)   :: model
  character(len=512) :: indir, outdir
  COMPLEX, allocatable :: ubuf(:,:,:,:,:)
  REAL, external       :: sg

  read(*,*) lat
  read(*,*) beta, kappa, csw, rho
  read(*,*) tau, ntau, m_scale, model
  read(*,*) rest, maxiter, nmre, cglog
  read(*,*) seed
  read(*,'(a)') indir
  read(*,'(a)') outdir

  para%L = lat
  para%NPE = (/1, 1, 1, 1/)
  para%bc_fermions = (/1, 1, 1, -1/)
  para%gamma_index = (/1, 2, 3, 4/)

  call init_common(para)
  call init_modules()
  call init_flip_bc()
  call init_xbound()
  call init_cg_para(rest, maxiter, cglog)
  call init_cg_stat()
  mre_n_vec = nmre

  switches%quenched = .false.
  switches%dynamical = .true.
  switches%clover = (csw /= ZERO)
  switches%h_ext = .false.
  switches%hasenbusch = (model /= "A")
  switches%tempering = .false.
  switches%hmc_test = .false.

  hp%beta = beta
  hp%kappa = kappa
  hp%csw = csw
  hp%csw_kappa = csw * kappa
  hp%h = ZERO
  hp%rho = rho
  hp%tau = tau
  hp%ntau = ntau
  hp%traj_length = tau * ntau
  hp%m_scale = m_scale
  hp%model = model

  nullify(conf%u, conf%phi, conf%phi2, conf%a, conf%i, conf%b)
  call allocate_g_field(conf%u)
  call allocate_sc_field(conf%phi)
  call allocate_sc_field(conf%phi2)
  if (switches%clover) then
     call allocate_clover_field_a(conf%a)
     call allocate_clover_field_a(conf%i)
     call allocate_clover_field_b(conf%b)
  endif

  allocate(ubuf(NCOL, NCOL, volh, EVEN:ODD, DIM))
  call rd(trim(indir)//'/u.bin', ubuf, 2 * size(ubuf))
  conf%u = ZERO
  do mu = 1, DIM
     do eo = EVEN, ODD
        conf%u(:, :, 1:volh, eo, mu) = ubuf(:, :, 1:volh, eo, mu)
     enddo
  enddo
  call xbound_g_field(conf%u)
  if (switches%clover) call clover_init(conf%a, conf%i, conf%b, conf%u, hp%csw_kappa)

  call ranset(seed, 0_8)

  call hmc(hp, conf, out, 1, 0)          ! force_accept, no reversibility test

  plaq = sg(conf%u)
  open(20, file=trim(outdir)//'/info.txt', status='replace')
  write(20, '(a, es25.17)') 'exp_dh ', out%exp_dh
  write(20, '(a, es25.17)') 'sg_new ', plaq
  write(20, '(a, i10)') 'cg_ncall ', out%cg_ncall
  write(20, '(a, i10)') 'cg_niter_tot ', out%cg_niter_tot
  write(20, '(a, i10)') 'cg_niter_max ', out%cg_niter_max
  call ranget(seed)
  write(20, '(a, i20)') 'seed_after ', seed
  close(20)
  do mu = 1, DIM
     do eo = EVEN, ODD
        ubuf(:, :, 1:volh, eo, mu) = conf%u(:, :, 1:volh, eo, mu)
     enddo
  enddo
  call wr(trim(outdir)//'/u_new.bin', ubuf, 2 * size(ubuf))

end program traj_driver

subroutine rd(fname, a, n)       ! reads n float64 words
  implicit none
  character(len=*) :: fname
  integer :: n
  REAL :: a(*)
  open(30, file=fname, access='stream', form='unformatted', status='old', action='read')
  read(30) a(1:n)
  close(30)
end

subroutine wr(fname, a, n)       ! writes n float64 words
  implicit none
  character(len=*) :: fname
  integer :: n
  REAL :: a(*)
  open(31, file=fname, access='stream', form='unformatted', status='replace', action='write')
  write(31) a(1:n)
  close(31)
end
