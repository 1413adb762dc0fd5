!===============================================================================
!
! golden_driver.F90 - drives the reference BQCD CPU path (single_pe, libd21)
! on fields read from raw files and writes every hot-path result as raw
! little-endian float64 arrays.  This program is test infrastructure: it runs
! only in the build container (and, prebuilt, as the CPU timing baseline).
!
! Control is read from stdin, one item per line:
!   Lx Ly Lz Lt
!   kappa csw rho
!   rest maxiter                 (rest: float32 absolute |r|^2 bound, as cg())
!   step beta
!   input directory              (u.bin, eta_e.bin, eta_o.bin, mom.bin)
!   output directory
!   bench_iters                  (0: golden mode; >0: time that many CG iters)
!
! Field file layouts follow BQCD, packed to volh sites:
!   u.bin    complex(3,3,volh,0:1,4)  physical links (no boundary sign)
!   eta_*.bin complex(4,3,volh)
!   mom.bin  real(8,volh,0:1,4)
!
!-------------------------------------------------------------------------------
# include "defs.h"

program golden_driver

  use typedef_para
  use typedef_hmc
  use module_vol
  use module_cg
  use module_switches
  use module_p_interface
  implicit none

  type(type_para) :: para
  type(hmc_para)  :: hp
  type(hmc_conf)  :: conf
  integer         :: lat(4), maxiter, bench_iters, iters, mu, eo, i
  REAL            :: kappa, csw, rho, step, beta, s, t0, t1
  real            :: rest
  character(len=512) :: indir, outdir
  COMPLEX, allocatable :: eta_e(:,:,:), eta_o(:,:,:), out(:,:,:), tmp(:,:,:), x(:,:,:)
  COMPLEX, allocatable :: ubuf(:,:,:,:,:)
  REAL, allocatable    :: pbuf(:,:,:,:)
  P_GENERATOR_FIELD    :: p
  REAL, external       :: clover_action, sc_norm2, global_sum, sc_dot
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
  call init_cg_para(rest, maxiter, 2)
  call init_cg_stat()

  switches%clover = .true.
  switches%hasenbusch = .false.

  hp%beta = beta
  hp%kappa = kappa
  hp%csw = csw
  hp%csw_kappa = csw * kappa
  hp%h = ZERO
  hp%rho = rho

  nullify(conf%u, conf%a, conf%i, conf%b, p)
  call allocate_g_field(conf%u)
  call allocate_clover_field_a(conf%a)
  call allocate_clover_field_a(conf%i)
  call allocate_clover_field_b(conf%b)
  call allocate_gen_field(p)

  allocate(ubuf(NCOL, NCOL, volh, EVEN:ODD, DIM))
  allocate(pbuf(NGEN, volh, EVEN:ODD, DIM))
  call read_raw(trim(indir)//'/u.bin', ubuf, 2 * size(ubuf))
  do mu = 1, DIM
     do eo = EVEN, ODD
        conf%u(:, :, 1:volh, eo, mu) = ubuf(:, :, 1:volh, eo, mu)
     enddo
  enddo

  allocate(eta_e(NDIRAC, NCOL, volh_tot), eta_o(NDIRAC, NCOL, volh_tot))
  allocate(out(NDIRAC, NCOL, volh_tot), tmp(NDIRAC, NCOL, volh_tot))
  allocate(x(NDIRAC, NCOL, volh_tot))
  eta_e = ZERO ; eta_o = ZERO ; out = ZERO ; tmp = ZERO ; x = ZERO
  call read_raw(trim(indir)//'/eta_e.bin', eta_e, 24 * volh)
  call read_raw(trim(indir)//'/eta_o.bin', eta_o, 24 * volh)

  ! clover term from physical links
  call clover_init(conf%a, conf%i, conf%b, conf%u, hp%csw_kappa)

  ! solver-side convention: antiperiodic time boundary folded into the links
  call flip_bc(conf%u)

  if (bench_iters > 0) then
     call bench()
     stop
  endif

  call write_raw(trim(outdir)//'/clover_a.bin', conf%a, 72 * volh * 2)
  call write_raw(trim(outdir)//'/clover_i.bin', conf%i, 72 * volh * 2)
  call write_raw(trim(outdir)//'/clover_b.bin', conf%b, 72 * volh * 2)

  ! --- hopping matrix -------------------------------------------------------
  call d(EVEN, ODD, out, eta_o, conf%u)
  call write_sc(trim(outdir)//'/d_eo.bin', out)
  call d(ODD, EVEN, out, eta_e, conf%u)
  call write_sc(trim(outdir)//'/d_oe.bin', out)
  call d_dag(EVEN, ODD, out, eta_o, conf%u)
  call write_sc(trim(outdir)//'/ddag_eo.bin', out)
  call d_dag(ODD, EVEN, out, eta_e, conf%u)
  call write_sc(trim(outdir)//'/ddag_oe.bin', out)

  ! --- clover multiplies ----------------------------------------------------
  call clover_mult_a(out, conf%a(1, 1, EVEN), eta_e, volh)
  call write_sc(trim(outdir)//'/cma_e.bin', out)
  call clover_mult_a(out, conf%a(1, 1, ODD), eta_o, volh)
  call write_sc(trim(outdir)//'/cma_o.bin', out)
  out = eta_o
  call clover_mult_ao(conf%i(1, 1, ODD), out, volh)
  call write_sc(trim(outdir)//'/cmao_o.bin', out)
  out = eta_e
  call clover_mult_ao(conf%i(1, 1, EVEN), out, volh)
  call write_sc(trim(outdir)//'/cmao_e.bin', out)
  out = eta_o
  call clover_mult_b(conf%b(1, 1, ODD), out, volh)
  call write_sc(trim(outdir)//'/cmb_o.bin', out)

  ! --- even-odd operators ---------------------------------------------------
  call mtil(out, eta_e, hp, conf)
  call write_sc(trim(outdir)//'/mtil.bin', out)
  call mtil_dag(out, eta_e, hp, conf)
  call write_sc(trim(outdir)//'/mtil_dag.bin', out)
  call mtdagmt(out, eta_e, hp, conf)
  call write_sc(trim(outdir)//'/mtdagmt.bin', out)
  call w_mult(out, eta_e, hp, conf)
  call write_sc(trim(outdir)//'/w_mult.bin', out)
  call w_dagger_w(out, eta_e, hp, conf)
  call write_sc(trim(outdir)//'/wdw.bin', out)

  ! --- CG on M~+M~ from a zero guess, and on W+W from guess eta_o-copy --------
  open(20, file=trim(outdir)//'/info.txt', status='replace')
  x = ZERO
  call cg(mtdagmt, x, eta_e, hp, conf, iters)
  call write_sc(trim(outdir)//'/cg_x.bin', x)
  call mtdagmt(out, x, hp, conf)
  out(:, :, 1:volh) = eta_e(:, :, 1:volh) - out(:, :, 1:volh)
  s = sc_norm2(out)
  write(20, '(a, i8)') 'cg_iters ', iters
  write(20, '(a, es25.17)') 'cg_true_rtr ', s

  x = ZERO
  x(:, :, 1:volh) = 0.01_8 * eta_o(:, :, 1:volh)
  call cg(w_dagger_w, x, eta_e, hp, conf, iters)
  call write_sc(trim(outdir)//'/cgw_x.bin', x)
  call w_dagger_w(out, x, hp, conf)
  out(:, :, 1:volh) = eta_e(:, :, 1:volh) - out(:, :, 1:volh)
  s = sc_norm2(out)
  write(20, '(a, i8)') 'cgw_iters ', iters
  write(20, '(a, es25.17)') 'cgw_true_rtr ', s

  ! --- fermion force with a = eta_e, b = M~ eta_e (links flipped on entry) ----
  call mtil(tmp, eta_e, hp, conf)
  call zero_gen(p)
  call dsf(p, conf, step, hp, eta_e, tmp)      ! dsf un-flips the links
  call write_gen(trim(outdir)//'/dsf.bin', p)

  ! --- clover determinant force and gauge force (physical links) ------------
  call zero_gen(p)
  call dsd(p, conf, step, hp)
  call write_gen(trim(outdir)//'/dsd.bin', p)

  call zero_gen(p)
  call dsg(p, conf%u, step, beta)
  call write_gen(trim(outdir)//'/dsg.bin', p)

  s = clover_action(conf%b(1, 1, ODD))
  write(20, '(a, es25.17)') 'clover_action ', s

  ! --- link update with momenta from mom.bin --------------------------------
  call read_raw(trim(indir)//'/mom.bin', pbuf, size(pbuf))
  do mu = 1, DIM
     do eo = EVEN, ODD
        p(:, 1:volh, eo, mu) = pbuf(:, 1:volh, eo, mu)
     enddo
  enddo
  call hmc_u(p, conf, step, hp)
  do mu = 1, DIM
     do eo = EVEN, ODD
        ubuf(:, :, 1:volh, eo, mu) = conf%u(:, :, 1:volh, eo, mu)
     enddo
  enddo
  call write_raw(trim(outdir)//'/u_new.bin', ubuf, 2 * size(ubuf))
  call write_raw(trim(outdir)//'/clover_a_new.bin', conf%a, 72 * volh * 2)

  close(20)

CONTAINS

  subroutine bench()
    ! Times the reference CG (fixed iteration count) and hopping matrix calls.
    integer :: k, c0, c1, rate
    x = ZERO
    call init_cg_para(1.0e-30, bench_iters, 2)
    call system_clock(c0, rate)
    call cg(mtdagmt, x, eta_e, hp, conf, iters)
    call system_clock(c1)
    write(*, '(a, i8, a, es16.8)') 'cg_iters ', iters, ' cg_seconds ', &
         real(c1 - c0, 8) / real(rate, 8)
    call system_clock(c0)
    do k = 1, 4 * bench_iters
       call d(EVEN, ODD, out, eta_o, conf%u)
    enddo
    call system_clock(c1)
    write(*, '(a, i8, a, es16.8)') 'd_calls ', 4 * bench_iters, ' d_seconds ', &
         real(c1 - c0, 8) / real(rate, 8)
  end subroutine bench

  subroutine zero_gen(q)
    P_GENERATOR_FIELD :: q
    q = ZERO
  end subroutine zero_gen

  subroutine write_gen(fname, q)
    character(len=*) :: fname
    P_GENERATOR_FIELD :: q
    do mu = 1, DIM
       do eo = EVEN, ODD
          pbuf(:, :, eo, mu) = q(:, 1:volh, eo, mu)
       enddo
    enddo
    call write_raw(fname, pbuf, size(pbuf))
  end subroutine write_gen

  subroutine write_sc(fname, f)
    character(len=*) :: fname
    COMPLEX :: f(NDIRAC, NCOL, volh_tot)
    call write_raw(fname, f, 24 * volh)
  end subroutine write_sc

end program golden_driver

subroutine read_raw(fname, a, n)       ! reads n float64 words
  implicit none
  character(len=*) :: fname
  integer :: n
  REAL :: a(*)
  open(30, file=fname, access='stream', form='unformatted', status='old', &
       action='read')
  read(30) a(1:n)
  close(30)
end

subroutine write_raw(fname, a, n)      ! writes n float64 words
  implicit none
  character(len=*) :: fname
  integer :: n
  REAL :: a(*)
  open(31, file=fname, access='stream', form='unformatted', status='replace', &
       action='write')
  write(31) a(1:n)
  close(31)
end
