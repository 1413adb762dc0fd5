!===============================================================================
!
! confio_driver.F90 - the reference's own configuration-file I/O (conf_write,
! conf_read: conf.F90, conf_info.F90, field_io, cksum.c) on a link field read
! from a raw file.  Test infrastructure: golden configuration files for the
! library's reader, and a reader for files the library writes.
!
! stdin, one item per line:
!   Lx Ly Lz Lt
!   mode                       write | read
!   directory                  write: u.bin in, BQCD files out (cwd);
!                              read:  u_read.bin out
!   info file                  read mode only: the .info file to read
!
!-------------------------------------------------------------------------------
# include "defs.h"

program confio_driver

  use typedef_para
  use typedef_hmc
  use module_vol
  use module_counter
  use module_switches
  use module_p_interface
  implicit none

  type(type_para) :: para
  type(hmc_conf), dimension(MAX_TEMPER) :: conf
  integer         :: lat(4), mu, eo
  character(len=8)   :: mode
  character(len=512) :: dir, info
  COMPLEX, allocatable :: ubuf(:,:,:,:,:)

  read(*,*) lat
  read(*,'(a)') mode
  read(*,'(a)') dir
  if (trim(mode) == 'read') read(*,'(a)') info

  para%run = 1
  para%L = lat
  para%NPE = (/1, 1, 1, 1/)
  para%bc_fermions = (/1, 1, 1, -1/)
  para%gamma_index = (/1, 2, 3, 4/)
  para%n_temper = 1
  para%c_hmc(1)%beta = '5.29'
  para%c_hmc(1)%kappa = '0.1355'
  para%c_hmc(1)%csw = '1.9'
  para%c_hmc(1)%csw_kappa = '0.25745'
  para%c_hmc(1)%h = '0'

  call init_common(para)
  call init_modules()
  call init_flip_bc()
  call init_xbound()

  switches%tempering = .false.
  switches%dynamical = .true.

  counter%run = 1
  counter%job = 1
  counter%traj = 0
  counter%j_traj = 0

  nullify(conf(1)%u, conf(1)%phi, conf(1)%phi2, conf(1)%a, conf(1)%i, conf(1)%b)
  call allocate_g_field(conf(1)%u)
  conf(1)%u = ZERO
  conf(1)%former = 1
  allocate(ubuf(NCOL, NCOL, volh, EVEN:ODD, DIM))

  if (trim(mode) == 'write') then
     call rd(trim(dir)//'/u.bin', ubuf, 2 * size(ubuf))
     do mu = 1, DIM
        do eo = EVEN, ODD
           conf(1)%u(:, :, 1:volh, eo, mu) = ubuf(:, :, 1:volh, eo, mu)
        enddo
     enddo
     call conf_write(.false., para, conf)
  else
     para%info_file(1) = info
     call conf_read(.false., para, conf)
     do mu = 1, DIM
        do eo = EVEN, ODD
           ubuf(:, :, 1:volh, eo, mu) = conf(1)%u(:, :, 1:volh, eo, mu)
        enddo
     enddo
     call wr(trim(dir)//'/u_read.bin', ubuf, 2 * size(ubuf))
  endif

end program confio_driver

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
