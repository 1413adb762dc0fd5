!===============================================================================
!
! hip_plugins.F90 - the routines BQCD links from its plugin libraries, with
! the names and arguments of the reference, computed by the MI355X library:
!
!   LIBD       (d/Makefile:113-129)    d, d_dag, version_of_d,
!                                      dsf_{x,y,z,t}{f,b}
!   LIBCLOVER  (clover/Makefile:44-71) clover_init, clover_mult_a,
!                                      clover_mult_ao, clover_mult_b,
!                                      clover_dsf, clover_dsd, clover_action,
!                                      allocate_clover_field_a/b
!
! BQCD selects them like its own libraries (Makefile:97, LIBS = d/$(LIBD)
! ... clover/$(LIBCLOVER)): an archive of this file and bqcd_hip.F90 takes
! the place of libd21.a and libclover.a, and libbqcd_hip.so is linked.
!
! Every call receives host arrays, uploads its operands (the whole gauge
! field for d, d_dag, clover_init, clover_dsf and clover_dsd) and downloads its
! result; the solver-level drop-ins of hip_solver.F90 keep the fields on the
! device instead.  The routines share one context of their own with periodic
! boundaries: links are used exactly as passed, so a boundary sign that
! flip_bc folded into them acts as in the reference.
!
!-------------------------------------------------------------------------------
# include "defs.h"

module module_hip_plugins

  use, intrinsic :: iso_c_binding
  use bqcd_hip
  implicit none
  private
  public :: plugin_ctx, plugin_device, plugin_check, plugin_init, plugin_links, plugin_dsf

  type(c_ptr), save    :: plugin_ctx = c_null_ptr
  integer(c_int), save :: plugin_device = 0

contains

  subroutine plugin_check(rc, where)
    integer(c_int), intent(in)   :: rc
    character(len=*), intent(in) :: where
    if (rc /= HB_OK) call die(where // ": " // trim(hb_last_error()))
  end subroutine

  ! the context of the plugins on the first call: the BQCD lattice, periodic boundaries
  subroutine plugin_init()
    use module_lattice
    integer(c_int) :: lat(4), bc(4)
    if (c_associated(plugin_ctx)) return
    if (any(NPE /= 1)) call die("hip plugins: one process per lattice only")
    if (any(gamma_index /= (/1, 2, 3, 4/))) call die("hip plugins: gamma_index must be 1 2 3 4")
    lat = N
    bc = 1
    call plugin_check(hb_init(lat, bc, plugin_device, plugin_ctx), "hip plugins: hb_init")
  end subroutine

  ! the links of u (complex(3,3,volh_tot,0:1,4)) as passed
  subroutine plugin_links(u, where)
    use module_vol
    COMPLEX, dimension(NCOL, NCOL, volh_tot, EVEN:ODD, DIM), intent(in) :: u
    character(len=*), intent(in) :: where
    call plugin_init()
    call plugin_check(hb_load_gauge(plugin_ctx, u, int(volh_tot, c_int), 0_c_int), where)
  end subroutine

  ! dsf_{x,y,z,t}{f,b}: direction mu (0..3), forward or backward
  subroutine plugin_dsf(mu, forward, p, b, a, s, u, nn, nv, where)
    use module_vol
    integer, intent(in)  :: mu, forward, nv
    REAL, dimension(NGEN, *), intent(inout)          :: p
    COMPLEX, dimension(NDIRAC, NCOL, *), intent(in)  :: b, a
    REAL, intent(in)                                 :: s
    COMPLEX, dimension(NCOL, NCOL, *), intent(in)    :: u
    integer, intent(in)                              :: nn(*)
    character(len=*), intent(in)                     :: where
    call plugin_init()
    if (nv /= volh) call die(where // ": volh differs from the lattice's")
    call plugin_check(hb_dsf_dir(plugin_ctx, int(mu, c_int), int(forward, c_int), p, b, a, &
                                 real(s, c_double), u, nn, int(volh_tot, c_int)), where)
  end subroutine

end module module_hip_plugins

!-------------------------------------------------------------------------------
! LIBD
!-------------------------------------------------------------------------------
integer function version_of_d()
  ! the hopping matrix of D21 (d/D21.F90), with the gamma order of the reference
  ! fixed (the plugins refuse a permuted gamma_index)
  implicit none
  version_of_d = 21
end

subroutine d(e, o, out, in, u)  ! out := D in  (d/D21.F90:29-76)
  use module_hip_plugins
  use bqcd_hip
  use module_vol
  implicit none
  integer       :: e, o
  SPINCOL_FIELD :: out, in
  GAUGE_FIELD   :: u
  TIMING_START(timing_bin_d)
  if (o /= 1 - e) call die("d(): in and out must have opposite parities")
  call plugin_links(u, "d()")
  call plugin_check(hb_dslash(plugin_ctx, int(e, c_int), 0_c_int, out, in), "d()")
  TIMING_STOP(timing_bin_d)
end

subroutine d_dag(e, o, out, in, u)  ! out := D^+ in
  use module_hip_plugins
  use bqcd_hip
  use module_vol
  implicit none
  integer       :: e, o
  SPINCOL_FIELD :: out, in
  GAUGE_FIELD   :: u
  TIMING_START(timing_bin_d_dag)
  if (o /= 1 - e) call die("d_dag(): in and out must have opposite parities")
  call plugin_links(u, "d_dag()")
  call plugin_check(hb_dslash(plugin_ctx, int(e, c_int), 1_c_int, out, in), "d_dag()")
  TIMING_STOP(timing_bin_d_dag)
end

subroutine dsf_xf(p, b, a, s, u, nn, volh)  ! d/DSFxyzt.F90:77-117, x forward
  use module_hip_plugins
  implicit none
  REAL, dimension(NGEN, *), intent(inout)         :: p
  REAL, intent(in)                                :: s
  COMPLEX, dimension(NDIRAC, NCOL, *), intent(in) :: b, a
  COMPLEX, dimension(NCOL, NCOL, *), intent(in)   :: u
  integer, intent(in)                             :: nn(*)
  integer                                         :: volh
  call plugin_dsf(0, 1, p, b, a, s, u, nn, volh, "dsf_xf()")
end

subroutine dsf_xb(p, b, a, s, u, nn, volh)  ! d/DSFxyzt.F90:77-117, x backward
  use module_hip_plugins
  implicit none
  REAL, dimension(NGEN, *), intent(inout)         :: p
  REAL, intent(in)                                :: s
  COMPLEX, dimension(NDIRAC, NCOL, *), intent(in) :: b, a
  COMPLEX, dimension(NCOL, NCOL, *), intent(in)   :: u
  integer, intent(in)                             :: nn(*)
  integer                                         :: volh
  call plugin_dsf(0, 0, p, b, a, s, u, nn, volh, "dsf_xb()")
end

subroutine dsf_yf(p, b, a, s, u, nn, volh)  ! d/DSFxyzt.F90:77-117, y forward
  use module_hip_plugins
  implicit none
  REAL, dimension(NGEN, *), intent(inout)         :: p
  REAL, intent(in)                                :: s
  COMPLEX, dimension(NDIRAC, NCOL, *), intent(in) :: b, a
  COMPLEX, dimension(NCOL, NCOL, *), intent(in)   :: u
  integer, intent(in)                             :: nn(*)
  integer                                         :: volh
  call plugin_dsf(1, 1, p, b, a, s, u, nn, volh, "dsf_yf()")
end

subroutine dsf_yb(p, b, a, s, u, nn, volh)  ! d/DSFxyzt.F90:77-117, y backward
  use module_hip_plugins
  implicit none
  REAL, dimension(NGEN, *), intent(inout)         :: p
  REAL, intent(in)                                :: s
  COMPLEX, dimension(NDIRAC, NCOL, *), intent(in) :: b, a
  COMPLEX, dimension(NCOL, NCOL, *), intent(in)   :: u
  integer, intent(in)                             :: nn(*)
  integer                                         :: volh
  call plugin_dsf(1, 0, p, b, a, s, u, nn, volh, "dsf_yb()")
end

subroutine dsf_zf(p, b, a, s, u, nn, volh)  ! d/DSFxyzt.F90:77-117, z forward
  use module_hip_plugins
  implicit none
  REAL, dimension(NGEN, *), intent(inout)         :: p
  REAL, intent(in)                                :: s
  COMPLEX, dimension(NDIRAC, NCOL, *), intent(in) :: b, a
  COMPLEX, dimension(NCOL, NCOL, *), intent(in)   :: u
  integer, intent(in)                             :: nn(*)
  integer                                         :: volh
  call plugin_dsf(2, 1, p, b, a, s, u, nn, volh, "dsf_zf()")
end

subroutine dsf_zb(p, b, a, s, u, nn, volh)  ! d/DSFxyzt.F90:77-117, z backward
  use module_hip_plugins
  implicit none
  REAL, dimension(NGEN, *), intent(inout)         :: p
  REAL, intent(in)                                :: s
  COMPLEX, dimension(NDIRAC, NCOL, *), intent(in) :: b, a
  COMPLEX, dimension(NCOL, NCOL, *), intent(in)   :: u
  integer, intent(in)                             :: nn(*)
  integer                                         :: volh
  call plugin_dsf(2, 0, p, b, a, s, u, nn, volh, "dsf_zb()")
end

subroutine dsf_tf(p, b, a, s, u, nn, volh)  ! d/DSFxyzt.F90:77-117, t forward
  use module_hip_plugins
  implicit none
  REAL, dimension(NGEN, *), intent(inout)         :: p
  REAL, intent(in)                                :: s
  COMPLEX, dimension(NDIRAC, NCOL, *), intent(in) :: b, a
  COMPLEX, dimension(NCOL, NCOL, *), intent(in)   :: u
  integer, intent(in)                             :: nn(*)
  integer                                         :: volh
  call plugin_dsf(3, 1, p, b, a, s, u, nn, volh, "dsf_tf()")
end

subroutine dsf_tb(p, b, a, s, u, nn, volh)  ! d/DSFxyzt.F90:77-117, t backward
  use module_hip_plugins
  implicit none
  REAL, dimension(NGEN, *), intent(inout)         :: p
  REAL, intent(in)                                :: s
  COMPLEX, dimension(NDIRAC, NCOL, *), intent(in) :: b, a
  COMPLEX, dimension(NCOL, NCOL, *), intent(in)   :: u
  integer, intent(in)                             :: nn(*)
  integer                                         :: volh
  call plugin_dsf(3, 0, p, b, a, s, u, nn, volh, "dsf_tb()")
end

!-------------------------------------------------------------------------------
! LIBCLOVER
!-------------------------------------------------------------------------------
subroutine allocate_clover_field_a(a)
  use typedef_clover
  use module_vol
  implicit none
  P_CLOVER_FIELD_A :: a
  allocate(a(2, volh, EVEN:ODD))
end

subroutine allocate_clover_field_b(b)
  use typedef_clover
  use module_vol
  implicit none
  P_CLOVER_FIELD_B :: b
  allocate(b(2, volh, EVEN:ODD))
end

! A/2, A^-1/2 and the LDL^+ factors from physical links (clover/clover_init.F90:28-78)
subroutine clover_init(a, ainv, b, u, csw_kappa)
  use typedef_clover
  use module_hip_plugins
  use bqcd_hip
  use module_vol
  implicit none
  CLOVER_FIELD_A, intent(out) :: a, ainv
  CLOVER_FIELD_B, intent(out) :: b
  GAUGE_FIELD,    intent(in)  :: u
  REAL,           intent(in)  :: csw_kappa
  TIMING_START(timing_bin_clover_init)
  call plugin_links(u, "clover_init()")
  call plugin_check(hb_build_clover(plugin_ctx, real(csw_kappa, c_double)), "clover_init()")
  call plugin_check(hb_save_clover(plugin_ctx, a, ainv, b), "clover_init()")
  TIMING_STOP(timing_bin_clover_init)
end

subroutine clover_mult_a(out, a, in, volh)  ! out := A in
  use module_hip_plugins
  use bqcd_hip
  implicit none
  COMPLEX, dimension(18, 2, *)        :: a
  COMPLEX, dimension(NDIRAC, NCOL, *) :: out, in
  integer                             :: volh
  TIMING_START(timing_bin_clover_mult_a)
  call plugin_init()
  if (volh /= hb_volh(plugin_ctx)) call die("clover_mult_a(): volh differs from the lattice's")
  call plugin_check(hb_clover_mult_field(plugin_ctx, 0_c_int, a, out, in), "clover_mult_a()")
  TIMING_STOP(timing_bin_clover_mult_a)
end

subroutine clover_mult_ao(a, x, volh)  ! x := A x
  use module_hip_plugins
  use bqcd_hip
  implicit none
  COMPLEX, dimension(18, 2, *)        :: a
  COMPLEX, dimension(NDIRAC, NCOL, *) :: x
  integer                             :: volh
  TIMING_START(timing_bin_clover_mult_ao)
  call plugin_init()
  if (volh /= hb_volh(plugin_ctx)) call die("clover_mult_ao(): volh differs from the lattice's")
  call plugin_check(hb_clover_mult_field(plugin_ctx, 0_c_int, a, x, x), "clover_mult_ao()")
  TIMING_STOP(timing_bin_clover_mult_ao)
end

subroutine clover_mult_b(b, x, volh)  ! x := B x, the solve with the LDL^+ factors
  use module_hip_plugins
  use bqcd_hip
  implicit none
  COMPLEX, dimension(18, 2, *)        :: b
  COMPLEX, dimension(NDIRAC, NCOL, *) :: x
  integer                             :: volh
  TIMING_START(timing_bin_clover_mult_b)
  call plugin_init()
  if (volh /= hb_volh(plugin_ctx)) call die("clover_mult_b(): volh differs from the lattice's")
  call plugin_check(hb_clover_mult_field(plugin_ctx, 1_c_int, b, x, x), "clover_mult_b()")
  TIMING_STOP(timing_bin_clover_mult_b)
end

! clover part of dS_f/dU for spinors b, a of parity eo (clover/clover_d.F90:78-106)
subroutine clover_dsf(eo, p, b, a, s, u)
  use module_hip_plugins
  use bqcd_hip
  use module_vol
  implicit none
  integer         :: eo
  GENERATOR_FIELD :: p
  SPINCOL_FIELD   :: b, a
  REAL            :: s
  GAUGE_FIELD     :: u
  TIMING_START(timing_bin_clover_dsf)
  call plugin_links(u, "clover_dsf()")
  call plugin_check(hb_clover_dsf(plugin_ctx, int(eo, c_int), p, int(volh_tot, c_int), b, a, &
                                  real(s, c_double)), "clover_dsf()")
  TIMING_STOP(timing_bin_clover_dsf)
end

! force of the clover determinant with the LDL^+ field of parity eo (clover/clover_d.F90:40-75)
subroutine clover_dsd(eo, p, b, s, u)
  use typedef_clover
  use module_hip_plugins
  use bqcd_hip
  use module_vol
  implicit none
  integer         :: eo
  GENERATOR_FIELD :: p
  CLOVER_FIELD_B  :: b
  REAL            :: s
  GAUGE_FIELD     :: u
  TIMING_START(timing_bin_clover_dsd)
  call plugin_links(u, "clover_dsd()")
  call plugin_check(hb_load_clover(plugin_ctx, b=b), "clover_dsd()")
  call plugin_check(hb_clover_dsd(plugin_ctx, int(eo, c_int), p, int(volh_tot, c_int), &
                                  real(s, c_double)), "clover_dsd()")
  TIMING_STOP(timing_bin_clover_dsd)
end

! 2 sum log of the stored LDL^+ diagonals (clover/clover_action.F90:29-62)
REAL function clover_action(b)
  use typedef_clover
  use module_hip_plugins
  use bqcd_hip
  use module_vol
  implicit none
  type(type_clover_b) :: b(2, volh)
  real(c_double)      :: s
  call plugin_init()
  call plugin_check(hb_clover_action_field(plugin_ctx, b, s), "clover_action()")
  clover_action = s
end

!===============================================================================
