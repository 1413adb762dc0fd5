!===============================================================================
!
! bqcd_hip.F90 - Fortran 2018 bindings (iso_c_binding) of the C ABI declared
! in include/bqcd_hip.h.  Self-contained: needs no BQCD module, so any Fortran
! code can call the library.  The BQCD-side drop-ins for quda_solver / dsf /
! dsd / dsg / hmc_u (hip_solver.F90) are built on this module.
!
! Host arrays are passed by address in BQCD's own layouts (spinor
! complex(4,3,volh), gauge complex(3,3,site_stride,0:1,4), clover
! type_clover_a/b(2,volh,0:1), momenta real(8,site_stride,0:1,4)); the
! assumed-type dummies accept any of them without a copy.
!
!-------------------------------------------------------------------------------
module bqcd_hip

  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: HB_OK = 0, HB_E_ARG = -1, HB_E_HIP = -2, HB_E_STATE = -3, &
                               HB_E_NOCONV = -4, HB_E_NOMEM = -5
  integer(c_int), parameter :: HB_OP_MTIL = 0, HB_OP_MTIL_DAG = 1, HB_OP_MTDAGMT = 2, &
                               HB_OP_W = 3, HB_OP_W_DAG = 4, HB_OP_WDAGW = 5
  integer(c_int), parameter :: HB_CLOVER_A = 0, HB_CLOVER_AINV = 1, HB_CLOVER_B = 2
  integer(c_int), parameter :: HB_PREC_F64 = 64, HB_PREC_F32 = 32, HB_PREC_F16 = 16, &
                               HB_PREC_F16_GAUGE = 18
  integer(c_int), parameter :: HB_INT_LEAPFROG = 0, HB_INT_OMELYAN = 1

  ! hb_hmc_params / hb_hmc_out of include/bqcd_hip.h
  type, bind(C) :: hb_hmc_params
     real(c_double)         :: beta, kappa, csw, rho, tau
     integer(c_int)         :: ntau, m_scale
     character(kind=c_char) :: model
     integer(c_int)         :: integrator
     real(c_double)         :: lambda
     real(c_float)          :: rest
     integer(c_int)         :: maxiter, log, precision
     real(c_double)         :: delta
     integer(c_int)         :: max_inner, mre_vectors, batch, verbose
  end type hb_hmc_params

  type, bind(C) :: hb_hmc_out
     real(c_double) :: exp_dh, h_old, h_new, sg_old, sg_new, sp_old, sp_new, sd_old, sd_new
     real(c_double) :: sf1_old, sf1_new, sf2_old, sf2_new, plaquette
     integer(c_int) :: accepted, cg_ncall, cg_niter_tot, cg_niter_max, n_force_ir, n_force_uv
     real(c_double) :: seconds
  end type hb_hmc_out

  interface

     function hb_last_error_c() bind(C, name="hb_last_error") result(msg)
       import :: c_ptr
       type(c_ptr) :: msg
     end function

     integer(c_int) function hb_init(L, bc, device, ctx) bind(C, name="hb_init")
       import :: c_int, c_ptr
       integer(c_int), intent(in)        :: L(4), bc(4)
       integer(c_int), value             :: device
       type(c_ptr),    intent(out)       :: ctx
     end function

     integer(c_int) function hb_finalize(ctx) bind(C, name="hb_finalize")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function

     integer(c_int) function hb_volh(ctx) bind(C, name="hb_volh")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function

     integer(c_int) function hb_load_gauge(ctx, u, site_stride, flipped) bind(C, name="hb_load_gauge")
       import :: c_int, c_ptr
       type(c_ptr), value    :: ctx
       type(*), intent(in)   :: u(*)
       integer(c_int), value :: site_stride, flipped
     end function

     integer(c_int) function hb_save_gauge(ctx, u, site_stride, flipped) bind(C, name="hb_save_gauge")
       import :: c_int, c_ptr
       type(c_ptr), value    :: ctx
       type(*)               :: u(*)
       integer(c_int), value :: site_stride, flipped
     end function

     ! absent arguments are passed as NULL (that field is not uploaded)
     integer(c_int) function hb_load_clover(ctx, a, ainv, b) bind(C, name="hb_load_clover")
       import :: c_int, c_ptr
       type(c_ptr), value            :: ctx
       type(*), intent(in), optional :: a(*), ainv(*), b(*)
     end function

     integer(c_int) function hb_save_clover(ctx, a, ainv, b) bind(C, name="hb_save_clover")
       import :: c_int, c_ptr
       type(c_ptr), value  :: ctx
       type(*), optional   :: a(*), ainv(*), b(*)
     end function

     integer(c_int) function hb_build_clover(ctx, csw_kappa) bind(C, name="hb_build_clover")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       real(c_double), value :: csw_kappa
     end function

     integer(c_int) function hb_clover_action(ctx, s_det) bind(C, name="hb_clover_action")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: s_det
     end function

     integer(c_int) function hb_dslash(ctx, out_parity, dagger, out, in) bind(C, name="hb_dslash")
       import :: c_int, c_ptr
       type(c_ptr), value    :: ctx
       integer(c_int), value :: out_parity, dagger
       type(*)               :: out(*)
       type(*), intent(in)   :: in(*)
     end function

     integer(c_int) function hb_clover_mult(ctx, which, parity, out, in) bind(C, name="hb_clover_mult")
       import :: c_int, c_ptr
       type(c_ptr), value    :: ctx
       integer(c_int), value :: which, parity
       type(*)               :: out(*)
       type(*), intent(in)   :: in(*)
     end function

     integer(c_int) function hb_apply(ctx, op, kappa, rho, out, in) bind(C, name="hb_apply")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       integer(c_int), value :: op
       real(c_double), value :: kappa, rho
       type(*)               :: out(*)
       type(*), intent(in)   :: in(*)
     end function

     integer(c_int) function hb_cg(ctx, op, kappa, rho, x, b, rest, maxiter, log, iters, true_rtr) &
          bind(C, name="hb_cg")
       import :: c_int, c_ptr, c_double, c_float
       type(c_ptr), value                    :: ctx
       integer(c_int), value                 :: op, maxiter, log
       real(c_double), value                 :: kappa, rho
       type(*)                               :: x(*)
       type(*), intent(in)                   :: b(*)
       real(c_float), value                  :: rest
       integer(c_int), intent(out)           :: iters
       real(c_double), intent(out), optional :: true_rtr
     end function

     integer(c_int) function hb_cg_mixed(ctx, op, kappa, rho, x, b, rest, maxiter, log, precision, &
          delta, max_inner, iters, n_reliable, true_rtr) bind(C, name="hb_cg_mixed")
       import :: c_int, c_ptr, c_double, c_float
       type(c_ptr), value                    :: ctx
       integer(c_int), value                 :: op, maxiter, log, precision, max_inner
       real(c_double), value                 :: kappa, rho, delta
       type(*)                               :: x(*)
       type(*), intent(in)                   :: b(*)
       real(c_float), value                  :: rest
       integer(c_int), intent(out)           :: iters
       integer(c_int), intent(out), optional :: n_reliable
       real(c_double), intent(out), optional :: true_rtr
     end function

     ! x(k), b(k) = c_loc of the k-th spinor
     integer(c_int) function hb_cg_multi(ctx, op, nrhs, kappa, rho, csw, x, b, rest, maxiter, log, &
          precision, block, iters, true_rtr) bind(C, name="hb_cg_multi")
       import :: c_int, c_ptr, c_double, c_float
       type(c_ptr), value          :: ctx
       integer(c_int), value       :: op, nrhs, maxiter, log, precision, block
       real(c_double), intent(in)  :: kappa(*), rho(*)
       real(c_double), value       :: csw
       type(c_ptr), intent(in)     :: x(*), b(*)
       real(c_float), intent(in)   :: rest(*)
       integer(c_int), intent(out) :: iters(*)
       real(c_double), intent(out) :: true_rtr(*)
     end function

     integer(c_int) function hb_multi_block_size(ctx) bind(C, name="hb_multi_block_size")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function hb_multi_block_limit(volh, precision) bind(C, name="hb_multi_block_limit")
       import :: c_int
       integer(c_int), value :: volh, precision
     end function
     integer(c_int) function hb_multi_tile(ctx) bind(C, name="hb_multi_tile")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function hb_time_dslash_lp(ctx, precision, reps, in_d, ms) bind(C, name="hb_time_dslash_lp")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, in_d
       integer(c_int), value :: precision, reps
       real(c_double) :: ms
     end function
     integer(c_int) function hb_multi_set_tile(ctx, tile) bind(C, name="hb_multi_set_tile")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: tile
     end function

     integer(c_int) function hb_dsf(ctx, p, site_stride, step, kappa, csw_kappa, a, b) &
          bind(C, name="hb_dsf")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       type(*)               :: p(*)
       integer(c_int), value :: site_stride
       real(c_double), value :: step, kappa, csw_kappa
       type(*), intent(in)   :: a(*), b(*)
     end function

     integer(c_int) function hb_dsd(ctx, p, site_stride, step, csw_kappa) bind(C, name="hb_dsd")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       type(*)               :: p(*)
       integer(c_int), value :: site_stride
       real(c_double), value :: step, csw_kappa
     end function

     integer(c_int) function hb_dsf_dir(ctx, mu, forward, p, b, a, s, u, nn, n_tot) &
          bind(C, name="hb_dsf_dir")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       integer(c_int), value :: mu, forward
       type(*)               :: p(*)
       type(*), intent(in)   :: b(*), a(*)
       real(c_double), value :: s
       type(*), intent(in)   :: u(*)
       integer(c_int), intent(in) :: nn(*)
       integer(c_int), value :: n_tot
     end function

     integer(c_int) function hb_clover_mult_field(ctx, mode, clov, out, in) bind(C, name="hb_clover_mult_field")
       import :: c_int, c_ptr
       type(c_ptr), value    :: ctx
       integer(c_int), value :: mode
       type(*), intent(in)   :: clov(*)
       type(*)               :: out(*)
       type(*), intent(in)   :: in(*)
     end function

     integer(c_int) function hb_clover_dsf(ctx, eo, p, site_stride, b, a, s) bind(C, name="hb_clover_dsf")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       integer(c_int), value :: eo
       type(*)               :: p(*)
       integer(c_int), value :: site_stride
       type(*), intent(in)   :: b(*), a(*)
       real(c_double), value :: s
     end function

     integer(c_int) function hb_clover_dsd(ctx, eo, p, site_stride, s) bind(C, name="hb_clover_dsd")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       integer(c_int), value :: eo
       type(*)               :: p(*)
       integer(c_int), value :: site_stride
       real(c_double), value :: s
     end function

     integer(c_int) function hb_clover_action_field(ctx, b, s_det) bind(C, name="hb_clover_action_field")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx
       type(*), intent(in)         :: b(*)
       real(c_double), intent(out) :: s_det
     end function

     integer(c_int) function hb_dsg(ctx, p, site_stride, step, beta) bind(C, name="hb_dsg")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       type(*)               :: p(*)
       integer(c_int), value :: site_stride
       real(c_double), value :: step, beta
     end function

     integer(c_int) function hb_hmc_u(ctx, p, site_stride, step, csw_kappa) bind(C, name="hb_hmc_u")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       type(*), intent(in)   :: p(*)
       integer(c_int), value :: site_stride
       real(c_double), value :: step, csw_kappa
     end function

     integer(c_int) function hb_plaquette(ctx, re_tr_sum) bind(C, name="hb_plaquette")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: re_tr_sum
     end function

     integer(c_int) function hb_gauge_backup(ctx, slot) bind(C, name="hb_gauge_backup")
       import :: c_int, c_ptr
       type(c_ptr), value    :: ctx
       integer(c_int), value :: slot
     end function

     integer(c_int) function hb_gauge_restore(ctx, slot) bind(C, name="hb_gauge_restore")
       import :: c_int, c_ptr
       type(c_ptr), value    :: ctx
       integer(c_int), value :: slot
     end function

     ! one HMC trajectory (hmc.F90:28-180); p, eta1, eta2 in BQCD layouts
     integer(c_int) function hb_hmc_trajectory(ctx, prm, p, site_stride, eta1, eta2, accept_random, out) &
          bind(C, name="hb_hmc_trajectory")
       import :: c_int, c_ptr, c_double, hb_hmc_params, hb_hmc_out
       type(c_ptr), value            :: ctx
       type(hb_hmc_params), intent(in) :: prm
       type(*), intent(in)           :: p(*)
       integer(c_int), value         :: site_stride
       type(*), intent(in)           :: eta1(*)
       type(*), intent(in), optional :: eta2(*)
       real(c_double), value         :: accept_random
       type(hb_hmc_out), intent(out) :: out
     end function

     ! hmc() in its reversibility-test modes (hmc.F90:124-210); p in and out
     integer(c_int) function hb_hmc_test(ctx, prm, direction, p, site_stride, eta1, eta2, out) &
          bind(C, name="hb_hmc_test")
       import :: c_int, c_ptr, hb_hmc_params, hb_hmc_out
       type(c_ptr), value            :: ctx
       type(hb_hmc_params), intent(in) :: prm
       integer(c_int), value         :: direction
       type(*), intent(inout)        :: p(*)
       integer(c_int), value         :: site_stride
       type(*), intent(in), optional :: eta1(*)
       type(*), intent(in), optional :: eta2(*)
       type(hb_hmc_out), intent(out) :: out
     end function

     ! BQCD configuration files (conf.F90:305-470), host only; strings end in c_null_char
     integer(c_int) function hb_conf_write(L, bc, u, site_stride, dir, run, traj, ensemble, former, &
          beta, kappa, csw, h, plaq_energy, info_path, info_path_len) bind(C, name="hb_conf_write")
       import :: c_int, c_double, c_char
       integer(c_int), intent(in)         :: L(4), bc(4)
       type(*), intent(in)                :: u(*)
       integer(c_int), value              :: site_stride, run, traj, ensemble, former, info_path_len
       character(kind=c_char), intent(in) :: dir(*)
       real(c_double), value              :: beta, kappa, csw, h, plaq_energy
       character(kind=c_char)             :: info_path(*)
     end function

     integer(c_int) function hb_conf_read(info_file, L, u, site_stride, plaq_energy) &
          bind(C, name="hb_conf_read")
       import :: c_int, c_double, c_char
       character(kind=c_char), intent(in)    :: info_file(*)
       integer(c_int), intent(in)            :: L(4)
       type(*)                               :: u(*)
       integer(c_int), value                 :: site_stride
       real(c_double), intent(out), optional :: plaq_energy
     end function

     integer(c_int) function hb_last_timing(ctx, solve_ms, n_matvec) bind(C, name="hb_last_timing")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: solve_ms
       integer(c_int), intent(out) :: n_matvec
     end function

  end interface

contains

  ! message of the last failed call, as a Fortran string
  function hb_last_error() result(msg)
    character(len=256) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n

    interface
       integer(c_size_t) function c_strlen(str) bind(C, name="strlen")
         import :: c_ptr, c_size_t
         type(c_ptr), value :: str
       end function
    end interface

    msg = ' '
    p = hb_last_error_c()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [int(c_strlen(p))])
    do n = 1, min(size(s), len(msg))
       msg(n:n) = s(n)
    enddo
  end function

end module bqcd_hip
