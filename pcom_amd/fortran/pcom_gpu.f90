! pcom_gpu.f90 - ISO_C_BINDING interface of libpcom_gpu for Fortran callers.
!
! One bind(C) interface per function of include/pcom_gpu.h, with the
! argument meaning of that header; derived types mirror its structs field
! for field.  This module is the layer through which a Fortran driver like
! the reference's main.f90 (main.f90:158-238) reaches the MI355X time
! stepper instead of calling den_prho, den_alpha, int_pgra, int_readyc,
! int_trop, int_clin, int_ssh and int_ts itself.
module pcom_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: PCOM_OK = 0, PCOM_ERR_ARG = -1, PCOM_ERR_HIP = -2, &
                               PCOM_ERR_NAN = -3, PCOM_ERR_STATE = -4, PCOM_ERR_NOGPU = -5
  integer(c_int), parameter :: PCOM_FP64 = 0, PCOM_MIXED = 1, PCOM_MIXED_EOS32 = 2
  integer(c_int), parameter :: PCOM_TROP_AUTO = 0, PCOM_TROP_PERSISTENT = 1, PCOM_TROP_LAUNCHES = 2

  ! field ids, in the order of the enum in include/pcom_gpu.h
  enum, bind(c)
    enumerator :: PCOM_F_TC = 0, PCOM_F_SC, PCOM_F_TP, PCOM_F_SP, &
                  PCOM_F_UC, PCOM_F_VC, PCOM_F_UP, PCOM_F_VP, &
                  PCOM_F_AUC, PCOM_F_AVC, PCOM_F_AUP, PCOM_F_AVP, PCOM_F_AUPP, PCOM_F_AVPP, &
                  PCOM_F_FX, PCOM_F_FY, PCOM_F_AM, PCOM_F_ALPHA, PCOM_F_ADP, PCOM_F_WM, &
                  PCOM_F_PRHO_T, PCOM_F_PRHO_S, &
                  PCOM_F_ACT, PCOM_F_ACS, PCOM_F_ACR, PCOM_F_ACU, PCOM_F_ACV, PCOM_F_ACW, &
                  PCOM_F_CHC, PCOM_F_CHP, PCOM_F_MCH, PCOM_F_MCH2, PCOM_F_ACCH, PCOM_F_ACSSH, &
                  PCOM_F_BUC, PCOM_F_BVC, PCOM_F_BUP, PCOM_F_BVP, PCOM_F_BUT, PCOM_F_BVT, &
                  PCOM_F_PAX, PCOM_F_PAY, &
                  PCOM_F_BPHI_XN, PCOM_F_BPHI_XS, PCOM_F_BPHI_YE, PCOM_F_BPHI_YW, &
                  PCOM_F_BGRA_XN, PCOM_F_BGRA_XS, PCOM_F_BGRA_YE, PCOM_F_BGRA_YW, &
                  PCOM_F_BND_TAUX, PCOM_F_BND_TAUY, PCOM_F_BND_T, PCOM_F_BND_S, &
                  PCOM_F_BND_PA, PCOM_F_BND_FW, PCOM_F_SSH, &
                  PCOM_F_GRAPHIHX, PCOM_F_GRAPHIHY, PCOM_F_COR, &
                  PCOM_F_PHIH_T, PCOM_F_PRH_T, PCOM_F_PHIH_U, PCOM_F_PRH_U, &
                  PCOM_F_KM, PCOM_F_KH, PCOM_F_COUNT
  end enum

  type, bind(C) :: pcom_inputs
    integer(c_int) :: glo_ni, glo_nj, nk
    type(c_ptr) :: lon, lat, z, itn, tmn, smn, pt, sa, vmix
    type(c_ptr) :: taux, tauy, bct, bcs, pa, fw
  end type pcom_inputs

  type, bind(C) :: pcom_opts
    integer(c_int) :: dtbt, dtbc, year, month, day, hour
    integer(c_int) :: precision, restart_run, device
    integer(c_int) :: trop_tile_i, trop_tile_j, trop_substeps
    integer(c_int) :: trop_mode
  end type pcom_opts

  type, bind(C) :: pcom_means
    type(c_ptr) :: pt, sa, rho, u, v, w, ch, ssh
  end type pcom_means

  real(c_float), parameter :: PCOM_MISSING_FLOAT = -9.99999e30

  ! forcing records of one refresh (c_loc of a (glo_ni, glo_nj) array, or
  ! c_null_ptr to keep int_bnd's value)
  type, bind(C) :: pcom_forcing
    type(c_ptr) :: taux, tauy, pa, bct
  end type pcom_forcing

  type, bind(C) :: pcom_counters
    integer(c_long_long) :: step
    integer(c_int) :: n_ts, n_uv, n_w, n_ch, n_ssh
    integer(c_int) :: year, month, day, hour, minute, second
  end type pcom_counters

  interface
    integer(c_int) function pcom_create(inp, opts, ctx) bind(C, name='pcom_create')
      import :: c_int, c_ptr, pcom_inputs, pcom_opts
      type(pcom_inputs), intent(in) :: inp
      type(pcom_opts), intent(in) :: opts
      type(c_ptr), intent(out) :: ctx
    end function pcom_create

    integer(c_int) function pcom_step(ctx, nsteps) bind(C, name='pcom_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function pcom_step

    integer(c_int) function pcom_get_dims(ctx, ni, nj, nk) bind(C, name='pcom_get_dims')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: ni, nj, nk
    end function pcom_get_dims

    integer(c_size_t) function pcom_field_size(ctx, id) bind(C, name='pcom_field_size')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
    end function pcom_field_size

    integer(c_int) function pcom_get_field(ctx, id, host, n) bind(C, name='pcom_get_field')
      import :: c_int, c_ptr, c_size_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), intent(out) :: host(*)
      integer(c_size_t), value :: n
    end function pcom_get_field

    integer(c_int) function pcom_set_field(ctx, id, host, n) bind(C, name='pcom_set_field')
      import :: c_int, c_ptr, c_size_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), intent(in) :: host(*)
      integer(c_size_t), value :: n
    end function pcom_set_field

    integer(c_int) function pcom_get_counters(ctx, c) bind(C, name='pcom_get_counters')
      import :: c_int, c_ptr, pcom_counters
      type(c_ptr), value :: ctx
      type(pcom_counters), intent(out) :: c
    end function pcom_get_counters

    integer(c_int) function pcom_set_clock(ctx, step, cdate, pdate) bind(C, name='pcom_set_clock')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), value :: step
      integer(c_int), intent(in) :: cdate(6), pdate(6)
    end function pcom_set_clock

    integer(c_int) function pcom_forcing_due(ctx, due) bind(C, name='pcom_forcing_due')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: due(2)
    end function pcom_forcing_due

    integer(c_int) function pcom_set_forcing(ctx, f, interpolate) bind(C, name='pcom_set_forcing')
      import :: c_int, c_ptr, pcom_forcing
      type(c_ptr), value :: ctx
      type(pcom_forcing), intent(in) :: f
      integer(c_int), value :: interpolate
    end function pcom_set_forcing

    integer(c_int) function pcom_output_means(ctx, out) bind(C, name='pcom_output_means')
      import :: c_int, c_ptr, pcom_means
      type(c_ptr), value :: ctx
      type(pcom_means), intent(in) :: out
    end function pcom_output_means

    integer(c_int) function pcom_get_metric(ctx, g, which, out, n) bind(C, name='pcom_get_metric')
      import :: c_int, c_ptr, c_size_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: g, which
      real(c_double), intent(out) :: out(*)
      integer(c_size_t), value :: n
    end function pcom_get_metric

    integer(c_int) function pcom_global_sums(ctx, out) bind(C, name='pcom_global_sums')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(3)
    end function pcom_global_sums

    integer(c_int) function pcom_check_finite(ctx) bind(C, name='pcom_check_finite')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function pcom_check_finite

    integer(c_int) function pcom_sync(ctx) bind(C, name='pcom_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function pcom_sync

    integer(c_int) function pcom_phase_times(ctx, enable, out_ms, nsteps) bind(C, name='pcom_phase_times')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      real(c_double), intent(out) :: out_ms(3)
      integer(c_int), intent(out) :: nsteps
    end function pcom_phase_times

    integer(c_int) function pcom_destroy(ctx) bind(C, name='pcom_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function pcom_destroy

    type(c_ptr) function pcom_last_error() bind(C, name='pcom_last_error')
      import :: c_ptr
    end function pcom_last_error

    integer(c_int) function pcom_op_ter(ctx, ga, gb, nd3, dft, in, out) bind(C, name='pcom_op_ter')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ga, gb, nd3
      real(c_double), value :: dft
      real(c_double), intent(in) :: in(*)
      real(c_double), intent(out) :: out(*)
    end function pcom_op_ter

    integer(c_int) function pcom_op_px1(ctx, ga, gb, nd3, in, out) bind(C, name='pcom_op_px1')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ga, gb, nd3
      real(c_double), intent(in) :: in(*)
      real(c_double), intent(out) :: out(*)
    end function pcom_op_px1

    integer(c_int) function pcom_op_px2(ctx, ga, gb, nd3, in, out) bind(C, name='pcom_op_px2')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ga, gb, nd3
      real(c_double), intent(in) :: in(*)
      real(c_double), intent(out) :: out(*)
    end function pcom_op_px2

    integer(c_int) function pcom_op_div(ctx, ga, gb, gc, nd3, va, vb, out) bind(C, name='pcom_op_div')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ga, gb, gc, nd3
      real(c_double), intent(in) :: va(*), vb(*)
      real(c_double), intent(out) :: out(*)
    end function pcom_op_div

    integer(c_int) function pcom_op_lap(ctx, ga, gb, in, out) bind(C, name='pcom_op_lap')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ga, gb
      real(c_double), intent(in) :: in(*)
      real(c_double), intent(out) :: out(*)
    end function pcom_op_lap

    integer(c_int) function pcom_den_rho(ctx, n, t, s, p, rho) bind(C, name='pcom_den_rho')
      import :: c_int, c_ptr, c_size_t, c_double
      type(c_ptr), value :: ctx
      integer(c_size_t), value :: n
      real(c_double), intent(in) :: t(*), s(*), p(*)
      real(c_double), intent(out) :: rho(*)
    end function pcom_den_rho

    ! ensembles and the subcycle autotuner
    integer(c_int) function pcom_ensemble_create(members, count, ens) bind(C, name='pcom_ensemble_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: members(*)
      integer(c_int), value :: count
      type(c_ptr), intent(out) :: ens
    end function pcom_ensemble_create

    integer(c_int) function pcom_ensemble_step(ens, nsteps) bind(C, name='pcom_ensemble_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), value :: nsteps
    end function pcom_ensemble_step

    integer(c_int) function pcom_ensemble_packs(ens, npacks, pack_of_member) bind(C, name='pcom_ensemble_packs')
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
      integer(c_int), intent(out) :: npacks, pack_of_member(*)
    end function pcom_ensemble_packs

    integer(c_int) function pcom_ensemble_destroy(ens) bind(C, name='pcom_ensemble_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
    end function pcom_ensemble_destroy

    integer(c_int) function pcom_tune_subcycle(ctx, trials, tile, ms) bind(C, name='pcom_tune_subcycle')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: trials
      integer(c_int), intent(out) :: tile(3)
      real(c_double), intent(out) :: ms
    end function pcom_tune_subcycle

    integer(c_int) function pcom_subcycle_tiling(ctx, set, tile) bind(C, name='pcom_subcycle_tiling')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: set
      integer(c_int), intent(inout) :: tile(4)
    end function pcom_subcycle_tiling

    ! learned barotropic surrogate
    integer(c_int) function pcom_nn_create(ctx, capacity, seed) bind(C, name='pcom_nn_create')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: capacity
      integer(c_long_long), value :: seed
    end function pcom_nn_create

    integer(c_int) function pcom_nn_record(ctx, on) bind(C, name='pcom_nn_record')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function pcom_nn_record

    integer(c_int) function pcom_nn_train(ctx, iters, batch, lr, seed, loss) bind(C, name='pcom_nn_train')
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: iters, batch
      real(c_double), value :: lr
      integer(c_long_long), value :: seed
      real(c_double), intent(out) :: loss(*)
    end function pcom_nn_train

    integer(c_int) function pcom_nn_use(ctx, check_every, tol) bind(C, name='pcom_nn_use')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: check_every
      real(c_double), value :: tol
    end function pcom_nn_use

    integer(c_int) function pcom_nn_resume(ctx, plain_steps) bind(C, name='pcom_nn_resume')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: plain_steps
    end function pcom_nn_resume

    integer(c_int) function pcom_nn_eval(ctx, first, count, rmse) bind(C, name='pcom_nn_eval')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: first, count
      real(c_double), intent(out) :: rmse(8)
    end function pcom_nn_eval

    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function c_strlen
  end interface

contains

  ! message of the last failed call on this thread, as a Fortran string
  function pcom_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: cs(:)
    integer :: n, k
    p = pcom_last_error()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, cs, [n])
    allocate(character(len=n) :: msg)
    do k = 1, n
      msg(k:k) = cs(k)
    end do
  end function pcom_error_message

  ! stop with the library's message when a call failed (the reference
  ! stops on errors too, mod_io.f90:540-550)
  subroutine pcom_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= PCOM_OK) then
      write(*, '(a, a, a, i0, a, a)') 'pcom: ', what, ' failed (', rc, '): ', pcom_error_message()
      error stop 1
    end if
  end subroutine pcom_check

end module pcom_gpu
