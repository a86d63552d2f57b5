! pcom_main.f90 - Fortran driver of the MI355X time stepper.
!
! The caller side of the reference's main program (main.f90:85-238):
! reads the namelist (namelist time_ctrl / io_ctrl, main.f90:763-800 and
! the reference's `namelist` file), loads the raw inputs, creates the
! device context through the ISO_C_BINDING layer (pcom_gpu.f90), runs the
! baroclinic loop from bdate to edate, prints the reference's progress line
! and writes the time means and restart files at the output period.
!
! Files are raw little-endian streams (netCDF is not part of this build):
!   inputs   header int32 glo_ni, glo_nj, nk, has_vmix; then float64 lon,
!            lat, z, int32 itn, float64 tmn, smn, pt, sa, [vmix], taux,
!            tauy, bct, bcs, pa, fw (interior arrays, Fortran order)
!   means    means_<date>.bin: int32 glo_ni, glo_nj, nk, int32 date(6),
!            float64 pt, sa, rho, u, v, w (glo_ni, glo_nj, nk), ch, ssh
!   restart  restart_<date>.bin: the layout of pcom_amd/restart.py
!   forcing  ff_ws / ff_pa / ff_bct (optional, main.f90:167-202): float64
!            records of (glo_ni, glo_nj); ff_ws holds taux then tauy per
!            record.  Record nday on each daily refresh, and with
!            out_per = '6hour' record (nday-1)*4 + hour/6 + 1 every 6 hours
!
! Usage: pcom_main [namelist-file]      (default: ./namelist)
!        pcom_main --abi                 (print the sizes of the C structs)
program pcom_main
  use, intrinsic :: iso_c_binding
  use pcom_gpu
  implicit none

  ! namelist (the reference's time_ctrl / io_ctrl groups, plus gpu_ctrl)
  character(len=19) :: bdate = '0001-01-01 00:00:00', edate = '0001-01-02 00:00:00'
  integer :: dtbt = 60, dtbc = 3600
  character(len=256) :: fname_ini = 'input/inputs.bin', fname_rst = '', out_dir = 'output/'
  character(len=256) :: ff_ws = '', ff_pa = '', ff_bct = ''
  character(len=16) :: out_per = 'day'
  integer :: restart_per = 1, restart_run = 0
  character(len=12) :: precision = 'fp64'
  integer :: device = 0, tile_i = 0, tile_j = 0, tile_s = 0, trop_mode = 0
  namelist /time_ctrl/ bdate, edate, dtbt, dtbc
  namelist /io_ctrl/ fname_ini, fname_rst, out_dir, out_per, restart_per, restart_run, ff_ws, ff_pa, ff_bct
  namelist /gpu_ctrl/ precision, device, tile_i, tile_j, tile_s, trop_mode

  character(len=256) :: nml_file
  integer :: u, ios, has_vmix
  integer(c_int) :: glo_ni, glo_nj, nk
  real(c_double), allocatable, target :: lon(:), lat(:), z(:), tmn(:), smn(:)
  real(c_double), allocatable, target :: pt(:,:,:), sa(:,:,:), vmix(:,:,:)
  real(c_double), allocatable, target :: taux(:,:,:), tauy(:,:,:), bct(:,:,:), bcs(:,:,:), pa(:,:,:), fw(:,:,:)
  integer(c_int), allocatable, target :: itn(:,:)
  type(pcom_inputs) :: inp
  type(pcom_opts) :: opts
  type(c_ptr) :: ctx
  type(pcom_counters) :: cnt
  integer :: bd(6), ed(6), cd(6), pd(6)
  integer(c_long_long) :: i, nt, ist
  integer :: nrec, nday
  integer(c_int) :: due(2)
  real(c_double), allocatable, target :: ftx(:,:), fty(:,:), fpa(:,:), fbt(:,:)
  real(c_double) :: sums(3)
  integer(c_int64_t) :: t0, t1, rate
  type(pcom_means) :: means_probe
  type(pcom_forcing) :: forcing_probe

  nml_file = 'namelist'
  if (command_argument_count() >= 1) call get_command_argument(1, nml_file)
  if (trim(nml_file) == '--abi') then   ! sizes of the interoperable types
    write(*, '(a, 5(1x, i0))') 'abi', c_sizeof(inp), c_sizeof(opts), c_sizeof(cnt), c_sizeof(means_probe), &
      c_sizeof(forcing_probe)
    stop
  end if
  open(newunit=u, file=trim(nml_file), status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*, '(a, a)') 'pcom_main: cannot open namelist file ', trim(nml_file)
    error stop 1
  end if
  read(u, nml=time_ctrl)
  rewind(u)
  read(u, nml=io_ctrl)
  rewind(u)
  read(u, nml=gpu_ctrl, iostat=ios)
  close(u)

  ! ---- raw inputs (what mod_io reads from initial.nc / forcing.nc)
  open(newunit=u, file=trim(fname_ini), access='stream', form='unformatted', status='old', &
       action='read', iostat=ios)
  if (ios /= 0) then
    write(*, '(a, a)') 'pcom_main: cannot open input file ', trim(fname_ini)
    error stop 1
  end if
  read(u) glo_ni, glo_nj, nk, has_vmix
  allocate(lon(glo_ni), lat(glo_nj), z(nk), tmn(nk), smn(nk), itn(glo_ni, glo_nj))
  allocate(pt(glo_ni, glo_nj, nk), sa(glo_ni, glo_nj, nk))
  allocate(taux(glo_ni, glo_nj, 12), tauy(glo_ni, glo_nj, 12), bct(glo_ni, glo_nj, 12))
  allocate(bcs(glo_ni, glo_nj, 12), pa(glo_ni, glo_nj, 12), fw(glo_ni, glo_nj, 12))
  read(u) lon, lat, z, itn, tmn, smn, pt, sa
  if (has_vmix /= 0) then
    allocate(vmix(glo_ni, glo_nj, nk))
    read(u) vmix
  end if
  read(u) taux, tauy, bct, bcs, pa, fw
  close(u)

  inp%glo_ni = glo_ni; inp%glo_nj = glo_nj; inp%nk = nk
  inp%lon = c_loc(lon); inp%lat = c_loc(lat); inp%z = c_loc(z); inp%itn = c_loc(itn)
  inp%tmn = c_loc(tmn); inp%smn = c_loc(smn); inp%pt = c_loc(pt); inp%sa = c_loc(sa)
  inp%vmix = c_null_ptr
  if (has_vmix /= 0) inp%vmix = c_loc(vmix)
  inp%taux = c_loc(taux); inp%tauy = c_loc(tauy); inp%bct = c_loc(bct)
  inp%bcs = c_loc(bcs); inp%pa = c_loc(pa); inp%fw = c_loc(fw)

  call parse_date(bdate, bd)
  call parse_date(edate, ed)
  opts%dtbt = dtbt; opts%dtbc = dtbc
  opts%year = bd(1); opts%month = bd(2); opts%day = bd(3); opts%hour = bd(4)
  opts%precision = PCOM_FP64
  if (trim(precision) == 'mixed') opts%precision = PCOM_MIXED
  if (trim(precision) == 'mixed_eos32') opts%precision = PCOM_MIXED_EOS32
  opts%restart_run = restart_run
  opts%device = device
  opts%trop_tile_i = tile_i; opts%trop_tile_j = tile_j; opts%trop_substeps = tile_s
  opts%trop_mode = trop_mode
  call pcom_check(pcom_create(inp, opts, ctx), 'pcom_create')

  ! ---- time control (main.f90:123-141)
  nt = (date_sec(ed) - date_sec(bd)) / dtbc
  ist = 1
  if (restart_run == 1 .and. len_trim(fname_rst) > 0) then
    call read_restart(trim(fname_rst), cd, pd, i)
    if (date_sec(cd) > date_sec(bd)) then
      ist = (date_sec(cd) - date_sec(bd)) / dtbc + 1
      call pcom_check(pcom_set_clock(ctx, ist - 1, cd, pd), 'pcom_set_clock')
      write(*, '(a, a, a)') "I'm going for a restart-run from ", trim(fname_rst), '.'
    else
      write(*, '(a, a, a)') "I'm using the status of ", trim(fname_rst), ' as initial status.'
    end if
  end if

  ! ---- baroclinic loop (main.f90:156-238); the library advances the
  ! calendar and refreshes the daily forcing itself
  nrec = 0
  nday = 0
  allocate(ftx(glo_ni, glo_nj), fty(glo_ni, glo_nj), fpa(glo_ni, glo_nj), fbt(glo_ni, glo_nj))
  call system_clock(t0, rate)
  do i = ist, nt
    ! forcing files (main.f90:164-202): the library says which refreshes
    ! this step performs; the records replace int_bnd's fields before
    ! den_prho
    call pcom_check(pcom_forcing_due(ctx, due), 'pcom_forcing_due')
    if (due(1) /= 0) then
      nday = nday + 1
      if (len_trim(ff_ws) + len_trim(ff_pa) + len_trim(ff_bct) > 0) call apply_forcing(nday, 1_c_int)
    end if
    if (due(2) /= 0 .and. trim(out_per) == '6hour' .and. &
        len_trim(ff_ws) + len_trim(ff_pa) + len_trim(ff_bct) > 0) then
      call pcom_check(pcom_get_counters(ctx, cnt), 'pcom_get_counters')
      call apply_forcing((nday - 1) * 4 + cnt%hour / 6 + 1, 0_c_int)
    end if
    call pcom_check(pcom_step(ctx, 1_c_int), 'pcom_step')
    call pcom_check(pcom_get_counters(ctx, cnt), 'pcom_get_counters')
    cd = [cnt%year, cnt%month, cnt%day, cnt%hour, cnt%minute, cnt%second]
    if (output_due(cd, date_after(cd, -dtbc))) then
      call system_clock(t1)
      write(*, '(a, i4.4, a, i2.2, a, i2.2, a, i2.2, a, f8.2, a)') 'integrated to ', cd(1), '-', cd(2), &
        '-', cd(3), ' ', cd(4), ':00:00, used ', real(t1 - t0) / real(rate), ' seconds'
      t0 = t1
      call write_means(cd)
      nrec = nrec + 1
      if (restart_per > 0) then
        if (mod(nrec, restart_per) == 0) call write_restart(cd, date_after(cd, -dtbc), i)
      end if
    end if
  end do
  call pcom_check(pcom_check_finite(ctx), 'pcom_check_finite')
  call pcom_check(pcom_global_sums(ctx, sums), 'pcom_global_sums')
  write(*, '(a, i0, a, 3es24.16)') 'steps ', nt - ist + 1, ' heat/salt/mass sums ', sums
  call pcom_check(pcom_destroy(ctx), 'pcom_destroy')

contains

  ! record `rec` of each forcing file named in the namelist, handed to the
  ! library as one refresh (interpolate = 1: the daily one after int_bnd)
  subroutine apply_forcing(rec, interpolate)
    integer, intent(in) :: rec
    integer(c_int), intent(in) :: interpolate
    type(pcom_forcing) :: f
    integer(c_int64_t) :: rb
    rb = int(glo_ni, c_int64_t) * glo_nj * 8
    f%taux = c_null_ptr; f%tauy = c_null_ptr; f%pa = c_null_ptr; f%bct = c_null_ptr
    if (len_trim(ff_ws) > 0) then
      call read_record(ff_ws, 2 * rb * (rec - 1) + 1, ftx)
      call read_record(ff_ws, 2 * rb * (rec - 1) + rb + 1, fty)
      f%taux = c_loc(ftx); f%tauy = c_loc(fty)
    end if
    if (len_trim(ff_pa) > 0) then
      call read_record(ff_pa, rb * (rec - 1) + 1, fpa)
      f%pa = c_loc(fpa)
    end if
    if (len_trim(ff_bct) > 0) then
      call read_record(ff_bct, rb * (rec - 1) + 1, fbt)
      f%bct = c_loc(fbt)
    end if
    call pcom_check(pcom_set_forcing(ctx, f, interpolate), 'pcom_set_forcing')
  end subroutine apply_forcing

  subroutine read_record(fname, pos, a)
    character(len=*), intent(in) :: fname
    integer(c_int64_t), intent(in) :: pos
    real(c_double), intent(out) :: a(:,:)
    integer :: v, ios
    open(newunit=v, file=trim(fname), access='stream', form='unformatted', status='old', action='read', &
         iostat=ios)
    if (ios /= 0) then
      write(*, '(a, a)') 'pcom_main: cannot open forcing file ', trim(fname)
      stop 1
    end if
    read(v, pos=pos, iostat=ios) a
    if (ios /= 0) then
      write(*, '(a, a, a, i0)') 'pcom_main: forcing file ', trim(fname), ' has no record at byte ', pos
      stop 1
    end if
    close(v)
  end subroutine read_record

  subroutine parse_date(s, d)
    character(len=*), intent(in) :: s
    integer, intent(out) :: d(6)
    read(s, '(i4, 1x, i2, 1x, i2, 1x, i2, 1x, i2, 1x, i2)') d
  end subroutine parse_date

  ! days of month and seconds since 0001-01-01 (mod_type.f90:403-438)
  integer function days_month(y, m)
    integer, intent(in) :: y, m
    integer, parameter :: dm(12) = [31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]
    days_month = dm(m)
    if (m == 2) then
      if (mod(y, 100) == 0) then
        if (mod(y, 400) == 0) days_month = 29
      else
        if (mod(y, 4) == 0) days_month = 29
      end if
    end if
  end function days_month

  integer(c_long_long) function date_sec(d)
    integer, intent(in) :: d(6)
    integer :: y, m
    integer(c_long_long) :: days
    days = 0
    do y = 1, d(1) - 1
      days = days + 365
      if ((mod(y, 4) == 0 .and. mod(y, 100) /= 0) .or. mod(y, 400) == 0) days = days + 1
    end do
    do m = 1, d(2) - 1
      days = days + days_month(d(1), m)
    end do
    days = days + d(3) - 1
    date_sec = days * 86400_c_long_long + d(4) * 3600 + d(5) * 60 + d(6)
  end function date_sec

  function date_after(d, dt) result(r)
    integer, intent(in) :: d(6), dt
    integer :: r(6)
    integer(c_long_long) :: s, rem
    integer :: y, m, dim
    s = date_sec(d) + dt
    y = 1
    do
      dim = 365
      if ((mod(y, 4) == 0 .and. mod(y, 100) /= 0) .or. mod(y, 400) == 0) dim = 366
      if (s < int(dim, c_long_long) * 86400) exit
      s = s - int(dim, c_long_long) * 86400
      y = y + 1
    end do
    m = 1
    do while (s >= int(days_month(y, m), c_long_long) * 86400)
      s = s - int(days_month(y, m), c_long_long) * 86400
      m = m + 1
    end do
    rem = s
    r = [y, m, int(rem / 86400) + 1, int(mod(rem, 86400_c_long_long) / 3600), &
         int(mod(rem, 3600_c_long_long) / 60), int(mod(rem, 60_c_long_long))]
  end function date_after

  ! main.f90:228-236
  logical function output_due(c, p)
    integer, intent(in) :: c(6), p(6)
    select case (trim(out_per))
    case ('hour')
      output_due = c(4) /= p(4)
    case ('6hour')
      output_due = c(4) /= p(4) .and. mod(c(4), 6) == 0
    case ('day')
      output_due = c(3) /= p(3)
    case ('month')
      output_due = c(2) /= p(2)
    case ('year')
      output_due = c(1) /= p(1)
    case ('decade')
      output_due = c(1) /= p(1) .and. mod(c(1), 10) == 0
    case default
      output_due = .false.
    end select
  end function output_due

  function date_tag(d) result(s)
    integer, intent(in) :: d(6)
    character(len=19) :: s
    write(s, '(i4.4, a, i2.2, a, i2.2, a, i2.2, a, i2.2, a, i2.2)') d(1), '-', d(2), '-', d(3), '_', &
      d(4), '-', d(5), '-', d(6)
  end function date_tag

  subroutine write_means(d)
    integer, intent(in) :: d(6)
    real(c_double), allocatable, target :: m3(:,:,:,:), m2(:,:,:)
    type(pcom_means) :: mo
    integer :: v
    allocate(m3(glo_ni, glo_nj, nk, 6), m2(glo_ni, glo_nj, 2))
    mo%pt = c_loc(m3(1, 1, 1, 1)); mo%sa = c_loc(m3(1, 1, 1, 2)); mo%rho = c_loc(m3(1, 1, 1, 3))
    mo%u = c_loc(m3(1, 1, 1, 4)); mo%v = c_loc(m3(1, 1, 1, 5)); mo%w = c_loc(m3(1, 1, 1, 6))
    mo%ch = c_loc(m2(1, 1, 1)); mo%ssh = c_loc(m2(1, 1, 2))
    call pcom_check(pcom_output_means(ctx, mo), 'pcom_output_means')
    open(newunit=v, file=trim(out_dir)//'means_'//date_tag(d)//'.bin', access='stream', &
         form='unformatted', status='replace', action='write')
    write(v) glo_ni, glo_nj, nk, d
    write(v) m3, m2
    close(v)
  end subroutine write_means

  ! the 29 restart fields of mod_io.f90:80-113, interior only
  subroutine write_restart(c, p, step)
    integer, intent(in) :: c(6), p(6)
    integer(c_long_long), intent(in) :: step
    integer, parameter :: ids(29) = [PCOM_F_CHC, PCOM_F_CHP, PCOM_F_BND_TAUX, PCOM_F_BND_TAUY, &
      PCOM_F_BND_T, PCOM_F_BND_S, PCOM_F_BND_PA, PCOM_F_BND_FW, PCOM_F_BUC, PCOM_F_BVC, &
      PCOM_F_BUP, PCOM_F_BVP, PCOM_F_TC, PCOM_F_SC, PCOM_F_TP, PCOM_F_SP, PCOM_F_PRHO_T, &
      PCOM_F_PRHO_S, PCOM_F_UC, PCOM_F_VC, PCOM_F_UP, PCOM_F_VP, PCOM_F_AUC, PCOM_F_AVC, &
      PCOM_F_AUP, PCOM_F_AVP, PCOM_F_AUPP, PCOM_F_AVPP, PCOM_F_AM]
    real(c_double), allocatable :: a(:,:,:)
    integer :: v, n, nd
    open(newunit=v, file=trim(out_dir)//'restart_'//date_tag(c)//'.bin', access='stream', &
         form='unformatted', status='replace', action='write')
    write(v) int(z'50434F4D', c_int32_t), 1_c_int32_t, glo_ni, glo_nj, nk, step, c, p
    do n = 1, 29
      nd = nk
      if (n <= 12) nd = 1
      allocate(a(glo_ni + 2, glo_nj + 2, nd))
      call pcom_check(pcom_get_field(ctx, ids(n), a, size(a, kind=c_size_t)), 'pcom_get_field')
      write(v) a(2:glo_ni + 1, 2:glo_nj + 1, :)
      deallocate(a)
    end do
    close(v)
  end subroutine write_restart

  subroutine read_restart(fname, c, p, step)
    character(len=*), intent(in) :: fname
    integer, intent(out) :: c(6), p(6)
    integer(c_long_long), intent(out) :: step
    integer, parameter :: ids(29) = [PCOM_F_CHC, PCOM_F_CHP, PCOM_F_BND_TAUX, PCOM_F_BND_TAUY, &
      PCOM_F_BND_T, PCOM_F_BND_S, PCOM_F_BND_PA, PCOM_F_BND_FW, PCOM_F_BUC, PCOM_F_BVC, &
      PCOM_F_BUP, PCOM_F_BVP, PCOM_F_TC, PCOM_F_SC, PCOM_F_TP, PCOM_F_SP, PCOM_F_PRHO_T, &
      PCOM_F_PRHO_S, PCOM_F_UC, PCOM_F_VC, PCOM_F_UP, PCOM_F_VP, PCOM_F_AUC, PCOM_F_AVC, &
      PCOM_F_AUP, PCOM_F_AVP, PCOM_F_AUPP, PCOM_F_AVPP, PCOM_F_AM]
    real(c_double), allocatable :: a(:,:,:)
    integer(c_int32_t) :: magic, ver, gi, gj, gk
    integer :: v, n, nd
    open(newunit=v, file=fname, access='stream', form='unformatted', status='old', action='read')
    read(v) magic, ver, gi, gj, gk, step, c, p
    if (magic /= int(z'50434F4D', c_int32_t) .or. ver /= 1 .or. gi /= glo_ni .or. gj /= glo_nj .or. gk /= nk) then
      write(*, '(a, a)') 'pcom_main: restart file does not match the grid: ', fname
      error stop 1
    end if
    do n = 1, 29
      nd = nk
      if (n <= 12) nd = 1
      allocate(a(glo_ni + 2, glo_nj + 2, nd))
      call pcom_check(pcom_get_field(ctx, ids(n), a, size(a, kind=c_size_t)), 'pcom_get_field')
      read(v) a(2:glo_ni + 1, 2:glo_nj + 1, :)
      a(1, :, :) = a(glo_ni + 1, :, :)   ! one-rank halo swap (mod_mympi.f90:669-746)
      a(glo_ni + 2, :, :) = a(2, :, :)
      call pcom_check(pcom_set_field(ctx, ids(n), a, size(a, kind=c_size_t)), 'pcom_set_field')
      deallocate(a)
    end do
    close(v)
  end subroutine read_restart

end program pcom_main
