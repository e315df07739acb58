! A minimal Fortran host of the cm1gpu library: the call sequence the CM1
! driver makes around its time loop (src/cm1.F: allocation and init, the
! step loop of :3313-3476, output syncs), with the model state passed in
! and out as unformatted stream files in the reference's array shapes.
!
! usage: cm1g_host <state-in> <state-out>
!
! state-in:  nsteps (int32), cm1g_config (108 bytes), the columns zf, zh,
!            mh, mf, pi0, th0, qv0, u0, v0 (nk+3 floats each), then ua, va, wa,
!            tha, ppi, qa0, qa1, qa2 in their reference shapes (i fastest)
! state-out: cflmax of the last step, dry-air mass and total water (3 x
!            float64), then ua, va, wa, tha, ppi, qa0, qa1, qa2, rho, prs
program cm1g_host
  use, intrinsic :: iso_c_binding
  use cm1gpu
  implicit none

  type(cm1g_config) :: cfg
  type(c_ptr) :: m
  integer(c_int32_t) :: nsteps
  integer :: n, u, nk, ios
  integer :: ns, nu, nv, nw
  real(c_float), allocatable :: col(:), a(:)
  real(c_float) :: cflmax
  real(c_double) :: tmass, tmois
  character(len=1024) :: fin, fout
  integer(c_int), parameter :: col_ids(9) = [CM1G_K_ZF, CM1G_K_ZH, CM1G_K_MH, CM1G_K_MF, &
                                            CM1G_K_PI0, CM1G_K_TH0, CM1G_K_QV0, CM1G_K_U0, &
                                            CM1G_K_V0]
  integer(c_int), parameter :: in_ids(8) = [CM1G_UA, CM1G_VA, CM1G_WA, CM1G_THA, CM1G_PPI, &
                                           CM1G_QA0, CM1G_QA1, CM1G_QA2]
  integer(c_int), parameter :: in_stag(8) = [CM1G_STAG_U, CM1G_STAG_V, CM1G_STAG_W, CM1G_STAG_S, &
                                            CM1G_STAG_S, CM1G_STAG_S, CM1G_STAG_S, CM1G_STAG_S]
  integer(c_int), parameter :: out_ids(10) = [CM1G_UA, CM1G_VA, CM1G_WA, CM1G_THA, CM1G_PPI, &
                                             CM1G_QA0, CM1G_QA1, CM1G_QA2, CM1G_RHO, CM1G_PRS]
  integer(c_int), parameter :: out_stag(10) = [CM1G_STAG_U, CM1G_STAG_V, CM1G_STAG_W, &
                                              CM1G_STAG_S, CM1G_STAG_S, CM1G_STAG_S, &
                                              CM1G_STAG_S, CM1G_STAG_S, CM1G_STAG_S, CM1G_STAG_S]

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', iostat=ios)
  if (ios /= 0) stop 'cm1g_host: cannot open the input state'
  read(u) nsteps
  read(u) cfg
  nk = cfg%nk
  ns = (cfg%ni + 6) * (cfg%nj + 6) * (nk + 2)
  nu = (cfg%ni + 7) * (cfg%nj + 6) * (nk + 2)
  nv = (cfg%ni + 6) * (cfg%nj + 7) * (nk + 2)
  nw = (cfg%ni + 6) * (cfg%nj + 6) * (nk + 3)
  allocate(col(nk + 3), a(max(ns, nu, nv, nw)))

  call check(cm1g_create(cfg, m), 'cm1g_create')
  ! vertical grid, then metrics; base-state sounding, then derived columns
  do n = 1, 9
    read(u) col
    call check(cm1g_put_column(m, col_ids(n), col), 'cm1g_put_column')
    if (n == 4) call check(cm1g_metrics(m), 'cm1g_metrics')
  end do
  call check(cm1g_base_derive(m), 'cm1g_base_derive')
  do n = 1, 8
    read(u) a(1:count_of(in_stag(n)))
    call check(cm1g_put_array(m, in_ids(n), in_stag(n), a), 'cm1g_put_array')
  end do
  close(u)

  call check(cm1g_getset(m, 1), 'cm1g_getset')
  do n = 1, nsteps
    call check(cm1g_step(m, cflmax), 'cm1g_step')
  end do
  call check(cm1g_budgets(m, tmass, tmois), 'cm1g_budgets')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) real(cflmax, c_double), tmass, tmois
  do n = 1, 10
    call check(cm1g_get_array(m, out_ids(n), out_stag(n), a), 'cm1g_get_array')
    write(u) a(1:count_of(out_stag(n)))
  end do
  close(u)
  call check(cm1g_destroy(m), 'cm1g_destroy')

contains

  integer function count_of(stag)
    integer(c_int), intent(in) :: stag
    select case (stag)
    case (CM1G_STAG_U)
      count_of = nu
    case (CM1G_STAG_V)
      count_of = nv
    case (CM1G_STAG_W)
      count_of = nw
    case default
      count_of = ns
    end select
  end function

  ! the reference's reaction to a fatal condition (stopcm1)
  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= CM1G_OK) then
      write(*, '(a,": ",i0,": ",a)') what, rc, cm1g_error_message()
      error stop 1
    end if
  end subroutine

end program cm1g_host
