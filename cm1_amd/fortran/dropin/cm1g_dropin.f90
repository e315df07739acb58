! Shared state of the drop-in replacements of CM1's adv_module, sound_module,
! sounde_module, kessler_module, turbtend_module and turb_module (the other
! files of this directory): one
! device model for the run, created at the first call from the namelist
! scalars of module input (src/input.F:261-351) and the constants of module
! constants (src/constants.F), and the host<->device moves around each call.
!
! Each replacement routine copies in the arrays the reference routine reads
! and copies out the arrays it writes, so the caller's arrays hold exactly
! what the reference routine leaves, and all host code between the calls
! (solve2's own loops, the bc fills, pdefq, statistics, output, restart) runs
! unchanged on them.  A switch the GPU path does not implement stops the run
! through stopcm1, as the reference does for its own fatal conditions.
module cm1g_dropin
  use, intrinsic :: iso_c_binding
  use cm1gpu
  implicit none
  private
  public :: dg_model, dg_check, dg_need, dg_put, dg_get, dg_col_s, dg_col_w, dg_put2, dg_get2, &
            dg_set_turbulence

  type(c_ptr), save :: gm = c_null_ptr

contains

  ! the run's device model, created on first use
  function dg_model() result(m)
    use input, only : ni, nj, nk, wbc, ebc, sbc, nbc, imoist, ptype, psolver, advwenos, advwenov, &
                      apmasscon, irbc, nsound, iptra, npt, pdtra, dx, dy, dz, dtl, alph, kdiv,     &
                      irdamp, rdalpha, zd, terrain_flag, axisymm, stretch_x, stretch_y, adapt_dt,  &
                      weno_order, eqtset, do_ib, ihail
    use constants, only : csmax
    type(c_ptr) :: m
    type(cm1g_config) :: c
    if (.not. c_associated(gm)) then
      call dg_need(.not. terrain_flag, 'terrain (terrain_flag) is not on the GPU path')
      call dg_need(axisymm == 0, 'axisymmetric grids (axisymm) are not on the GPU path')
      call dg_need(stretch_x == 0 .and. stretch_y == 0, 'horizontal stretching is not on the GPU path')
      call dg_need(adapt_dt == 0, 'adaptive time steps (adapt_dt) are not on the GPU path')
      call dg_need(weno_order == 5 .or. (advwenos == 0 .and. advwenov == 0), &
                   'WENO orders other than 5 are not on the GPU path')
      call dg_need(eqtset == 2, 'eqtset other than 2 is not on the GPU path')
      call dg_need(.not. do_ib, 'immersed boundaries (do_ib) are not on the GPU path')
      c%ni = ni
      c%nj = nj
      c%nk = nk
      c%wbc = wbc
      c%ebc = ebc
      c%sbc = sbc
      c%nbc = nbc
      c%imoist = imoist
      c%ptype = ptype
      c%psolver = psolver
      c%advwenos = advwenos
      c%advwenov = advwenov
      c%apmasscon = apmasscon
      c%irbc = irbc
      c%nsound = nsound
      c%npt = 0
      if (iptra == 1) c%npt = npt
      c%pdtra = pdtra
      c%dx = dx
      c%dy = dy
      c%dz = dz
      c%dt = dtl
      c%alph = alph
      c%kdiv = kdiv
      c%csmax = csmax
      c%irdamp = irdamp
      c%rdalpha = rdalpha
      c%zd = zd
      call dg_check(cm1g_create(c, gm), 'cm1g_create')
      call dg_set_turbulence(.false.)
    endif
    m = gm
  end function dg_model

  ! the LES subgrid turbulence of the dry CBL branch set, when the namelist
  ! asks for it (cm1g_set_turbulence; again after the vertical columns are
  ! in, whose zh and rmf it reads).  With other subgrid switches the GPU
  ! drop-ins of the turbulence stop the run.  required: stop if the
  ! namelist's switches are outside the branch set.
  subroutine dg_set_turbulence(required)
    use input, only : cm1setup, sgsmodel, tconfig, bcturbs, horizturb, doimpl, bbc, tbc, isfcflx, &
                      sfcmodel, set_flx, set_znt, ramp_sgs, cnst_shflx, cnst_znt
    logical, intent(in) :: required
    type(cm1g_turb_config) :: t
    logical :: ok
    ok = cm1setup == 1 .and. sgsmodel == 1 .and. tconfig == 1 .and. bcturbs == 1 .and.      &
         horizturb == 0 .and. bbc == 3 .and. tbc == 1 .and. isfcflx == 1 .and. sfcmodel == 1 &
         .and. set_flx == 1 .and. set_znt == 1 .and. ramp_sgs == 0
    if (required) call dg_need(ok, 'subgrid turbulence: only cm1setup 1, sgsmodel 1, tconfig 1, bcturbs 1, '// &
                               'horizturb 0, bbc 3, tbc 1, isfcflx 1, sfcmodel 1, set_flx 1, set_znt 1, '// &
                               'ramp_sgs 0 are on the GPU path')
    if (.not. ok) return
    t%cm1setup = 1
    t%doimpl = doimpl
    t%bbc = bbc
    t%isfcflx = isfcflx
    t%cnst_shflx = cnst_shflx
    t%cnst_znt = cnst_znt
    call dg_check(cm1g_set_turbulence(gm, t), 'cm1g_set_turbulence')
  end subroutine dg_set_turbulence

  subroutine dg_check(ierr, what)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= CM1G_OK) then
      print *, ' cm1gpu error ', ierr, ' in ', what, ': ', cm1g_error_message()
      call stopcm1
    endif
  end subroutine dg_check

  subroutine dg_need(cond, msg)
    logical, intent(in) :: cond
    character(len=*), intent(in) :: msg
    if (.not. cond) then
      print *, ' cm1gpu: ', msg
      call stopcm1
    endif
  end subroutine dg_need

  ! whole 3-D arrays in the reference's shapes (stag: CM1G_STAG_*)
  subroutine dg_put(id, stag, a)
    integer(c_int), intent(in) :: id, stag
    real, intent(in) :: a(*)
    call dg_check(cm1g_put_array(dg_model(), id, stag, a), 'cm1g_put_array')
  end subroutine dg_put

  subroutine dg_get(id, stag, a)
    integer(c_int), intent(in) :: id, stag
    real, intent(out) :: a(*)
    call dg_check(cm1g_get_array(dg_model(), id, stag, a), 'cm1g_get_array')
  end subroutine dg_get

  ! 2-D arrays in the reference's shapes (CM1G_D_*)
  subroutine dg_put2(id, a)
    integer(c_int), intent(in) :: id
    real, intent(in) :: a(*)
    call dg_check(cm1g_put_2d_array(dg_model(), id, a), 'cm1g_put_2d_array')
  end subroutine dg_put2

  subroutine dg_get2(id, a)
    integer(c_int), intent(in) :: id
    real, intent(out) :: a(*)
    call dg_check(cm1g_get_2d_array(dg_model(), id, a), 'cm1g_get_2d_array')
  end subroutine dg_get2

  ! a vertical column of a horizontally uniform 3-D array (the reference
  ! keeps its metrics and base state as full arrays, src/cm1.F:779-830):
  ! scalar levels kb:ke, the model's column k = 0..nk+2 (the last level
  ! repeats ke)
  subroutine dg_col_s(id, a)
    use input, only : ib, ie, jb, je, kb, ke, nk
    integer(c_int), intent(in) :: id
    real, intent(in) :: a(ib:ie, jb:je, kb:ke)
    real(c_float) :: col(0:nk + 2)
    col(0:nk + 1) = a(1, 1, kb:ke)
    col(nk + 2) = a(1, 1, ke)
    call dg_check(cm1g_put_column(dg_model(), id, col), 'cm1g_put_column')
  end subroutine dg_col_s

  ! w levels kb:ke+1
  subroutine dg_col_w(id, a)
    use input, only : ib, ie, jb, je, kb, ke, nk
    integer(c_int), intent(in) :: id
    real, intent(in) :: a(ib:ie, jb:je, kb:ke + 1)
    real(c_float) :: col(0:nk + 2)
    col(0:nk + 2) = a(1, 1, kb:ke + 1)
    call dg_check(cm1g_put_column(dg_model(), id, col), 'cm1g_put_column')
  end subroutine dg_col_w

end module cm1g_dropin
