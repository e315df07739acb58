! The host<->device moves shared by the drop-in sound and sounde
! (sound_module.f90, sounde_module.f90): both read the same arrays of
! solve2's call (src/solve2.F:1103-1137) and write the same ones.
module cm1g_dropin_acoustic
  use, intrinsic :: iso_c_binding
  use cm1gpu
  use cm1g_dropin
  implicit none
  private
  public :: acoustic_columns, acoustic_in, acoustic_out

contains

  ! the vertical metric and base-state columns of the substeps
  subroutine acoustic_columns(mh, c1, c2, mf, pi0, rho0, rr0, rf0, rrf0, th0)
    use input, only : ib, ie, jb, je, kb, ke
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh, c1, c2, pi0, rho0, rr0, rf0, rrf0, th0
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    call dg_col_s(CM1G_K_MH, mh)
    call dg_col_s(CM1G_K_C1, c1)
    call dg_col_s(CM1G_K_C2, c2)
    call dg_col_w(CM1G_K_MF, mf)
    call dg_col_s(CM1G_K_PI0, pi0)
    call dg_col_s(CM1G_K_RHO0, rho0)
    call dg_col_s(CM1G_K_RR0, rr0)
    call dg_col_s(CM1G_K_RF0, rf0)
    call dg_col_s(CM1G_K_RRF0, rrf0)
    call dg_col_s(CM1G_K_TH0, th0)
  end subroutine acoustic_columns

  ! what the substeps read: the radiative boundary speeds, the a and 3d
  ! time levels, the fixed tendencies, thv (t11), ppterm (t22) and
  ! piadv (sadv) of solve2's call
  subroutine acoustic_in(radbcw, radbce, radbcs, radbcn, rru, ua, u3d, uten, rrv, va, v3d, vten, &
                         rrw, wa, w3d, wten, ppi, pp3d, piadv, ppten, ppx, thv, ppterm)
    use input, only : ib, ie, jb, je, kb, ke
    real, intent(in), dimension(jb:je,kb:ke) :: radbcw, radbce
    real, intent(in), dimension(ib:ie,kb:ke) :: radbcs, radbcn
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: rru, ua, u3d, uten
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: rrv, va, v3d, vten
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: rrw, wa, w3d, wten
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: ppi, pp3d, piadv, ppten, ppx, thv, ppterm
    call dg_put2(CM1G_D_RADBCW, radbcw)
    call dg_put2(CM1G_D_RADBCE, radbce)
    call dg_put2(CM1G_D_RADBCS, radbcs)
    call dg_put2(CM1G_D_RADBCN, radbcn)
    call dg_put(CM1G_RRU, CM1G_STAG_U, rru)
    call dg_put(CM1G_UA, CM1G_STAG_U, ua)
    call dg_put(CM1G_U3D, CM1G_STAG_U, u3d)
    call dg_put(CM1G_UTEN, CM1G_STAG_U, uten)
    call dg_put(CM1G_RRV, CM1G_STAG_V, rrv)
    call dg_put(CM1G_VA, CM1G_STAG_V, va)
    call dg_put(CM1G_V3D, CM1G_STAG_V, v3d)
    call dg_put(CM1G_VTEN, CM1G_STAG_V, vten)
    call dg_put(CM1G_RRW, CM1G_STAG_W, rrw)
    call dg_put(CM1G_WA, CM1G_STAG_W, wa)
    call dg_put(CM1G_W3D, CM1G_STAG_W, w3d)
    call dg_put(CM1G_WTEN, CM1G_STAG_W, wten)
    call dg_put(CM1G_PPI, CM1G_STAG_S, ppi)
    call dg_put(CM1G_PP3D, CM1G_STAG_S, pp3d)
    call dg_put(CM1G_SADV, CM1G_STAG_S, piadv)
    call dg_put(CM1G_PPTEN, CM1G_STAG_S, ppten)
    call dg_put(CM1G_PPX, CM1G_STAG_S, ppx)
    call dg_put(CM1G_T11, CM1G_STAG_S, thv)
    call dg_put(CM1G_T22, CM1G_STAG_S, ppterm)
  end subroutine acoustic_in

  ! what the substeps write, and dtu/dtv = dts*dtu0/dtv0 as the reference
  ! leaves them (src/sound.F:84-113), with dts of the stage: dt/(3 nloop),
  ! nloop = nint(nsound/3) (one more if that step would be too long), on
  ! the first stage, dt/nsound on the others
  subroutine acoustic_out(nrk, dt, dtu, dtu0, dtv, dtv0, rru, u3d, rrv, v3d, rrw, w3d, pp3d, ppx)
    use input, only : ib, ie, jb, je, kb, ke, ni, nj, nsound
    integer, intent(in) :: nrk
    real, intent(in) :: dt
    real, intent(inout), dimension(ib:ie,jb:je) :: dtu, dtv
    real, intent(in), dimension(ib:ie,jb:je) :: dtu0, dtv0
    real, intent(inout), dimension(ib:ie+1,jb:je,kb:ke) :: rru, u3d
    real, intent(inout), dimension(ib:ie,jb:je+1,kb:ke) :: rrv, v3d
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: rrw, w3d
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: pp3d, ppx
    integer :: nloop
    real :: dts
    call dg_get(CM1G_RRU, CM1G_STAG_U, rru)
    call dg_get(CM1G_U3D, CM1G_STAG_U, u3d)
    call dg_get(CM1G_RRV, CM1G_STAG_V, rrv)
    call dg_get(CM1G_V3D, CM1G_STAG_V, v3d)
    call dg_get(CM1G_RRW, CM1G_STAG_W, rrw)
    call dg_get(CM1G_W3D, CM1G_STAG_W, w3d)
    call dg_get(CM1G_PP3D, CM1G_STAG_S, pp3d)
    call dg_get(CM1G_PPX, CM1G_STAG_S, ppx)
    if (nrk == 1) then
      nloop = nint(float(nsound) / 3.0)
      dts = dt / (nloop * 3.0)
      if (dts > (dt / nsound)) then
        nloop = nloop + 1
        dts = dt / (nloop * 3.0)
      endif
    else
      dts = dt / nsound
    endif
    dtu(1:ni+1, 1:nj+1) = dts * dtu0(1:ni+1, 1:nj+1)
    dtv(1:ni+1, 1:nj+1) = dts * dtv0(1:ni+1, 1:nj+1)
  end subroutine acoustic_out

end module cm1g_dropin_acoustic
