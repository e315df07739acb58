! Drop-in replacement of CM1's adv_module (src/adv.F): advs, advu, advv and
! advw keep the reference's dummy-argument lists (src/adv.F:14-19, 911-912,
! 1847-1848, 2778-2781), so solve2.F and pdcomp.F compile against this module
! unchanged, and compute on the GPU (cm1g_advs, cm1g_advu/v/w): the arrays a
! routine reads are copied to the device model, the tendency it writes is
! copied back.  The work arrays of the reference (dumx..subs, advx..advz,
! mass, flag, the MPI buffers) stay untouched: the GPU keeps its own.
! advw also runs solve2's TKE call (iamw = .false.) through cm1g_advw_scalar.
! Branches outside the GPU path (budgets and diagnostics, hyperdiffusion in
! advs, advection orders other than 5th/WENO5) stop the run through stopcm1.
module adv_module
  implicit none
  private
  public :: advs, advu, advv, advw

contains

  subroutine advs(nrk,wflag,bflag,bsq,xh,rxh,arh1,arh2,uh,ruh,xf,vh,rvh,gz,rgz,mh,rmh,     &
                  rho0,rr0,rf0,rrf0,advx,advy,advz,dumx,dumy,dumz,mass,subs,divx,         &
                  rru,rrv,rrw,s0,s,sten,pdef,pdefweno,dt,weps,                            &
                  flag,sw31,sw32,se31,se32,ss31,ss32,sn31,sn32,rdsf,c1,c2,rho,ri,diffit,  &
                  dobud,ibd,ied,jbd,jed,kbd,ked,ndiag,diag,sd_hadv,sd_vadv,sd_lsw,        &
                  sd_hidiff,sd_vidiff,sd_hediff,wprof,dumk1,dumk2,hadvorder,vadvorder,    &
                  kbdy,bndy,hflxw,hflxe,hflxs,hflxn,out3d)
    use, intrinsic :: iso_c_binding
    use input
    use ib_module, only : ibib, ieib, jbib, jeib, kbib, keib, kmaxib
    use cm1gpu
    use cm1g_dropin
    implicit none
    integer, intent(in) :: nrk, wflag, bflag
    double precision, intent(inout) :: bsq
    real, intent(in), dimension(ib:ie) :: xh, rxh, arh1, arh2, uh, ruh
    real, intent(in), dimension(ib:ie+1) :: xf
    real, intent(in), dimension(jb:je) :: vh, rvh
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh, rmh, rho0, rr0, rf0, rrf0
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: advx, advy, advz, dumx, dumy, dumz, mass, subs
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: divx
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: rru
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: rrv
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: rrw
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: s0, s
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: sten
    integer, intent(in) :: pdef, pdefweno
    real, intent(in) :: dt
    double precision, intent(in) :: weps
    logical, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: flag
    real, intent(inout), dimension(cmp,jmp,kmp) :: sw31, sw32, se31, se32
    real, intent(inout), dimension(imp,cmp,kmp) :: ss31, ss32, sn31, sn32
    real, intent(in), dimension(kb:ke+1) :: rdsf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: c1, c2
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, ri
    integer, intent(in) :: diffit
    logical, intent(in) :: dobud
    integer, intent(in) :: ibd, ied, jbd, jed, kbd, ked, ndiag, sd_hadv, sd_vadv, sd_lsw, &
                           sd_hidiff, sd_vidiff, sd_hediff
    real, intent(inout), dimension(ibd:ied,jbd:jed,kbd:ked,ndiag) :: diag
    real, intent(in), dimension(kb:ke) :: wprof
    double precision, intent(inout), dimension(kb:ke) :: dumk1, dumk2
    integer, intent(in) :: hadvorder, vadvorder
    integer, intent(in), dimension(ibib:ieib,jbib:jeib) :: kbdy
    logical, intent(in), dimension(ibib:ieib,jbib:jeib,kbib:keib) :: bndy
    integer, intent(in), dimension(ibib:ieib,jbib:jeib,kmaxib) :: hflxw, hflxe, hflxs, hflxn
    real, intent(inout), dimension(ib3d:ie3d,jb3d:je3d,kb3d:ke3d,nout3d) :: out3d
    type(c_ptr) :: m

    call dg_need(bflag == 0, 'advs: boundary-flux budgets (bflag = 1) are not on the GPU path')
    call dg_need(diffit == 0, 'advs: sixth-order diffusion (diffit = 1) is not on the GPU path')
    call dg_need(.not. dobud, 'advs: tendency budgets (dobud) are not on the GPU path')
    call dg_need(hadvorder == 5 .and. vadvorder == 5, 'advs: advection orders other than 5')
    call dg_need(wflag == 1 .or. advwenos == 0, 'advs: wflag = 0 with WENO scalar advection')
    m = dg_model()
    call adv_columns(mh, rho0, rr0, rf0, rrf0, c1, c2)
    call dg_put(CM1G_RRU, CM1G_STAG_U, rru)
    call dg_put(CM1G_RRV, CM1G_STAG_V, rrv)
    call dg_put(CM1G_RRW, CM1G_STAG_W, rrw)
    call dg_put(CM1G_DIVX, CM1G_STAG_S, divx)
    call dg_put(CM1G_DUM1, CM1G_STAG_S, s0)
    call dg_put(CM1G_DUM2, CM1G_STAG_S, s)
    call dg_put(CM1G_DUM3, CM1G_STAG_S, sten)
    call dg_check(cm1g_advs(m, int(nrk, c_int), CM1G_DUM2, CM1G_DUM1, CM1G_DUM3, int(pdef, c_int), &
                            int(pdefweno, c_int), real(dt, c_float), real(weps, c_double)), 'cm1g_advs')
    call dg_get(CM1G_DUM3, CM1G_STAG_S, sten)
  end subroutine advs

  subroutine advu(nrk,arh1,arh2,uh,xf,rxf,arf1,arf2,uf,vh,gz,rgz,gzu,mh,rho0,rr0,rf0,rrf0,  &
                  dumx,dumy,dumz,advx,advy,advz,subs,divx,                                 &
                  rru,u3d,uten,rrv,rrw,rdsf,c1,c2,rho,dt,doubud,udiag,wprof,bndy,kbdy,     &
                  hflxw,hflxe,hflxs,hflxn,vf,mf,v3d,w3d)
    use, intrinsic :: iso_c_binding
    use input
    use ib_module, only : ibib, ieib, jbib, jeib, kbib, keib, kmaxib
    use cm1gpu
    use cm1g_dropin
    implicit none
    integer, intent(in) :: nrk
    real, intent(in), dimension(ib:ie) :: arh1, arh2, uh
    real, intent(in), dimension(ib:ie+1) :: xf, rxf, arf1, arf2, uf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh, rho0, rr0, rf0, rrf0
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: dumx, dumy, dumz, advx, advy, advz, subs
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: divx
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: rru, u3d
    real, intent(inout), dimension(ib:ie+1,jb:je,kb:ke) :: uten
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: rrv
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: rrw
    real, intent(in), dimension(kb:ke+1) :: rdsf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: c1, c2
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho
    real, intent(in) :: dt
    logical, intent(in) :: doubud
    real, intent(inout), dimension(ibdv:iedv,jbdv:jedv,kbdv:kedv,nudiag) :: udiag
    real, intent(in), dimension(kb:ke) :: wprof
    logical, intent(in), dimension(ibib:ieib,jbib:jeib,kbib:keib) :: bndy
    integer, intent(in), dimension(ibib:ieib,jbib:jeib) :: kbdy
    integer, intent(in), dimension(ibib:ieib,jbib:jeib,kmaxib) :: hflxw, hflxe, hflxs, hflxn
    real, intent(in), dimension(jb:je+1) :: vf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: v3d
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: w3d

    call dg_need(.not. doubud, 'advu: tendency budgets (doubud) are not on the GPU path')
    call dg_need(hadvordrv == 5 .and. vadvordrv == 5, 'advu: advection orders other than 5')
    call adv_columns(mh, rho0, rr0, rf0, rrf0, c1, c2)
    call dg_col_w(CM1G_K_MF, mf)
    call momentum_in(rru, rrv, rrw, divx, u3d, v3d, w3d, rho)
    call dg_put(CM1G_UTEN, CM1G_STAG_U, uten)
    call dg_check(cm1g_advu(dg_model(), int(nrk, c_int)), 'cm1g_advu')
    call dg_get(CM1G_UTEN, CM1G_STAG_U, uten)
  end subroutine advu

  subroutine advv(nrk,xh,rxh,arh1,arh2,uh,xf,vh,vf,gz,rgz,gzv,mh,rho0,rr0,rf0,rrf0,       &
                  dumx,dumy,dumz,advx,advy,advz,subs,divx,                                &
                  rru,rrv,v3d,vten,rrw,rdsf,c1,c2,rho,dt,dovbud,vdiag,wprof,bndy,kbdy,    &
                  hflxw,hflxe,hflxs,hflxn,uf,mf,u3d,w3d)
    use, intrinsic :: iso_c_binding
    use input
    use ib_module, only : ibib, ieib, jbib, jeib, kbib, keib, kmaxib
    use cm1gpu
    use cm1g_dropin
    implicit none
    integer, intent(in) :: nrk
    real, intent(in), dimension(ib:ie) :: xh, rxh, arh1, arh2, uh
    real, intent(in), dimension(ib:ie+1) :: xf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(jb:je+1) :: vf
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzv
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh, rho0, rr0, rf0, rrf0
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: dumx, dumy, dumz, advx, advy, advz, subs
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: divx
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: rru
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: rrv, v3d
    real, intent(inout), dimension(ib:ie,jb:je+1,kb:ke) :: vten
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: rrw
    real, intent(in), dimension(kb:ke+1) :: rdsf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: c1, c2
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho
    real, intent(in) :: dt
    logical, intent(in) :: dovbud
    real, intent(inout), dimension(ibdv:iedv,jbdv:jedv,kbdv:kedv,nvdiag) :: vdiag
    real, intent(in), dimension(kb:ke) :: wprof
    logical, intent(in), dimension(ibib:ieib,jbib:jeib,kbib:keib) :: bndy
    integer, intent(in), dimension(ibib:ieib,jbib:jeib) :: kbdy
    integer, intent(in), dimension(ibib:ieib,jbib:jeib,kmaxib) :: hflxw, hflxe, hflxs, hflxn
    real, intent(in), dimension(ib:ie+1) :: uf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: u3d
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: w3d

    call dg_need(.not. dovbud, 'advv: tendency budgets (dovbud) are not on the GPU path')
    call dg_need(hadvordrv == 5 .and. vadvordrv == 5, 'advv: advection orders other than 5')
    call adv_columns(mh, rho0, rr0, rf0, rrf0, c1, c2)
    call dg_col_w(CM1G_K_MF, mf)
    call momentum_in(rru, rrv, rrw, divx, u3d, v3d, w3d, rho)
    call dg_put(CM1G_VTEN, CM1G_STAG_V, vten)
    call dg_check(cm1g_advv(dg_model(), int(nrk, c_int)), 'cm1g_advv')
    call dg_get(CM1G_VTEN, CM1G_STAG_V, vten)
  end subroutine advv

  subroutine advw(nrk,xh,rxh,arh1,arh2,uh,xf,vh,gz,rgz,mh,mf,rho0,rr0,rf0,rrf0,           &
                  dumx,dumy,dumz,advx,advy,advz,subs,divx,                                &
                  rru,rrv,rrw,w3d,wten,rds,rdsf,c1,c2,rho,dt,                             &
                  dowbud,wdiag,hadvorder,vadvorder,advweno,bndy,kbdy,uf,vf,u3d,v3d,       &
                  hflxw,hflxe,hflxs,hflxn,iamw)
    use, intrinsic :: iso_c_binding
    use input
    use ib_module, only : ibib, ieib, jbib, jeib, kbib, keib, kmaxib
    use cm1gpu
    use cm1g_dropin
    implicit none
    integer, intent(in) :: nrk
    real, intent(in), dimension(ib:ie) :: xh, rxh, arh1, arh2, uh
    real, intent(in), dimension(ib:ie+1) :: xf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho0, rr0, rf0, rrf0
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: dumx, dumy, dumz, advx, advy, advz, subs
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: divx
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: rru
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: rrv
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: rrw, w3d
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: wten
    real, intent(in), dimension(kb:ke) :: rds
    real, intent(in), dimension(kb:ke+1) :: rdsf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: c1, c2
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho
    real, intent(in) :: dt
    logical, intent(in) :: dowbud
    real, intent(inout), dimension(ibdv:iedv,jbdv:jedv,kbdv:kedv,nwdiag) :: wdiag
    integer, intent(in) :: hadvorder, vadvorder, advweno
    logical, intent(in), dimension(ibib:ieib,jbib:jeib,kbib:keib) :: bndy
    integer, intent(in), dimension(ibib:ieib,jbib:jeib) :: kbdy
    real, intent(in), dimension(ib:ie+1) :: uf
    real, intent(in), dimension(jb:je+1) :: vf
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: u3d
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: v3d
    integer, intent(in), dimension(ibib:ieib,jbib:jeib,kmaxib) :: hflxw, hflxe, hflxs, hflxn
    logical, intent(in) :: iamw

    call dg_need(.not. dowbud, 'advw: tendency budgets (dowbud) are not on the GPU path')
    call dg_need(hadvorder == 5 .and. vadvorder == 5, 'advw: advection orders other than 5')
    call adv_columns(mh, rho0, rr0, rf0, rrf0, c1, c2)
    call dg_col_w(CM1G_K_MF, mf)
    if (.not. iamw) then
      ! a w-level scalar, the subgrid TKE (src/solve2.F:1746-1749): the mass
      ! fluxes as getdivx left them, the scalars' WENO switch
      call dg_put(CM1G_RRU, CM1G_STAG_U, rru)
      call dg_put(CM1G_RRV, CM1G_STAG_V, rrv)
      call dg_put(CM1G_RRW, CM1G_STAG_W, rrw)
      call dg_put(CM1G_DIVX, CM1G_STAG_S, divx)
      call dg_put(CM1G_DUM9, CM1G_STAG_W, w3d)
      call dg_put(CM1G_WTEN, CM1G_STAG_W, wten)
      call dg_check(cm1g_advw_scalar(dg_model(), int(nrk, c_int), int(advweno, c_int), CM1G_DUM9, CM1G_WTEN), &
                    'cm1g_advw_scalar')
      call dg_get(CM1G_WTEN, CM1G_STAG_W, wten)
      return
    endif
    call dg_need(advweno == advwenov, 'advw: a WENO switch other than advwenov')
    call momentum_in(rru, rrv, rrw, divx, u3d, v3d, w3d, rho)
    call dg_put(CM1G_WTEN, CM1G_STAG_W, wten)
    call dg_check(cm1g_advw(dg_model(), int(nrk, c_int)), 'cm1g_advw')
    call dg_get(CM1G_WTEN, CM1G_STAG_W, wten)
  end subroutine advw

  ! the vertical metric and base-state columns an advection routine reads
  subroutine adv_columns(mh, rho0, rr0, rf0, rrf0, c1, c2)
    use input, only : ib, ie, jb, je, kb, ke
    use cm1gpu
    use cm1g_dropin
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh, rho0, rr0, rf0, rrf0, c1, c2
    call dg_col_s(CM1G_K_MH, mh)
    call dg_col_s(CM1G_K_RHO0, rho0)
    call dg_col_s(CM1G_K_RR0, rr0)
    call dg_col_s(CM1G_K_RF0, rf0)
    call dg_col_s(CM1G_K_RRF0, rrf0)
    call dg_col_s(CM1G_K_C1, c1)
    call dg_col_s(CM1G_K_C2, c2)
  end subroutine adv_columns

  ! the mass fluxes, divergence and velocities of a momentum advection
  subroutine momentum_in(rru, rrv, rrw, divx, u3d, v3d, w3d, rho)
    use input, only : ib, ie, jb, je, kb, ke
    use cm1gpu
    use cm1g_dropin
    real, intent(in) :: rru(ib:ie+1,jb:je,kb:ke), rrv(ib:ie,jb:je+1,kb:ke), rrw(ib:ie,jb:je,kb:ke+1)
    real, intent(in) :: divx(ib:ie,jb:je,kb:ke), rho(ib:ie,jb:je,kb:ke)
    real, intent(in) :: u3d(ib:ie+1,jb:je,kb:ke), v3d(ib:ie,jb:je+1,kb:ke), w3d(ib:ie,jb:je,kb:ke+1)
    call dg_put(CM1G_RRU, CM1G_STAG_U, rru)
    call dg_put(CM1G_RRV, CM1G_STAG_V, rrv)
    call dg_put(CM1G_RRW, CM1G_STAG_W, rrw)
    call dg_put(CM1G_DIVX, CM1G_STAG_S, divx)
    call dg_put(CM1G_U3D, CM1G_STAG_U, u3d)
    call dg_put(CM1G_V3D, CM1G_STAG_V, v3d)
    call dg_put(CM1G_W3D, CM1G_STAG_W, w3d)
    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
  end subroutine momentum_in

end module adv_module
