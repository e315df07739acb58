! Drop-in replacement of CM1's turbtend_module (src/turbtend.F): turbs,
! turbt, turbu, turbv and turbw keep the reference's dummy-argument lists
! (src/turbtend.F:18-21, 700-702, 1052-1055, 1816-1819, 2473-2475), so
! solve1.F compiles against this module unchanged, and compute on the GPU
! (cm1g_turbs, cm1g_turbt, cm1g_turbu/v/w) for the subgrid-turbulence branch
! set cm1g_set_turbulence supports: the arrays a routine reads are copied to
! the device model, the tendency it adds to is copied back.  The reference's
! work arrays (turbx..dumz, dum1..dum9, the two-part-model arrays) stay
! untouched: the GPU keeps its own.  Branches outside the GPU path (budgets,
! the LES subdomain of cm1setup = 4, DNS, horizturb) stop the run through
! stopcm1; turbsz, turbuz and turbvz are only called from inside the
! routines above and stop if called directly.
module turbtend_module
  implicit none
  private
  public :: turbs, turbt, turbu, turbv, turbw, turbsz, turbuz, turbvz

contains

  subroutine turbs(iflux,dt,dosfcflx,xh,rxh,arh1,arh2,uh,xf,arf1,arf2,uf,vh,vf,sflux,  &
                   rds,sigma,rdsf,sigmaf,mh,mf,gz,rgz,gzu,rgzu,gzv,rgzv,gx,gxu,gy,gyv, &
                   turbx,turby,turbz,dumx,dumy,dumz,rho,rr,rf,s,sten,khh,khv,cm0,dum7,dum8, &
                   dobud,ibd,ied,jbd,jed,kbd,ked,ndiag,diag,sd_hturb,sd_vturb,ivar)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    integer, intent(in) :: iflux
    real, intent(in) :: dt
    logical, intent(in) :: dosfcflx
    real, intent(in), dimension(ib:ie) :: xh, rxh, arh1, arh2, uh
    real, intent(in), dimension(ib:ie+1) :: xf, arf1, arf2, uf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(jb:je+1) :: vf
    real, intent(in), dimension(ib:ie,jb:je) :: sflux
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(in), dimension(kb:ke+1) :: rdsf, sigmaf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu, rgzu, gzv, rgzv
    real, intent(in), dimension(itb:ite,jtb:jte,ktb:kte) :: gx, gxu, gy, gyv
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: turbx, turby, turbz, dumx, dumy, dumz, sten
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, rr, rf, s
    real, intent(in), dimension(ibc:iec,jbc:jec,kbc:kec) :: khh, khv
    real, intent(in), dimension(ib:ie,jb:je) :: cm0
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: dum7, dum8
    logical, intent(in) :: dobud
    integer, intent(in) :: ibd, ied, jbd, jed, kbd, ked, ndiag, sd_hturb, sd_vturb
    real, intent(inout), dimension(ibd:ied,jbd:jed,kbd:ked,ndiag) :: diag
    integer, intent(in) :: ivar
    integer(c_int) :: fl

    call dg_need(.not. dobud, 'turbs: tendency budgets (dobud) are not on the GPU path')
    call turb_columns(mh, mf)
    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
    call dg_put(CM1G_RR, CM1G_STAG_S, rr)
    call dg_put(CM1G_RF, CM1G_STAG_S, rf)
    call dg_put(CM1G_KHH, CM1G_STAG_W, khh)
    call dg_put(CM1G_KHV, CM1G_STAG_W, khv)
    call dg_put(CM1G_SADV, CM1G_STAG_S, s)
    call dg_put(CM1G_STEN, CM1G_STAG_S, sten)
    call dg_put2(CM1G_D_THFLUX, sflux)
    fl = 0
    if (dosfcflx) fl = 1
    call dg_check(cm1g_turbs(dg_model(), int(iflux, c_int), fl, CM1G_SADV, CM1G_STEN), 'cm1g_turbs')
    call dg_get(CM1G_STEN, CM1G_STAG_S, sten)
  end subroutine turbs

  subroutine turbt(dt,xh,rxh,uh,xf,uf,vh,vf,mh,mf,rho,rr,rf,          &
                   rds,sigma,gz,rgz,gzu,rgzu,gzv,rgzv,                &
                   turbx,turby,turbz,dumx,dumy,dumz,t,tten,kmh,kmv)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    real, intent(in) :: dt
    real, intent(in), dimension(ib:ie) :: xh, rxh, uh
    real, intent(in), dimension(ib:ie+1) :: xf, uf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(jb:je+1) :: vf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, rr, rf
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu, rgzu, gzv, rgzv
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: turbx, turby, turbz, dumx, dumy, dumz
    real, intent(in), dimension(ibt:iet,jbt:jet,kbt:ket) :: t
    real, intent(inout), dimension(ibt:iet,jbt:jet,kbt:ket) :: tten
    real, intent(in), dimension(ibc:iec,jbc:jec,kbc:kec) :: kmh, kmv

    call turb_columns(mh, mf)
    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
    call dg_put(CM1G_RF, CM1G_STAG_S, rf)
    call dg_put(CM1G_KMH, CM1G_STAG_W, kmh)
    call dg_put(CM1G_KMV, CM1G_STAG_W, kmv)
    call dg_put(CM1G_TKEA, CM1G_STAG_W, t)
    call dg_put(CM1G_TKETEN, CM1G_STAG_W, tten)
    call dg_check(cm1g_turbt(dg_model()), 'cm1g_turbt')
    call dg_get(CM1G_TKETEN, CM1G_STAG_W, tten)
  end subroutine turbt

  subroutine turbu(dt,xh,ruh,xf,rxf,arf1,arf2,uf,vh,mh,mf,rmf,rho,rf,  &
                   zs,gz,rgz,gzu,gzv,rds,sigma,rdsf,sigmaf,gxu,     &
                   turbx,turby,turbz,dum1,dum2,dum3,dum7,dum8,dum9,u,uten,w,t11,t12,t13,t22,kmv,cm0, &
                   kmw,ufw,u1b,u2pt,ufwk,doubud,udiag)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    real, intent(in) :: dt
    real, intent(in), dimension(ib:ie) :: xh, ruh
    real, intent(in), dimension(ib:ie+1) :: xf, rxf, arf1, arf2, uf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf, rmf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, rf
    real, intent(in), dimension(ib:ie,jb:je) :: zs
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu, gzv
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(in), dimension(kb:ke+1) :: rdsf, sigmaf
    real, intent(in), dimension(itb:ite,jtb:jte,ktb:kte) :: gxu
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: turbx, turby, turbz, dum1, dum2, dum3, dum7, dum8, dum9
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: u
    real, intent(inout), dimension(ib:ie+1,jb:je,kb:ke) :: uten
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: w
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: t11, t12, t13, t22
    real, intent(in), dimension(ibc:iec,jbc:jec,kbc:kec) :: kmv
    real, intent(in), dimension(ib:ie,jb:je) :: cm0
    real, intent(in), dimension(kb:ke) :: kmw, ufw, u1b
    real, intent(inout), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: u2pt
    real, intent(in), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: ufwk
    logical, intent(in) :: doubud
    real, intent(inout), dimension(ibdv:iedv,jbdv:jedv,kbdv:kedv,nudiag) :: udiag

    call dg_need(.not. doubud, 'turbu: tendency budgets (doubud) are not on the GPU path')
    call turb_columns(mh, mf)
    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
    call dg_put(CM1G_RF, CM1G_STAG_S, rf)
    call dg_put(CM1G_KMV, CM1G_STAG_W, kmv)
    call dg_put(CM1G_UA, CM1G_STAG_U, u)
    call dg_put(CM1G_WA, CM1G_STAG_W, w)
    call dg_put(CM1G_T11, CM1G_STAG_S, t11)
    call dg_put(CM1G_T12, CM1G_STAG_S, t12)
    call dg_put(CM1G_T13, CM1G_STAG_S, t13)
    call dg_put(CM1G_UTEN1, CM1G_STAG_U, uten)
    call dg_check(cm1g_turbu(dg_model()), 'cm1g_turbu')
    call dg_get(CM1G_UTEN1, CM1G_STAG_U, uten)
  end subroutine turbu

  subroutine turbv(dt,xh,rxh,arh1,arh2,uh,xf,rvh,vf,mh,mf,rho,rr,rf,   &
                   zs,gz,rgz,gzu,gzv,rds,sigma,rdsf,sigmaf,gyv,  &
                   turbx,turby,turbz,dum1,dum2,dum3,dum7,dum8,dum9,v,vten,w,t12,t22,t23,kmv,cm0, &
                   kmw,vfw,v1b,v2pt,vfwk,dovbud,vdiag)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    real, intent(in) :: dt
    real, intent(in), dimension(ib:ie) :: xh, rxh, arh1, arh2, uh
    real, intent(in), dimension(ib:ie+1) :: xf
    real, intent(in), dimension(jb:je) :: rvh
    real, intent(in), dimension(jb:je+1) :: vf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, rr, rf
    real, intent(in), dimension(ib:ie,jb:je) :: zs
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu, gzv
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(in), dimension(kb:ke+1) :: rdsf, sigmaf
    real, intent(in), dimension(itb:ite,jtb:jte,ktb:kte) :: gyv
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: turbx, turby, turbz, dum1, dum2, dum3, dum7, dum8, dum9
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: v
    real, intent(inout), dimension(ib:ie,jb:je+1,kb:ke) :: vten
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: w
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: t12, t22, t23
    real, intent(in), dimension(ibc:iec,jbc:jec,kbc:kec) :: kmv
    real, intent(in), dimension(ib:ie,jb:je) :: cm0
    real, intent(in), dimension(kb:ke) :: kmw, vfw, v1b
    real, intent(inout), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: v2pt
    real, intent(in), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: vfwk
    logical, intent(in) :: dovbud
    real, intent(inout), dimension(ibdv:iedv,jbdv:jedv,kbdv:kedv,nvdiag) :: vdiag

    call dg_need(.not. dovbud, 'turbv: tendency budgets (dovbud) are not on the GPU path')
    call turb_columns(mh, mf)
    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
    call dg_put(CM1G_RF, CM1G_STAG_S, rf)
    call dg_put(CM1G_KMV, CM1G_STAG_W, kmv)
    call dg_put(CM1G_VA, CM1G_STAG_V, v)
    call dg_put(CM1G_WA, CM1G_STAG_W, w)
    call dg_put(CM1G_T12, CM1G_STAG_S, t12)
    call dg_put(CM1G_T22, CM1G_STAG_S, t22)
    call dg_put(CM1G_T23, CM1G_STAG_S, t23)
    call dg_put(CM1G_VTEN1, CM1G_STAG_V, vten)
    call dg_check(cm1g_turbv(dg_model()), 'cm1g_turbv')
    call dg_get(CM1G_VTEN1, CM1G_STAG_V, vten)
  end subroutine turbv

  subroutine turbw(dt,xh,rxh,arh1,arh2,uh,xf,vh,mh,mf,rho,rf,gz,rgzu,rgzv,rds,sigma,   &
                   turbx,turby,turbz,dum1,dum2,dum3,w,wten,t13,t23,t33,t22,kmh,cm0,  &
                   dowbud,wdiag)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    real, intent(in) :: dt
    real, intent(in), dimension(ib:ie) :: xh, rxh, arh1, arh2, uh
    real, intent(in), dimension(ib:ie+1) :: xf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, rf
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgzu, rgzv
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: turbx, turby, turbz, dum1, dum2, dum3
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: w
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: wten
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: t13, t23, t33, t22
    real, intent(in), dimension(ibc:iec,jbc:jec,kbc:kec) :: kmh
    real, intent(in), dimension(ib:ie,jb:je) :: cm0
    logical, intent(in) :: dowbud
    real, intent(inout), dimension(ibdv:iedv,jbdv:jedv,kbdv:kedv,nwdiag) :: wdiag

    call dg_need(.not. dowbud, 'turbw: tendency budgets (dowbud) are not on the GPU path')
    call turb_columns(mh, mf)
    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
    call dg_put(CM1G_RF, CM1G_STAG_S, rf)
    call dg_put(CM1G_KMH, CM1G_STAG_W, kmh)
    call dg_put(CM1G_WA, CM1G_STAG_W, w)
    call dg_put(CM1G_T13, CM1G_STAG_S, t13)
    call dg_put(CM1G_T23, CM1G_STAG_S, t23)
    call dg_put(CM1G_T33, CM1G_STAG_S, t33)
    call dg_put(CM1G_WTEN1, CM1G_STAG_W, wten)
    call dg_check(cm1g_turbw(dg_model()), 'cm1g_turbw')
    call dg_get(CM1G_WTEN1, CM1G_STAG_W, wten)
  end subroutine turbw

  subroutine turbsz(iflux,dt,dosfcflx,sflux,mh,mf,  &
                    turbz,dum1,dum2,dumz,rho,rr,rf,s,khv,dum7,dum8,ivar)
    use input
    use cm1g_dropin
    implicit none
    integer, intent(in) :: iflux
    real, intent(in) :: dt
    logical, intent(in) :: dosfcflx
    real, intent(in), dimension(ib:ie,jb:je) :: sflux
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: turbz, dum1, dum2, dumz
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, rr, rf, s
    real, intent(in), dimension(ibc:iec,jbc:jec,kbc:kec) :: khv
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: dum7, dum8
    integer, intent(in) :: ivar
    call dg_need(.false., 'turbsz: only the vertical part of turbs, which runs it on the GPU')
  end subroutine turbsz

  subroutine turbuz(dt,xh,ruh,xf,rxf,arf1,arf2,uf,vh,mh,mf,rmf,rho,rf,  &
                    zs,gz,rgz,gzu,gzv,rds,sigma,rdsf,sigmaf,gxu,     &
                    turbz,dum1,dum2,dum3,dum7,dum9,u,w,t11,t12,t13,t22,kmv, &
                    kmw,ufw,u1b,u2pt,ufwk)
    use input
    use cm1g_dropin
    implicit none
    real, intent(in) :: dt
    real, intent(in), dimension(ib:ie) :: xh, ruh
    real, intent(in), dimension(ib:ie+1) :: xf, rxf, arf1, arf2, uf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf, rmf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, rf
    real, intent(in), dimension(ib:ie,jb:je) :: zs
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu, gzv
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(in), dimension(kb:ke+1) :: rdsf, sigmaf
    real, intent(in), dimension(itb:ite,jtb:jte,ktb:kte) :: gxu
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: turbz, dum1, dum2, dum3, dum7, dum9
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: u
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: w
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: t11, t12, t13, t22
    real, intent(in), dimension(ibc:iec,jbc:jec,kbc:kec) :: kmv
    real, intent(in), dimension(kb:ke) :: kmw, ufw, u1b
    real, intent(inout), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: u2pt
    real, intent(in), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: ufwk
    call dg_need(.false., 'turbuz: only the vertical part of turbu, which runs it on the GPU')
  end subroutine turbuz

  subroutine turbvz(dt,xh,rxh,arh1,arh2,uh,xf,rvh,vf,mh,mf,rho,rr,rf,   &
                    zs,gz,rgz,gzu,gzv,rds,sigma,rdsf,sigmaf,gyv,  &
                    turbz,dum1,dum2,dum3,dum7,dum9,v,w,t12,t22,t23,kmv, &
                    kmw,vfw,v1b,v2pt,vfwk)
    use input
    use cm1g_dropin
    implicit none
    real, intent(in) :: dt
    real, intent(in), dimension(ib:ie) :: xh, rxh, arh1, arh2, uh
    real, intent(in), dimension(ib:ie+1) :: xf
    real, intent(in), dimension(jb:je) :: rvh
    real, intent(in), dimension(jb:je+1) :: vf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, rr, rf
    real, intent(in), dimension(ib:ie,jb:je) :: zs
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu, gzv
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(in), dimension(kb:ke+1) :: rdsf, sigmaf
    real, intent(in), dimension(itb:ite,jtb:jte,ktb:kte) :: gyv
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: turbz, dum1, dum2, dum3, dum7, dum9
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: v
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: w
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: t12, t22, t23
    real, intent(in), dimension(ibc:iec,jbc:jec,kbc:kec) :: kmv
    real, intent(in), dimension(kb:ke) :: kmw, vfw, v1b
    real, intent(inout), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: v2pt
    real, intent(in), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: vfwk
    call dg_need(.false., 'turbvz: only the vertical part of turbv, which runs it on the GPU')
  end subroutine turbvz

  ! the vertical metric columns the turbulence kernels read
  subroutine turb_columns(mh, mf)
    use input, only : ib, ie, jb, je, kb, ke
    use cm1gpu
    use cm1g_dropin
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    call dg_col_s(CM1G_K_MH, mh)
    call dg_col_w(CM1G_K_MF, mf)
  end subroutine turb_columns

end module turbtend_module
