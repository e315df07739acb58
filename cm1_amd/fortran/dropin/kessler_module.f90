! Drop-in replacement of CM1's kessler_module (src/kessler.F): kessler,
! satadj, k_fallout, fallout, getefall and geterain keep the reference's
! dummy-argument lists (src/kessler.F:8-10, 146-147, 368, 404-405, 485,
! 530), so mp_driver.F compiles against this module unchanged, and compute
! on the GPU (cm1g_kessler, cm1g_satadj, cm1g_k_fallout, cm1g_fallout,
! cm1g_geterain): the arrays a routine reads are copied to the device
! model, the ones it writes are copied back.  The rain budget (train of
! fallout) and the rain-energy budget (erain of geterain) are formed on the
! device as the reference forms them; the process budgets of kessler and
! satadj (tauto, taccr, tevar, tcond, tevac, which feed only statpack's
! stat_qsrc output) are not, so a run with stat_qsrc = 1 stops.  Branches
! outside the GPU path (efall = 1 through getefall, satadj outside the
! nrk = 4 form, translating rain with imove = 1) stop the run through
! stopcm1.
module kessler_module
  implicit none
  private
  public :: kessler, satadj, k_fallout, fallout, getefall, geterain

contains

  subroutine kessler(dt,tauto,taccr,tevar,ruh,rvh,rmh,pi0,th0,tmp,   &
                     rho,rr,pp3d,th3d,prs,                           &
                     qv3d,qc3d,qr3d)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    real :: dt
    double precision :: tauto, taccr, tevar
    real, dimension(ib:ie) :: ruh
    real, dimension(jb:je) :: rvh
    real, dimension(ib:ie,jb:je,kb:ke) :: rmh, pi0, th0
    real, dimension(ib:ie,jb:je,kb:ke) :: tmp, rho, rr, pp3d, th3d, prs, qv3d, qc3d, qr3d

    call dg_need(stat_qsrc == 0, 'kessler: the process budgets (stat_qsrc = 1) are not on the GPU path')
    call dg_need(dt == dtl, 'kessler: dt differs from the namelist dtl')
    call dg_col_s(CM1G_K_PI0, pi0)
    call dg_col_s(CM1G_K_TH0, th0)
    call thermo_in(rho, pp3d, th3d, prs, qv3d, qc3d, qr3d)
    call dg_check(cm1g_kessler(dg_model()), 'cm1g_kessler')
    call thermo_out(pp3d, th3d, prs, qv3d, qc3d, qr3d)
  end subroutine kessler

  subroutine satadj(nrk,dt,tcond,tevac,ruh,rvh,rmh,pi0,th0,   &
                    rho,rr,pp3d,prs,th3d,q3d)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    integer nrk
    real, intent(in) :: dt
    double precision :: tcond, tevac
    real, dimension(ib:ie) :: ruh
    real, dimension(jb:je) :: rvh
    real, dimension(ib:ie,jb:je,kb:ke) :: rmh, pi0, th0
    real, dimension(ib:ie,jb:je,kb:ke) :: rho, rr, pp3d, prs, th3d
    real, dimension(ibm:iem,jbm:jem,kbm:kem,numq) :: q3d

    call dg_need(nrk == 4, 'satadj: only the end-of-step form (nrk = 4) is on the GPU path')
    call dg_need(stat_qsrc == 0, 'satadj: the process budgets (stat_qsrc = 1) are not on the GPU path')
    call dg_need(ibm == ib .and. iem == ie .and. jbm == jb .and. jem == je .and. kbm == kb .and. kem == ke, &
                 'satadj: moisture arrays with bounds other than the scalars''')
    call dg_col_s(CM1G_K_PI0, pi0)
    call dg_col_s(CM1G_K_TH0, th0)
    call thermo_in(rho, pp3d, th3d, prs, q3d(:,:,:,nqv), q3d(:,:,:,nqc), q3d(:,:,:,nqr))
    call dg_check(cm1g_satadj(dg_model()), 'cm1g_satadj')
    call thermo_out(pp3d, th3d, prs, q3d(:,:,:,nqv), q3d(:,:,:,nqc), q3d(:,:,:,nqr))
  end subroutine satadj

  subroutine k_fallout(rho,qr3d,vr)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    real, dimension(ib:ie,jb:je,kb:ke) :: rho, qr3d, vr

    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
    call dg_put(CM1G_Q3D2, CM1G_STAG_S, qr3d)
    call dg_put(CM1G_QTEN2, CM1G_STAG_S, vr)
    call dg_check(cm1g_k_fallout(dg_model()), 'cm1g_k_fallout')
    call dg_get(CM1G_QTEN2, CM1G_STAG_S, vr)
  end subroutine k_fallout

  subroutine fallout(dt,train,ruh,rvh,zh,mh,mf,rain,prate,rr,rho,   &
                     q3d,vq)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    real :: dt
    double precision :: train
    real, dimension(ib:ie) :: ruh
    real, dimension(jb:je) :: rvh
    real, dimension(ib:ie,jb:je,kb:ke) :: zh, mh
    real, dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, dimension(ib:ie,jb:je,nrain) :: rain
    real, intent(inout), dimension(ib:ie,jb:je) :: prate
    real, dimension(ib:ie,jb:je,kb:ke) :: rr, rho, q3d, vq
    real(c_double) :: tr

    call dg_need(nrain == 1, 'fallout: translated rain (nrain = 2) is not on the GPU path')
    call dg_need(dt == dtl, 'fallout: dt differs from the namelist dtl')
    call dg_col_s(CM1G_K_MH, mh)
    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
    call dg_put(CM1G_Q3D2, CM1G_STAG_S, q3d)
    call dg_put(CM1G_QTEN2, CM1G_STAG_S, vq)
    call dg_put2(CM1G_D_RAIN, rain)
    call dg_put2(CM1G_D_PRATE, prate)
    tr = train
    call dg_check(cm1g_fallout(dg_model(), tr), 'cm1g_fallout')
    train = tr
    call dg_get(CM1G_Q3D2, CM1G_STAG_S, q3d)
    call dg_get(CM1G_QTEN2, CM1G_STAG_S, vq)
    call dg_get2(CM1G_D_RAIN, rain)
    call dg_get2(CM1G_D_PRATE, prate)
  end subroutine fallout

  subroutine getefall(setzero,cpx,mf,t,cvm,tten,q3d,vr)
    use input
    use cm1g_dropin
    implicit none
    integer, intent(in) :: setzero
    real, intent(in) :: cpx
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: t, cvm
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: tten
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: q3d, vr

    call dg_need(.false., 'getefall: the energy flux of falling rain (efall = 1) is not on the GPU path')
  end subroutine getefall

  subroutine geterain(dt,cpx,lx1,erain,ruh,rvh,t,rr,q3d,vr)
    use, intrinsic :: iso_c_binding
    use input
    use cm1gpu
    use cm1g_dropin
    implicit none
    real :: dt, cpx, lx1
    double precision :: erain
    real, dimension(ib:ie) :: ruh
    real, dimension(jb:je) :: rvh
    real, dimension(ib:ie,jb:je,kb:ke) :: t, rr
    real, dimension(ib:ie,jb:je,kb:ke) :: q3d, vr
    real(c_double) :: e

    call dg_put(CM1G_DUM1, CM1G_STAG_S, t)
    call dg_put(CM1G_DUM3, CM1G_STAG_S, rr)
    call dg_put(CM1G_DUM4, CM1G_STAG_S, q3d)
    call dg_put(CM1G_QTEN2, CM1G_STAG_S, vr)
    e = erain
    call dg_check(cm1g_geterain(dg_model(), real(dt, c_float), real(cpx, c_float), real(lx1, c_float), &
                                CM1G_DUM1, CM1G_DUM3, CM1G_DUM4, e), 'cm1g_geterain')
    erain = e
  end subroutine geterain

  ! the thermodynamic state kessler and satadj read
  subroutine thermo_in(rho, pp3d, th3d, prs, qv, qc, qr)
    use input, only : ib, ie, jb, je, kb, ke
    use cm1gpu
    use cm1g_dropin
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, pp3d, th3d, prs, qv, qc, qr
    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
    call dg_put(CM1G_PP3D, CM1G_STAG_S, pp3d)
    call dg_put(CM1G_TH3D, CM1G_STAG_S, th3d)
    call dg_put(CM1G_PRS, CM1G_STAG_S, prs)
    call dg_put(CM1G_Q3D0, CM1G_STAG_S, qv)
    call dg_put(CM1G_Q3D1, CM1G_STAG_S, qc)
    call dg_put(CM1G_Q3D2, CM1G_STAG_S, qr)
  end subroutine thermo_in

  ! and write
  subroutine thermo_out(pp3d, th3d, prs, qv, qc, qr)
    use input, only : ib, ie, jb, je, kb, ke
    use cm1gpu
    use cm1g_dropin
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: pp3d, th3d, prs, qv, qc, qr
    call dg_get(CM1G_PP3D, CM1G_STAG_S, pp3d)
    call dg_get(CM1G_TH3D, CM1G_STAG_S, th3d)
    call dg_get(CM1G_PRS, CM1G_STAG_S, prs)
    call dg_get(CM1G_Q3D0, CM1G_STAG_S, qv)
    call dg_get(CM1G_Q3D1, CM1G_STAG_S, qc)
    call dg_get(CM1G_Q3D2, CM1G_STAG_S, qr)
  end subroutine thermo_out

end module kessler_module
