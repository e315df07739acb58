! Drop-in replacement of CM1's sound_module (src/sound.F): sound keeps the
! reference's dummy-argument list (src/sound.F:10-24), so solve2.F compiles
! against this module unchanged, and runs the acoustic substeps of one RK
! stage on the GPU (cm1g_sound, psolver = 3): the arrays the substeps read
! are copied to the device model and the arrays they write (u3d, v3d, w3d,
! pp3d, the time-averaged rru/rrv/rrw and, on the last stage, ppx) are
! copied back.  The reference's work arrays (dum1..dum4, mm, bk, ppd, fpk,
! qk, pk1, pk2, wtens and the MPI buffers) stay untouched: the GPU keeps its
! own.
module sound_module
  implicit none
  private
  public :: sound

contains

  subroutine sound(dt,xh,arh1,arh2,uh,ruh,xf,uf,yh,vh,rvh,yf,vf,     &
                   rds,sigma,rdsf,sigmaf,zh,mh,rmh,c1,c2,mf,zf,      &
                   pi0,rho0,rr0,rf0,rrf0,th0,rth0,zs,                &
                   gz,rgz,gzu,rgzu,gzv,rgzv,                         &
                   dzdx,dzdy,gx,gxu,gy,gyv,                          &
                   radbcw,radbce,radbcs,radbcn,dtu,dtu0,dtv,dtv0,    &
                   dum1,dum2,dum3,dum4,mm,bk,                        &
                   ppd,fpk,qk,pk1,pk2,wtens,                         &
                   u0,rru,ua,u3d,uten,                               &
                   v0,rrv,va,v3d,vten,                               &
                   rrw,wa,w3d,wten,                                  &
                   ppi,pp3d,piadv,ppten,ppx,                         &
                   thv,ppterm,nrk,dttmp,rtime,mtime,get_time_avg,    &
                   bndy,kbdy,                                        &
                   pw1,pw2,pe1,pe2,ps1,ps2,pn1,pn2,reqs_p)
    use, intrinsic :: iso_c_binding
    use input
    use ib_module, only : ibib, ieib, jbib, jeib, kbib, keib
    use cm1gpu
    use cm1g_dropin
    use cm1g_dropin_acoustic
    implicit none
    real, intent(in) :: dt
    real, intent(in), dimension(ib:ie) :: xh, arh1, arh2, uh, ruh
    real, intent(in), dimension(ib:ie+1) :: xf, uf
    real, intent(in), dimension(jb:je) :: yh, vh, rvh
    real, intent(in), dimension(jb:je+1) :: yf, vf
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(in), dimension(kb:ke+1) :: rdsf, sigmaf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: zh, mh, rmh, c1, c2
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf, zf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: pi0, rho0, rr0, rf0, rrf0, th0, rth0
    real, intent(in), dimension(ib:ie,jb:je) :: zs
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu, rgzu, gzv, rgzv, dzdx, dzdy
    real, intent(in), dimension(itb:ite,jtb:jte,ktb:kte) :: gx, gxu, gy, gyv
    real, intent(inout), dimension(jb:je,kb:ke) :: radbcw, radbce
    real, intent(inout), dimension(ib:ie,kb:ke) :: radbcs, radbcn
    real, intent(inout), dimension(ib:ie,jb:je) :: dtu, dtv
    real, intent(in), dimension(ib:ie,jb:je) :: dtu0, dtv0
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: dum1, dum2, dum3, dum4, mm, bk
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: ppd, fpk, qk, pk1, pk2, wtens
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: u0
    real, intent(inout), dimension(ib:ie+1,jb:je,kb:ke) :: rru, ua, u3d, uten
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: v0
    real, intent(inout), dimension(ib:ie,jb:je+1,kb:ke) :: rrv, va, v3d, vten
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: rrw, wa, w3d, wten
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: ppi, pp3d, piadv, ppten, ppx
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: thv, ppterm
    integer, intent(in) :: nrk
    real, intent(in) :: dttmp, rtime
    double precision, intent(in) :: mtime
    logical, intent(in) :: get_time_avg
    logical, intent(in), dimension(ibib:ieib,jbib:jeib,kbib:keib) :: bndy
    integer, intent(in), dimension(ibib:ieib,jbib:jeib) :: kbdy
    real, intent(inout), dimension(jmp,kmp) :: pw1, pw2, pe1, pe2
    real, intent(inout), dimension(imp,kmp) :: ps1, ps2, pn1, pn2
    integer, intent(inout), dimension(rmp) :: reqs_p

    call dg_need(psolver == 3, 'sound: called with psolver other than 3')
    call dg_need(get_time_avg, 'sound: the GPU substeps always form the time averages')
    call dg_need(dt == dtl, 'sound: dt differs from the namelist dtl')
    call acoustic_columns(mh, c1, c2, mf, pi0, rho0, rr0, rf0, rrf0, th0)
    call acoustic_in(radbcw, radbce, radbcs, radbcn, rru, ua, u3d, uten, rrv, va, v3d, vten, &
                     rrw, wa, w3d, wten, ppi, pp3d, piadv, ppten, ppx, thv, ppterm)
    call dg_check(cm1g_sound(dg_model(), int(nrk, c_int), real(dttmp, c_float)), 'cm1g_sound')
    call acoustic_out(nrk, dt, dtu, dtu0, dtv, dtv0, rru, u3d, rrv, v3d, rrw, w3d, pp3d, ppx)
  end subroutine sound

end module sound_module
