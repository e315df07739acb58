! Drop-in replacement of CM1's turb_module (src/turb.F): sfc_and_turb keeps
! the reference's dummy-argument list (src/turb.F:12-67), so cm1.F compiles
! against this module unchanged, and prepares the subgrid turbulence of the
! next step on the GPU (cm1g_sfc_and_turb) for the branch set
! cm1g_set_turbulence supports: it copies in the state the set-up reads and
! copies back what the rest of the step and the output read (km, kh, tkea
! with its surface value, the t_ij stresses, N^2, the deformation, the
! dissipation and length scale, u1, v1, s1, ust and the surface heat flux;
! zntmp and the boundary levels of tke3d are the host copies the reference
! makes).  The surface diagnostics of other surface models and PBL schemes
! (psfc, hfx, u10, cd, ch, ...) are not formed.  getepst, tkebc, calcnm,
! calcdef and turbsmag, called from domaindiag.F and init3d.F only for
! domain diagnostics and re-initialised restarts, keep their dummy lists
! and stop the run if called.
module turb_module
  implicit none
  private
  public :: sfc_and_turb, getepst, tkebc, calcnm, calcdef, turbsmag

contains

  subroutine sfc_and_turb(getsfc,getpbl,nstep,dt,dosfcflx,cloudvar,qbudget,    &
                   avgsfcu,avgsfcv,avgsfcs,avgsfcsu,avgsfcsv,avgsfct,avgsfcq,avgsfcp, &
                   xh,rxh,arh1,arh2,uh,ruh,xf,rxf,arf1,arf2,uf,ruf,  &
                   yh,vh,rvh,yf,vf,rvf,                              &
                   rds,sigma,rdsf,sigmaf,zh,mh,rmh,c1,c2,zf,mf,rmf,  &
                   pi0s,rth0s,pi0,rho0,prs0,thv0,th0,qv0,rr0,rf0,rrf0, &
                   zs,gz,rgz,gzu,rgzu,gzv,rgzv,gx,gxu,gy,gyv,        &
                   tsk,thflux,qvflux,cd,ch,cq,u1,v1,s1,t1,tlh,f2d,ustt,ut,vt,st,cm0,   &
                   dum1,dum2,dum3,dum4,dum5,dum6,dum7,dum8,dum9,     &
                   divx,rho,rr,rf,prs,                               &
                   t11,t12,t13,t22,t23,t33,                          &
                   m11,m12,m13,m22,m23,m33,                          &
                   u0,rru,ua ,ugr ,fsu  ,                            &
                   v0,rrv,va ,vgr ,fsv  ,                            &
                   rrw,wa ,dumw1,dumw2,                              &
                   ppi ,ppten,sten,sadv,                             &
                   tha ,thten,thten1,thterm,qa ,                     &
                   kmh,kmv,khh,khv,cme,csm,ce1,ce2,tkea,tke3d,       &
                   nm,defv,defh,lenscl,dissten,epst,epsd1,epsd2,radsw,radswnet,radlwin, &
                   thpten,qvpten,qcpten,qipten,upten,vpten,qnipten,qncpten,qrpten,qspten,qgpten, &
                   xkzh,xkzq,xkzm,                                   &
                   tsq,qsq,cov,sh3d,sm3d,el_pbl,qc_bl,qi_bl,cldfra_bl,       &
                   qWT,qSHEAR,qBUOY,qDISS,dqke,qke_adv,qke,qke3d,            &
                   edmf_a,edmf_w,edmf_qt,edmf_thl,edmf_ent,edmf_qc,          &
                   sub_thl3D,sub_sqv3D,det_thl3D,det_sqv3D,                  &
                   vdfg,maxmf,nupdraft,ktop_plume,                           &
                   lwten,swten,thraten,tdiag,raincv,pratecv,                 &
                   thcuten,qvcuten,qccuten,qicuten,ucuten,vcuten,thften,qvften, &
                   lu_index,kpbl2d,psfc,u10,v10,s10,hfx,qfx,         &
                   xland,znt,rznt,ust,stau,tst,qst,z0t,z0q,          &
                   hpbl,wspd,phim,phih,psim,psih,psiq,gz1oz0,br,brcr, &
                   CHS,CHS2,CQS2,CPMM,ZOL,MAVAIL,                    &
                   MOL,RMOL,REGIME,LH,FLHC,FLQC,QGH,                 &
                   CK,CKA,CDA,USTM,QSFC,T2,Q2,TH2,EMISS,THC,ALBD,    &
                   gsw,glw,chklowq,capg,snowc,snowh,qcg,dsxy,wstar,delta,prkpp,fm,fh,  &
                   charn,msang,scurx,scury,zkmax,cd_out,ch_out,wscale,wscaleu, &
                   zntmp,mznt,swspd,smois,taux,tauy,hpbl2d,evap2d,heat2d,  &
                   mixht,akhs,akms,ct,snow,sice,thz0,qz0,uz0,vz0,u10e,v10e,th10,q10,tshltr,qshltr,pshltr,z0base,zntmyj,lowlyr,ivgtyp,tke_myj,el_myj,tmp2d,tmp1d,  &
                   num_soil_layers,slab_zs,slab_dzs,tslb,tmn,        &
                   tml,t0ml,hml,h0ml,huml,hvml,tmoml,                &
                   cavg,uavg,vavg,savg,gamk,gamwall,kmw,ufw,vfw,u1b,v1b,l2p,s2p,s2b,t2pm1,t2pm2,t2pm3,t2pm4,u2pt,v2pt,kmwk,ufwk,vfwk, &
                   bndy,kbdy,timavg,sfctimavg,                       &
                   reqs_u,reqs_v,reqs_w,reqs_s,reqs_p,reqs_x,reqs_y, &
                   nw1,nw2,ne1,ne2,sw1,sw2,se1,se2,                  &
                   pw1,pw2,pe1,pe2,ps1,ps2,pn1,pn2,                  &
                   vw1,vw2,ve1,ve2,vs1,vs2,vn1,vn2,                  &
                   uw31,uw32,ue31,ue32,us31,us32,un31,un32,          &
                   sw31,sw32,se31,se32,ss31,ss32,sn31,sn32,          &
                   kw1,kw2,ke1,ke2,ks1,ks2,kn1,kn2,reqk,             &
                   iamsat,out2d,out3d,rtime,update_sfc,dotbud,dotdwrite,restarted,  &
                   dowriteout,doazimwrite,dostat)
    use, intrinsic :: iso_c_binding
    use input
    use ib_module, only : ibib, ieib, jbib, jeib, kbib, keib
    use cm1gpu
    use cm1g_dropin
    implicit none
    logical, intent(in) :: getsfc, getpbl
    integer, intent(in) :: nstep
    real, intent(inout) :: dt
    logical, intent(in) :: dosfcflx
    logical, intent(in), dimension(maxq) :: cloudvar
    double precision, intent(inout), dimension(nbudget) :: qbudget
    double precision, intent(inout) :: avgsfcu, avgsfcv, avgsfcs, avgsfcsu, avgsfcsv, avgsfct, avgsfcq, avgsfcp
    real, intent(in), dimension(ib:ie) :: xh, rxh, arh1, arh2, uh, ruh
    real, intent(in), dimension(ib:ie+1) :: xf, rxf, arf1, arf2, uf, ruf
    real, intent(in), dimension(jb:je) :: yh, vh, rvh
    real, intent(in), dimension(jb:je+1) :: yf, vf, rvf
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(in), dimension(kb:ke+1) :: rdsf, sigmaf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: zh, mh, rmh, c1, c2
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: zf, mf, rmf
    real, intent(in), dimension(ib:ie,jb:je) :: pi0s, rth0s
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: pi0, rho0, prs0, thv0, th0, qv0, rr0, rf0, rrf0
    real, intent(in), dimension(ib:ie,jb:je) :: zs
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu, rgzu, gzv, rgzv
    real, intent(in), dimension(itb:ite,jtb:jte,ktb:kte) :: gx, gxu, gy, gyv
    real, intent(inout), dimension(ib:ie,jb:je) :: tsk, znt, rznt, zntmp, ust, stau, tst, qst, z0t, z0q, &
                                                   thflux, qvflux, cd, ch, cq, u1, v1, s1, t1, psfc, tlh, &
                                                   ustt, ut, vt, st, cm0
    real, intent(in), dimension(ib:ie,jb:je) :: xland, f2d
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: dum1, dum2, dum3, dum4, dum5, dum6, dum7, dum8, dum9, &
                                                         divx, rho, rr, rf, prs
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: t11, t12, t13, t22, t23, t33
    real, intent(inout), dimension(ibnba:ienba,jbnba:jenba,kbnba:kenba) :: m11, m12, m13, m22, m23, m33
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: u0
    real, intent(inout), dimension(ib:ie+1,jb:je,kb:ke) :: rru, ua, ugr, fsu
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: v0
    real, intent(inout), dimension(ib:ie,jb:je+1,kb:ke) :: rrv, va, vgr, fsv
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: rrw, wa, dumw1, dumw2
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: ppi, ppten, sten, sadv
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: tha, thten, thten1, thterm
    real, intent(inout), dimension(ibm:iem,jbm:jem,kbm:kem,numq) :: qa
    real, intent(inout), dimension(ibc:iec,jbc:jec,kbc:kec) :: kmh, kmv, khh, khv
    real, intent(inout), dimension(ibc:iec,jbc:jec,kbc:kec) :: cme, csm, ce1, ce2
    real, intent(inout), dimension(ibt:iet,jbt:jet,kbt:ket) :: tkea, tke3d
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: nm, defv, defh, lenscl, dissten, epst, epsd1, epsd2
    real, intent(inout), dimension(ni,nj) :: radsw, radswnet, radlwin
    real, intent(inout), dimension(ibb:ieb,jbb:jeb,kbb:keb) :: thpten, qvpten, qcpten, qipten, upten, vpten, &
                                                               qnipten, qncpten, qrpten, qspten, qgpten
    real, intent(inout), dimension(ibb:ieb,jbb:jeb,kbb:keb) :: xkzh, xkzq, xkzm
    real, intent(inout), dimension(ibmynn:iemynn,jbmynn:jemynn,kbmynn:kemynn) :: tsq, qsq, cov, sh3d, sm3d, &
         el_pbl, qc_bl, qi_bl, cldfra_bl, qWT, qSHEAR, qBUOY, qDISS, dqke, qke_adv, qke, qke3d, edmf_a, edmf_w, &
         edmf_qt, edmf_thl, edmf_ent, edmf_qc, sub_thl3D, sub_sqv3D, det_thl3D, det_sqv3D
    real, intent(inout), dimension(ibmynn:iemynn,jbmynn:jemynn) :: vdfg, maxmf
    integer, intent(inout), dimension(ibmynn:iemynn,jbmynn:jemynn) :: nupdraft, ktop_plume
    real, intent(inout), dimension(ibr:ier,jbr:jer,kbr:ker) :: swten, lwten, thraten
    real, intent(inout), dimension(ibdt:iedt,jbdt:jedt,kbdt:kedt,ntdiag) :: tdiag
    real, intent(inout), dimension(ib:ie,jb:je,nrain) :: raincv
    real, intent(inout), dimension(ib:ie,jb:je) :: pratecv
    real, intent(inout), dimension(ibcu:iecu,jbcu:jecu,kbcu:kecu) :: thcuten, qvcuten, qccuten, qicuten, &
                                                                     ucuten, vcuten, thften, qvften
    integer, intent(inout), dimension(ibl:iel,jbl:jel) :: lu_index, kpbl2d
    real, intent(inout), dimension(ibl:iel,jbl:jel) :: u10, v10, s10, hfx, qfx, hpbl, wspd, phim, phih, psim, &
         psih, psiq, gz1oz0, br, brcr, CHS, CHS2, CQS2, CPMM, ZOL, MAVAIL, MOL, RMOL, REGIME, LH, FLHC, FLQC, &
         QGH, CK, CKA, CDA, USTM, QSFC, T2, Q2, TH2, EMISS, THC, ALBD, gsw, glw, chklowq, capg, snowc, snowh, &
         qcg, dsxy, wstar, delta, prkpp, fm, fh
    real, intent(inout), dimension(ibl:iel,jbl:jel) :: charn, msang, scurx, scury, zkmax, cd_out, ch_out, &
                                                       wscale, wscaleu
    real, intent(inout), dimension(ibl:iel,jbl:jel) :: mznt, swspd, smois, taux, tauy, hpbl2d, evap2d, heat2d
    real, intent(inout), dimension(ibmyj:iemyj,jbmyj:jemyj) :: mixht, akhs, akms, ct, snow, sice, thz0, qz0, &
         uz0, vz0, u10e, v10e, th10, q10, tshltr, qshltr, pshltr, z0base, zntmyj
    integer, intent(inout), dimension(ibmyj:iemyj,jbmyj:jemyj) :: lowlyr, ivgtyp
    real, intent(inout), dimension(ibmyj:iemyj,jbmyj:jemyj,kbmyj:kemyj) :: tke_myj, el_myj
    real, intent(inout), dimension(ibtmp:ietmp,kbtmp:ketmp,ntmp2d) :: tmp2d
    real, intent(inout), dimension(ibtmp:ietmp,ntmp1d) :: tmp1d
    integer, intent(in) :: num_soil_layers
    real, intent(in), dimension(num_soil_layers) :: slab_zs, slab_dzs
    real, intent(inout), dimension(ibl:iel,jbl:jel,num_soil_layers) :: tslb
    real, intent(inout), dimension(ibl:iel,jbl:jel) :: tmn, tml, t0ml, hml, h0ml, huml, hvml, tmoml
    double precision, intent(inout), dimension(kb:ke,3+numq) :: cavg
    real, intent(inout), dimension(kb:ke) :: uavg, vavg, savg, gamk, gamwall, kmw, ufw, vfw, u1b, v1b, l2p, &
                                             s2p, s2b, t2pm1, t2pm2, t2pm3, t2pm4
    real, intent(inout), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: u2pt, v2pt
    real, intent(inout), dimension(ib2pt:ie2pt,jb2pt:je2pt,kb2pt:ke2pt) :: kmwk, ufwk, vfwk
    integer, intent(inout), dimension(rmp) :: reqs_u, reqs_v, reqs_w, reqs_s, reqs_p, reqs_x, reqs_y
    real, intent(inout), dimension(kmt) :: nw1, nw2, ne1, ne2, sw1, sw2, se1, se2
    real, intent(inout), dimension(jmp,kmp) :: pw1, pw2, pe1, pe2
    real, intent(inout), dimension(imp,kmp) :: ps1, ps2, pn1, pn2
    real, intent(inout), dimension(jmp,kmp) :: vw1, vw2, ve1, ve2
    real, intent(inout), dimension(imp,kmp) :: vs1, vs2, vn1, vn2
    real, intent(inout), dimension(cmp,jmp,kmp) :: uw31, uw32, ue31, ue32
    real, intent(inout), dimension(imp+1,cmp,kmp) :: us31, us32, un31, un32
    real, intent(inout), dimension(cmp,jmp,kmp) :: sw31, sw32, se31, se32
    real, intent(inout), dimension(imp,cmp,kmp) :: ss31, ss32, sn31, sn32
    real, intent(inout), dimension(jmp,kmt,4) :: kw1, kw2, ke1, ke2
    real, intent(inout), dimension(imp,kmt,4) :: ks1, ks2, kn1, kn2
    integer, intent(inout) :: reqk
    logical, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: iamsat
    real, intent(inout), dimension(ib2d:ie2d,jb2d:je2d,nout2d) :: out2d
    real, intent(inout), dimension(ib3d:ie3d,jb3d:je3d,kb3d:ke3d,nout3d) :: out3d
    logical, intent(in), dimension(ibib:ieib,jbib:jeib,kbib:keib) :: bndy
    integer, intent(in), dimension(ibib:ieib,jbib:jeib) :: kbdy
    real, intent(in), dimension(ibta:ieta,jbta:jeta,kbta:keta,ntavr) :: timavg
    real, intent(in), dimension(ibta:ieta,jbta:jeta,nsfctavr) :: sfctimavg
    real, intent(in) :: rtime
    logical, intent(in) :: update_sfc, dotbud, dotdwrite, restarted, dowriteout, doazimwrite, dostat
    type(c_ptr) :: m
    integer(c_int) :: gs
    integer :: i, j, k

    if (cm1setup == 0 .and. ipbl == 0 .and. horizturb == 0 .and. idiss == 0 .and. output_dissten == 0) then
      ! no subgrid model: the reference's set-up then only forms the
      ! surface pressure and the ground-relative winds (src/turb.F:290-333);
      ! N^2 and the deformation it forms for output only are not formed
      call dg_need(bbc /= 3, 'sfc_and_turb: a surface layer without a subgrid model is not on the GPU path')
      call dg_need(output_nm == 0 .and. output_def == 0, &
                   'sfc_and_turb: output of N^2 or deformation without a subgrid model is not on the GPU path')
      do j = 1, nj
        do i = 1, ni
          psfc(i, j) = cgs1 * prs(i, j, 1) + cgs2 * prs(i, j, 2) + cgs3 * prs(i, j, 3)
        enddo
      enddo
      do k = 1, nk
        do j = jb, je
          do i = ib, ie
            if (imove == 1) then
              ugr(i, j, k) = ua(i, j, k) + umove
              vgr(i, j, k) = va(i, j, k) + vmove
            elseif (pertflx == 1) then
              ugr(i, j, k) = ua(i, j, k) - u0(i, j, k)
              vgr(i, j, k) = va(i, j, k) - v0(i, j, k)
            else
              ugr(i, j, k) = ua(i, j, k)
              vgr(i, j, k) = va(i, j, k)
            endif
          enddo
        enddo
      enddo
      return
    endif
    call dg_need(imoist == 0, 'sfc_and_turb: moist subgrid turbulence is not on the GPU path')
    call dg_need(imove == 0 .and. pertflx == 0, 'sfc_and_turb: imove or pertflx with turbulence')
    call dg_need(testcase < 1 .or. testcase > 7, &
                 'sfc_and_turb: the PBL depth of the LES test cases (testcase 1-7) is not on the GPU path')
    call dg_need(.not. dotdwrite .or. .not. dodomaindiag, &
                 'sfc_and_turb: the turbulence diagnostics of domaindiag are not on the GPU path')
    m = dg_model()
    ! the vertical columns, then the turbulence set-up that reads zh and rmf
    call dg_col_s(CM1G_K_ZH, zh)
    call dg_col_w(CM1G_K_RMF, rmf)
    call dg_col_s(CM1G_K_MH, mh)
    call dg_col_w(CM1G_K_MF, mf)
    call dg_col_s(CM1G_K_C1, c1)
    call dg_col_s(CM1G_K_C2, c2)
    call dg_col_s(CM1G_K_TH0, th0)
    call dg_set_turbulence(.true.)
    call dg_put(CM1G_UA, CM1G_STAG_U, ua)
    call dg_put(CM1G_VA, CM1G_STAG_V, va)
    call dg_put(CM1G_WA, CM1G_STAG_W, wa)
    call dg_put(CM1G_THA, CM1G_STAG_S, tha)
    call dg_put(CM1G_RHO, CM1G_STAG_S, rho)
    call dg_put(CM1G_RF, CM1G_STAG_S, rf)
    call dg_put(CM1G_TKEA, CM1G_STAG_W, tkea)
    gs = 1
    if (.not. getsfc) then
      gs = 0
      call dg_put2(CM1G_D_UST, ust)
      call dg_put2(CM1G_D_U1, u1)
      call dg_put2(CM1G_D_V1, v1)
      call dg_put2(CM1G_D_S1, s1)
      call dg_put2(CM1G_D_THFLUX, thflux)
    endif
    call dg_check(cm1g_sfc_and_turb(m, gs), 'cm1g_sfc_and_turb')
    call dg_get(CM1G_KMH, CM1G_STAG_W, kmh)
    call dg_get(CM1G_KMV, CM1G_STAG_W, kmv)
    call dg_get(CM1G_KHH, CM1G_STAG_W, khh)
    call dg_get(CM1G_KHV, CM1G_STAG_W, khv)
    call dg_get(CM1G_TKEA, CM1G_STAG_W, tkea)
    call dg_get(CM1G_T11, CM1G_STAG_S, t11)
    call dg_get(CM1G_T12, CM1G_STAG_S, t12)
    call dg_get(CM1G_T13, CM1G_STAG_S, t13)
    call dg_get(CM1G_T22, CM1G_STAG_S, t22)
    call dg_get(CM1G_T23, CM1G_STAG_S, t23)
    call dg_get(CM1G_T33, CM1G_STAG_S, t33)
    call dg_get(CM1G_NM, CM1G_STAG_W, nm)
    call dg_get(CM1G_DEFV, CM1G_STAG_W, defv)
    call dg_get(CM1G_DEFH, CM1G_STAG_W, defh)
    call dg_get(CM1G_LENSCL, CM1G_STAG_W, lenscl)
    call dg_get(CM1G_DISSTEN, CM1G_STAG_W, dissten)
    ! with getsfc false the halos of ust, u1, v1 and s1 are still filled
    ! (src/turb.F:1134-1160)
    call dg_get2(CM1G_D_UST, ust)
    call dg_get2(CM1G_D_U1, u1)
    call dg_get2(CM1G_D_V1, v1)
    call dg_get2(CM1G_D_S1, s1)
    if (getsfc) call dg_get2(CM1G_D_THFLUX, thflux)
    ! the host copies of src/turb.F:1236 and :1383-1388, the surface
    ! pressure and the ground-relative winds (:290-333)
    zntmp = znt
    do j = 1, nj
      do i = 1, ni
        psfc(i, j) = cgs1 * prs(i, j, 1) + cgs2 * prs(i, j, 2) + cgs3 * prs(i, j, 3)
      enddo
    enddo
    do k = 1, nk
      do j = jb, je
        do i = ib, ie
          ugr(i, j, k) = ua(i, j, k)
          vgr(i, j, k) = va(i, j, k)
        enddo
      enddo
    enddo
    do j = 0, nj + 1
      do i = 0, ni + 1
        tke3d(i, j, 1) = tkea(i, j, 1)
        tke3d(i, j, nk + 1) = tkea(i, j, nk + 1)
      enddo
    enddo
  end subroutine sfc_and_turb

  subroutine getepst(xh,rxh,uh,xf,rxf,uf,vh,vf,mh,c1,c2,mf,ua ,va ,wa ,  &
                     t11,t12,t13,t22,t23,t33,rf,                         &
                     dum1,dum2,epst,epstp,epstn)
    use input
    use cm1g_dropin
    implicit none
    real, intent(in), dimension(ib:ie) :: xh, rxh, uh
    real, intent(in), dimension(ib:ie+1) :: xf, rxf, uf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(jb:je+1) :: vf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh, c1, c2
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: ua
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: va
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: wa
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: t11, t12, t13, t22, t23, t33, rf
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: dum1, dum2
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: epst
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke), optional :: epstp, epstn
    call dg_need(.false., 'getepst: the subgrid dissipation diagnostic is not on the GPU path')
  end subroutine getepst

  subroutine calcnm(c1,c2,mf,pi0,thv0,th0,cloudvar,nm,t,qt,thv,cloud,rh,qvci,   &
                    prs,pp,th,qa,iamsat)
    use input
    use cm1g_dropin
    implicit none
    logical, intent(in), dimension(maxq) :: cloudvar
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: c1, c2
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: pi0, thv0, th0
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: nm
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: t, qt, thv, cloud, rh, qvci
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: prs, pp, th
    real, intent(in), dimension(ibm:iem,jbm:jem,kbm:kem,numq) :: qa
    logical, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: iamsat
    call dg_need(.false., 'calcnm outside sfc_and_turb (re-initialised restarts) is not on the GPU path')
  end subroutine calcnm

  subroutine calcdef(rds,sigma,rdsf,sigmaf,zs,gz,rgz,gzu,rgzu,gzv,rgzv,                &
                     xh,rxh,arh1,arh2,uh,xf,rxf,arf1,arf2,uf,vh,vf,mh,c1,c2,mf,defv,defh,  &
                     dum1,dum2,ua,va,wa,s11,s12,s13,s22,s23,s33,gx,gy,rho,rr,rf)
    use input
    use cm1g_dropin
    implicit none
    real, intent(in), dimension(kb:ke) :: rds, sigma
    real, intent(in), dimension(kb:ke+1) :: rdsf, sigmaf
    real, intent(in), dimension(ib:ie,jb:je) :: zs
    real, intent(in), dimension(itb:ite,jtb:jte) :: gz, rgz, gzu, rgzu, gzv, rgzv
    real, intent(in), dimension(ib:ie) :: xh, rxh, arh1, arh2, uh
    real, intent(in), dimension(ib:ie+1) :: xf, rxf, arf1, arf2, uf
    real, intent(in), dimension(jb:je) :: vh
    real, intent(in), dimension(jb:je+1) :: vf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: mh, c1, c2
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: defv, defh
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: dum1, dum2
    real, intent(in), dimension(ib:ie+1,jb:je,kb:ke) :: ua
    real, intent(in), dimension(ib:ie,jb:je+1,kb:ke) :: va
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: wa
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: s11, s12, s13, s22, s23, s33
    real, intent(in), dimension(itb:ite,jtb:jte,ktb:kte) :: gx, gy
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rho, rr, rf
    call dg_need(.false., 'calcdef outside sfc_and_turb (re-initialised restarts) is not on the GPU path')
  end subroutine calcdef

  subroutine turbsmag(nstep,dt,ruh,rvh,rmh,mf,rmf,th0,rf,              &
                      nm,defv,defh,dum4,grdscl,lenh  ,zf,zntmp,ust,csm, &
                      kmh,kmv,khh,khv,lenscl,dissten,                  &
                      nw1,nw2,ne1,ne2,sw1,sw2,se1,se2,                 &
                      khcw1,khcw2,khce1,khce2,khcs1,khcs2,khcn1,khcn2, &
                      kvcw1,kvcw2,kvce1,kvce2,kvcs1,kvcs2,kvcn1,kvcn2)
    use input
    use cm1g_dropin
    implicit none
    integer, intent(in) :: nstep
    real, intent(in) :: dt
    real, intent(in), dimension(ib:ie) :: ruh
    real, intent(in), dimension(jb:je) :: rvh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rmh
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: mf, rmf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: th0
    real, intent(in), dimension(ib:ie,jb:je,kb:ke) :: rf
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: nm, defv, defh
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke) :: dum4, grdscl, lenh
    real, intent(in), dimension(ib:ie,jb:je) :: zntmp, ust
    real, intent(in), dimension(ibc:iec,jbc:jec,kbc:kec) :: csm
    real, intent(in), dimension(ib:ie,jb:je,kb:ke+1) :: zf
    real, intent(inout), dimension(ibc:iec,jbc:jec,kbc:kec) :: kmh, kmv, khh, khv
    real, intent(inout), dimension(ib:ie,jb:je,kb:ke+1) :: lenscl, dissten
    real, intent(inout), dimension(kmt) :: nw1, nw2, ne1, ne2, sw1, sw2, se1, se2
    real, intent(inout), dimension(jmp,kmt) :: khcw1, khcw2, khce1, khce2
    real, intent(inout), dimension(imp,kmt) :: khcs1, khcs2, khcn1, khcn2
    real, intent(inout), dimension(jmp,kmt) :: kvcw1, kvcw2, kvce1, kvce2
    real, intent(inout), dimension(imp,kmt) :: kvcs1, kvcs2, kvcn1, kvcn2
    call dg_need(.false., 'turbsmag (Smagorinsky) is not on the GPU path')
  end subroutine turbsmag

  subroutine tkebc(tkea,kmh,kmv,khh,khv,zntmp,ust,ustt,              &
                   khcw1,khcw2,khce1,khce2,khcs1,khcs2,khcn1,khcn2,  &
                   kvcw1,kvcw2,kvce1,kvce2,kvcs1,kvcs2,kvcn1,kvcn2)
    use input
    use cm1g_dropin
    implicit none
    real, intent(inout), dimension(ibt:iet,jbt:jet,kbt:ket) :: tkea
    real, intent(inout), dimension(ibc:iec,jbc:jec,kbc:kec) :: kmh, kmv, khh, khv
    real, intent(in), dimension(ib:ie,jb:je) :: zntmp, ust, ustt
    real, intent(inout), dimension(jmp,kmt) :: khcw1, khcw2, khce1, khce2
    real, intent(inout), dimension(imp,kmt) :: khcs1, khcs2, khcn1, khcn2
    real, intent(inout), dimension(jmp,kmt) :: kvcw1, kvcw2, kvce1, kvce2
    real, intent(inout), dimension(imp,kmt) :: kvcs1, kvcs2, kvcn1, kvcn2
    call dg_need(.false., 'tkebc outside sfc_and_turb is not on the GPU path')
  end subroutine tkebc

end module turb_module
