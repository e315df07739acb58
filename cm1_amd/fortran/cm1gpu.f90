! ISO_C_BINDING interface of the cm1gpu C ABI (include/cm1gpu.h) for the
! CM1 Fortran host.  A host that keeps its own namelist, allocation and
! output code calls these in place of solve1/solve2/mp_driver/solve3
! (src/cm1.F:3313-3476) or of the individual routines (advs, sound, ...);
! every entry returns 0 or a CM1G_E_* code (message: cm1g_error_message).
!
! Array arguments of cm1g_put_array/cm1g_get_array use the reference's own
! shapes: scalars (ib:ie,jb:je,kb:ke), u (ib:ie+1,...), v (...,jb:je+1,...),
! w (...,kb:ke+1) with ngxy=3, ngz=1 (src/cm1.F:747-777).
module cm1gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_strlen_

  type, bind(C) :: cm1g_config
    integer(c_int32_t) :: ni, nj, nk
    integer(c_int32_t) :: wbc, ebc, sbc, nbc
    integer(c_int32_t) :: imoist, ptype, psolver, advwenos, advwenov
    integer(c_int32_t) :: apmasscon, irbc, nsound, npt, pdtra
    real(c_float) :: dx, dy, dz, dt, alph, kdiv, csmax
    integer(c_int32_t) :: irdamp
    real(c_float) :: rdalpha, zd
  end type cm1g_config

  ! LES subgrid-turbulence switches (cm1g_set_turbulence)
  type, bind(C) :: cm1g_turb_config
    integer(c_int32_t) :: cm1setup, doimpl, bbc, isfcflx
    real(c_float) :: cnst_shflx, cnst_znt
  end type cm1g_turb_config

  ! return codes
  integer(c_int), parameter :: CM1G_OK = 0, CM1G_E_ARG = 1, CM1G_E_HIP = 2, &
                               CM1G_E_NOMEM = 3, CM1G_E_UNSUPPORTED = 4, &
                               CM1G_E_SATADJ = 5, CM1G_E_CFL = 6, CM1G_E_IO = 7
  ! GrADS output switches of cm1g_write_grads (the namelist's output_* names)
  integer(c_int32_t), parameter :: CM1G_OUT_RAIN      = 1
  integer(c_int32_t), parameter :: CM1G_OUT_ZH        = 2
  integer(c_int32_t), parameter :: CM1G_OUT_TH        = 4
  integer(c_int32_t), parameter :: CM1G_OUT_THPERT    = 8
  integer(c_int32_t), parameter :: CM1G_OUT_PRS       = 16
  integer(c_int32_t), parameter :: CM1G_OUT_PRSPERT   = 32
  integer(c_int32_t), parameter :: CM1G_OUT_PI        = 64
  integer(c_int32_t), parameter :: CM1G_OUT_PIPERT    = 128
  integer(c_int32_t), parameter :: CM1G_OUT_RHO       = 256
  integer(c_int32_t), parameter :: CM1G_OUT_RHOPERT   = 512
  integer(c_int32_t), parameter :: CM1G_OUT_QV        = 1024
  integer(c_int32_t), parameter :: CM1G_OUT_QVPERT    = 2048
  integer(c_int32_t), parameter :: CM1G_OUT_Q         = 4096
  integer(c_int32_t), parameter :: CM1G_OUT_BUOYANCY  = 8192
  integer(c_int32_t), parameter :: CM1G_OUT_UINTERP   = 16384
  integer(c_int32_t), parameter :: CM1G_OUT_VINTERP   = 32768
  integer(c_int32_t), parameter :: CM1G_OUT_WINTERP   = 65536
  integer(c_int32_t), parameter :: CM1G_OUT_BASESTATE = 131072
  integer(c_int32_t), parameter :: CM1G_OUT_U         = 262144
  integer(c_int32_t), parameter :: CM1G_OUT_UPERT     = 524288
  integer(c_int32_t), parameter :: CM1G_OUT_V         = 1048576
  integer(c_int32_t), parameter :: CM1G_OUT_VPERT     = 2097152
  integer(c_int32_t), parameter :: CM1G_OUT_W         = 4194304
  integer(c_int32_t), parameter :: CM1G_OUT_ALL       = 8388607
  integer(c_int32_t), parameter :: CM1G_OUT_TKE       = 8388608
  integer(c_int32_t), parameter :: CM1G_OUT_KM        = 16777216
  integer(c_int32_t), parameter :: CM1G_OUT_KH        = 33554432
  integer(c_int32_t), parameter :: CM1G_OUT_DISSTEN   = 67108864
  integer(c_int32_t), parameter :: CM1G_OUT_NM        = 134217728
  integer(c_int32_t), parameter :: CM1G_OUT_DEF       = 268435456
  ! statistics switches of cm1g_statpack (the namelist's stat_* names)
  integer(c_int32_t), parameter :: CM1G_STAT_W       = 1
  integer(c_int32_t), parameter :: CM1G_STAT_WLEVS   = 2
  integer(c_int32_t), parameter :: CM1G_STAT_U       = 4
  integer(c_int32_t), parameter :: CM1G_STAT_V       = 8
  integer(c_int32_t), parameter :: CM1G_STAT_PIPERT  = 16
  integer(c_int32_t), parameter :: CM1G_STAT_PRSPERT = 32
  integer(c_int32_t), parameter :: CM1G_STAT_THPERT  = 64
  integer(c_int32_t), parameter :: CM1G_STAT_Q       = 128
  integer(c_int32_t), parameter :: CM1G_STAT_SFCPRS  = 256
  integer(c_int32_t), parameter :: CM1G_STAT_WSP     = 512
  integer(c_int32_t), parameter :: CM1G_STAT_CFL     = 1024
  integer(c_int32_t), parameter :: CM1G_STAT_TMASS   = 2048
  integer(c_int32_t), parameter :: CM1G_STAT_TMOIS   = 4096
  integer(c_int32_t), parameter :: CM1G_STAT_QMASS   = 8192
  integer(c_int32_t), parameter :: CM1G_STAT_ALL    = 16383
  integer(c_int32_t), parameter :: CM1G_STAT_TKE     = 16384
  integer(c_int32_t), parameter :: CM1G_STAT_KM      = 32768
  integer(c_int32_t), parameter :: CM1G_STAT_KH      = 65536
  integer(c_int), parameter :: CM1G_STAT_MAXN = 128
  ! staggering of host arrays
  integer(c_int), parameter :: CM1G_STAG_S = 0, CM1G_STAG_U = 1, CM1G_STAG_V = 2, &
                               CM1G_STAG_W = 3

  ! 3-D field ids
  integer(c_int), parameter :: CM1G_UA       = 0
  integer(c_int), parameter :: CM1G_U3D      = 1
  integer(c_int), parameter :: CM1G_UTEN     = 2
  integer(c_int), parameter :: CM1G_UTEN1    = 3
  integer(c_int), parameter :: CM1G_VA       = 4
  integer(c_int), parameter :: CM1G_V3D      = 5
  integer(c_int), parameter :: CM1G_VTEN     = 6
  integer(c_int), parameter :: CM1G_VTEN1    = 7
  integer(c_int), parameter :: CM1G_WA       = 8
  integer(c_int), parameter :: CM1G_W3D      = 9
  integer(c_int), parameter :: CM1G_WTEN     = 10
  integer(c_int), parameter :: CM1G_WTEN1    = 11
  integer(c_int), parameter :: CM1G_RRU      = 12
  integer(c_int), parameter :: CM1G_RRV      = 13
  integer(c_int), parameter :: CM1G_RRW      = 14
  integer(c_int), parameter :: CM1G_PPI      = 15
  integer(c_int), parameter :: CM1G_PP3D     = 16
  integer(c_int), parameter :: CM1G_PPTEN    = 17
  integer(c_int), parameter :: CM1G_STEN     = 18
  integer(c_int), parameter :: CM1G_SADV     = 19
  integer(c_int), parameter :: CM1G_PPX      = 20
  integer(c_int), parameter :: CM1G_THA      = 21
  integer(c_int), parameter :: CM1G_TH3D     = 22
  integer(c_int), parameter :: CM1G_THTEN    = 23
  integer(c_int), parameter :: CM1G_THTEN1   = 24
  integer(c_int), parameter :: CM1G_THTERM   = 25
  integer(c_int), parameter :: CM1G_QPTEN    = 26
  integer(c_int), parameter :: CM1G_QTTEN    = 27
  integer(c_int), parameter :: CM1G_QVTEN    = 28
  integer(c_int), parameter :: CM1G_QCTEN    = 29
  integer(c_int), parameter :: CM1G_RHO      = 30
  integer(c_int), parameter :: CM1G_RR       = 31
  integer(c_int), parameter :: CM1G_RF       = 32
  integer(c_int), parameter :: CM1G_PRS      = 33
  integer(c_int), parameter :: CM1G_DIVX     = 34
  integer(c_int), parameter :: CM1G_DUM1     = 35
  integer(c_int), parameter :: CM1G_DUM2     = 36
  integer(c_int), parameter :: CM1G_DUM3     = 37
  integer(c_int), parameter :: CM1G_DUM4     = 38
  integer(c_int), parameter :: CM1G_DUM5     = 39
  integer(c_int), parameter :: CM1G_DUM6     = 40
  integer(c_int), parameter :: CM1G_DUM7     = 41
  integer(c_int), parameter :: CM1G_DUM8     = 42
  integer(c_int), parameter :: CM1G_DUM9     = 43
  integer(c_int), parameter :: CM1G_T11      = 44
  integer(c_int), parameter :: CM1G_T12      = 45
  integer(c_int), parameter :: CM1G_T13      = 46
  integer(c_int), parameter :: CM1G_T22      = 47
  integer(c_int), parameter :: CM1G_T23      = 48
  integer(c_int), parameter :: CM1G_T33      = 49
  integer(c_int), parameter :: CM1G_QA0      = 50
  integer(c_int), parameter :: CM1G_QA1      = 51
  integer(c_int), parameter :: CM1G_QA2      = 52
  integer(c_int), parameter :: CM1G_Q3D0     = 53
  integer(c_int), parameter :: CM1G_Q3D1     = 54
  integer(c_int), parameter :: CM1G_Q3D2     = 55
  integer(c_int), parameter :: CM1G_QTEN0    = 56
  integer(c_int), parameter :: CM1G_QTEN1    = 57
  integer(c_int), parameter :: CM1G_QTEN2    = 58
  integer(c_int), parameter :: CM1G_PTA0     = 59
  integer(c_int), parameter :: CM1G_PTA1     = 60
  integer(c_int), parameter :: CM1G_PTA2     = 61
  integer(c_int), parameter :: CM1G_PTA3     = 62
  integer(c_int), parameter :: CM1G_PTA4     = 63
  integer(c_int), parameter :: CM1G_PTA5     = 64
  integer(c_int), parameter :: CM1G_PTA6     = 65
  integer(c_int), parameter :: CM1G_PTA7     = 66
  integer(c_int), parameter :: CM1G_PT3D0    = 67
  integer(c_int), parameter :: CM1G_PT3D1    = 68
  integer(c_int), parameter :: CM1G_PT3D2    = 69
  integer(c_int), parameter :: CM1G_PT3D3    = 70
  integer(c_int), parameter :: CM1G_PT3D4    = 71
  integer(c_int), parameter :: CM1G_PT3D5    = 72
  integer(c_int), parameter :: CM1G_PT3D6    = 73
  integer(c_int), parameter :: CM1G_PT3D7    = 74
  integer(c_int), parameter :: CM1G_PTTEN0   = 75
  integer(c_int), parameter :: CM1G_PTTEN1   = 76
  integer(c_int), parameter :: CM1G_PTTEN2   = 77
  integer(c_int), parameter :: CM1G_PTTEN3   = 78
  integer(c_int), parameter :: CM1G_PTTEN4   = 79
  integer(c_int), parameter :: CM1G_PTTEN5   = 80
  integer(c_int), parameter :: CM1G_PTTEN6   = 81
  integer(c_int), parameter :: CM1G_PTTEN7   = 82
  ! LES subgrid turbulence (allocated by cm1g_set_turbulence)
  integer(c_int), parameter :: CM1G_TKEA      = 83
  integer(c_int), parameter :: CM1G_TKE3D     = 84
  integer(c_int), parameter :: CM1G_TKETEN    = 85
  integer(c_int), parameter :: CM1G_KMH       = 86
  integer(c_int), parameter :: CM1G_KMV       = 87
  integer(c_int), parameter :: CM1G_KHH       = 88
  integer(c_int), parameter :: CM1G_KHV       = 89
  integer(c_int), parameter :: CM1G_NM        = 90
  integer(c_int), parameter :: CM1G_DEFV      = 91
  integer(c_int), parameter :: CM1G_DEFH      = 92
  integer(c_int), parameter :: CM1G_DISSTEN   = 93
  integer(c_int), parameter :: CM1G_LENSCL    = 94

  ! 1-D column ids (k = 0..nk+2)
  integer(c_int), parameter :: CM1G_K_MH    = 0
  integer(c_int), parameter :: CM1G_K_RMH   = 1
  integer(c_int), parameter :: CM1G_K_MF    = 2
  integer(c_int), parameter :: CM1G_K_RMF   = 3
  integer(c_int), parameter :: CM1G_K_C1    = 4
  integer(c_int), parameter :: CM1G_K_C2    = 5
  integer(c_int), parameter :: CM1G_K_ZH    = 6
  integer(c_int), parameter :: CM1G_K_ZF    = 7
  integer(c_int), parameter :: CM1G_K_PI0   = 8
  integer(c_int), parameter :: CM1G_K_TH0   = 9
  integer(c_int), parameter :: CM1G_K_PRS0  = 10
  integer(c_int), parameter :: CM1G_K_QV0   = 11
  integer(c_int), parameter :: CM1G_K_RHO0  = 12
  integer(c_int), parameter :: CM1G_K_RR0   = 13
  integer(c_int), parameter :: CM1G_K_RF0   = 14
  integer(c_int), parameter :: CM1G_K_RRF0  = 15
  integer(c_int), parameter :: CM1G_K_THV0  = 16
  integer(c_int), parameter :: CM1G_K_RTH0  = 17
  integer(c_int), parameter :: CM1G_K_QC0   = 18
  integer(c_int), parameter :: CM1G_K_QI0   = 19
  integer(c_int), parameter :: CM1G_K_U0    = 20
  integer(c_int), parameter :: CM1G_K_V0    = 21
  integer(c_int), parameter :: CM1G_K_TAUH  = 22
  integer(c_int), parameter :: CM1G_K_TAUF  = 23

  ! 2-D field ids
  integer(c_int), parameter :: CM1G_D_RAIN    = 0
  integer(c_int), parameter :: CM1G_D_PRATE   = 1
  integer(c_int), parameter :: CM1G_D_RADBCW  = 2
  integer(c_int), parameter :: CM1G_D_RADBCE  = 3
  integer(c_int), parameter :: CM1G_D_RADBCS  = 4
  integer(c_int), parameter :: CM1G_D_RADBCN  = 5
  integer(c_int), parameter :: CM1G_D_UST     = 6
  integer(c_int), parameter :: CM1G_D_U1      = 7
  integer(c_int), parameter :: CM1G_D_V1      = 8
  integer(c_int), parameter :: CM1G_D_S1      = 9
  integer(c_int), parameter :: CM1G_D_THFLUX  = 10
  integer(c_int), parameter :: CM1G_D_ZNT     = 11

  interface
    integer(c_int) function cm1g_create(cfg, m) bind(C, name="cm1g_create")
      import :: c_int, c_ptr, cm1g_config
      type(cm1g_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: m
    end function
    integer(c_int) function cm1g_destroy(m) bind(C, name="cm1g_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    type(c_ptr) function cm1g_last_error() bind(C, name="cm1g_last_error")
      import :: c_ptr
    end function
    integer(c_int) function cm1g_dtsmall(cfg, min_dz, nsound) bind(C, name="cm1g_dtsmall")
      import :: c_int, c_int32_t, c_float, cm1g_config
      type(cm1g_config), intent(in) :: cfg
      real(c_float), value :: min_dz
      integer(c_int32_t), intent(out) :: nsound
    end function
    integer(c_int) function cm1g_put_column(m, id, col) bind(C, name="cm1g_put_column")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id
      real(c_float), intent(in) :: col(*)
    end function
    integer(c_int) function cm1g_get_column(m, id, col) bind(C, name="cm1g_get_column")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id
      real(c_float), intent(out) :: col(*)
    end function
    integer(c_int) function cm1g_metrics(m) bind(C, name="cm1g_metrics")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_base_derive(m) bind(C, name="cm1g_base_derive")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_put_array(m, id, stag, a) bind(C, name="cm1g_put_array")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id, stag
      real(c_float), intent(in) :: a(*)
    end function
    integer(c_int) function cm1g_get_array(m, id, stag, a) bind(C, name="cm1g_get_array")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id, stag
      real(c_float), intent(out) :: a(*)
    end function
    integer(c_int) function cm1g_put_2d_array(m, id, a) bind(C, name="cm1g_put_2d_array")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id
      real(c_float), intent(in) :: a(*)
    end function
    integer(c_int) function cm1g_get_2d_array(m, id, a) bind(C, name="cm1g_get_2d_array")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id
      real(c_float), intent(out) :: a(*)
    end function
    integer(c_int) function cm1g_k_fallout(m) bind(C, name="cm1g_k_fallout")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_fallout(m, train) bind(C, name="cm1g_fallout")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), intent(inout) :: train
    end function
    integer(c_int) function cm1g_geterain(m, dt, cpx, lx1, t_id, rr_id, q_id, erain) &
        bind(C, name="cm1g_geterain")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: m
      real(c_float), value :: dt, cpx, lx1
      integer(c_int), value :: t_id, rr_id, q_id
      real(c_double), intent(inout) :: erain
    end function
    integer(c_int) function cm1g_kessler(m) bind(C, name="cm1g_kessler")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_satadj(m) bind(C, name="cm1g_satadj")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_ensemble_set_order(e, mode) bind(C, name="cm1g_ensemble_set_order")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: mode
    end function
    integer(c_int) function cm1g_ensemble_tune_schedule(e, steps, ty_best, order_best, ms9) &
        bind(C, name="cm1g_ensemble_tune_schedule")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: e
      integer(c_int), value :: steps
      integer(c_int), intent(out) :: ty_best, order_best
      real(c_float), intent(out) :: ms9(9)
    end function
    integer(c_int) function cm1g_sync(m) bind(C, name="cm1g_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_set_turbulence(m, t) bind(C, name="cm1g_set_turbulence")
      import :: c_int, c_ptr, cm1g_turb_config
      type(c_ptr), value :: m
      type(cm1g_turb_config), intent(in) :: t
    end function
    integer(c_int) function cm1g_sfc_and_turb(m, getsfc) bind(C, name="cm1g_sfc_and_turb")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: getsfc
    end function
    integer(c_int) function cm1g_turbs(m, iflux, dosfcflx, s_id, sten_id) bind(C, name="cm1g_turbs")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: iflux, dosfcflx, s_id, sten_id
    end function
    integer(c_int) function cm1g_turbt(m) bind(C, name="cm1g_turbt")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_turbu(m) bind(C, name="cm1g_turbu")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_turbv(m) bind(C, name="cm1g_turbv")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_turbw(m) bind(C, name="cm1g_turbw")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_advw_scalar(m, nrk, advweno, a_id, ten_id) bind(C, name="cm1g_advw_scalar")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: nrk, advweno, a_id, ten_id
    end function
    integer(c_int) function cm1g_getset(m, fresh_start) bind(C, name="cm1g_getset")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: fresh_start
    end function
    integer(c_int) function cm1g_step(m, cflmax) bind(C, name="cm1g_step")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      real(c_float), intent(out) :: cflmax
    end function
    integer(c_int) function cm1g_set_graphs(m, on) bind(C, name="cm1g_set_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: on
    end function
    integer(c_int) function cm1g_set_tiles(m, ty) bind(C, name="cm1g_set_tiles")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: ty
    end function
    ! ensembles: members(n) are model handles
    integer(c_int) function cm1g_ensemble_create(members, n, e) bind(C, name="cm1g_ensemble_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: members(*)
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: e
    end function
    integer(c_int) function cm1g_ensemble_destroy(e) bind(C, name="cm1g_ensemble_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function cm1g_ensemble_step(e) bind(C, name="cm1g_ensemble_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function cm1g_ensemble_sync(e) bind(C, name="cm1g_ensemble_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function cm1g_ensemble_set_tiles(e, ty) bind(C, name="cm1g_ensemble_set_tiles")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: ty
    end function
    integer(c_int) function cm1g_ensemble_tune(e, steps, ty_best, ms_per_step) &
        bind(C, name="cm1g_ensemble_tune")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: e
      integer(c_int), value :: steps
      integer(c_int), intent(out) :: ty_best
      real(c_float), intent(out) :: ms_per_step(3)
    end function
    ! learned Kessler surrogate; x, y, t, buf are device addresses
    integer(c_int) function cm1g_mlp_nparams(n) bind(C, name="cm1g_mlp_nparams")
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(out) :: n
    end function
    integer(c_int) function cm1g_mlp_create(seed, net) bind(C, name="cm1g_mlp_create")
      import :: c_int, c_ptr, c_int64_t
      integer(c_int64_t), value :: seed
      type(c_ptr), intent(out) :: net
    end function
    integer(c_int) function cm1g_mlp_destroy(net) bind(C, name="cm1g_mlp_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: net
    end function
    integer(c_int) function cm1g_mlp_get_params(net, host) bind(C, name="cm1g_mlp_get_params")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: net
      real(c_float), intent(out) :: host(*)
    end function
    integer(c_int) function cm1g_mlp_set_params(net, host) bind(C, name="cm1g_mlp_set_params")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: net
      real(c_float), intent(in) :: host(*)
    end function
    integer(c_int) function cm1g_mlp_set_norm(net, norm) bind(C, name="cm1g_mlp_set_norm")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: net
      real(c_float), intent(in) :: norm(40)
    end function
    integer(c_int) function cm1g_mlp_forward(net, x, rows, y) bind(C, name="cm1g_mlp_forward")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: net, x, y
      integer(c_int64_t), value :: rows
    end function
    integer(c_int) function cm1g_mlp_loss_grad(net, x, t, rows, loss, grad) bind(C, name="cm1g_mlp_loss_grad")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: net, x, t, loss, grad
      integer(c_int64_t), value :: rows
    end function
    integer(c_int) function cm1g_mlp_adam(net, lr) bind(C, name="cm1g_mlp_adam")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: net
      real(c_float), value :: lr
    end function
    integer(c_int) function cm1g_mlp_sync(net) bind(C, name="cm1g_mlp_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: net
    end function
    integer(c_int) function cm1g_set_surrogate(m, net) bind(C, name="cm1g_set_surrogate")
      import :: c_int, c_ptr
      type(c_ptr), value :: m, net
    end function
    integer(c_int) function cm1g_set_harvest(m, buf, capacity, keep_one_in) bind(C, name="cm1g_set_harvest")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: m, buf
      integer(c_int64_t), value :: capacity
      integer(c_int), value :: keep_one_in
    end function
    integer(c_int) function cm1g_surrogate_counts(m, out6) bind(C, name="cm1g_surrogate_counts")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: m
      integer(c_int64_t), intent(out) :: out6(6)
    end function
    integer(c_int) function cm1g_solve1(m) bind(C, name="cm1g_solve1")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_solve2(m) bind(C, name="cm1g_solve2")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_mp_driver(m) bind(C, name="cm1g_mp_driver")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_solve3(m) bind(C, name="cm1g_solve3")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_write_grads(m, dir, nwrite, mask, iorigin, stretch_z, mtime, nstep) &
        bind(C, name="cm1g_write_grads")
      import :: c_int, c_ptr, c_char, c_int32_t, c_double
      type(c_ptr), value :: m
      character(kind=c_char), dimension(*), intent(in) :: dir
      integer(c_int), value :: nwrite, iorigin, stretch_z, nstep
      integer(c_int32_t), value :: mask
      real(c_double), value :: mtime
    end function
    integer(c_int) function cm1g_statpack(m, dir, nrec, mask, mtime, rstat, nstat) &
        bind(C, name="cm1g_statpack")
      import :: c_int, c_ptr, c_char, c_int32_t, c_double, c_float
      type(c_ptr), value :: m
      character(kind=c_char), dimension(*), intent(in) :: dir
      integer(c_int), value :: nrec
      integer(c_int32_t), value :: mask
      real(c_double), value :: mtime
      real(c_float), dimension(*), intent(out) :: rstat
      integer(c_int32_t), intent(out) :: nstat
    end function
    integer(c_int) function cm1g_budgets(m, tmass, tmois) bind(C, name="cm1g_budgets")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), intent(out) :: tmass, tmois
    end function
    integer(c_int) function cm1g_dims(m, px, py, pz) bind(C, name="cm1g_dims")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: m
      integer(c_int64_t), intent(out) :: px, py, pz
    end function
    integer(c_int) function cm1g_put_field(m, id, a) bind(C, name="cm1g_put_field")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id
      real(c_float), intent(in) :: a(*)
    end function
    integer(c_int) function cm1g_get_field(m, id, a) bind(C, name="cm1g_get_field")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id
      real(c_float), intent(out) :: a(*)
    end function
    integer(c_int) function cm1g_put_2d(m, id, a) bind(C, name="cm1g_put_2d")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id
      real(c_float), intent(in) :: a(*)
    end function
    integer(c_int) function cm1g_get_2d(m, id, a) bind(C, name="cm1g_get_2d")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: id
      real(c_float), intent(out) :: a(*)
    end function
    integer(c_int) function cm1g_field_ptr(m, id, dptr) bind(C, name="cm1g_field_ptr")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: id
      type(c_ptr), intent(out) :: dptr
    end function
    ! routine-level entries (the reference routine each replaces is named
    ! in include/cm1gpu.h)
    integer(c_int) function cm1g_advs(m, nrk, s_id, s0_id, sten_id, pdef, pdefweno, dt, weps) &
        bind(C, name="cm1g_advs")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: m
      integer(c_int), value :: nrk, s_id, s0_id, sten_id, pdef, pdefweno
      real(c_float), value :: dt
      real(c_double), value :: weps
    end function
    integer(c_int) function cm1g_advu(m, nrk) bind(C, name="cm1g_advu")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: nrk
    end function
    integer(c_int) function cm1g_advv(m, nrk) bind(C, name="cm1g_advv")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: nrk
    end function
    integer(c_int) function cm1g_advw(m, nrk) bind(C, name="cm1g_advw")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: nrk
    end function
    integer(c_int) function cm1g_sound(m, nrk, dttmp) bind(C, name="cm1g_sound")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      integer(c_int), value :: nrk
      real(c_float), value :: dttmp
    end function
    integer(c_int) function cm1g_kessler_branch(m) bind(C, name="cm1g_kessler_branch")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_pdefq(m, rmax, q_id) bind(C, name="cm1g_pdefq")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      real(c_float), value :: rmax
      integer(c_int), value :: q_id
    end function
    integer(c_int) function cm1g_bc(m, id, kind) bind(C, name="cm1g_bc")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: id, kind
    end function
    integer(c_int) function cm1g_calccflquick(m, dt, u_id, v_id, w_id, cflmax) &
        bind(C, name="cm1g_calccflquick")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      real(c_float), value :: dt
      integer(c_int), value :: u_id, v_id, w_id
      real(c_float), intent(out) :: cflmax
    end function
    integer(c_int) function cm1g_getdiv(m, div_id) bind(C, name="cm1g_getdiv")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: div_id
    end function
    integer(c_int) function cm1g_getdivx(m) bind(C, name="cm1g_getdivx")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function cm1g_maxmin(m, id, stag, vmax, locmax, vmin, locmin) bind(C, name="cm1g_maxmin")
      import :: c_int, c_ptr, c_float, c_int32_t
      type(c_ptr), value :: m
      integer(c_int), value :: id, stag
      real(c_float), intent(out) :: vmax, vmin
      integer(c_int32_t), intent(out) :: locmax(3), locmin(3)
    end function
    ! no counterpart in the reference: maxmin of a device field without allocation
    integer(c_int) function cm1g_maxmin_noalloc(field, ni, nj, nk, stag, stream, vmax, locmax, vmin, locmin) &
        bind(C, name="cm1g_maxmin_noalloc")
      import :: c_int, c_ptr, c_float, c_int32_t
      type(c_ptr), value :: field, stream
      integer(c_int), value :: ni, nj, nk, stag
      real(c_float), intent(out) :: vmax, vmin
      integer(c_int32_t), intent(out) :: locmax(3), locmin(3)
    end function
    integer(c_int) function cm1g_scalars(m, out4) bind(C, name="cm1g_scalars")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), intent(out) :: out4(4)
    end function
    ! device-state checkpoint
    integer(c_int) function cm1g_state_bytes(m, nbytes) bind(C, name="cm1g_state_bytes")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: m
      integer(c_int64_t), intent(out) :: nbytes
    end function
    integer(c_int) function cm1g_state_save(m, buf, nbytes) bind(C, name="cm1g_state_save")
      import :: c_int, c_ptr, c_int64_t, c_int8_t
      type(c_ptr), value :: m
      integer(c_int8_t), intent(out) :: buf(*)
      integer(c_int64_t), value :: nbytes
    end function
    integer(c_int) function cm1g_state_load(m, buf, nbytes) bind(C, name="cm1g_state_load")
      import :: c_int, c_ptr, c_int64_t, c_int8_t
      type(c_ptr), value :: m
      integer(c_int8_t), intent(in) :: buf(*)
      integer(c_int64_t), value :: nbytes
    end function
    ! kernel timer (measurement hook)
    integer(c_int) function cm1g_ktimer_start(m, which) bind(C, name="cm1g_ktimer_start")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: which
    end function
    integer(c_int) function cm1g_ktimer_stop(m, total_ms, launches, bytes_per_launch, name, &
                                             name_len) bind(C, name="cm1g_ktimer_stop")
      import :: c_int, c_ptr, c_double, c_int64_t, c_char
      type(c_ptr), value :: m
      real(c_double), intent(out) :: total_ms, bytes_per_launch
      integer(c_int64_t), intent(out) :: launches
      character(kind=c_char), intent(out) :: name(*)
      integer(c_int), value :: name_len
    end function
  end interface

contains

  ! the message of the last failed call, as a Fortran string
  function cm1g_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: n, i
    cp = cm1g_last_error()
    n = c_strlen_(cp)
    call c_f_pointer(cp, p, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = p(i)
    end do
  end function

  integer function c_strlen_(cp)
    type(c_ptr), intent(in) :: cp
    character(kind=c_char), pointer :: p(:)
    integer :: i
    c_strlen_ = 0
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [4096])
    do i = 1, 4096
      if (p(i) == c_null_char) exit
      c_strlen_ = i
    end do
  end function

end module cm1gpu
