C  harness_qeph.F -- test harness, not product code.
C
C  Drives the reference's own QEPH shell routines, compiled from their sources
C  under the reference tree by oracle/build_ref.py, in the order of CZFORC3
C  (engine/source/elements/shell/coquez/czforc3.F:436-808) for Ishell=24,
C  IGTYP=1, law 2, Ismstr=2, ISROT=0, IORTH=0, NPT points, PARITH/ON.  As in
C  harness_shell.F, two pieces are restated because the routines that hold them
C  dispatch over every element type and law: the element strain rate
C  (czforc3.F:554-561) and the law-2 point loop of MULAWC
C  (materials/mat_share/mulawc.F90:530-3046) with its QEPH factors
C  (mulawc.F90:509-510,565-569,2787-2794), calling the reference SIGEPS02C.
C
C  ref_qeph_group: one group of NEL <= 128 quadrilaterals, corner k of element e
C  is node 4*(e-1)+k, FSKY slot of corner k is 4*(e-1)+k.
      SUBROUTINE REF_QEPH_GROUP(NEL_C, XG, VG, VRG, THKE, FOR, MOM,
     .   THK, EINT, EPSDE, HOURG, OFFG, SIG, PLA, EPSDP, DT1_C, MPAR, IPAR,
     .   PPAR, FSKYO, DTE_O, EHOU_O, DT2T_O) BIND(C, NAME='ref_qeph_group')
      USE ISO_C_BINDING
      USE MATPARAM_DEF_MOD
      USE ELBUFDEF_MOD
      USE PROP_PARAM_MOD , only : n_var_igeo
      USE ELEMENT_MOD , only : NIXC
#include      "implicit_f.inc"
#include      "mvsiz_p.inc"
#include      "param_c.inc"
#include      "com01_c.inc"
#include      "com04_c.inc"
#include      "com06_c.inc"
#include      "com08_c.inc"
#include      "com20_c.inc"
#include      "impl1_c.inc"
#include      "scr02_c.inc"
#include      "scr05_c.inc"
#include      "scr06_c.inc"
#include      "scr14_c.inc"
#include      "scr16_c.inc"
#include      "scr17_c.inc"
#include      "scr18_c.inc"
#include      "parit_c.inc"
#include      "remesh_c.inc"
#include      "sms_c.inc"
      INTEGER(C_INT), VALUE :: NEL_C
      REAL(C_DOUBLE) :: XG(4,3,*), VG(4,3,*), VRG(4,3,*), THKE(*),
     .   FOR(NEL_C,5), MOM(NEL_C,3), THK(*), EINT(NEL_C,2), EPSDE(*),
     .   HOURG(NEL_C,12), OFFG(*), SIG(NEL_C,5,*), PLA(NEL_C,*),
     .   EPSDP(NEL_C,*), MPAR(16), PPAR(9), FSKYO(8,*), DTE_O(*),
     .   EHOU_O(*), DT2T_O
      REAL(C_DOUBLE), VALUE :: DT1_C
      INTEGER(C_INT) :: IPAR(6)
C     ---- locals
      INTEGER NEL, I, K, NN, IT, IPLA, NPTT, IOFC, NELTST, ITYPTST, G_DT
      INTEGER IXC(NIXC,MVSIZ), NGL(MVSIZ), MAT(MVSIZ), PID(MVSIZ),
     .   IADC(4,MVSIZ), IPARTC(MVSIZ), INDXOF(MVSIZ),
     .   IOFF_DUCT(MVSIZ), IGEO(n_var_igeo,1)
      LOGICAL PLAT(MVSIZ)
      my_real
     .   X(3,4*MVSIZ), V(3,4*MVSIZ), VR(3,4*MVSIZ), GEO(1000,1),
     .   PM(250,1), OFF(MVSIZ), SIGY(MVSIZ), DT1C(MVSIZ), THK0(MVSIZ),
     .   DIR_A(2), DIR_B(2), AREA(MVSIZ), A_I(MVSIZ),
     .   VL13(MVSIZ,3), VL24(MVSIZ,3), VLHI(MVSIZ,3), RLXYZ(MVSIZ,2,4),
     .   VQN(MVSIZ,3,4), VQ(MVSIZ,3,3), LL(MVSIZ), L13(MVSIZ),
     .   L24(MVSIZ), X13(MVSIZ), X24(MVSIZ), Y13(MVSIZ), Y24(MVSIZ),
     .   MX13(MVSIZ), MX23(MVSIZ), MX34(MVSIZ), MY13(MVSIZ),
     .   MY23(MVSIZ), MY34(MVSIZ), Z1(MVSIZ), LXYZ(MVSIZ,2,4),
     .   DD(MVSIZ,6), DB(MVSIZ,3,4), FACN(MVSIZ,2), PY1(MVSIZ),
     .   PX2(MVSIZ), PY2(MVSIZ), R11(MVSIZ), R12(MVSIZ), R13(MVSIZ),
     .   R21(MVSIZ), R22(MVSIZ), R23(MVSIZ), R31(MVSIZ), R32(MVSIZ),
     .   R33(MVSIZ), RLZ(MVSIZ,4), DIZ(MVSIZ,3)
      my_real
     .   VX1(MVSIZ),VX2(MVSIZ),VX3(MVSIZ),VX4(MVSIZ),
     .   VY1(MVSIZ),VY2(MVSIZ),VY3(MVSIZ),VY4(MVSIZ),
     .   VZ1(MVSIZ),VZ2(MVSIZ),VZ3(MVSIZ),VZ4(MVSIZ),
     .   VRX1(MVSIZ),VRX2(MVSIZ),VRX3(MVSIZ),VRX4(MVSIZ),
     .   VRY1(MVSIZ),VRY2(MVSIZ),VRY3(MVSIZ),VRY4(MVSIZ),
     .   VRZ1(MVSIZ),VRZ2(MVSIZ),VRZ3(MVSIZ),VRZ4(MVSIZ),
     .   X1G(MVSIZ),X2G(MVSIZ),X3G(MVSIZ),X4G(MVSIZ),
     .   Y1G(MVSIZ),Y2G(MVSIZ),Y3G(MVSIZ),Y4G(MVSIZ),
     .   Z1G(MVSIZ),Z2G(MVSIZ),Z3G(MVSIZ),Z4G(MVSIZ),
     .   UX1(MVSIZ),UX2(MVSIZ),UX3(MVSIZ),UX4(MVSIZ),
     .   UY1(MVSIZ),UY2(MVSIZ),UY3(MVSIZ),UY4(MVSIZ),
     .   XL2(MVSIZ),XL3(MVSIZ),XL4(MVSIZ),YL2(MVSIZ),YL3(MVSIZ),
     .   YL4(MVSIZ),Z2(MVSIZ),VL1(MVSIZ,3),VL2(MVSIZ,3),VL3(MVSIZ,3),
     .   VL4(MVSIZ,3)
      DOUBLE PRECISION SMSTR(6*MVSIZ)
      my_real
     .   SHF(MVSIZ),THK02(MVSIZ),NU(MVSIZ),G(MVSIZ),YM(MVSIZ),
     .   A11(MVSIZ),A12(MVSIZ),SSP(MVSIZ),RHO(MVSIZ),VOL0(MVSIZ),
     .   GS(MVSIZ),AMU(MVSIZ),GSR(MVSIZ),A11SR(MVSIZ),A12SR(MVSIZ),
     .   NUSR(MVSIZ),SHFSR(MVSIZ),KRZ(MVSIZ),A11R(MVSIZ),
     .   PM_STACK(20,1),UPARAM(12),DIRA(MVSIZ,2),DIRB(MVSIZ,2),
     .   UVAR(MVSIZ,12),FAC58(MVSIZ,2),ZOFFSET(MVSIZ),ALPE(MVSIZ),
     .   ZCFAC(MVSIZ,2),FAC1(MVSIZ)
      my_real
     .   VDEF(MVSIZ,8),DHG(MVSIZ,6),
     .   EXX(MVSIZ),EYY(MVSIZ),EXY(MVSIZ),EXZ(MVSIZ),EYZ(MVSIZ),
     .   KXX(MVSIZ),KYY(MVSIZ),KXY(MVSIZ),TANI(6,MVSIZ),
     .   WXY(MVSIZ),GSTR(MVSIZ,8),F_DEF(MVSIZ,8),GSTRW(MVSIZ)
      my_real
     .   DEGMB(MVSIZ),DEGFX(MVSIZ),THKN(MVSIZ),OFF_OLD(MVSIZ),
     .   EPCHK(MVSIZ),VISCMX(MVSIZ),EPSD_PG(MVSIZ),THKLYL(MVSIZ),
     .   WMC(MVSIZ),DEPSXX(MVSIZ),DEPSYY(MVSIZ),DEPSXY(MVSIZ),
     .   DEPSYZ(MVSIZ),DEPSZX(MVSIZ),EPSPXX(MVSIZ),EPSPYY(MVSIZ),
     .   EPSPXY(MVSIZ),SIGOXX(MVSIZ),SIGOYY(MVSIZ),SIGOXY(MVSIZ),
     .   SIGOYZ(MVSIZ),SIGOZX(MVSIZ),SIGNXX(MVSIZ),SIGNYY(MVSIZ),
     .   SIGNXY(MVSIZ),SIGNYZ(MVSIZ),SIGNZX(MVSIZ),LPLA(MVSIZ),
     .   LEPSD(MVSIZ),TEMPEL(MVSIZ),FHEAT(MVSIZ),
     .   SIGKSI(MVSIZ,5),VOLPT(MVSIZ),ETSE(MVSIZ),G_IMP(MVSIZ),
     .   DPLA(MVSIZ),TSTAR(MVSIZ),HARDM(MVSIZ),LOFF(MVSIZ),
     .   SIGBX(MVSIZ),SIGBY(MVSIZ),SIGBXY(MVSIZ),DPLANL(MVSIZ),
     .   WML(11,11)
      my_real
     .   STI(MVSIZ),STIR(MVSIZ),PARTSAV(29,MVSIZ),DT2T,BID(1),
     .   MSC(MVSIZ),DMELC(MVSIZ),DTEL(MVSIZ),SSP_EQ(MVSIZ),
     .   VF(MVSIZ,3,4),VM(MVSIZ,2,4),VMZ(MVSIZ,4)
      my_real
     .   F11(MVSIZ),F12(MVSIZ),F13(MVSIZ),F14(MVSIZ),
     .   F21(MVSIZ),F22(MVSIZ),F23(MVSIZ),F24(MVSIZ),
     .   F31(MVSIZ),F32(MVSIZ),F33(MVSIZ),F34(MVSIZ),
     .   M11(MVSIZ),M12(MVSIZ),M13(MVSIZ),M14(MVSIZ),
     .   M21(MVSIZ),M22(MVSIZ),M23(MVSIZ),M24(MVSIZ),
     .   M31(MVSIZ),M32(MVSIZ),M33(MVSIZ),M34(MVSIZ),
     .   FZERO(3,4,1), FSKY(8,4*MVSIZ), FSKYV(1,8), THEM(MVSIZ,4),
     .   FTHESKY(1), CONDNSKY(1), CONDE(MVSIZ)
      my_real DT1L, DTINV, EPS_K2, EPS_M2, ASRATE, FACT, VISC, VOL2, ZT
      TYPE(MATPARAM_STRUCT_) :: MATP
      TYPE(ELBUF_STRUCT_) :: ELB
C=======================================================================
      NEL = NEL_C
      NPTT = IPAR(5)
      IPLA = IPAR(4)
      NPROPM = 250
      NPROPG = 1000
      NPARG = 100
      NPSAV = 29
      NPROPMI = 300
      NPROPGI = n_var_igeo
      DT1 = DT1_C
      TT = ONE
      NCYCLE = 1
      IRUN = 1
      ISIGI = 0
      IMPL_S = 0
      IMP_LR = 0
      IKPROJ = 0
      ISMDISP = 0
      NUMELS = 0
      NUMMAT = 1
      NUMGEO = 1
      NODADT = 0
      IDT1SH = 0
      IDTMINS = 0
      NADMESH = 0
      IEPSDOT = 0
      IVECTOR = 0
      IRESP = 0
      LSKY = 4*NEL
      IDTMIN(3) = 2
      DTFAC1(3) = PPAR(6)
      DTMIN1(3) = PPAR(7)
      ELB%IDRAPE = 0
      CALL COQINI
      CALL COQINI_WM(WML)
C     ---- material (hm_read_mat02_jc.F90:234-339), resol_init.F:689-696
      ALLOCATE(MATP%UPARAM(12))
      ALLOCATE(MATP%IPARAM(4))
      MATP%RHO   = MPAR(1)
      MATP%YOUNG = MPAR(2)
      MATP%NU    = MPAR(3)
      MATP%SHEAR = MPAR(4)
      MATP%UPARAM(:) = ZERO
      MATP%UPARAM(1) = MPAR(8)
      MATP%UPARAM(2) = MPAR(9)
      MATP%UPARAM(3) = MPAR(10)
      MATP%UPARAM(4) = MPAR(11)
      MATP%UPARAM(5) = MPAR(12)
      MATP%UPARAM(6) = MPAR(13)
      MATP%UPARAM(7) = MPAR(14)
      MATP%UPARAM(8) = ZERO
      MATP%UPARAM(9) = MPAR(16)
      MATP%UPARAM(10)= MPAR(15)
      MATP%IPARAM(1) = 0
      MATP%IPARAM(2) = IPAR(1)
      MATP%IPARAM(3) = IPAR(2)
      MATP%IPARAM(4) = IPAR(3)
      MATP%THERM%TREF  = 300.0D0
      MATP%THERM%TMELT = EP20
      MATP%THERM%RHOCP = ZERO
      MATP%THERM%EFRAC = ZERO
      PM(:,1) = ZERO
      PM(1,1)  = MPAR(1)
      PM(9,1)  = MPAR(16)
      PM(20,1) = MPAR(2)
      PM(21,1) = MPAR(3)
      PM(22,1) = MPAR(4)
      PM(24,1) = MPAR(5)
      PM(25,1) = MPAR(6)
      PM(27,1) = MPAR(7)
      PM(12,1) = SQRT(ABS(PM(22,1)))
      PM(13,1) = SQRT(ABS(PM(24,1)))
      PM(14,1) = SQRT(ABS(PM(25,1)))
      PM(190,1)= SQRT(ABS(PM(21,1)))
      GEO(:,1) = ZERO
      GEO(13,1) = PPAR(8)
      GEO(17,1) = PPAR(9)
      GEO(38,1) = PPAR(5)
      GEO(100,1) = SQRT(GEO(38,1))
      IGEO(:,1) = 0
      IGEO(11,1) = 1
      PM_STACK(:,1) = ZERO
      UPARAM(:) = ZERO
      DO I=1,NEL
        IXC(1,I) = 1
        DO K=1,4
          NN = 4*(I-1)+K
          IXC(K+1,I) = NN
          X(1,NN) = XG(K,1,I)
          X(2,NN) = XG(K,2,I)
          X(3,NN) = XG(K,3,I)
          V(1,NN) = VG(K,1,I)
          V(2,NN) = VG(K,2,I)
          V(3,NN) = VG(K,3,I)
          VR(1,NN) = VRG(K,1,I)
          VR(2,NN) = VRG(K,2,I)
          VR(3,NN) = VRG(K,3,I)
          IADC(K,I) = NN
        ENDDO
        IXC(6,I) = 1
        IXC(7,I) = I
        MAT(I) = 1
        PID(I) = 1
        NGL(I) = I
        IPARTC(I) = I
        PARTSAV(8,I) = ZERO
        TEMPEL(I) = ZERO
        FHEAT(I) = ZERO
        SIGBX(I) = ZERO
        SIGBY(I) = ZERO
        SIGBXY(I) = ZERO
        DPLANL(I) = ZERO
        IOFF_DUCT(I) = 0
        G_IMP(I) = ZERO
        HARDM(I) = ZERO
        GSTR(I,:) = ZERO
        GSTRW(I) = ZERO
        SIGY(I) = EP30
        ZCFAC(I,1) = ONE
        ZCFAC(I,2) = ONE
        ALPE(I) = ONE
        FAC1(I) = GEO(17,1)
        VF(I,:,:) = ZERO
        VM(I,:,:) = ZERO
        SSP_EQ(I) = ZERO
        MSC(I) = ZERO
        DMELC(I) = ZERO
        TANI(:,I) = ZERO
      ENDDO
      DT2T = EP20
      NELTST = 0
      ITYPTST = 0
      IOFC = 0
      G_DT = 1
C     ---- CZFORC3 chain
      CALL CZCORC1(4*NEL, NEL, ELB,
     1   1, NEL, X, V, VR,
     2   IXC, PM, PLAT, AREA,
     3   A_I, VL13, VL24, VLHI, RLXYZ,
     4   VQN, VQ, LL, L13, L24,
     5   X13, X24, Y13, Y24, MX13,
     6   MX23, MX34, MY13, MY23, MY34,
     7   Z1, LXYZ, DD, DB, SMSTR,
     9   0, NPTT, 1, 2,
     A   DIR_A, DIR_B, OFFG, RLXYZ, LXYZ,
     B   FACN, PY1, PX2, PY2, R11,
     C   R12, R13, R21, R22, R23,
     D   R31, R32, R33, RLZ, 0,
     E   0, VX1, VX2, VX3, VX4,
     F   VY1, VY2, VY3, VY4, VZ1,
     G   VZ2, VZ3, VZ4, VRX1, VRX2,
     H   VRX3, VRX4, VRY1, VRY2, VRY3,
     I   VRY4, VRZ1, VRZ2, VRZ3, VRZ4,
     J   X1G, X2G, X3G, X4G, Y1G,
     K   Y2G, Y3G, Y4G, Z1G, Z2G,
     L   Z3G, Z4G, THKE, DIZ, UX1,
     M   UX2, UX3, UX4, UY1, UY2,
     N   UY3, UY4, XL2, XL3, XL4,
     O   YL2, YL3, YL4, VL1, VL2,
     P   VL3, VL4, NEL, Z2)
      CALL CNCOEF3B(1, NEL, PM, MAT, GEO,
     2   PID, AREA, SHF, THK0,
     3   THK02, NU, G, YM,
     4   A11, A12, THK, THKE, SSP,
     5   RHO, VOL0, GS, 2, 0,
     6   NPTT, DT1C, DT1, 23, AMU,
     7   GSR, A11SR, A12SR, NUSR, SHFSR,
     8   KRZ, IGEO, A11R, 1, PM_STACK,
     9   UPARAM, DIRA, DIRB, UVAR, FAC58,
     A   NEL, ZOFFSET)
      CALL CZDEF(1, NEL, AREA, A_I, VL13, VL24,
     2   VLHI, RLXYZ, VDEF, DHG, X13,
     3   X24, Y13, Y24, MX13, MX23, MX34,
     4   MY13, MY23, MY34, Z1, DT1, OFF,
     5   OFFG, RLXYZ)
      CALL CZSTRA3(1, NEL, 0, VDEF, GSTR,
     2   EXX, EYY, EXY, EXZ, EYZ,
     3   KXX, KYY, KXY, DT1C, TANI,
     4   0, 0, UX1, UX2, UX3,
     5   UX4, UY1, UY2, UY3, UY4,
     6   Y24, PX2, PY1, PY2, AREA,
     7   2, 2, WXY, F_DEF, GSTRW,
     8   NEL)
C     ---- element strain rate (czforc3.F:554-561), restated
      DT1L = DT1
      DTINV = DT1L / MAX(DT1L**2,EM20)
      DO I=1,NEL
        EPS_K2 = (KXX(I)**2+KYY(I)**2+KXX(I)*KYY(I)+FOURTH*KXY(I)**2)
     .         * ONE_OVER_9*THK(I)**2
        EPS_M2 = FOUR_OVER_3*(EXX(I)**2+EYY(I)**2+EXX(I)*EYY(I)
     .         + FOURTH*EXY(I)**2)
        EPSD_PG(I) = SQRT(EPS_K2 + EPS_M2)*DTINV
      ENDDO
      EPSDE(1:NEL) = ONE * EPSD_PG(1:NEL) + (ONE - ONE) * EPSDE(1:NEL)
C     ---- MULAWC law-2 point loop with the QEPH factors, restated
      DO I=1,NEL
        DEGMB(I) = FOR(I,1)*EXX(I)+FOR(I,2)*EYY(I)
     .           + FOR(I,3)*EXY(I)+FOR(I,4)*EYZ(I)
     .           + FOR(I,5)*EXZ(I)
        DEGFX(I) = MOM(I,1)*KXX(I)+MOM(I,2)*KYY(I)
     .           + MOM(I,3)*KXY(I)
        VOL0(I) = AREA(I)*THK0(I)
        THKN(I) = THK(I)
        FOR(I,1:5) = ZERO
        MOM(I,1:3) = ZERO
        SIGY(I) = ZERO
        ZCFAC(I,1) = ZERO
        ZCFAC(I,2) = ONE
        ETSE(I) = ONE
        OFF_OLD(I) = OFF(I)
        EPCHK(I) = ZERO
        VISCMX(I) = ZERO
        DEPSYZ(I) = EYZ(I)
        DEPSZX(I) = EXZ(I)
      ENDDO
      IF (IPAR(3) > 0) THEN
        ASRATE = MIN(ONE, PM(9,1)*DT1)
      ELSE
        ASRATE = ONE
      END IF
      DO IT=1,NPTT
        DO I=1,NEL
          THKLYL(I) = WF(IT,NPTT)*THK0(I)
          WMC(I) = WML(IT,NPTT)
          ZT = Z0(IT,NPTT)*THK0(I)
          DEPSXX(I) = EXX(I)+ZT*KXX(I)
          DEPSYY(I) = EYY(I)+ZT*KYY(I)
          DEPSXY(I) = EXY(I)+ZT*KXY(I)
          EPSPXX(I) = DEPSXX(I)*DTINV
          EPSPYY(I) = DEPSYY(I)*DTINV
          EPSPXY(I) = DEPSXY(I)*DTINV
          DPLA(I) = ZERO
          LPLA(I) = PLA(I,IT)
          LEPSD(I) = EPSDP(I,IT)
          LOFF(I) = ONE
          SIGOXX(I) = SIG(I,1,IT)
          SIGOYY(I) = SIG(I,2,IT)
          SIGOXY(I) = SIG(I,3,IT)
          SIGOYZ(I) = SIG(I,4,IT)
          SIGOZX(I) = SIG(I,5,IT)
          TSTAR(I) = ZERO
          VOLPT(I) = ZERO
        ENDDO
        CALL SIGEPS02C(MATP,
     .     NEL, EINT, THKN, TEMPEL, FHEAT,
     .     OFF, SIGY, DT1, IPLA, SIGKSI,
     .     VOLPT, GS, THKLYL, ETSE, G_IMP,
     .     DPLA, TSTAR, 0, HARDM, EPCHK,
     .     NPTT, LPLA, OFF_OLD, LOFF, IOFF_DUCT,
     .     SIGOXX, SIGOYY, SIGOXY, SIGOYZ, SIGOZX,
     .     SIGNXX, SIGNYY, SIGNXY, SIGNYZ, SIGNZX,
     .     DEPSXX, DEPSYY, DEPSXY, DEPSYZ, DEPSZX,
     .     EPSPXX, EPSPYY, EPSPXY, EPSD_PG, LEPSD,
     .     ASRATE, SIGBX, SIGBY, SIGBXY,
     .     0, DPLANL)
        DO I=1,NEL
          VISCMX(I) = MAX(PPAR(4),VISCMX(I))
          PLA(I,IT) = LPLA(I)
          EPSDP(I,IT) = LEPSD(I)
          SIG(I,1,IT) = SIGNXX(I) * ONE
          SIG(I,2,IT) = SIGNYY(I) * ONE
          SIG(I,3,IT) = SIGNXY(I) * ONE
          SIG(I,4,IT) = SIGNYZ(I) * ONE
          SIG(I,5,IT) = SIGNZX(I) * ONE
          FOR(I,1) = FOR(I,1) + WF(IT,NPTT)*SIGNXX(I)
          FOR(I,2) = FOR(I,2) + WF(IT,NPTT)*SIGNYY(I)
          FOR(I,3) = FOR(I,3) + WF(IT,NPTT)*SIGNXY(I)
          FOR(I,4) = FOR(I,4) + WF(IT,NPTT)*SIGNYZ(I)
          FOR(I,5) = FOR(I,5) + WF(IT,NPTT)*SIGNZX(I)
          MOM(I,1) = MOM(I,1) + WMC(I)*SIGNXX(I)
          MOM(I,2) = MOM(I,2) + WMC(I)*SIGNYY(I)
          MOM(I,3) = MOM(I,3) + WMC(I)*SIGNXY(I)
          ZCFAC(I,1) = ZCFAC(I,1) + ETSE(I) / NPTT
          ZCFAC(I,2) = MIN(ETSE(I),ZCFAC(I,2))
        ENDDO
      ENDDO
      DO I=1,NEL
        THK(I) = MAX(THKN(I),EM30)
        FACT = ONEP414*PPAR(4)
        VISC = FACT*SSP(I)*SQRT(AREA(I))*DTINV*RHO(I)
        FOR(I,1)=FOR(I,1)+VISC*(EXX(I)+HALF*EYY(I))
        FOR(I,2)=FOR(I,2)+VISC*(EYY(I)+HALF*EXX(I))
        FOR(I,3)=FOR(I,3)+VISC* EXY(I) *THIRD
        DO K=1,5
          FOR(I,K)=FOR(I,K)*OFF(I)*ONE
        ENDDO
        DO K=1,3
          MOM(I,K)=MOM(I,K)*OFF(I)*ONE
        ENDDO
        DEGMB(I) = DEGMB(I)+ FOR(I,1)*EXX(I)+FOR(I,2)*EYY(I)
     .           + FOR(I,3)*EXY(I)+FOR(I,4)*EYZ(I)+ FOR(I,5)*EXZ(I)
        DEGFX(I) = DEGFX(I)+ MOM(I,1)*KXX(I)+MOM(I,2)*KYY(I)
     .           + MOM(I,3)*KXY(I)
        VOL2 = HALF*VOL0(I)
        EINT(I,1) = EINT(I,1) + DEGMB(I)*VOL2
        EINT(I,2) = EINT(I,2) + DEGFX(I)*THK0(I)*VOL2
      ENDDO
      CALL CNDT3(
     1   1, NEL, OFF, DT2T, AMU,
     2   NELTST, ITYPTST, STI, STIR, OFFG,
     3   SSP, VISCMX, RHO, VOL0, THK0, THK02,
     4   A11, LL, ALPE, NGL, 2,
     5   IOFC, 4, AREA, G, SHF,
     6   MSC, DMELC, 0, BID, 1,
     7   0, A11R, G_DT, DTEL, 2,
     8   PM, MAT(1), NEL, ZOFFSET, SSP_EQ)
      CALL CZFINTCE(1, NEL, THK0, THK02, A_I, X13,
     2   X24, Y13, Y24, Z1, MX23, MX13,
     3   MX34, MY13, MY23, MY34, FOR, MOM,
     4   VF, VM, NEL)
      CALL CZFINTN1(
     1   1, NEL, THK0, THK02, A_I, DHG,
     2   X13, X24, Y13, Y24, Z1, MX23,
     3   MX13, MX34, MY13, MY23, MY34, HOURG,
     4   FOR, MOM, VF, VM, ZCFAC, A11,
     5   A12, G, SHF, SIGY, OFF, FAC1,
     6   RHO, AREA, DT1, EINT, AMU, VLHI,
     7   NPTT, IPARTC, PARTSAV, 0, GSR, NEL,
     8   A11SR, A12SR, NUSR, SHFSR, 2, FAC58)
      CALL CZPROJ1(
     1   1, NEL, VQN, VQ, VF,
     2   VM, PLAT,
     3   F11, F12, F13, F14, F21,
     4   F22, F23, F24, F31, F32,
     5   F33, F34, M11, M12, M13,
     6   M14, M21, M22, M23, M24,
     7   M31, M32, M33, M34, FZERO,
     8   Z1, LXYZ, DD, DB, LXYZ,
     9   0, DIZ, VMZ)
      CALL CUPDTN3P(1, NEL, OFFG, OFF, STI,
     2   STIR, FSKY, FSKYV, IADC,
     4   F11, F12, F13, F14, F21,
     5   F22, F23, F24, F31, F32,
     6   F33, F34, M11, M12, M13,
     7   M14, M21, M22, M23, M24,
     8   M31, M32, M33, M34, IXC,
     A   EINT, PARTSAV, MAT, IPARTC, PM,
     B   AREA, THK, FACN, 0, THEM,
     C   FTHESKY, CONDNSKY, CONDE, 0)
      DO I=1,NEL
        DO K=1,4
          NN = 4*(I-1)+K
          FSKYO(1:8,NN) = FSKY(1:8,NN)
        ENDDO
        DTE_O(I) = DTEL(I)
        EHOU_O(I) = PARTSAV(8,I)
      ENDDO
      DT2T_O = DT2T
      DEALLOCATE(MATP%UPARAM)
      DEALLOCATE(MATP%IPARAM)
      RETURN
      END
