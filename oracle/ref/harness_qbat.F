C  harness_qbat.F -- test harness, not product code.
C
C  Drives the reference's own QBAT shell routines, compiled from their sources
C  under the reference tree by oracle/build_ref.py, in the order of CBAFORC3
C  (engine/source/elements/shell/coqueba/cbaforc3.F:440-1140) for Ishell=12
C  (IHBE=11), IGTYP=1, law 2, Ismstr=2, ISROT=0, IDRIL=0, no pinching, NPT >= 2
C  points, 2 x 2 Gauss points, PARITH/ON.  As in harness_qeph.F, two pieces are
C  restated because the routines that hold them dispatch over every element type
C  and law: the Gauss point's element strain rate (cbaforc3.F:759-767) and
C  MULAWC's law-2 point loop for NPG > 1 (materials/mat_share/mulawc.F90:
C  530-560,2968-3020: the energy increments scaled by OFF, the Gauss point's
C  volume, the thickness carried from one Gauss point to the next), calling the
C  reference SIGEPS02C.
C
C  ref_qbat_group: one group of NEL <= 128 quadrilaterals, corner k of element e
C  is node 4*(e-1)+k, FSKY slot of corner k is 4*(e-1)+k.  Per Gauss point
C  arrays are indexed (element, component, point, Gauss point).
      SUBROUTINE REF_QBAT_GROUP(NEL_C, XG, VG, VRG, THKE, FOR, MOM,
     .   FORPG, MOMPG, THK, EINT, EPSDE, OFFG, SIG, PLA, EPSDP, DT1_C,
     .   MPAR, IPAR, PPAR, FSKYO, DTE_O, EVIS_O, DT2T_O)
     .   BIND(C, NAME='ref_qbat_group')
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
     .   FOR(NEL_C,5), MOM(NEL_C,3), FORPG(NEL_C,5,4), MOMPG(NEL_C,3,4),
     .   THK(*), EINT(NEL_C,2), EPSDE(*), OFFG(*),
     .   SIG(NEL_C,5,5,4), PLA(NEL_C,5,4), EPSDP(NEL_C,5,4),
     .   MPAR(16), PPAR(9), FSKYO(8,*), DTE_O(*), EVIS_O(*), DT2T_O
      REAL(C_DOUBLE), VALUE :: DT1_C
      INTEGER(C_INT) :: IPAR(6)
C     ---- locals
      INTEGER NEL, I, K, NN, IT, IPLA, NPTT, IOFC, NELTST, ITYPTST,
     .   G_DT, NG, IR, IS, NPLAT
      INTEGER IXC(NIXC,MVSIZ), NGL(MVSIZ), MAT(MVSIZ), PID(MVSIZ),
     .   IADC(4,MVSIZ), IPARTC(MVSIZ), IPLAT(MVSIZ), INOD(1),
     .   IOFF_DUCT(MVSIZ), IGEO(n_var_igeo,1)
      my_real
     .   X(3,4*MVSIZ), V(3,4*MVSIZ), VR(3,4*MVSIZ), GEO(1000,1),
     .   PM(250,1), OFF(MVSIZ), SIGY(MVSIZ), DT1C(MVSIZ), THK0(MVSIZ),
     .   DIR_A(MVSIZ,2), DIR_B(MVSIZ,2), AREA(MVSIZ), LL(MVSIZ),
     .   VXYZ(MVSIZ,3,4), RXYZ(MVSIZ,2,4), VCORE(MVSIZ,3,4),
     .   JAC(MVSIZ,4), HX(MVSIZ,4), HY(MVSIZ,4), VKSI(4,4), VETA(4,4),
     .   VQN(MVSIZ,9,4), VQG(MVSIZ,9,4), VQ(MVSIZ,3,3), VJFI(MVSIZ,6,4),
     .   VNRM(MVSIZ,3,4), VASTN(MVSIZ,4,4), X13(MVSIZ), X24(MVSIZ),
     .   Y13(MVSIZ), Y24(MVSIZ), DD(MVSIZ,6), FACN(MVSIZ,2), ZL(MVSIZ),
     .   R11(MVSIZ), R12(MVSIZ), R13(MVSIZ), R21(MVSIZ), R22(MVSIZ),
     .   R23(MVSIZ), R31(MVSIZ), R32(MVSIZ), R33(MVSIZ), RLZ(MVSIZ,4),
     .   UX1(MVSIZ),UX2(MVSIZ),UX3(MVSIZ),UX4(MVSIZ),
     .   UY1(MVSIZ),UY2(MVSIZ),UY3(MVSIZ),UY4(MVSIZ),
     .   VL1(MVSIZ,3),VL2(MVSIZ,3),VL3(MVSIZ,3),VL4(MVSIZ,3),
     .   XL2(MVSIZ),XL3(MVSIZ),XL4(MVSIZ),YL2(MVSIZ),YL3(MVSIZ),
     .   YL4(MVSIZ),XLCOR(MVSIZ,2,3)
      DOUBLE PRECISION SMSTR(6*MVSIZ)
      my_real
     .   SHF(MVSIZ),THK2(MVSIZ),NU(MVSIZ),G(MVSIZ),YM(MVSIZ),
     .   A11(MVSIZ),A12(MVSIZ),SSP(MVSIZ),RHO(MVSIZ),VOLG(MVSIZ),
     .   GS(MVSIZ),AMU(MVSIZ),KRZ(MVSIZ),A11R(MVSIZ),
     .   PM_STACK(20,1),ZOFFSET(MVSIZ),ALPE(MVSIZ),
     .   CDET(MVSIZ),VOL0(MVSIZ)
      my_real
     .   VDEF(MVSIZ,8),BM(MVSIZ,36),BMF(MVSIZ,36),BF(MVSIZ,24),
     .   BC(MVSIZ,40),TC(MVSIZ,2,2),BRZ(MVSIZ,4,4),
     .   EXX(MVSIZ),EYY(MVSIZ),EXY(MVSIZ),EXZ(MVSIZ),EYZ(MVSIZ),
     .   KXX(MVSIZ),KYY(MVSIZ),KXY(MVSIZ),TANI(6,MVSIZ),
     .   WXY(MVSIZ),GSTR(MVSIZ,8),GSTRPG(MVSIZ,8),F_DEF(MVSIZ,8),
     .   GSTRW(MVSIZ)
      my_real
     .   DEGMB(MVSIZ),DEGFX(MVSIZ),THKN(MVSIZ),OFF_OLD(MVSIZ),
     .   EPCHK(MVSIZ),VISCMX(MVSIZ),EPSD_PG(MVSIZ),EPSD_GLOB(MVSIZ),
     .   THKLYL(MVSIZ),
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
C     the Gauss point's resultants, shaped (NEL,.) as CBAFORC3 passes
C     GBUF%FORPG(PTF) and GBUF%MOMPG(PTM) (cbaforc3.F:574-575,868-881)
      my_real FORL(NEL_C,5), MOML(NEL_C,3)
      my_real
     .   STI(MVSIZ),STIR(MVSIZ),PARTSAV(29,MVSIZ),DT2T,BID(1),
     .   MSC(MVSIZ),DMELC(MVSIZ),DTEL(MVSIZ),SSP_MAX(MVSIZ),
     .   VF(MVSIZ,3,4),VM(MVSIZ,2,4),VMZ(MVSIZ,4)
      my_real
     .   F11(MVSIZ),F12(MVSIZ),F13(MVSIZ),F14(MVSIZ),
     .   F21(MVSIZ),F22(MVSIZ),F23(MVSIZ),F24(MVSIZ),
     .   F31(MVSIZ),F32(MVSIZ),F33(MVSIZ),F34(MVSIZ),
     .   M11(MVSIZ),M12(MVSIZ),M13(MVSIZ),M14(MVSIZ),
     .   M21(MVSIZ),M22(MVSIZ),M23(MVSIZ),M24(MVSIZ),
     .   M31(MVSIZ),M32(MVSIZ),M33(MVSIZ),M34(MVSIZ),
     .   FSKY(8,4*MVSIZ), FSKYV(1,8), THEM(MVSIZ,4),
     .   FTHESKY(1), CONDNSKY(1), CONDE(MVSIZ)
      my_real DTINV, EPS_K2, EPS_M2, ASRATE, FACT, VISC, VOL2, ZT
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
      ANIM_PLY = 0
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
      GEO(:,1) = ZERO
      GEO(13,1) = PPAR(8)
      GEO(38,1) = PPAR(5)
      IGEO(:,1) = 0
      IGEO(11,1) = 1
      PM_STACK(:,1) = ZERO
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
        GSTRPG(I,:) = ZERO
        GSTRW(I) = ZERO
        SIGY(I) = EP30
        ALPE(I) = ONE
        VF(I,:,:) = ZERO
        VM(I,:,:) = ZERO
        VMZ(I,:) = ZERO
        SSP_MAX(I) = ZERO
        MSC(I) = ZERO
        DMELC(I) = ZERO
        TANI(:,I) = ZERO
        EPSD_GLOB(I) = ZERO
        WXY(I) = ZERO
      ENDDO
      DT2T = EP20
      NELTST = 0
      ITYPTST = 0
      IOFC = 0
      G_DT = 1
C     ---- CBAFORC3 chain: CBACOOR, CNCOEF3, CBADEFSH, CBAENERS
      CALL CBACOOR(ELB, 1, NEL, X, V,
     .   VR, IXC, PM, OFFG, LL,
     1   AREA, VXYZ, RXYZ, VCORE, JAC, HX, HY, VKSI, VETA,
     2   VQN, VQG, VQ, VJFI, VNRM, VASTN, NPLAT, IPLAT,
     3   X13, X24, Y13, Y24, OFF, DD, 1,
     4   0, NPTT, 2, NEL, 0,
     5   SMSTR, DIR_A, DIR_B, FACN, ZL,
     6   R11, R12, R13, R21, R22, R23,
     7   R31, R32, R33, INOD, RLZ,
     8   THKE, 0, UX1, UX2, UX3,
     9   UX4, UY1, UY2, UY3, UY4,
     A   VL1, VL2, VL3, VL4, XL2,
     B   XL3, XL4, YL2, YL3, YL4, XLCOR, 0)
      CALL CNCOEF3(1, NEL, PM, MAT, GEO,
     2   PID, OFF, AREA, SHF, THK0,
     3   THK2, NU, G, YM,
     4   A11, A12, THK, THKE, SSP,
     5   RHO, VOLG, GS, 2, 0,
     6   NPTT, DT1C, DT1, 11, AMU,
     7   KRZ, IGEO, A11R, 1, PM_STACK,
     8   NEL, ZOFFSET)
      CALL CBADEFSH(1, NEL, X13, X24, Y13, Y24, BM, VDEF, VXYZ,
     .   NPLAT, IPLAT)
      CALL CBAENERS(1, NEL, OFF, AREA, THK0,
     .   VDEF, FOR, EINT, DT1, NEL)
      IF (IPAR(3) > 0) THEN
        ASRATE = MIN(ONE, PM(9,1)*DT1)
      ELSE
        ASRATE = ONE
      END IF
C     ---- the Gauss points
      DO IS=1,2
       DO IR=1,2
        NG = 2*(IS-1) + IR
        DO I=1,NEL
          CDET(I) = JAC(I,NG)
          VOL0(I) = THK0(I)*CDET(I)
        ENDDO
        CALL CBADEF(1, NEL, NG, VCORE, AREA, CDET, VQN, VQG, VJFI,
     1     VXYZ, RXYZ, VDEF, VNRM, VASTN,
     2     HX, HY, VETA, VKSI, BM, BMF, BF, BC, TC, NPLAT, IPLAT,
     3     0, BRZ)
        CALL CBASTRA3(GSTR, GSTRPG,
     1     1, NEL, 0, 4, VDEF,
     2     EXX, EYY, EXY, EXZ, EYZ,
     3     KXX, KYY, KXY, DT1C, TANI,
     4     0, 0, UX1, UX2, UX3,
     6     UX4, UY1, UY2, UY3, UY4,
     7     X13, X24, Y13, Y24, BM,
     8     2, 2, NPLAT, IPLAT, 0,
     9     WXY, F_DEF, GSTRW, NEL)
        DO K=1,5
          FORL(1:NEL,K) = FORPG(1:NEL,K,NG)
        ENDDO
        DO K=1,3
          MOML(1:NEL,K) = MOMPG(1:NEL,K,NG)
        ENDDO
        CALL CBAENER(FORL, EINT, 1, NEL, OFF, VOL0, EXY, NEL)
C       ---- Gauss point's element strain rate (cbaforc3.F:759-767), restated
        DTINV = DT1 / MAX(DT1**2,EM20)
        DO I=1,NEL
          EPS_K2 = (KXX(I)**2+KYY(I)**2+KXX(I)*KYY(I)+FOURTH*KXY(I)**2)
     .           * ONE_OVER_9*THK(I)**2
          EPS_M2 = FOUR_OVER_3*(EXX(I)**2+EYY(I)**2+EXX(I)*EYY(I)
     .           + FOURTH*EXY(I)**2)
          EPSD_PG(I) = SQRT(EPS_K2 + EPS_M2)*DTINV
          EPSD_GLOB(I) = EPSD_GLOB(I) + EPSD_PG(I) / 4
        ENDDO
C       ---- MULAWC law-2 point loop, NPG > 1, restated
        DO I=1,NEL
          DEGMB(I) = FORL(I,1)*EXX(I)+FORL(I,2)*EYY(I)
     .             + FORL(I,3)*EXY(I)+FORL(I,4)*EYZ(I)
     .             + FORL(I,5)*EXZ(I)
          DEGFX(I) = MOML(I,1)*KXX(I)+MOML(I,2)*KYY(I)
     .             + MOML(I,3)*KXY(I)
          DEGMB(I) = DEGMB(I)*OFF(I)
          DEGFX(I) = DEGFX(I)*OFF(I)
          VOL0(I) = CDET(I)*THK0(I)
          THKN(I) = THK(I)
          FORL(I,1:5) = ZERO
          MOML(I,1:3) = ZERO
          SIGY(I) = ZERO
          ETSE(I) = ONE
          OFF_OLD(I) = OFF(I)
          EPCHK(I) = ZERO
          VISCMX(I) = ZERO
          DEPSYZ(I) = EYZ(I)
          DEPSZX(I) = EXZ(I)
        ENDDO
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
            LPLA(I) = PLA(I,IT,NG)
            LEPSD(I) = EPSDP(I,IT,NG)
            LOFF(I) = ONE
            SIGOXX(I) = SIG(I,1,IT,NG)
            SIGOYY(I) = SIG(I,2,IT,NG)
            SIGOXY(I) = SIG(I,3,IT,NG)
            SIGOYZ(I) = SIG(I,4,IT,NG)
            SIGOZX(I) = SIG(I,5,IT,NG)
            TSTAR(I) = ZERO
            VOLPT(I) = ZERO
          ENDDO
          CALL SIGEPS02C(MATP,
     .       NEL, EINT, THKN, TEMPEL, FHEAT,
     .       OFF, SIGY, DT1, IPLA, SIGKSI,
     .       VOLPT, GS, THKLYL, ETSE, G_IMP,
     .       DPLA, TSTAR, 0, HARDM, EPCHK,
     .       NPTT, LPLA, OFF_OLD, LOFF, IOFF_DUCT,
     .       SIGOXX, SIGOYY, SIGOXY, SIGOYZ, SIGOZX,
     .       SIGNXX, SIGNYY, SIGNXY, SIGNYZ, SIGNZX,
     .       DEPSXX, DEPSYY, DEPSXY, DEPSYZ, DEPSZX,
     .       EPSPXX, EPSPYY, EPSPXY, EPSD_PG, LEPSD,
     .       ASRATE, SIGBX, SIGBY, SIGBXY,
     .       0, DPLANL)
          DO I=1,NEL
            VISCMX(I) = MAX(PPAR(4),VISCMX(I))
            PLA(I,IT,NG) = LPLA(I)
            EPSDP(I,IT,NG) = LEPSD(I)
            SIG(I,1,IT,NG) = SIGNXX(I) * ONE
            SIG(I,2,IT,NG) = SIGNYY(I) * ONE
            SIG(I,3,IT,NG) = SIGNXY(I) * ONE
            SIG(I,4,IT,NG) = SIGNYZ(I) * ONE
            SIG(I,5,IT,NG) = SIGNZX(I) * ONE
            FORL(I,1) = FORL(I,1) + WF(IT,NPTT)*SIGNXX(I)
            FORL(I,2) = FORL(I,2) + WF(IT,NPTT)*SIGNYY(I)
            FORL(I,3) = FORL(I,3) + WF(IT,NPTT)*SIGNXY(I)
            FORL(I,4) = FORL(I,4) + WF(IT,NPTT)*SIGNYZ(I)
            FORL(I,5) = FORL(I,5) + WF(IT,NPTT)*SIGNZX(I)
            MOML(I,1) = MOML(I,1) + WMC(I)*SIGNXX(I)
            MOML(I,2) = MOML(I,2) + WMC(I)*SIGNYY(I)
            MOML(I,3) = MOML(I,3) + WMC(I)*SIGNXY(I)
          ENDDO
        ENDDO
        DO I=1,NEL
          THK(I) = MAX(THKN(I),EM30)
          FACT = ONEP414*PPAR(4)
          VISC = FACT*SSP(I)*SQRT(CDET(I))*DTINV*RHO(I)
          FORL(I,1)=FORL(I,1)+VISC*(EXX(I)+HALF*EYY(I))
          FORL(I,2)=FORL(I,2)+VISC*(EYY(I)+HALF*EXX(I))
          FORL(I,3)=FORL(I,3)+VISC* EXY(I) *THIRD
          DO K=1,5
            FORL(I,K)=FORL(I,K)*OFF(I)*ONE
          ENDDO
          DO K=1,3
            MOML(I,K)=MOML(I,K)*OFF(I)*ONE
          ENDDO
          DEGMB(I) = DEGMB(I)+ FORL(I,1)*EXX(I)+FORL(I,2)*EYY(I)
     .             + FORL(I,3)*EXY(I)+FORL(I,4)*EYZ(I)+ FORL(I,5)*EXZ(I)
          DEGFX(I) = DEGFX(I)+ MOML(I,1)*KXX(I)+MOML(I,2)*KYY(I)
     .             + MOML(I,3)*KXY(I)
          VOL2 = HALF*VOL0(I)
          EINT(I,1) = EINT(I,1) + DEGMB(I)*VOL2
          EINT(I,2) = EINT(I,2) + DEGFX(I)*THK0(I)*VOL2
        ENDDO
        CALL CBAENER(FORL, EINT, 1, NEL, OFF, VOL0, EXY, NEL)
        CALL CBAVISC(1, NEL, VDEF, AMU, OFF,
     2     SHF, NU, RHO, SSP, CDET,
     3     THK0, FORL, MOML, NPTT, 2,
     4     IPARTC, PARTSAV, DT1, NEL)
        CALL CBAFORI(1, NEL, NG, CDET, THK0,
     2     THK2, FORL, MOML, NEL, BM,
     3     BMF, BF, BC, TC, VF,
     4     VM, NPLAT, IPLAT, VOL0)
        DO K=1,5
          FORPG(1:NEL,K,NG) = FORL(1:NEL,K)
        ENDDO
        DO K=1,3
          MOMPG(1:NEL,K,NG) = MOML(1:NEL,K)
        ENDDO
       ENDDO
      ENDDO
      DO I=1,NEL
        EPSDE(I) = ONE * EPSD_GLOB(I) + (ONE - ONE) * EPSDE(I)
        DO K=1,5
          FOR(I,K) = FOURTH*(FORPG(I,K,1) + FORPG(I,K,2)
     .                     + FORPG(I,K,3) + FORPG(I,K,4))
        ENDDO
        DO K=1,3
          MOM(I,K) = FOURTH*(MOMPG(I,K,1) + MOMPG(I,K,2)
     .                     + MOMPG(I,K,3) + MOMPG(I,K,4))
        ENDDO
      ENDDO
      CALL CBAFORCT(1, NEL, VOLG, X13, X24,
     2   Y13, Y24, FOR, VF, NPLAT,
     3   IPLAT, OFF, NEL)
      CALL CBAENERS(1, NEL, OFF, AREA, THK0,
     .   VDEF, FOR, EINT, DT1, NEL)
      CALL CBAPROJ(
     1   1, NEL, VQN, VQ, VF,
     2   VM, NPLAT, IPLAT,
     3   F11, F12, F13, F14, F21,
     4   F22, F23, F24, F31, F32,
     5   F33, F34, M11, M12, M13,
     6   M14, M21, M22, M23, M24,
     7   M31, M32, M33, M34, VCORE,
     8   DD, VMZ, 0, OFF)
      CALL CNDT3(
     1   1, NEL, OFF, DT2T, AMU,
     2   NELTST, ITYPTST, STI, STIR, OFFG,
     3   SSP, VISCMX, RHO, VOLG, THK0, THK2,
     4   A11, LL, ALPE, NGL, 2,
     5   IOFC, 4, AREA, G, SHF,
     6   MSC, DMELC, 0, BID, 1,
     7   0, A11R, G_DT, DTEL, 2,
     8   PM, MAT(1), NEL, ZOFFSET, SSP_MAX)
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
        EVIS_O(I) = PARTSAV(8,I)
      ENDDO
      DT2T_O = DT2T
      DEALLOCATE(MATP%UPARAM)
      DEALLOCATE(MATP%IPARAM)
      RETURN
      END
