C  harness_shell.F -- test harness, not product code.
C
C  Drives the reference's own shell routines, compiled from their sources under
C  the reference tree by oracle/build_ref.py, in the order of CFORC3
C  (engine/source/elements/shell/coque/cforc3.F:398-746) for the default
C  4-node shell: Ishell=1 (IHBE=1), Ismstr=2, ISHFRAM=0, law 2, NPT points,
C  PARITH/ON.  Two pieces of CFORC3/MULAWC are restated here because the
C  routines that hold them dispatch over every element type and law: the
C  element strain rate (cforc3.F:535-546) and the law-2 integration-point loop
C  of MULAWC (materials/mat_share/mulawc.F90:530-3046), which calls the
C  reference SIGEPS02C for each point.
C
C  ref_shell_group: one group of NEL <= 128 elements, corner k of element e is
C  node 4*(e-1)+k, FSKY slot of corner k is 4*(e-1)+k, outputs as CUPDT3P
C  writes them.
      SUBROUTINE REF_SHELL_GROUP(NEL_C, XG, VG, VRG, THKE, ITRI, FOR, MOM,
     .   THK, EINT, EPSDE, HOUR, OFFG, SIG, PLA, EPSDP, DT1_C, MPAR, IPAR,
     .   PPAR, FSKYO, DTE_O, EHOU_O) BIND(C, NAME='ref_shell_group')
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
     .   HOUR(NEL_C,5), OFFG(*), SIG(NEL_C,5,*), PLA(NEL_C,*),
     .   EPSDP(NEL_C,*), MPAR(16), PPAR(7), FSKYO(8,*), DTE_O(*),
     .   EHOU_O(*)
      REAL(C_DOUBLE), VALUE :: DT1_C
      INTEGER(C_INT) :: ITRI(*), IPAR(6)
C     ---- locals
      INTEGER NEL, I, K, NN, IT, IPLA, NPTT, IOFC, NELTST, ITYPTST, G_DT
      INTEGER IXC(NIXC,MVSIZ), NGL(MVSIZ), MAT(MVSIZ), PID(MVSIZ),
     .   IADC(4,MVSIZ), IPARTC(MVSIZ), NDT(1), INDXOF(MVSIZ),
     .   IOFF_DUCT(MVSIZ), IGEO(n_var_igeo,1)
      my_real
     .   X(3,4*MVSIZ), V(3,4*MVSIZ), VR(3,4*MVSIZ), GEO(1000,1),
     .   PM(250,1), OFF(MVSIZ), SIGY(MVSIZ), DT1C(MVSIZ), THK0(MVSIZ),
     .   VL1(MVSIZ,3),VL2(MVSIZ,3),VL3(MVSIZ,3),VL4(MVSIZ,3),
     .   VRL1(MVSIZ,3),VRL2(MVSIZ,3),VRL3(MVSIZ,3),VRL4(MVSIZ,3),
     .   X1G(MVSIZ),X2G(MVSIZ),X3G(MVSIZ),X4G(MVSIZ),
     .   Y1G(MVSIZ),Y2G(MVSIZ),Y3G(MVSIZ),Y4G(MVSIZ),
     .   Z1G(MVSIZ),Z2G(MVSIZ),Z3G(MVSIZ),Z4G(MVSIZ),
     .   E1X(MVSIZ),E1Y(MVSIZ),E1Z(MVSIZ),E2X(MVSIZ),E2Y(MVSIZ),
     .   E2Z(MVSIZ),E3X(MVSIZ),E3Y(MVSIZ),E3Z(MVSIZ),
     .   DIR_A(2), DIR_B(2)
      my_real
     .   STI(MVSIZ),STIR(MVSIZ),AREA(MVSIZ),PX1(MVSIZ),PX2(MVSIZ),
     .   PY1(MVSIZ),PY2(MVSIZ),X2(MVSIZ),X3(MVSIZ),X4(MVSIZ),
     .   Y2(MVSIZ),Y3(MVSIZ),Y4(MVSIZ),Z2(MVSIZ),VHX(MVSIZ),VHY(MVSIZ),
     .   A_I(MVSIZ),UX1(MVSIZ),UX2(MVSIZ),UX3(MVSIZ),UX4(MVSIZ),
     .   UY1(MVSIZ),UY2(MVSIZ),UY3(MVSIZ),UY4(MVSIZ),
     .   SHF(MVSIZ),THK02(MVSIZ),NU(MVSIZ),G(MVSIZ),YM(MVSIZ),
     .   A11(MVSIZ),A12(MVSIZ),SSP(MVSIZ),RHO(MVSIZ),H1(MVSIZ),
     .   H2(MVSIZ),H3(MVSIZ),VOL0(MVSIZ),VOL00(MVSIZ),ALPE(MVSIZ),
     .   GS(MVSIZ),SRH1(MVSIZ),SRH2(MVSIZ),SRH3(MVSIZ),A11R(MVSIZ),
     .   PM_STACK(20,1),ALDT(MVSIZ)
      DOUBLE PRECISION SMSTR(6*MVSIZ)
      my_real
     .   EXX(MVSIZ),EYY(MVSIZ),EXY(MVSIZ),EXZ(MVSIZ),EYZ(MVSIZ),
     .   KXX(MVSIZ),KYY(MVSIZ),KXY(MVSIZ),
     .   VX1(MVSIZ),VX2(MVSIZ),VX3(MVSIZ),VX4(MVSIZ),
     .   VY1(MVSIZ),VY2(MVSIZ),VY3(MVSIZ),VY4(MVSIZ),
     .   VZ1(MVSIZ),VZ2(MVSIZ),VZ3(MVSIZ),VZ4(MVSIZ),
     .   RX1(MVSIZ),RX2(MVSIZ),RX3(MVSIZ),RX4(MVSIZ),
     .   RY1(MVSIZ),RY2(MVSIZ),RY3(MVSIZ),RY4(MVSIZ),
     .   WXY(MVSIZ),GSTR(MVSIZ,8),EPSDOT(6,MVSIZ),F_DEF(MVSIZ,8),
     .   GSTRW(MVSIZ)
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
     .   H11(MVSIZ),H12(MVSIZ),H13(MVSIZ),H21(MVSIZ),H22(MVSIZ),
     .   H23(MVSIZ),H31(MVSIZ),H32(MVSIZ),H33(MVSIZ),
     .   B11(MVSIZ),B12(MVSIZ),B13(MVSIZ),B14(MVSIZ),
     .   B21(MVSIZ),B22(MVSIZ),B23(MVSIZ),B24(MVSIZ),
     .   EANI(2*MVSIZ),PARTSAV(29,MVSIZ),DT2T,DTC(1),DT2C(1),
     .   MSC(MVSIZ),DMELC(MVSIZ),DTEL(MVSIZ)
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
C     ---- array sizes of /PARAM/
      NPROPM = 250
      NPROPG = 1000
      NPARG = 100
      NPSAV = 29
      NPROPMI = 300
      NPROPGI = n_var_igeo
C     ---- engine globals at their explicit-run values
      DT1 = DT1_C
      TT = ONE
      NCYCLE = 1
      IRUN = 1
      ISIGI = 0
      ISHFRAM = 0
      IMPL_S = 0
      INVSTR = 2026
      NUMELS = 0
      NODADT = 0
      IDT1SH = 0
      IDTMINS = 0
      NADMESH = 0
      IEPSDOT = 0
      IVECTOR = 0
      LSKY = 4*NEL
      IDTMIN(3) = 2
      DTFAC1(3) = PPAR(6)
      DTMIN1(3) = PPAR(7)
      HELAS = HALF
      HVISC = HALF
      HVLIN = ZERO
      EHOUR = ZERO
      ELB%IDRAPE = 0
      CALL COQINI
      CALL COQINI_WM(WML)
C     ---- material (hm_read_mat02_jc.F90:234-339)
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
      GEO(13,1) = PPAR(1)
      GEO(14,1) = PPAR(2)
      GEO(15,1) = PPAR(3)
      GEO(37,1) = IPAR(6)
      GEO(38,1) = PPAR(5)
      IGEO(:,1) = 0
      IGEO(11,1) = 1
      PM_STACK(:,1) = ZERO
C     ---- local nodes: corner k of element I is node 4*(I-1)+K
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
        IF (ITRI(I) /= 0) IXC(5,I) = IXC(4,I)
        IXC(6,I) = 1
        IXC(7,I) = I
        IPARTC(I) = I
        PARTSAV(8,I) = ZERO
        EANI(I) = ZERO
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
      ENDDO
      DT2T = EP20
      NELTST = 0
      ITYPTST = 0
      IOFC = 0
      G_DT = 1
C     ---- CFORC3 chain
      CALL CCOOR3(1, NEL, X, IXC, GEO, OFFG,
     .     OFF, SIGY, PID, V, VR, VL1, VL2, VL3, VL4,
     .     VRL1, VRL2, VRL3, VRL4, MAT, DT1C, THKE, THK0, NGL,
     .     X1G, X2G, X3G, X4G, Y1G, Y2G,
     .     Y3G, Y4G, Z1G, Z2G, Z3G, Z4G)
      CALL CNVEC3(ELB, DIR_A, DIR_B,
     .     1, NEL, 0, 1, 1,
     .     X1G, X2G, X3G, X4G, Y1G, Y2G,
     .     Y3G, Y4G, Z1G, Z2G, Z3G, Z4G,
     .     E1X, E1Y, E1Z, E2X, E2Y, E2Z,
     .     E3X, E3Y, E3Z, NEL)
      CALL CDERI3(
     1   1,         NEL,       SMSTR,     OFFG,
     2   STI,       STIR,      AREA,      PX1,
     3   PX2,       PY1,       PY2,       X2,
     4   X3,        X4,        Y2,        Y3,
     5   Y4,        Z2,        X1G,       X2G,
     6   X3G,       X4G,       Y1G,       Y2G,
     7   Y3G,       Y4G,       Z1G,       Z2G,
     8   Z3G,       Z4G,       E1X,       E1Y,
     9   E1Z,       E2X,       E2Y,       E2Z,
     A   E3X,       E3Y,       E3Z,       VHX,
     B   VHY,       A_I,       UX1,       UX2,
     C   UX3,       UX4,       UY1,       UY2,
     D   UY3,       UY4,       NEL,       2)
      CALL CCOEF3(1, NEL, PM, MAT, GEO,
     2            PID, OFF, AREA, STI, STIR,
     3            SHF, THK0, THK02, NU,
     4            G, YM, A11, A12, THK,
     5            SSP, RHO, H1, H2, H3,
     6            VOL0, VOL00, ALPE, GS, 2,
     7            0, 2, NPTT, 0,
     8            SRH1, SRH2, SRH3, IGEO,
     9            A11R, 1, PM_STACK)
      CALL CDLEN3(1, NEL, PM, OFF, AREA,
     2             X2, X3, X4, Y2, Y3,
     3             Y4, ALDT, MAT, GEO, PID,
     4             1)
      CALL CDEFO3(1, NEL, VL1, VL2, VL3, VL4, DT1C, PX1, PX2, PY1, PY2, AREA,
     2            EXX, EYY, EXY, EXZ, EYZ, X2, X3, X4, Y2, Y3,
     3            Y4, Z2, VX1, VX2, VX3, VX4, VY1, VY2, VY3, VY4,
     4            VZ1, VZ2, VZ3, VZ4, E1X, E1Y, E1Z, E2X, E2Y, E2Z,
     5            E3X, E3Y, E3Z, 1)
      CALL CCURV3(1, NEL, VRL1, VRL2, VRL3, VRL4, PX1,
     1            PX2, PY1, PY2, AREA,
     2            RX1, RX2, RX3, RX4, RY1, RY2, RY3, RY4,
     3            E1X, E1Y, E1Z, E2X, E2Y, E2Z, E3X, E3Y,
     4            E3Z, KXX, KYY, KXY, EXZ, EYZ,
     5            WXY, 2)
      CALL CSTRA3(1, NEL, GSTR, SHF, AREA,
     2            EXX, EYY, EXY, EXZ, EYZ,
     3            KXX, KYY, KXY, DT1C, EPSDOT,
     4            FOR, MOM, 2, 2,
     6            1, 0, 0, UX1, UX2,
     7            UX3, UX4, UY1, UY2, UY3,
     8            UY4, PX1, PX2, PY1, PY2,
     9            WXY, GSTRW, F_DEF, NEL)
C     ---- element strain rate (cforc3.F:535-546), restated
      DT1L = DT1C(1)
      DTINV = DT1L / MAX(DT1L**2,EM20)
      DO I=1,NEL
        EPS_K2 = (KXX(I)**2+KYY(I)**2+KXX(I)*KYY(I)+FOURTH*KXY(I)**2)
     .         * ONE_OVER_9*THK(I)**2
        EPS_M2 = FOUR_OVER_3*(EXX(I)**2+EYY(I)**2+EXX(I)*EYY(I)
     .         + FOURTH*EXY(I)**2)
        EPSD_PG(I) = SQRT(EPS_K2 + EPS_M2)*DTINV
      ENDDO
      EPSDE(1:NEL) = ONE * EPSD_PG(1:NEL) + (ONE - ONE) * EPSDE(1:NEL)
C     ---- MULAWC law-2 point loop (mulawc.F90:530-3046), restated
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
C     ---- back in CFORC3 (cforc3.F:587-589)
      DO I=1,NEL
        VISCMX(I) = SQRT(ONE + VISCMX(I)*VISCMX(I)) - VISCMX(I)
      ENDDO
      CALL CHVIS3(
     2   1, NEL, THK, HOUR, OFF, PX1, PX2, PY1, PY2,
     3   IXC, DT1C, SSP, RHO, STI, VX1, VX2, VX3, VX4, VY1,
     4   VY2, VY3, VY4, VZ1, VZ2, VZ3, VZ4, AREA, THK0, VHX,
     5   VHY, SHF, Z2, EANI, STIR, VISCMX, G, A11,
     6   H1, H2, H3, YM, NU, THK02, ALPE, H11,
     7   H12, H13, H21, H22, H23, H31, H32, H33,
     8   B11, B12, B13, B14, B21, B22, B23, B24,
     9   RX1, RX2, RX3, RX4, RY1, RY2, RY3, RY4,
     A   IPARTC, PARTSAV,
     B   1, 0, 2, 0,
     C   SRH1, SRH2, SRH3, 1,
     D   0, A11R, NEL)
      CALL CDT3(1, NEL, YM, OFF, DT2T,
     2          NELTST, ITYPTST, STI, STIR, OFFG,
     3          DTC, NDT, DT2C, IXC, SSP,
     4          VISCMX, PX1, PX2, PY1, PY2,
     5          VOL0, VOL00, RHO, ALDT, ALPE,
     6          INDXOF, NGL, 2, IOFC, MSC,
     7          DMELC, 0, G_DT, DTEL)
      CALL CFINT3(1, NEL, FOR, MOM, THK0, THK02,
     2            PX1, PX2, PY1, PY2, AREA, Z2,
     3            F11, F12, F13, F14, F21, F22,
     4            F23, F24, F31, F32, F33, F34,
     5            H11, H12, H13, H21, H22, H23,
     6            H31, H32, H33, B11, B12, B13,
     7            B14, B21, B22, B23, B24, NEL,
     8            M11, M12, M13, M14, M21, M22,
     9            M23, M24, M31, M32, M33, M34,
     A            E1X, E1Y, E1Z, E2X, E2Y, E2Z,
     B            E3X, E3Y, E3Z, 1, NPTT, FZERO)
      CALL CUPDT3P(1, NEL, OFFG, OFF, STI,
     2             STIR, FSKY, FSKYV, IADC, IXC,
     3             F11, F12, F13, F14, F21,
     4             F22, F23, F24, F31, F32,
     5             F33, F34, M11, M12, M13,
     6             M14, M21, M22, M23, M24,
     7             M31, M32, M33, M34,
     8             EINT, PARTSAV, MAT, IPARTC, PM,
     9             AREA, THK, 0, THEM, FTHESKY,
     A             CONDNSKY, CONDE, 0)
      DO I=1,NEL
        DO K=1,4
          NN = 4*(I-1)+K
          FSKYO(1:8,NN) = FSKY(1:8,NN)
        ENDDO
        DTE_O(I) = DTEL(I)
        EHOU_O(I) = PARTSAV(8,I)
      ENDDO
      DEALLOCATE(MATP%UPARAM)
      DEALLOCATE(MATP%IPARAM)
      RETURN
      END

C  The shell-path constants of constant_mod, in the order of
C  oracle_shell_constants, and the COQINI tables for NPT points.
      SUBROUTINE REF_SHELL_CONSTANTS(OUT) BIND(C, NAME='ref_shell_constants')
      USE ISO_C_BINDING
#include      "implicit_f.inc"
      REAL(C_DOUBLE) :: OUT(16)
      OUT(1) = ZEP81
      OUT(2) = ZEP072169
      OUT(3) = ONEP414
      OUT(4) = FIVE_OVER_6
      OUT(5) = FOUR_OVER_3
      OUT(6) = ONE_OVER_9
      OUT(7) = ONE_OVER_12
      OUT(8) = FOUR_OVER_5
      OUT(9) = EM30
      OUT(10) = SQRT(TWO)* HALF
      OUT(11) = EM15
      OUT(12) = EM20
      OUT(13) = EM07
      OUT(14) = EP20
      OUT(15) = THREE_OVER_4
      OUT(16) = THIRD
      RETURN
      END

      SUBROUTINE REF_SHELL_TABLES(NPT, Z, W, M) BIND(C, NAME='ref_shell_tables')
      USE ISO_C_BINDING
#include      "implicit_f.inc"
#include      "com20_c.inc"
      INTEGER(C_INT), VALUE :: NPT
      REAL(C_DOUBLE) :: Z(*), W(*), M(*)
      my_real WM1(11,11)
      INTEGER I
      CALL COQINI
      CALL COQINI_WM(WM1)
      DO I=1,NPT
        Z(I) = Z0(I,NPT)
        W(I) = WF(I,NPT)
        M(I) = WM1(I,NPT)
      ENDDO
      RETURN
      END
