C  harness_brick.F -- test harness, not product code.
C
C  Drives the reference's own brick leaf kernels, compiled from their sources
C  under the reference tree by oracle/build_ref.py, in the order of SFORC3
C  (engine/source/elements/solid/solide/sforc3.F:466-1733) for the default
C  solid: Isolid=1 (JHBE=1), JCVT=1, Ismstr=4, law 2, PARITH/ON.
C  The MMAIN law-2 prologue (materials/mat_share/mmain.F90:678-737) is
C  restated here because MMAIN itself dispatches over every law.
C
C  Entry point (C binding) ref_brick_group: one group of NEL <= 128 elements,
C  inputs gathered per element corner (slot of corner k of element e is
C  8*(e-1)+k), outputs the FSKY slots exactly as SCUMU3P writes them.
      SUBROUTINE REF_BRICK_GROUP(NEL_C, XG, VG, SIG, EINT, RHO, VOL0,
     .   PLA, QVIS, OFFG, EPSD, DMG, DT1_C, MPAR, IPAR, FSKYO, DT2T_O,
     .   EHOU_O, VOLN_O, MSTOP_O) BIND(C, NAME='ref_brick_group')
      USE ISO_C_BINDING
      USE MATPARAM_DEF_MOD
      USE GLOB_THERM_MOD
      USE ALEANIM_MOD, only : FANI_CELL_
      USE PROP_PARAM_MOD , only : n_var_igeo
      USE ELEMENT_MOD , only : NIXS
#include      "implicit_f.inc"
#include      "mvsiz_p.inc"
#include      "param_c.inc"
#include      "vect01_c.inc"
#include      "com01_c.inc"
#include      "com04_c.inc"
#include      "com06_c.inc"
#include      "com08_c.inc"
#include      "impl1_c.inc"
#include      "scr02_c.inc"
#include      "scr05_c.inc"
#include      "scr07_c.inc"
#include      "scr17_c.inc"
#include      "scr18_c.inc"
#include      "parit_c.inc"
#include      "sms_c.inc"
#include      "inter22.inc"
#include      "cong1_c.inc"
      INTEGER(C_INT), VALUE :: NEL_C
      REAL(C_DOUBLE) :: XG(8,3,*), VG(8,3,*), SIG(NEL_C,6), EINT(*),
     .   RHO(*), VOL0(*), PLA(*), QVIS(*), OFFG(*), EPSD(*), DMG(*),
     .   MPAR(16), FSKYO(8,*), DT2T_O, EHOU_O(*), VOLN_O(*)
      REAL(C_DOUBLE), VALUE :: DT1_C
      INTEGER(C_INT) :: IPAR(4), MSTOP_O
C     ---- locals
      INTEGER NEL, I, K, NN
      INTEGER, PARAMETER :: NUMG = 1
      INTEGER IXS(NIXS,MVSIZ), NGL(MVSIZ), MXT(MVSIZ), NGEO(MVSIZ),
     .   NC1(MVSIZ),NC2(MVSIZ),NC3(MVSIZ),NC4(MVSIZ),
     .   NC5(MVSIZ),NC6(MVSIZ),NC7(MVSIZ),NC8(MVSIZ),
     .   IPARG(100), IADS(8,MVSIZ), IPARTS(MVSIZ), IPARG1(100),
     .   IPM(300,1), NELTST, ITYPTST, PID(MVSIZ), G_DT
      INTEGER IGEO(n_var_igeo,1)
      my_real
     .   X(3,8*MVSIZ), V(3,8*MVSIZ), W(3,1), VIS(MVSIZ), VD2(MVSIZ),
     .   OFF(MVSIZ), RHOO(MVSIZ),
     .   GAMA0(MVSIZ,6), GAMA(MVSIZ,6), GAMA_R(MVSIZ,6),
     .   R11(MVSIZ),R12(MVSIZ),R13(MVSIZ),R21(MVSIZ),R22(MVSIZ),
     .   R23(MVSIZ),R31(MVSIZ),R32(MVSIZ),R33(MVSIZ),
     .   VGAX(MVSIZ),VGAY(MVSIZ),VGAZ(MVSIZ),VGA2(MVSIZ),
     .   XGAX(MVSIZ),XGAY(MVSIZ),XGAZ(MVSIZ),
     .   XGXA2(MVSIZ),XGYA2(MVSIZ),XGZA2(MVSIZ),
     .   XGXYA(MVSIZ),XGYZA(MVSIZ),XGZXA(MVSIZ)
      my_real
     .   X1(MVSIZ),X2(MVSIZ),X3(MVSIZ),X4(MVSIZ),
     .   X5(MVSIZ),X6(MVSIZ),X7(MVSIZ),X8(MVSIZ),
     .   Y1(MVSIZ),Y2(MVSIZ),Y3(MVSIZ),Y4(MVSIZ),
     .   Y5(MVSIZ),Y6(MVSIZ),Y7(MVSIZ),Y8(MVSIZ),
     .   Z1(MVSIZ),Z2(MVSIZ),Z3(MVSIZ),Z4(MVSIZ),
     .   Z5(MVSIZ),Z6(MVSIZ),Z7(MVSIZ),Z8(MVSIZ),
     .   VX1(MVSIZ),VX2(MVSIZ),VX3(MVSIZ),VX4(MVSIZ),
     .   VX5(MVSIZ),VX6(MVSIZ),VX7(MVSIZ),VX8(MVSIZ),
     .   VY1(MVSIZ),VY2(MVSIZ),VY3(MVSIZ),VY4(MVSIZ),
     .   VY5(MVSIZ),VY6(MVSIZ),VY7(MVSIZ),VY8(MVSIZ),
     .   VZ1(MVSIZ),VZ2(MVSIZ),VZ3(MVSIZ),VZ4(MVSIZ),
     .   VZ5(MVSIZ),VZ6(MVSIZ),VZ7(MVSIZ),VZ8(MVSIZ)
      DOUBLE PRECISION
     .   XD1(MVSIZ),XD2(MVSIZ),XD3(MVSIZ),XD4(MVSIZ),
     .   XD5(MVSIZ),XD6(MVSIZ),XD7(MVSIZ),XD8(MVSIZ),
     .   YD1(MVSIZ),YD2(MVSIZ),YD3(MVSIZ),YD4(MVSIZ),
     .   YD5(MVSIZ),YD6(MVSIZ),YD7(MVSIZ),YD8(MVSIZ),
     .   ZD1(MVSIZ),ZD2(MVSIZ),ZD3(MVSIZ),ZD4(MVSIZ),
     .   ZD5(MVSIZ),ZD6(MVSIZ),ZD7(MVSIZ),ZD8(MVSIZ),
     .   XDP(3,1), X0(MVSIZ,8), Y0(MVSIZ,8), Z0(MVSIZ,8),
     .   SAV(MVSIZ,21), VOLDP(MVSIZ), VOL0DP(MVSIZ)
      my_real
     .   VOLN(MVSIZ), DELTAX(MVSIZ),
     .   PX1(MVSIZ),PX2(MVSIZ),PX3(MVSIZ),PX4(MVSIZ),
     .   PY1(MVSIZ),PY2(MVSIZ),PY3(MVSIZ),PY4(MVSIZ),
     .   PZ1(MVSIZ),PZ2(MVSIZ),PZ3(MVSIZ),PZ4(MVSIZ),
     .   PX5(MVSIZ),PX6(MVSIZ),PX7(MVSIZ),PX8(MVSIZ),
     .   PY5(MVSIZ),PY6(MVSIZ),PY7(MVSIZ),PY8(MVSIZ),
     .   PZ5(MVSIZ),PZ6(MVSIZ),PZ7(MVSIZ),PZ8(MVSIZ),
     .   PX1H1(MVSIZ),PX1H2(MVSIZ),PX1H3(MVSIZ),
     .   PX2H1(MVSIZ),PX2H2(MVSIZ),PX2H3(MVSIZ),
     .   PX3H1(MVSIZ),PX3H2(MVSIZ),PX3H3(MVSIZ),
     .   PX4H1(MVSIZ),PX4H2(MVSIZ),PX4H3(MVSIZ),
     .   JAC1(MVSIZ),JAC2(MVSIZ),JAC3(MVSIZ),
     .   JAC4(MVSIZ),JAC5(MVSIZ),JAC6(MVSIZ),
     .   N1X(MVSIZ),N2X(MVSIZ),N3X(MVSIZ),N4X(MVSIZ),N5X(MVSIZ),
     .   N6X(MVSIZ),N1Y(MVSIZ),N2Y(MVSIZ),N3Y(MVSIZ),N4Y(MVSIZ),
     .   N5Y(MVSIZ),N6Y(MVSIZ),N1Z(MVSIZ),N2Z(MVSIZ),N3Z(MVSIZ),
     .   N4Z(MVSIZ),N5Z(MVSIZ),N6Z(MVSIZ)
      my_real
     .   DXX(MVSIZ),DXY(MVSIZ),DXZ(MVSIZ),DYX(MVSIZ),DYY(MVSIZ),
     .   DYZ(MVSIZ),DZX(MVSIZ),DZY(MVSIZ),DZZ(MVSIZ),
     .   D4(MVSIZ),D5(MVSIZ),D6(MVSIZ),WXX(MVSIZ),WYY(MVSIZ),
     .   WZZ(MVSIZ),VXR(MVSIZ),VYR(MVSIZ),VZR(MVSIZ),DIVDE(MVSIZ),
     .   FLUX(MVSIZ,6), FLU1(MVSIZ), DVOL(MVSIZ), TAG22(MVSIZ),
     .   AMU(MVSIZ), S1(MVSIZ),S2(MVSIZ),S3(MVSIZ),S4(MVSIZ),
     .   S5(MVSIZ),S6(MVSIZ), VOL_AVG(MVSIZ), RHOREF(MVSIZ),
     .   PM(250,1), GEO(1000,1)
      my_real
     .   STI(MVSIZ), MUMAX(MVSIZ), CXX(MVSIZ), AIRE(MVSIZ),
     .   PNEW(MVSIZ), PSH(MVSIZ), QNEW(MVSIZ), SSP_EQ(MVSIZ),
     .   SIGY(MVSIZ), DEFP(MVSIZ), DPLA(MVSIZ), EPSP(MVSIZ),
     .   TSTAR(MVSIZ), ET(MVSIZ), MSSA(MVSIZ), DMELS(MVSIZ),
     .   TEMPEL(MVSIZ), SIGB(MVSIZ,6), AL_IMP(MVSIZ),
     .   SIGNOR(MVSIZ,6), CONDE(MVSIZ), DTEL(MVSIZ), RHOSP(MVSIZ),
     .   DPDM(MVSIZ), FHEAT(MVSIZ), DT2T, EANI(MVSIZ),
     .   PARTSAV(29,MVSIZ), DFE(MVSIZ,3), PP(MVSIZ), SVIS(MVSIZ,6)
      my_real
     .   F11(MVSIZ),F21(MVSIZ),F31(MVSIZ),F12(MVSIZ),F22(MVSIZ),
     .   F32(MVSIZ),F13(MVSIZ),F23(MVSIZ),F33(MVSIZ),F14(MVSIZ),
     .   F24(MVSIZ),F34(MVSIZ),F15(MVSIZ),F25(MVSIZ),F35(MVSIZ),
     .   F16(MVSIZ),F26(MVSIZ),F36(MVSIZ),F17(MVSIZ),F27(MVSIZ),
     .   F37(MVSIZ),F18(MVSIZ),F28(MVSIZ),F38(MVSIZ),
     .   FSKY(8,8*MVSIZ), FSKYV(1,8), AR(3,1), FR_WAVE(1), FR_WAV(1),
     .   MDUM(MVSIZ), THEM(MVSIZ,8), FTHESKY(1), CONDNSKY(1)
      TYPE(MATPARAM_STRUCT_) :: MATP
      TYPE(GLOB_THERM_) :: GTH
      TYPE(FANI_CELL_) :: FANI
C=======================================================================
      NEL = NEL_C
C     ---- array sizes of /PARAM/ (starter0.F:529-531, rdresa.F:327)
      NPROPM = 250
      NPROPG = 1000
      NPARG = 100
      NPSAV = 29
      NPROPMI = 300
      NPROPGI = n_var_igeo
C     ---- group context (INITBUF /VECT01/, sgrtails.F:1449-1481)
      LFT = 1
      LLT = NEL
      NFT = 0
      MTN = 2
      ITY = 1
      NPT = 1
      JALE = 0
      JEUL = 0
      JTUR = 0
      JTHE = 0
      JLAG = 1
      ISMSTR = 4
      JHBE = 1
      JCVT = 1
      JSPH = 0
      JCLOSE = 0
      IREP = 0
      IINT = 1
      IGTYP = 14
      ISORTH = 0
      ISRAT = 0
      ISROT = 0
      JSMS = 0
      IPARTSPH = 0
C     ---- engine globals at their explicit-run values
      DT1 = DT1_C
      IRESP = 0
      IMPL_S = 0
      IDYNA = 0
      INCONV = 1
      IMCONV = 1
      ISCAU = 0
      IMP_LR = 0
      ISMDISP = 0
      ILINE = 0
      INVSTR = 2026
      N2D = 0
      INTEG8 = 0
      NODADT = 0
      IDTMINS = 0
      INEG_V = 0
      INT22 = 0
      IMPL = 0
      IVECTOR = 0
      IDTMIN(1) = 0
      KDTSMSTR = 1
      DTFAC1(1) = MPAR(14)
      MSTOP = 0
      LSKY = 8*NEL
      NUMMAT = 1
      NUMGEO = 1
      FANI%IS_VORT_X_REQUESTED = .FALSE.
      FANI%IS_VORT_Y_REQUESTED = .FALSE.
      FANI%IS_VORT_Z_REQUESTED = .FALSE.
C     ---- material (hm_read_mat02_jc.F90:319-339)
      MATP%RHO   = MPAR(1)
      MATP%SHEAR = MPAR(2)
      MATP%BULK  = MPAR(3)
      MATP%IEOS  = 0
      ALLOCATE(MATP%UPARAM(12))
      ALLOCATE(MATP%IPARAM(4))
      MATP%UPARAM(:) = ZERO
      MATP%UPARAM(1) = MPAR(4)
      MATP%UPARAM(2) = MPAR(5)
      MATP%UPARAM(3) = MPAR(6)
      MATP%UPARAM(4) = MPAR(7)
      MATP%UPARAM(5) = MPAR(8)
      MATP%UPARAM(6) = MPAR(9)
      MATP%UPARAM(7) = MPAR(10)
      MATP%UPARAM(8) = ZERO
      MATP%UPARAM(9) = EP20
      MATP%UPARAM(10)= MPAR(11)
      MATP%IPARAM(1) = 0
      MATP%IPARAM(2) = IPAR(1)
      MATP%IPARAM(3) = IPAR(2)
      MATP%IPARAM(4) = IPAR(3)
      MATP%THERM%TREF  = 300.0D0
      MATP%THERM%TMELT = EP20
      MATP%THERM%RHOCP = ZERO
      PM(:,1) = ZERO
      PM(1,1) = MPAR(1)
      IPM(:,1) = 0
      GEO(:,1) = ZERO
      GEO(13,1) = MPAR(15)
      GEO(14,1) = MPAR(12)
      GEO(15,1) = MPAR(13)
      IGEO(:,1) = 0
      IPARG(:) = 0
      IPARG1(:) = 0
      GTH%ITHERM_FE = 0
      GTH%NODADT_THERM = 0
C     ---- local node arrays: corner k of element I is node 8*(I-1)+K
      DO I=1,NEL
        IXS(1,I) = 1
        DO K=1,8
          NN = 8*(I-1)+K
          IXS(K+1,I) = NN
          X(1,NN) = XG(K,1,I)
          X(2,NN) = XG(K,2,I)
          X(3,NN) = XG(K,3,I)
          V(1,NN) = VG(K,1,I)
          V(2,NN) = VG(K,2,I)
          V(3,NN) = VG(K,3,I)
          IADS(K,I) = NN
        ENDDO
        IXS(10,I) = 1
        IXS(11,I) = I
        IPARTS(I) = I
        PID(I) = 1
        PARTSAV(8,I) = ZERO
        EANI(I) = ZERO
        SAV(I,:) = ZERO
        TAG22(I) = ZERO
        FLU1(I) = ZERO
        FLUX(I,:) = ZERO
        SVIS(I,:) = ZERO
        TSTAR(I) = ZERO
        TEMPEL(I) = ZERO
        SIGB(I,:) = ZERO
        DPDM(I) = ZERO
        FHEAT(I) = ZERO
        VOL0DP(I) = VOL0(I)
      ENDDO
      DT2T = EP20
      NELTST = 0
      ITYPTST = 0
C     ---- SFORC3 chain
      CALL SRCOOR3(X,IXS,V,W,GAMA0,GAMA,
     .     X1, X2, X3, X4, X5, X6, X7, X8,
     .     Y1, Y2, Y3, Y4, Y5, Y6, Y7, Y8,
     .     Z1, Z2, Z3, Z4, Z5, Z6, Z7, Z8,
     .     VX1, VX2, VX3, VX4, VX5, VX6, VX7, VX8,
     .     VY1, VY2, VY3, VY4, VY5, VY6, VY7, VY8,
     .     VZ1, VZ2, VZ3, VZ4, VZ5, VZ6, VZ7, VZ8,
     .     VD2,VIS,OFFG,OFF,SAV,RHO,RHOO,
     .     R11, R12, R13, R21, R22, R23, R31, R32, R33,
     .     NC1,NC2,NC3,NC4,NC5,NC6,NC7,NC8,NGL,MXT,NGEO,
     .     0, VGAX, VGAY, VGAZ, VGA2,
     .     XD1, XD2, XD3, XD4, XD5, XD6, XD7, XD8,
     .     YD1, YD2, YD3, YD4, YD5, YD6, YD7, YD8,
     .     ZD1, ZD2, ZD3, ZD4, ZD5, ZD6, ZD7, ZD8,
     .     XDP, X0 , Y0 , Z0 , NEL, XGAX, XGAY, XGAZ,
     .     XGXA2,XGYA2  ,XGZA2 ,XGXYA ,XGYZA ,XGZXA,IPARG,
     .     GAMA_R)
      CALL SDERI3(
     1   OFF,       VOLN,      NGL,       XD1,
     2   XD2,       XD3,       XD4,       XD5,
     3   XD6,       XD7,       XD8,       YD1,
     4   YD2,       YD3,       YD4,       YD5,
     5   YD6,       YD7,       YD8,       ZD1,
     6   ZD2,       ZD3,       ZD4,       ZD5,
     7   ZD6,       ZD7,       ZD8,       PX1,
     8   PX2,       PX3,       PX4,       PY1,
     9   PY2,       PY3,       PY4,       PZ1,
     A   PZ2,       PZ3,       PZ4,       PX1H1,
     B   PX1H2,     PX1H3,     PX2H1,     PX2H2,
     C   PX2H3,     PX3H1,     PX3H2,     PX3H3,
     D   PX4H1,     PX4H2,     PX4H3,     JAC1,
     E   JAC2,      JAC3,      JAC4,      JAC5,
     F   JAC6,      SAV,       OFFG,      NEL,
     G   VOLDP,     JHBE,      ISMSTR,    JLAG)
      CALL SDLEN3(
     1   VOLN,    DELTAX,  X1,      X2,
     2   X3,      X4,      X5,      X6,
     3   X7,      X8,      Y1,      Y2,
     4   Y3,      Y4,      Y5,      Y6,
     5   Y7,      Y8,      Z1,      Z2,
     6   Z3,      Z4,      Z5,      Z6,
     7   Z7,      Z8,      N1X,     N2X,
     8   N3X,     N4X,     N5X,     N6X,
     9   N1Y,     N2Y,     N3Y,     N4Y,
     A   N5Y,     N6Y,     N1Z,     N2Z,
     B   N3Z,     N4Z,     N5Z,     N6Z,
     C   NEL,     MTN,     JALE,    JEUL)
      CALL SDEFO3(FANI,
     1   PX1,     PX2,     PX3,     PX4,
     2   PY1,     PY2,     PY3,     PY4,
     3   PZ1,     PZ2,     PZ3,     PZ4,
     4   PX5,     PX6,     PX7,     PX8,
     5   PY5,     PY6,     PY7,     PY8,
     6   PZ5,     PZ6,     PZ7,     PZ8,
     7   VX1,     VX2,     VX3,     VX4,
     8   VX5,     VX6,     VX7,     VX8,
     9   VY1,     VY2,     VY3,     VY4,
     A   VY5,     VY6,     VY7,     VY8,
     B   VZ1,     VZ2,     VZ3,     VZ4,
     C   VZ5,     VZ6,     VZ7,     VZ8,
     D   DXX,     DXY,     DXZ,     DYX,
     E   DYY,     DYZ,     DZX,     DZY,
     F   DZZ,     D4,      D5,      D6,
     G   WXX,     WYY,     WZZ,     VXR,
     H   VYR,     VZR,     NEL,
     I   NFT,     ISMSTR,  JEUL,    JHBE,
     J   JCVT,    ISROT)
      DIVDE(1:NEL) = DT1*(DXX(1:NEL)+ DYY(1:NEL)+ DZZ(1:NEL))
      CALL SRHO3(
     1   PM,         VOL0,       RHO,        EINT,
     2   DIVDE,      FLUX,       FLU1,       VOLN,
     3   DVOL,       NGL,        MXT,        OFF,
     4   0,          TAG22,      VOLDP,      VOL0DP,
     5   AMU,        OFFG,       NEL,        MTN,
     6   JALE,       ISMSTR,     JEUL,       JLAG,
     7   MVSIZ, 6, 0)
      CALL SRMALLA3(SIG,S1,S2,S3,S4,S5,S6,OFFG,OFF,NEL)
C     ---- MMAIN law-2 prologue (mmain.F90:678-737)
      DO I=1,NEL
        VOL_AVG(I) = VOLN(I) - HALF*DVOL(I)
        AMU(I) = RHO(I)/PM(1,1) - ONE
        RHOREF(I) = RHO(I)
        TSTAR(I) = ZERO
        STI(I) = ZERO
        VIS(I) = ZERO
      ENDDO
      CALL M2LAW(MATP,
     1   PM,      OFF,        SIG,     EINT,
     2   RHO,     QVIS,       PLA,     EPSD,
     3   VOL0,    STI,        DT2T,    NELTST,
     4   ITYPTST, OFFG,       GEO,     PID,
     5   AMU,     VOL_AVG,    MUMAX,   MXT,
     6   NGL,     CXX,        DVOL,    AIRE,
     7   VOLN,    VD2,        DELTAX,  VIS,
     8   DXX,     DYY,        DZZ,     D4,
     9   D5,      D6,         PNEW,    PSH,
     A   QNEW,    SSP_EQ,     S1,      S2,
     B   S3,      S4,         S5,      S6,
     C   SIGY,    DEFP,       DPLA,    JLAG,
     D   EPSP,    TSTAR,      ET,      MSSA,
     E   DMELS,   TEMPEL,     SIGB,    AL_IMP,
     F   SIGNOR,  CONDE,      DTEL,    G_DT,
     G   NEL,     IPM,        RHOREF,  RHOSP,
     H   1,       DMG,        ITY,     JTUR,
     I   JTHE,    JSPH,       ISMSTR,  JSMS,
     J   1,       DPDM,       FHEAT,   GTH,
     H   1,       IGEO)
      CALL SHVIS3(
     1   PM,       RHO,      OFF,      VX1,
     2   VX2,      VX3,      VX4,      VX5,
     3   VX6,      VX7,      VX8,      VY1,
     4   VY2,      VY3,      VY4,      VY5,
     5   VY6,      VY7,      VY8,      VZ1,
     6   VZ2,      VZ3,      VZ4,      VZ5,
     7   VZ6,      VZ7,      VZ8,      F11,
     8   F21,      F31,      F12,      F22,
     9   F32,      F13,      F23,      F33,
     A   F14,      F24,      F34,      F15,
     B   F25,      F35,      F16,      F26,
     C   F36,      F17,      F27,      F37,
     D   F18,      F28,      F38,      PX1H1,
     E   PX1H2,    PX1H3,    PX2H1,    PX2H2,
     F   PX2H3,    PX3H1,    PX3H2,    PX3H3,
     G   PX4H1,    PX4H2,    PX4H3,    VOLN,
     H   MXT,      CXX,      VIS,      VD2,
     I   DELTAX,   EANI,     NGEO,     GEO,
     J   PARTSAV,  IPARTS,   OFFG,     VOL0,
     K   IPARG1,   0,        NEL,      NFT,
     L   MTN,      ISMSTR,   JLAG,     JHBE)
      CALL SFINT3(
     1   SIG,     PX1,     PX2,     PX3,
     2   PX4,     PY1,     PY2,     PY3,
     3   PY4,     PZ1,     PZ2,     PZ3,
     4   PZ4,     PX5,     PX6,     PX7,
     5   PX8,     PY5,     PY6,     PY7,
     6   PY8,     PZ5,     PZ6,     PZ7,
     7   PZ8,     F11,     F21,     F31,
     8   F12,     F22,     F32,     F13,
     9   F23,     F33,     F14,     F24,
     A   F34,     F15,     F25,     F35,
     B   F16,     F26,     F36,     F17,
     C   F27,     F37,     F18,     F28,
     D   F38,     VOLN,    QNEW,    N1X,
     E   N2X,     N3X,     N4X,     N5X,
     F   N6X,     N1Y,     N2Y,     N3Y,
     G   N4Y,     N5Y,     N6Y,     N1Z,
     H   N2Z,     N3Z,     N4Z,     N5Z,
     I   N6Z,     DFE,     PP,      IXS,
     J   NEL,     NFT,     JALE,    JEUL,
     K   SVIS)
      CALL SRROTA3(
     1   R11,     R21,     R31,     R12,
     2   R22,     R32,     R13,     R23,
     3   R33,     F11,     F12,     F13,
     4   F14,     F15,     F16,     F17,
     5   F18,     F21,     F22,     F23,
     6   F24,     F25,     F26,     F27,
     7   F28,     F31,     F32,     F33,
     8   F34,     F35,     F36,     F37,
     9   F38,     NEL)
      FSKY(:,1:8*NEL) = ZERO
      CALL SCUMU3P(
     1   OFFG,    STI,     FSKY,    FSKYV,
     2   IADS,    F11,     F21,     F31,
     3   F12,     F22,     F32,     F13,
     4   F23,     F33,     F14,     F24,
     5   F34,     F15,     F25,     F35,
     6   F16,     F26,     F36,     F17,
     7   F27,     F37,     F18,     F28,
     8   F38,     NC1,     NC2,     NC3,
     9   NC4,     NC5,     NC6,     NC7,
     A   NC8,     AR,      FR_WAVE, FR_WAV,
     B   MDUM,    MDUM,    MDUM,    MDUM,
     C   MDUM,    MDUM,    MDUM,    MDUM,
     D   MDUM,    MDUM,    MDUM,    MDUM,
     E   MDUM,    MDUM,    MDUM,    MDUM,
     F   MDUM,    MDUM,    MDUM,    MDUM,
     G   MDUM,    MDUM,    MDUM,    MDUM,
     H   THEM,    FTHESKY, CONDNSKY,CONDE,
     I   NEL,     NFT,     JTHE,    ISROT,
     J   0 ,      0)
C     ---- outputs
      DO I=1,8*NEL
        DO K=1,8
          FSKYO(K,I) = FSKY(K,I)
        ENDDO
      ENDDO
      DO I=1,NEL
        EHOU_O(I) = PARTSAV(8,I)
        VOLN_O(I) = VOLN(I)
      ENDDO
      DT2T_O = DT2T
      MSTOP_O = MSTOP
      DEALLOCATE(MATP%UPARAM)
      DEALLOCATE(MATP%IPARAM)
      RETURN
      END
C
C  Reference constants exported for the oracle's bit-pattern check.
      SUBROUTINE REF_CONSTANTS(C) BIND(C, NAME='ref_constants')
      USE ISO_C_BINDING
#include      "implicit_f.inc"
      REAL(C_DOUBLE) :: C(16)
      C(1) = ONEP333
      C(2) = THIRD
      C(3) = TWO_THIRD
      C(4) = ONE_OVER_64
      C(5) = EM15
      C(6) = EM20
      C(7) = ONEP1
      C(8) = FOURTH
      C(9) = EP20
      C(10) = 1.1*ONE
      C(11) = EP06
      C(12) = ZERO
      RETURN
      END
