C  harness_int7.F -- test harness, not product code.
C
C  Drives the reference's own /INTER/TYPE7 force routines, compiled from their
C  sources under the reference tree by oracle/build_ref.py, over a candidate
C  list in batches of NVSIZ, in I7MAINF's order
C  (engine/source/interfaces/int07/i7mainf.F:529-737):
C    I7COR3   gather of the secondary node and the main segment, STIF, GAPV
C    I7DST3   projection on the segment's four triangles, penetration test,
C             compression of the batch to the penetrating pairs
C    I7FOR3   penalty, damping and friction forces, contact time step
C    I7ASS2   PARITH/ON skyline entries FSKYI(·,1:4) and their nodes ISKY
C             (the routine I7ASS33 calls for IPARIT>0, KDTINT=0,
C             i7ass3.F:195-203)
C  Restated (not called): FRICTIONPARTS_MODEL_ISOT's INTFRIC=0 assignments
C  FRICC = FRIC, VISCFFRIC = VISCF (frictionparts_model.F:394-399).
C  The path: IGAP=0, IGSTI=1, INACTI=0, ICURV=0, IADM=0, no heat exchange,
C  no tied contact, IVIS2=0, IFQ=0, MFROT=0, isotropic friction, one domain.
C
C  Entry point (C binding) ref_int7_batch.  PARAMS = GAP, FRIC, VISC
C  (critical damping), VISCF (friction damping); DTS = DT1, DT12.  NSV,
C  IRECT, CAND_N (index into NSV) and CAND_E (index into IRECT) are 1-based.
C  Out: NSKY skyline entries, FSKYO(4,NSKY) and ISKYO(NSKY) in FSKYI order;
C  OUT(1) = DT2T (in and out), OUT(2) = ECONT, OUT(3) = ECONTV,
C  OUT(4) = ECONTD added by this call, OUT(5) = penetrating pairs;
C  PEN(NC) = 1 for the candidates I7DST3 keeps, 0 otherwise.
      SUBROUTINE REF_INT7_BATCH(NN_C, X, V, MS, NSN_C, NSV, NRTM_C,
     .   IRECT, STFM, STFNS, NC_C, CAND_N, CAND_E, PARAMS, DTS,
     .   NSKY, FSKYO, ISKYO, OUT, PEN) BIND(C, NAME='ref_int7_batch')
      USE ISO_C_BINDING
      USE TRI7BOX
      USE H3D_MOD
      USE OUTPUT_MOD
#include      "implicit_f.inc"
#include      "mvsiz_p.inc"
#include      "com01_c.inc"
#include      "com04_c.inc"
#include      "com06_c.inc"
#include      "com08_c.inc"
#include      "scr05_c.inc"
#include      "scr11_c.inc"
#include      "scr14_c.inc"
#include      "scr16_c.inc"
#include      "scr18_c.inc"
#include      "sms_c.inc"
#include      "parit_c.inc"
#include      "param_c.inc"
#include      "impl1_c.inc"
      INTEGER(C_INT), VALUE :: NN_C, NSN_C, NRTM_C, NC_C
      REAL(C_DOUBLE) :: X(3,*), V(3,*), MS(*), STFM(*), STFNS(*),
     .   PARAMS(4), DTS(2), FSKYO(4,*), OUT(5)
      INTEGER(C_INT) :: NSV(*), IRECT(4,*), CAND_N(*), CAND_E(*),
     .   NSKY, ISKYO(*), PEN(*)
C     ---- batch arrays
      INTEGER IX1(MVSIZ), IX2(MVSIZ), IX3(MVSIZ), IX4(MVSIZ),
     .   NSVG(MVSIZ), KINI(MVSIZ), CN_LOC(MVSIZ), CE_LOC(MVSIZ),
     .   INDEX2(MVSIZ), IELECI(MVSIZ), NSMS(MVSIZ),
     .   IPARTFRICSI(MVSIZ), IPARTFRICMI(MVSIZ), IREP_FRICMI(MVSIZ),
     .   INDEXORTH(MVSIZ), INDEXISOT(MVSIZ)
      my_real
     .   X1(MVSIZ), X2(MVSIZ), X3(MVSIZ), X4(MVSIZ),
     .   Y1(MVSIZ), Y2(MVSIZ), Y3(MVSIZ), Y4(MVSIZ),
     .   Z1(MVSIZ), Z2(MVSIZ), Z3(MVSIZ), Z4(MVSIZ),
     .   XI(MVSIZ), YI(MVSIZ), ZI(MVSIZ), STIF(MVSIZ), GAPV(MVSIZ),
     .   VXI(MVSIZ), VYI(MVSIZ), VZI(MVSIZ), MSI(MVSIZ),
     .   NX1(MVSIZ), NX2(MVSIZ), NX3(MVSIZ), NX4(MVSIZ),
     .   NY1(MVSIZ), NY2(MVSIZ), NY3(MVSIZ), NY4(MVSIZ),
     .   NZ1(MVSIZ), NZ2(MVSIZ), NZ3(MVSIZ), NZ4(MVSIZ),
     .   LB1(MVSIZ), LB2(MVSIZ), LB3(MVSIZ), LB4(MVSIZ),
     .   LC1(MVSIZ), LC2(MVSIZ), LC3(MVSIZ), LC4(MVSIZ),
     .   P1(MVSIZ), P2(MVSIZ), P3(MVSIZ), P4(MVSIZ),
     .   NNX1(MVSIZ), NNX2(MVSIZ), NNX3(MVSIZ), NNX4(MVSIZ),
     .   NNY1(MVSIZ), NNY2(MVSIZ), NNY3(MVSIZ), NNY4(MVSIZ),
     .   NNZ1(MVSIZ), NNZ2(MVSIZ), NNZ3(MVSIZ), NNZ4(MVSIZ),
     .   RCURVI(MVSIZ), ANGLMI(MVSIZ), TEMPI(MVSIZ), PHI(MVSIZ),
     .   AREASI(MVSIZ), CMAJ(MVSIZ), DIR_FRICMI(MVSIZ,2),
     .   FRICC(MVSIZ), VISCFFRIC(MVSIZ), FRIC_COEFS(MVSIZ,10),
     .   FRIC_COEFS2(MVSIZ,10), FRICC2(MVSIZ), VISCFFRIC2(MVSIZ),
     .   DIR1(MVSIZ,3), DIR2(MVSIZ,3), EFRICT(MVSIZ), FNI(MVSIZ),
     .   H1(MVSIZ), H2(MVSIZ), H3(MVSIZ), H4(MVSIZ),
     .   KS(MVSIZ), KT(MVSIZ), K1(MVSIZ), K2(MVSIZ), K3(MVSIZ),
     .   K4(MVSIZ), C1(MVSIZ), C2(MVSIZ), C3(MVSIZ), C4(MVSIZ),
     .   CS(MVSIZ), C(MVSIZ), CF(MVSIZ),
     .   FXI(MVSIZ), FYI(MVSIZ), FZI(MVSIZ),
     .   FX1(MVSIZ), FY1(MVSIZ), FZ1(MVSIZ), FX2(MVSIZ), FY2(MVSIZ),
     .   FZ2(MVSIZ), FX3(MVSIZ), FY3(MVSIZ), FZ3(MVSIZ), FX4(MVSIZ),
     .   FY4(MVSIZ), FZ4(MVSIZ),
     .   PHI1(MVSIZ), PHI2(MVSIZ), PHI3(MVSIZ), PHI4(MVSIZ),
     .   CONDINT(MVSIZ)
C     ---- interface-wide arrays the path passes through untouched
      INTEGER, ALLOCATABLE :: KINET(:), ICODT(:), ITAB(:), ICONTACT(:),
     .   IFPEN(:), NSTRF(:), NPC(:), TAGNCONT(:,:), IELEC(:),
     .   NODNX_SMS(:), IPARTFRICS(:), IPARTFRICM(:), IREP_FRICM(:),
     .   CANDN(:), CANDE(:)
      my_real, ALLOCATABLE :: FSKYI(:,:), A(:,:), STIFN(:), FCONT(:,:),
     .   VISCN(:), FNCONT(:,:), FTCONT(:,:), CAND_P(:), CAND_F(:,:),
     .   CAND_FX(:), CAND_FY(:), CAND_FZ(:), NOD_NORMAL(:,:),
     .   RCONTACT(:), ACONTACT(:), PCONTACT(:), TEMP(:), TF(:),
     .   GAP_S(:), GAP_M(:), GAP_SL(:), GAP_ML(:), RCURV(:), ANGLM(:),
     .   AREAS(:), DIR_FRICM(:,:), FTHESKYI(:), CONDNSKYI(:)
      INTEGER ISKY(1), ISKYW(1)
      INTEGER NSN, NRTM, NCAND, NFT, JLT, JLT_NEW, JLT_TIED, I, K,
     .   NELTST, ITYPTST, NEWFRONT, ISECIN, NISKYFI, ICURV3(3),
     .   ISENSINT(1), LISUB(1), ADDSUBS(1), ADDSUBM(1), LISUBS(1),
     .   LISUBM(1), TYPSUB(1), INFLG_SUBS(1), INFLG_SUBM(1),
     .   KLOADPINTER(2), LOADPINTER(1), LOADP_HYD_INTER(1)
      my_real GAP, FRIC, VISC, VISCF, STIGLO, DT2T, FSAV(100),
     .   FSAVSUB(1,1), FSAVPARIT(1,11,1), DGAPLOADINT(1), QFRIC,
     .   KMIN, KMAX, GAPMAX, GAPMIN, E0, EV0, ED0, ZEROR
      INTEGER, ALLOCATABLE :: ISKYA(:)
      TYPE(OUTPUT_) :: OUTPUT
      TYPE(H3D_DATABASE) :: H3D_DATA
C
      NSN = NSN_C
      NRTM = NRTM_C
      NCAND = NC_C
      NUMNOD = NN_C
      GAP = PARAMS(1)
      FRIC = PARAMS(2)
      VISC = PARAMS(3)
      VISCF = PARAMS(4)
      DT1 = DTS(1)
      DT12 = DTS(2)
      ZEROR = 0.0D0
C     engine flags of the path (explicit, one domain, no output requests,
C     no /DT/INTER, no SMS, no /ANIM or /H3D contact vectors)
      NVSIZ = 128
      NSPMD = 1
      ISTAMPING = 0
      IDAMP_RDOF = 0
      IRESP = 0
      INCONV = 1
      IMPL_S = 0
      IMP_INT7 = 0
      IDTMINS = 0
      IDTMINS_INT = 0
      KDTINT = 0
      IDTMIN(10) = 0
      DTMIN1(10) = 0.0D0
      ANIM_V(4) = 0
      ANIM_V(12) = 0
      OUTP_V(4) = 0
      OUTP_V(12) = 0
      TT = 0.0D0
      TOUTP = 1.0D30
      NINTER = 1
      NLOADP_HYD_INTER = 0
      NLOADP_HYD = 0
      IPARIT = 1
      NTHVKI = 1
      H3D_DATA%N_VECT_PCONT = 0
      H3D_DATA%N_VECT_CONT = 0
      H3D_DATA%N_SCAL_CSE_FRIC = 0
      H3D_DATA%MH3D = 0
      OUTPUT%TANIM = 1.0D30
      OUTPUT%TANIM_STOP = 1.0D30
      H3D_DATA%TH3D = 1.0D30
      H3D_DATA%TH3D_STOP = 1.0D30
C     skyline: every pair writes at most five entries
      LSKYI = 5 * MAX(NCAND, 1)
      NFSKYI = 4
      NISKY = 0
      NISKYFI = 0
      IF (ALLOCATED(NLSKYFI)) DEALLOCATE(NLSKYFI)
      ALLOCATE(NLSKYFI(1))
      NLSKYFI(1) = 0
      ALLOCATE(FSKYI(LSKYI, NFSKYI), ISKYA(LSKYI), FTHESKYI(LSKYI),
     .   CONDNSKYI(LSKYI))
      FSKYI = 0.0D0
      ISKYA = 0
      ALLOCATE(KINET(NUMNOD), ICODT(NUMNOD), ITAB(NUMNOD),
     .   ICONTACT(NUMNOD), A(3,NUMNOD), STIFN(NUMNOD), FCONT(3,NUMNOD),
     .   VISCN(NUMNOD), FNCONT(3,NUMNOD), FTCONT(3,NUMNOD),
     .   RCONTACT(NUMNOD), ACONTACT(NUMNOD), PCONTACT(NUMNOD),
     .   TEMP(NUMNOD), NOD_NORMAL(3,NUMNOD), NODNX_SMS(NUMNOD),
     .   TAGNCONT(1,NUMNOD))
      KINET = 0
      ICODT = 0
      ICONTACT = 0
      A = 0.0D0
      STIFN = 0.0D0
      DO I = 1, NUMNOD
        ITAB(I) = I
      ENDDO
      ALLOCATE(IFPEN(MAX(NCAND,1)), CAND_P(MAX(NCAND,1)),
     .   CAND_F(8,MAX(NCAND,1)), CAND_FX(MAX(NCAND,1)),
     .   CAND_FY(MAX(NCAND,1)), CAND_FZ(MAX(NCAND,1)),
     .   CANDN(MAX(NCAND,1)), CANDE(MAX(NCAND,1)))
      CAND_P = 0.0D0
      CAND_F = 0.0D0
      CAND_FX = 0.0D0
      CAND_FY = 0.0D0
      CAND_FZ = 0.0D0
      IFPEN = 0
      ALLOCATE(NSTRF(1), NPC(1), TF(1), IELEC(MAX(NSN,1)),
     .   IPARTFRICS(MAX(NSN,1)), IPARTFRICM(MAX(NRTM,1)),
     .   IREP_FRICM(MAX(NRTM,1)), DIR_FRICM(2,MAX(NRTM,1)),
     .   GAP_S(MAX(NSN,1)), GAP_M(MAX(NRTM,1)), GAP_SL(MAX(NSN,1)),
     .   GAP_ML(MAX(NRTM,1)), RCURV(MAX(NRTM,1)), ANGLM(MAX(NRTM,1)),
     .   AREAS(MAX(NSN,1)))
      ICURV3 = 0
      ISENSINT = 0
      KLOADPINTER = 0
      STIGLO = -1.0D0
      KMIN = 0.0D0
      KMAX = 1.0D30
      GAPMAX = 1.0D30
      GAPMIN = 0.0D0
      NELTST = 0
      ITYPTST = 0
      NEWFRONT = 0
      ISECIN = 0
      QFRIC = 0.0D0
      DT2T = OUT(1)
      E0 = ECONT
      EV0 = ECONTV
      ED0 = ECONTD
      ECONT = 0.0D0
      ECONTV = 0.0D0
      ECONTD = 0.0D0
      ECONT_CUMU = 0.0D0
      OUT(5) = 0.0D0
      DO I = 1, NCAND
        PEN(I) = 0
      ENDDO
C
      DO NFT = 0, NCAND - 1, NVSIZ
        JLT = MIN(NVSIZ, NCAND - NFT)
        DO I = 1, JLT
          CANDN(I) = CAND_N(NFT + I)
          CANDE(I) = CAND_E(NFT + I)
          INDEX2(I) = NFT + I
        ENDDO
        CALL I7COR3(
     1    JLT, X, IRECT, NSV, CANDE,
     2    CANDN, STFM, STFNS, X1, X2,
     3    X3, X4, Y1, Y2, Y3,
     4    Y4, Z1, Z2, Z3, Z4,
     5    XI, YI, ZI, STIF, IX1,
     6    IX2, IX3, IX4, NSVG, 0,
     7    GAP, GAP_S, GAP_M, GAPV,
     9    MS, VXI, VYI,
     A    VZI, MSI, NSN, V, KINET,
     B    KINI, 7, 1, 1, KMIN,
     C    KMAX, GAPMAX, GAPMIN, 0, RCURV,
     D    RCURVI, ANGLM, ANGLMI, 0, TEMP,
     E    TEMPI, PHI, AREAS, IELEC, AREASI,
     F    IELECI, NODNX_SMS, NSMS, GAP_SL, GAP_ML,
     G    0, IPARTFRICS, IPARTFRICSI, IPARTFRICM, IPARTFRICMI,
     H    0, IREP_FRICM, DIR_FRICM, IREP_FRICMI, DIR_FRICMI)
        JLT_NEW = 0
        JLT_TIED = 0
        CALL I7DST3(
     1    JLT, CANDN, CANDE, CN_LOC, CE_LOC,
     2    X1, X2, X3, X4, Y1,
     3    Y2, Y3, Y4, Z1, Z2,
     4    Z3, Z4, XI, YI, ZI,
     5    NX1, NX2, NX3, NX4, NY1,
     6    NY2, NY3, NY4, NZ1, NZ2,
     7    NZ3, NZ4, LB1, LB2, LB3,
     8    LB4, LC1, LC2, LC3, LC4,
     9    P1, P2, P3, P4, IX1,
     A    IX2, IX3, IX4, NSVG, STIF,
     B    JLT_NEW, GAPV, 0, CAND_P,
     C    INDEX2, VXI, VYI,
     D    VZI, MSI, KINI, 0, IRECT,
     E    NNX1, NNX2, NNX3, NNX4, NNY1,
     F    NNY2, NNY3, NNY4, NNZ1, NNZ2,
     G    NNZ3, NNZ4, NOD_NORMAL, 0, RCURVI,
     H    ANGLMI, 0, TEMPI, PHI, AREASI,
     I    IELECI, NSMS, CMAJ, ZEROR,
     J    0, IPARTFRICSI,
     K    IPARTFRICMI, 0, JLT_TIED, CAND_F, 0,
     L    IREP_FRICMI, DIR_FRICMI, ZEROR)
        JLT = JLT_NEW
        OUT(5) = OUT(5) + JLT
        DO I = 1, JLT
          PEN(INDEX2(I)) = 1
        ENDDO
        IF (JLT == 0) CYCLE
        DO I = 1, JLT
          FRICC(I) = FRIC
          VISCFFRIC(I) = VISCF
          INDEXISOT(I) = I
        ENDDO
        CALL I7FOR3(OUTPUT,
     1    JLT, A, V, 0, ICODT,
     2    FSAV, GAP, FRIC, MS, VISC,
     3    VISCF, 1, STFNS, ITAB, CN_LOC,
     4    STIGLO, STIFN, STIF, FSKYI, ISKYA,
     5    NX1, NX2, NX3, NX4, NY1,
     6    NY2, NY3, NY4, NZ1, NZ2,
     7    NZ3, NZ4, LB1, LB2, LB3,
     8    LB4, LC1, LC2, LC3, LC4,
     9    P1, P2, P3, P4, FCONT,
     A    IX1, IX2, IX3, IX4, NSVG,
     B    0, NELTST, ITYPTST, DT2T, GAPV,
     C    0, CAND_P, INDEX2, KINET, NEWFRONT,
     D    ISECIN, NSTRF, X, IRECT, CE_LOC,
     E    0, 0, CAND_FX, CAND_FY,
     F    CAND_FZ, ZEROR, IFPEN, 0,
     H    ICONTACT, VISCN, VXI, VYI, VZI,
     I    MSI, KINI, 1, 0, LISUB,
     J    ADDSUBS, ADDSUBM, LISUBS,
     .    LISUBM, FSAVSUB,
     K    CANDN, 0, ICURV3, NOD_NORMAL, FNCONT,
     L    FTCONT, X1, X2, X3, X4,
     M    Y1, Y2, Y3, Y4, Z1,
     N    Z2, Z3, Z4, XI, YI,
     O    ZI, 0, RCURVI, RCONTACT, ACONTACT,
     P    PCONTACT, ANGLMI, ZEROR, 0, TEMP,
     Q    TEMPI, 1, NPC, TF, CMAJ,
     R    ZEROR, ZEROR, ZEROR, ZEROR, EFRICT,
     S    QFRIC, FNI, 0, 1, H1,
     U    H2, H3, H4, KS, KT,
     V    K1, K2, K3, K4, C1,
     W    C2, C3, C4, CS, C,
     X    CF, ZEROR, ZEROR, FXI, FYI,
     Y    FZI, FX1, FY1, FZ1, FX2,
     Z    FY2, FZ2, FX3, FY3, FZ3,
     1    FX4, FY4, FZ4, ISENSINT, FSAVPARIT,
     5    NFT, 0, H3D_DATA, FRICC, VISCFFRIC,
     6    FRIC_COEFS, 0, JLT_TIED, CAND_F, 0,
     7    FRIC_COEFS2, FRICC2, VISCFFRIC2, 0, JLT,
     8    INDEXORTH, INDEXISOT, DIR1, DIR2, TAGNCONT,
     9    KLOADPINTER, LOADPINTER, LOADP_HYD_INTER, TYPSUB, INFLG_SUBS,
     A    INFLG_SUBM, 0, DGAPLOADINT, 1, ZEROR,
     B    0)
        CALL I7ASS2(JLT, IX1, IX2, IX3, IX4,
     2    NSVG, H1, H2, H3, H4, STIF,
     3    FX1, FY1, FZ1, FX2, FY2, FZ2,
     4    FX3, FY3, FZ3, FX4, FY4, FZ4,
     5    FXI, FYI, FZI, FSKYI, ISKYA, NISKYFI,
     6    1, 1, 0, PHI, FTHESKYI, PHI1,
     7    PHI2, PHI3, PHI4, CONDNSKYI, CONDINT,
     A    1, 0)
      ENDDO
C
      NSKY = NISKY
      DO K = 1, NISKY
        FSKYO(1,K) = FSKYI(K,1)
        FSKYO(2,K) = FSKYI(K,2)
        FSKYO(3,K) = FSKYI(K,3)
        FSKYO(4,K) = FSKYI(K,4)
        ISKYO(K) = ISKYA(K)
      ENDDO
      OUT(1) = DT2T
      OUT(2) = ECONT
      OUT(3) = ECONTV
      OUT(4) = ECONTD
      ECONT = E0
      ECONTV = EV0
      ECONTD = ED0
      DEALLOCATE(FSKYI, ISKYA, FTHESKYI, CONDNSKYI, KINET, ICODT, ITAB,
     .   ICONTACT, A, STIFN, FCONT, VISCN, FNCONT, FTCONT, RCONTACT,
     .   ACONTACT, PCONTACT, TEMP, NOD_NORMAL, NODNX_SMS, TAGNCONT,
     .   IFPEN, CAND_P, CAND_F, CAND_FX, CAND_FY, CAND_FZ, CANDN, CANDE,
     .   NSTRF, NPC, TF, IELEC, IPARTFRICS, IPARTFRICM, IREP_FRICM,
     .   DIR_FRICM, GAP_S, GAP_M, GAP_SL, GAP_ML, RCURV, ANGLM, AREAS)
      RETURN
      END
