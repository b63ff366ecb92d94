C  harness_nodal.F -- test harness, not product code.
C
C  Drives the reference's own nodal-update routines, compiled from their
C  sources under the reference tree by oracle/build_ref.py, for one explicit
C  cycle in RESOL order (engine/source/engine/resol.F):
C    time step      resol.F:2639-2640, 5857-5861, 6080, 6224-6226 (restated:
C                   RESOL's own lines, the only part of this cycle not called)
C    ACCELE         assembly/accele.F          (resol.F:6639)
C    BCS10 / BCS1V  constraints/general/bcs/bcs10.F, bcs1.F   (resol.F:7022)
C    FIXVEL         constraints/general/impvel/fixvel.F       (resol.F:7294)
C    time           resol.F:8256-8266 (restated: TT_DOUBLE = TT_DOUBLE + DT2)
C    VELOCITY       assembly/velocity.F        (resol.F:8599)
C    DEPLA          assembly/displacement.F    (resol.F:8692)
C  Input A (and AR) is the assembled force (moment) of the cycle, i.e. what
C  ASSPAR4 leaves in NODES%A; DT2T is the element time-step minimum.
C
C  Entry point (C binding) ref_nodal_cycle.  TIMES(6) = TT, DT1, DT2, DT12,
C  DT2OLD, DTMX in and out; NCYC = NCYCLE in, incremented.  FIXVEL entries:
C  FXNODE (1-based), FXDIR (1..3 translation, 4..6 rotation), FXVAL the imposed
C  (constant) velocity; FXMEM(2,NFX) carries VEL(4,N) (the energy memory DT2*DW)
C  and IBFV(5,N) (the curve position) from cycle to cycle; WFEXT in and out.
C  On return A (AR) hold the accelerations after FIXVEL, as VELOCITY sees them
C  (VELOCITY itself resets NODES%A to zero afterwards; the copy is taken just
C  before that call).
      SUBROUTINE REF_NODAL_CYCLE(NN_C, IROT_C, X, V, VR, D, A, AR, MS,
     .   XIN, ICODT_C, ICODR_C, NFX_C, FXNODE, FXDIR, FXVAL, FXMEM,
     .   DT2T_C, TIMES, NCYC, WFEXT_C) BIND(C, NAME='ref_nodal_cycle')
      USE ISO_C_BINDING
      USE NODAL_ARRAYS_MOD
      USE SENSOR_MOD
      USE PYTHON_FUNCT_MOD
#include      "implicit_f.inc"
#include      "com01_c.inc"
#include      "com04_c.inc"
#include      "com08_c.inc"
#include      "param_c.inc"
#include      "parit_c.inc"
#include      "sms_c.inc"
#include      "scr03_c.inc"
#include      "scr05_c.inc"
#include      "scr16_c.inc"
#include      "inter22.inc"
#include      "tabsiz_c.inc"
      INTEGER(C_INT), VALUE :: NN_C, IROT_C, NFX_C
      REAL(C_DOUBLE) :: X(3,*), V(3,*), VR(3,*), D(3,*), A(3,*),
     .   AR(3,*), MS(*), XIN(*), FXVAL(*), FXMEM(2,*), TIMES(6),
     .   WFEXT_C
      REAL(C_DOUBLE), VALUE :: DT2T_C
      INTEGER(C_INT) :: ICODT_C(*), ICODR_C(*), FXNODE(*), FXDIR(*),
     .   NCYC
C     ---- locals
      INTEGER I, N, NFX
      INTEGER, ALLOCATABLE :: ISKEW(:), NALE(:), NPBY(:,:), IBFV(:,:),
     .   WEIGHT(:), IFRAME(:,:), NODNX(:), ITAB(:)
      my_real, ALLOCATABLE :: SKEW(:,:), MS2D(:), VEL(:,:), RBY(:,:),
     .   XFRAME(:,:), DR(:,:), FZERO(:,:), XDP(:,:), DDP(:,:)
      my_real TF(2,2), DTMX
      INTEGER NPC(4)
      DOUBLE PRECISION TTD, WFEXT
      TYPE(NODAL_ARRAYS_) :: NODES
      TYPE(PYTHON_) :: PY
      TYPE(SENSOR_STR_), DIMENSION(1) :: SENSTAB
C
      NUMNOD = NN_C
      NFX = NFX_C
C     /PARAM/ record lengths as the starter sets them (starter0.F:534-604)
      NVSIZ = 128
      LFXVELR = 9
      NRBY = 30
      NNPBY = 20
      NXFRAME = 36
      LSKEW = 12
      LISKN = 6
      NIFV = 16
C     flags of the path: 3-D, Lagrangian, no multi-material, no SMS, no
C     /INTER/TYPE22, double precision positions, no imposed rotation output
      N2D = 0
      NMULT = 0
      ISIGI = 0
      INT22 = 0
      IVECTOR = 0
      IDTMINS = 0
      IDTMINS_INT = 0
      IRESP = 0
      ISECUT = 0
      IISROT = 0
      IMPOSE_DR = 0
      IDROT = 0
      NRBODY = 0
      IRODDL = IROT_C
      NFXVEL = NFX
      NFUNCT = 1
      NCYCLE = NCYC
      SISKEW = NUMNOD
      SICODT = NUMNOD
      SICODR = NUMNOD
      SSKEW = LSKEW
C
      ALLOCATE(ISKEW(NUMNOD), NALE(NUMNOD), NPBY(NNPBY,1),
     .   IBFV(NIFV,MAX(NFX,1)), WEIGHT(NUMNOD), IFRAME(LISKN,1),
     .   NODNX(NUMNOD), ITAB(NUMNOD))
      ALLOCATE(SKEW(LSKEW,1), MS2D(NUMNOD), VEL(LFXVELR,MAX(NFX,1)),
     .   RBY(NRBY,1), XFRAME(NXFRAME,1), DR(3,NUMNOD),
     .   FZERO(3,NUMNOD), XDP(3,NUMNOD), DDP(3,NUMNOD))
      ISKEW = 1
      NALE = 0
      NPBY = 0
      WEIGHT = 1
      IFRAME = 0
      NODNX = 0
      DO I = 1, NUMNOD
        ITAB(I) = I
      ENDDO
      SKEW = ZERO
      SKEW(1,1) = ONE
      SKEW(5,1) = ONE
      SKEW(9,1) = ONE
      MS2D = ZERO
      RBY = ZERO
      XFRAME = ZERO
      DR = ZERO
      FZERO = ZERO
      XDP = ZERO
      DDP = ZERO
C     one constant velocity curve f(t) = 1 (two points), scaled per entry
C     by VEL(1,N); NPC(L) / NPC(L+1) bound its points, NPC(2*NFUNCT+L+1) = 0
C     means no smoothing (fixvel.F:231,292-293)
      TF(1,1) = ZERO
      TF(2,1) = ONE
      TF(1,2) = EP30
      TF(2,2) = ONE
      NPC(1) = 0
      NPC(2) = 4
      NPC(3) = 1
      NPC(4) = 0
      IBFV = 0
      VEL = ZERO
      DO N = 1, NFX
        IBFV(1,N) = FXNODE(N)
        IBFV(2,N) = FXDIR(N)
        IBFV(3,N) = 1
        IBFV(5,N) = NINT(FXMEM(2,N))
        IBFV(7,N) = 1
        VEL(1,N) = FXVAL(N)
        VEL(2,N) = ZERO
        VEL(3,N) = EP30
        VEL(4,N) = FXMEM(1,N)
        VEL(5,N) = ONE
      ENDDO
      TT = TIMES(1)
      DT1 = TIMES(2)
      DT2 = TIMES(3)
      DT12 = TIMES(4)
      DT2OLD = TIMES(5)
      DTMX = TIMES(6)
      TTD = TT
      WFEXT = WFEXT_C
C     ---- time step (resol.F:2639-2640, 5857-5861, 6080, 6224-6226)
      DT1=DT2
      DT2=EP06
      IF(DT2T_C<DT2)THEN
        DT2    = DT2T_C
      ENDIF
      DT2= MIN(DT2,1.1*DT2OLD,DTMX)
      DT2OLD=DT2
      DT12=HALF*(DT1+DT2)
C     ---- ACCELE (resol.F:6639)
      CALL ACCELE(A, AR, V, MS, XIN, NUMNOD, NALE, MS2D, 1, NPBY)
C     ---- BCS10 (resol.F:7022)
      CALL BCS10(1, NUMNOD, ICODT_C, ICODR_C, ISKEW, SKEW, A, AR, MS,
     .           V, VR)
C     ---- FIXVEL (resol.F:7294)
      IF (NFX > 0)
     .  CALL FIXVEL(IBFV, A, V, NPC, TF, VEL, MS, X, SKEW, AR, VR, XIN,
     .              0, SENSTAB, WEIGHT, D, RBY, IFRAME, XFRAME, DR,
     .              NODNX, NODES, TTD, DDP, PY, WFEXT)
C     ---- time (resol.F:8256-8266)
      TTD = TTD + DT2
      TT = TTD
C     ---- VELOCITY (resol.F:8599); keep the accelerations it consumes
      FZERO(1:3,1:NUMNOD) = A(1:3,1:NUMNOD)
      IF (IROT_C /= 0) XDP(1:3,1:NUMNOD) = AR(1:3,1:NUMNOD)
      CALL VELOCITY(A, AR, V, VR, DR, ITAB, NALE)
      A(1:3,1:NUMNOD) = FZERO(1:3,1:NUMNOD)
      IF (IROT_C /= 0) AR(1:3,1:NUMNOD) = XDP(1:3,1:NUMNOD)
      XDP = ZERO
C     ---- DEPLA (resol.F:8692)
      CALL DEPLA(V, D, X, VR, DR, XDP, DDP, NUMNOD)
C
      DO N = 1, NFX
        FXMEM(1,N) = VEL(4,N)
        FXMEM(2,N) = IBFV(5,N)
      ENDDO
      TIMES(1) = TT
      TIMES(2) = DT1
      TIMES(3) = DT2
      TIMES(4) = DT12
      TIMES(5) = DT2OLD
      TIMES(6) = DTMX
      NCYC = NCYC + 1
      WFEXT_C = WFEXT
      DEALLOCATE(ISKEW, NALE, NPBY, IBFV, WEIGHT, IFRAME, NODNX, ITAB)
      DEALLOCATE(SKEW, MS2D, VEL, RBY, XFRAME, DR, FZERO, XDP, DDP)
      RETURN
      END
