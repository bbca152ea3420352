! A Fortran host for the binding orctm_nh_c: fills the test stand-ins of MO_PARAM1 / MO_COMMO1 /
! MO_OBCS / MO_PARA2 from one stream file, then makes the calls ORCTM makes, in the model's order:
!   CALL petsc_start                   (orctm.f90:970)
!   CALL petsc_solve_3D, twice         (ocvtot_nh.f90:150; the second warm-starts, mo_petsc.f90:232)
!   CALL ocvtot_nh_device              (the whole of OCVTOT_NH on the device)
!   CALL troneu_device                 (orctm.f90:1414, CALL TRONEU)
!   CALL petsc_end                     (orctm.f90:1585)
! and writes every array they produce to a second stream file.
!
! usage: orctm_nh_fdriver <in.bin> <out.bin>
! in.bin : int32 IE, JE, I_start, J_start, West, East, North, South; float64 DT, DTI, G, CONN, CONO,
!          STABN, SORPAR; then, each in Fortran order with the bounds of mo_commo1.f90:105-154:
!          DLXP DLYP DLXU DLYV DDPO WETO DZW DIVG | AMSUO DDUO DTDXUO DLYU AMSUE DDUE DPYE DLXV DTDXPO
!          DTDYO | UOO UOO_1 VOE VOE_1 WNO WNO_1 ZO | DEUTO DEUTE U1O UZO V1E VZE B1O
! out.bin: PNH after each solve, then DIVG PNH UOO VOE WNO after ocvtot_nh_device, [PNH_bot
!          DIVG_bot_surf with CG2D4NONHY], then Z1O B1O after troneu_device.
PROGRAM orctm_nh_fdriver
  USE MO_PARAM1
  USE MO_COMMO1
  USE MO_OBCS
  USE MO_PARA2
  USE orctm_nh_c, ONLY: petsc_start, petsc_solve_3D, petsc_end, ocvtot_nh_device, troneu_device
  IMPLICIT NONE
  INTEGER(4) :: hdr(8)
  REAL(8) :: sc(7)
  CHARACTER(LEN=1024) :: fin, fout
  INTEGER :: ui, uo

  IF (COMMAND_ARGUMENT_COUNT() /= 2) THEN
    PRINT *, 'usage: orctm_nh_fdriver <in.bin> <out.bin>'
    STOP 2
  ENDIF
  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(NEWUNIT=ui, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
  READ(ui) hdr, sc
  IE = hdr(1); JE = hdr(2); I_start = hdr(3); J_start = hdr(4); KEP = KE + 1
  West = hdr(5) /= 0; East = hdr(6) /= 0; North = hdr(7) /= 0; South = hdr(8) /= 0
  DT = sc(1); DTI = sc(2); G = sc(3); CONN = sc(4); CONO = sc(5); STABN = sc(6); SORPAR = sc(7)

  ALLOCATE(DLXP(IE,JE), DLYP(IE,JE), DLXU(I_start:IE,JE), DLYV(IE,J_start:JE), DDPO(IE,JE,KE), WETO(IE,JE,KE))
  ALLOCATE(DIVG(IE,JE,KE), PNH(IE,JE,KE))
  ALLOCATE(AMSUO(I_start:IE,JE,KE), DDUO(I_start:IE,JE,KE), DTDXUO(I_start:IE,JE), DLYU(I_start:IE,JE))
  ALLOCATE(AMSUE(IE,J_start:JE,KE), DDUE(IE,J_start:JE,KE), DPYE(IE,J_start:JE), DLXV(IE,J_start:JE))
  ALLOCATE(DTDXPO(IE,JE), DTDYO(IE,JE))
  ALLOCATE(UOO(I_start:IE,JE,KE), UOO_1(I_start:IE,JE,KE), VOE(IE,J_start:JE,KE), VOE_1(IE,J_start:JE,KE))
  ALLOCATE(WNO(IE,JE,KEP), WNO_1(IE,JE,KEP), WO(IE,JE,KEP), ZO(IE,JE))
  ALLOCATE(DEUTO(I_start:IE,JE), DEUTE(IE,J_start:JE), U1O(I_start:IE,JE), UZO(I_start:IE,JE))
  ALLOCATE(V1E(IE,J_start:JE), VZE(IE,J_start:JE), B1O(IE,JE), Z1O(IE,JE))
  ALLOCATE(PNH_bot(IE,JE), DIVG_bot_surf(IE,JE))

  READ(ui) DLXP, DLYP, DLXU, DLYV, DDPO, WETO, DZW, DIVG
  READ(ui) AMSUO, DDUO, DTDXUO, DLYU, AMSUE, DDUE, DPYE, DLXV, DTDXPO, DTDYO
  READ(ui) UOO, UOO_1, VOE, VOE_1, WNO, WNO_1, ZO
  READ(ui) DEUTO, DEUTE, U1O, UZO, V1E, VZE, B1O
  CLOSE(ui)
  PNH = 0.0; WO = 0.0; Z1O = 0.0; PNH_bot = 0.0; DIVG_bot_surf = 0.0

  OPEN(NEWUNIT=uo, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
  CALL petsc_start
  CALL petsc_solve_3D
  WRITE(uo) PNH
  CALL petsc_solve_3D
  WRITE(uo) PNH
  CALL ocvtot_nh_device
  WRITE(uo) DIVG, PNH, UOO, VOE, WNO
#ifdef CG2D4NONHY
  WRITE(uo) PNH_bot, DIVG_bot_surf
#endif
  CALL troneu_device
  WRITE(uo) Z1O, B1O
  CALL petsc_end
  CLOSE(uo)
  PRINT *, 'orctm_nh_fdriver: done'
END PROGRAM orctm_nh_fdriver
