! A Fortran host for the restart path of the binding orctm_nh_c: the resident pressure saved with the
! restart files (aufw_nh) and read back by a new run (aufr_nh), against the same steps uninterrupted.
!   run 1: petsc_start; step 1 (DIVG); aufw_nh to <restart>; step 2 (DIVG*1.01, warm-started); petsc_end
!   run 2: petsc_start; aufr_nh from <restart>; step 2; petsc_end
!   run 3: petsc_start; step 2 from zero (what a restart without the record gives); petsc_end
! usage: orctm_nh_frestart <in.bin> <restart> <out.bin>
! in.bin : the input file of orctm_nh_fdriver (only the header and the arrays up to DIVG are read)
! out.bin: PNH of step 2 from runs 1, 2 and 3
PROGRAM orctm_nh_frestart
  USE MO_PARAM1
  USE MO_COMMO1
  USE MO_OBCS
  USE MO_PARA2
  USE orctm_nh_c, ONLY: petsc_start, petsc_solve_3D, petsc_end, aufw_nh, aufr_nh
  IMPLICIT NONE
  INTEGER(4) :: hdr(8)
  REAL(8) :: sc(7)
  REAL(8), ALLOCATABLE :: d1(:,:,:), d2(:,:,:)
  CHARACTER(LEN=1024) :: fin, frst, fout
  INTEGER :: ui, uo, ur

  IF (COMMAND_ARGUMENT_COUNT() /= 3) THEN
    PRINT *, 'usage: orctm_nh_frestart <in.bin> <restart> <out.bin>'
    STOP 2
  ENDIF
  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, frst)
  CALL GET_COMMAND_ARGUMENT(3, fout)
  OPEN(NEWUNIT=ui, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
  READ(ui) hdr, sc
  IE = hdr(1); JE = hdr(2); I_start = hdr(3); J_start = hdr(4); KEP = KE + 1
  West = hdr(5) /= 0; East = hdr(6) /= 0; North = hdr(7) /= 0; South = hdr(8) /= 0
  ALLOCATE(DLXP(IE,JE), DLYP(IE,JE), DLXU(I_start:IE,JE), DLYV(IE,J_start:JE), DDPO(IE,JE,KE), WETO(IE,JE,KE))
  ALLOCATE(DIVG(IE,JE,KE), PNH(IE,JE,KE), d1(IE,JE,KE), d2(IE,JE,KE))
  READ(ui) DLXP, DLYP, DLXU, DLYV, DDPO, WETO, DZW, DIVG
  CLOSE(ui)
  d1 = DIVG
  d2 = DIVG * 1.01
  OPEN(NEWUNIT=uo, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')

  CALL petsc_start
  PNH = 0.0
  DIVG = d1
  CALL petsc_solve_3D
  OPEN(NEWUNIT=ur, FILE=TRIM(frst), ACCESS='SEQUENTIAL', FORM='UNFORMATTED', STATUS='REPLACE')
  CALL aufw_nh(ur, 20261203)
  CLOSE(ur)
  DIVG = d2
  CALL petsc_solve_3D
  WRITE(uo) PNH
  CALL petsc_end

  CALL petsc_start
  OPEN(NEWUNIT=ur, FILE=TRIM(frst), ACCESS='SEQUENTIAL', FORM='UNFORMATTED', STATUS='OLD')
  CALL aufr_nh(ur)
  CLOSE(ur)
  PNH = 0.0
  DIVG = d2
  CALL petsc_solve_3D
  WRITE(uo) PNH
  CALL petsc_end

  CALL petsc_start
  PNH = 0.0
  DIVG = d2
  CALL petsc_solve_3D
  WRITE(uo) PNH
  CALL petsc_end
  CLOSE(uo)
  PRINT *, 'orctm_nh_frestart: done'
END PROGRAM orctm_nh_frestart
