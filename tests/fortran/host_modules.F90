! Test stand-ins for the ORCTM modules that orctm_nh_c reads: the names, kinds and array bounds of
! ORCTM's own declarations, so the binding compiles and runs as it would inside the model.
!   MO_PARAM1  src/mo_param1.f90:32-77   KE, ICYCLI_X, ICYCLI_Y are compile-time PARAMETERs there,
!                                         here set by -DTEST_KE, -DTEST_CYCLIC_X, -DTEST_CYCLIC_Y
!   MO_COMMO1  src/mo_commo1.f90:11-154  REAL scalars, DZW(KE) a fixed array, every field a REAL
!                                         POINTER allocated with the bounds of :105-154
!   MO_OBCS    src/mo_obcs.f90:9-10      the open-boundary flags
!   MO_PARA2   src/mo_para2.f90:10       SORPAR
! Built like the model (-r8: REAL is 8 bytes), with -fdefault-real-8.
MODULE MO_PARAM1
  IMPLICIT NONE
  INTEGER, PARAMETER :: KE = TEST_KE
  INTEGER, PARAMETER :: ICYCLI_X = TEST_CYCLIC_X, ICYCLI_Y = TEST_CYCLIC_Y
  INTEGER IE, JE, I_start, J_start, KEP
END MODULE MO_PARAM1

MODULE MO_COMMO1
  USE MO_PARAM1, ONLY: KE
  IMPLICIT NONE
  REAL DT, DTI, G, CONN, CONO, STABN
  REAL DZW(KE)
  REAL, POINTER :: DLXP(:,:), DLYP(:,:), DLXU(:,:), DLYV(:,:), DLXV(:,:), DLYU(:,:), DTDXUO(:,:), &
                   DTDYO(:,:), DTDXPO(:,:), DPYE(:,:), DEUTE(:,:), DEUTO(:,:)
  REAL, POINTER :: DDPO(:,:,:), DDUE(:,:,:), DDUO(:,:,:), WETO(:,:,:), AMSUE(:,:,:), AMSUO(:,:,:)
  REAL, POINTER :: UOO(:,:,:), VOE(:,:,:), WO(:,:,:), UOO_1(:,:,:), VOE_1(:,:,:), WNO(:,:,:), &
                   WNO_1(:,:,:), DIVG(:,:,:), PNH(:,:,:)
  REAL, POINTER :: ZO(:,:), Z1O(:,:), U1O(:,:), V1E(:,:), UZO(:,:), VZE(:,:), B1O(:,:)
  REAL, POINTER :: PNH_bot(:,:), DIVG_bot_surf(:,:)
END MODULE MO_COMMO1

MODULE MO_OBCS
  IMPLICIT NONE
  LOGICAL :: West = .TRUE., East = .TRUE.
  LOGICAL :: North = .FALSE., South = .FALSE.
END MODULE MO_OBCS

MODULE MO_PARA2
  IMPLICIT NONE
  REAL SORPAR
END MODULE MO_PARA2
