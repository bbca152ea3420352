! orctm_nh_c.f90 -- ISO_C_BINDING interface to liborctm_nh.so (include/orctm_nh.h) and a drop-in
! replacement for the three MO_PETSC entry points of ORCTM (ORCTM/src_orctm/mo_petsc.f90):
!
!   petsc_start     (mo_petsc.f90:41-128, called from orctm.f90:970)
!   petsc_solve_3D  (mo_petsc.f90:211-233, called from ocvtot_nh.f90:150)
!   petsc_end       (mo_petsc.f90:131-143, called from orctm.f90:1585,1651)
!
! In ocvtot_nh.f90 and orctm.f90 replace `use mo_petsc` by `use orctm_nh_c`; nothing else changes.
! Optionally, ocvtot_nh_device replaces the body of OCVTOT_NH (ocvtot_nh.f90:11-249, single process):
! the divergence, with CG2D4NONHY the 2-D bottom-pressure solve and the corrections of DIVG, the solve
! and the projection then all run on the device; and
! troneu_device replaces CALL TRONEU (orctm.f90:1414): the barotropic SOR solve on the device.
! aufw_nh / aufr_nh save and restore the solver's resident pressure, from which the next solve
! warm-starts, with the restart files: call aufw_nh after AUFW writes its records and aufr_nh after
! AUFR and petsc_start (the reference's restart holds no NH state, aufw_fragment.f90:40-47, and
! beleg_zero.f90:440-506 zeroes it, so its first solve after a restart starts from zero).
! The module reads the same module globals compute_matrix reads (MO_PARAM1 sizes, MO_COMMO1 metrics
! and masks, MO_OBCS flags) and hands the caller's arrays, halo included, to the C ABI.
! One GPU holds the whole domain (nprocx = nprocy = 1). Build (with the model's -r8 or its equivalent):
! amdflang|gfortran|ifort -cpp -c orctm_nh_c.f90 [-DCG2D4NONHY], link -L<repo>/orctm_amd -lorctm_nh.
MODULE orctm_nh_c
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: petsc_start, petsc_solve_3D, petsc_end, ocvtot_nh_device, troneu_device, aufw_nh, aufr_nh
  PUBLIC :: orctm_nh_grid_t, orctm_nh_options_t, orctm_nh_stats_t, orctm_nh_vgrid_t

  INTEGER(C_INT), PARAMETER, PUBLIC :: ORCTM_NH_OK = 0, ORCTM_NH_FP64 = 0, ORCTM_NH_MIXED = 1
  ! code of the resident-pressure record in a restart file (EXTRA-style header: IDATE, ICODE, KCODE, IEDIM)
  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: NH_RESTART_CODE = 594

  TYPE, BIND(C) :: orctm_nh_grid_t
    INTEGER(C_INT32_T) :: ie, je, ke, i_start, j_start
    TYPE(C_PTR) :: dlxp, dlyp, dlxu, dlyv, ddpo, weto, dzw
    INTEGER(C_INT32_T) :: cyclic_x, cyclic_y, open_w, open_e, open_n, open_s
  END TYPE

  TYPE, BIND(C) :: orctm_nh_options_t
    INTEGER(C_INT32_T) :: precision, maxit, max_outer, warm_start
    REAL(C_DOUBLE) :: rtol, rtol_inner
    INTEGER(C_INT32_T) :: check_every, device, profile, tile_rows
    INTEGER(C_INT32_T) :: bottom_solve, zw_table, buffer_addressing, reserved(1)
  END TYPE

  TYPE, BIND(C) :: orctm_nh_stats_t
    INTEGER(C_INT32_T) :: status, outer_iters, inner_iters, tile_rows
    REAL(C_DOUBLE) :: rel_res, bnorm, solve_ms, total_ms
    INTEGER(C_INT32_T) :: zw_table, buffer_addressing
    REAL(C_DOUBLE) :: reserved(3)
  END TYPE

  TYPE, BIND(C) :: orctm_nh_sor_grid_t
    INTEGER(C_INT32_T) :: ie, je, i_start, j_start, cyclic_x, cyclic_y, iterations, device
    INTEGER(C_INT32_T) :: half_sweeps, reserved
    TYPE(C_PTR) :: dlxp, dlyp, dtdxpo, dtdyo, weto, dlxu, dlyu, deuto, dlxv, dlyv, deute
    REAL(C_DOUBLE) :: g, conn, cono, stabn, dt, sorpar
  END TYPE

  TYPE, BIND(C) :: orctm_nh_vgrid_t
    TYPE(C_PTR) :: dlxp, dlyp, ddpo, weto, amsuo, dduo, dtdxuo, dlyu, amsue, ddue, dpye, dlxv
    TYPE(C_PTR) :: dtdxpo, dtdyo
    REAL(C_DOUBLE) :: dt, dti
  END TYPE

  INTERFACE
    INTEGER(C_INT) FUNCTION orctm_nh_sor_create(sg, sor) BIND(C, name="orctm_nh_sor_create")
      IMPORT :: C_INT, C_PTR, orctm_nh_sor_grid_t
      TYPE(orctm_nh_sor_grid_t), INTENT(IN) :: sg
      TYPE(C_PTR), INTENT(OUT) :: sor
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_troneu(sor, u1o, uzo, v1e, vze, b1o, z1o, device_ms) &
        BIND(C, name="orctm_nh_troneu")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: sor
      REAL(C_DOUBLE), INTENT(IN) :: u1o(*), uzo(*), v1e(*), vze(*)
      REAL(C_DOUBLE), INTENT(INOUT) :: b1o(*), z1o(*)
      TYPE(C_PTR), VALUE :: device_ms
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_attach_vgrid(handle, vg) BIND(C, name="orctm_nh_attach_vgrid")
      IMPORT :: C_INT, C_PTR, orctm_nh_vgrid_t
      TYPE(C_PTR), VALUE :: handle
      TYPE(orctm_nh_vgrid_t), INTENT(IN) :: vg
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_divergence(handle, uoo, uoo_1, voe, voe_1, wno, wno_1, zo, divg) &
        BIND(C, name="orctm_nh_divergence")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), INTENT(IN) :: uoo(*), uoo_1(*), voe(*), voe_1(*), wno(*), wno_1(*), zo(*)
      REAL(C_DOUBLE), INTENT(INOUT) :: divg(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_get_bottom(handle, pnh_bot, divg_bot_surf, iterations, rel_res, &
        device_ms) BIND(C, name="orctm_nh_get_bottom")
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), INTENT(OUT) :: pnh_bot(*)
      REAL(C_DOUBLE), INTENT(INOUT) :: divg_bot_surf(*)
      INTEGER(C_INT32_T), INTENT(OUT) :: iterations
      REAL(C_DOUBLE), INTENT(OUT) :: rel_res, device_ms
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_project(handle, pnh, uoo, voe, wno) BIND(C, name="orctm_nh_project")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), INTENT(INOUT) :: pnh(*), uoo(*), voe(*), wno(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_solve_resident(handle, stats) BIND(C, name="orctm_nh_solve_resident")
      IMPORT :: C_INT, C_PTR, orctm_nh_stats_t
      TYPE(C_PTR), VALUE :: handle
      TYPE(orctm_nh_stats_t), INTENT(OUT) :: stats
    END FUNCTION
    SUBROUTINE orctm_nh_default_options(opts) BIND(C, name="orctm_nh_default_options")
      IMPORT :: orctm_nh_options_t
      TYPE(orctm_nh_options_t), INTENT(OUT) :: opts
    END SUBROUTINE
    INTEGER(C_INT) FUNCTION orctm_nh_create(grid, opts, handle) BIND(C, name="orctm_nh_create")
      IMPORT :: C_INT, C_PTR, orctm_nh_grid_t, orctm_nh_options_t
      TYPE(orctm_nh_grid_t), INTENT(IN) :: grid
      TYPE(orctm_nh_options_t), INTENT(IN) :: opts
      TYPE(C_PTR), INTENT(OUT) :: handle
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_solve(handle, divg, pnh, stats) BIND(C, name="orctm_nh_solve")
      IMPORT :: C_INT, C_PTR, C_DOUBLE, orctm_nh_stats_t
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), INTENT(IN) :: divg(*)
      REAL(C_DOUBLE), INTENT(INOUT) :: pnh(*)
      TYPE(orctm_nh_stats_t), INTENT(OUT) :: stats
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_get_solution(handle, pnh) BIND(C, name="orctm_nh_get_solution")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), INTENT(INOUT) :: pnh(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_set_guess(handle, pnh) BIND(C, name="orctm_nh_set_guess")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), INTENT(IN) :: pnh(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION orctm_nh_destroy(handle) BIND(C, name="orctm_nh_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
    END FUNCTION
  END INTERFACE

  TYPE(C_PTR), SAVE :: nh_handle = C_NULL_PTR
  TYPE(C_PTR), SAVE :: sor_handle = C_NULL_PTR
  LOGICAL, SAVE :: nh_vgrid_attached = .FALSE.

CONTAINS

  ! mo_petsc.f90:41-128 -- sizes, metrics and masks come from the module globals, as in compute_matrix.
  SUBROUTINE petsc_start
    USE MO_PARAM1, ONLY: IE, JE, KE, ICYCLI_X, ICYCLI_Y, I_start, J_start
    USE MO_COMMO1, ONLY: DLXP, DLYP, DLXU, DLYV, DDPO, WETO, DZW
    USE MO_OBCS, ONLY: West, East, North, South
    TYPE(orctm_nh_grid_t) :: g
    TYPE(orctm_nh_options_t) :: o
    INTEGER(C_INT) :: ierr
    ! DZW is a plain array in MO_COMMO1 (neither POINTER nor TARGET), so C_LOC may not take it; the
    ! library reads every grid array during orctm_nh_create only, so a copy that lives for the call serves
    REAL(C_DOUBLE), TARGET :: dzw_c(KE)
    dzw_c(:) = DZW(:)
    g%ie = IE; g%je = JE; g%ke = KE
    g%i_start = I_start; g%j_start = J_start
    g%dlxp = C_LOC(DLXP); g%dlyp = C_LOC(DLYP)
    g%dlxu = C_LOC(DLXU); g%dlyv = C_LOC(DLYV)
    g%ddpo = C_LOC(DDPO); g%weto = C_LOC(WETO); g%dzw = C_LOC(dzw_c)
    g%cyclic_x = ICYCLI_X; g%cyclic_y = ICYCLI_Y
    g%open_w = MERGE(1, 0, West);  g%open_e = MERGE(1, 0, East)
    g%open_n = MERGE(1, 0, North); g%open_s = MERGE(1, 0, South)
    CALL orctm_nh_default_options(o)
#ifdef CG2D4NONHY
    o%bottom_solve = 1
#endif
    ierr = orctm_nh_create(g, o, nh_handle)
    IF (ierr /= ORCTM_NH_OK) THEN
      PRINT *, 'orctm_nh_create failed, status ', ierr
      STOP
    ENDIF
  END SUBROUTINE petsc_start

  ! mo_petsc.f90:211-233 -- reads DIVG, writes the PNH interior, warm-starts from the last solution.
  SUBROUTINE petsc_solve_3D
    USE MO_COMMO1, ONLY: DIVG, PNH
    TYPE(orctm_nh_stats_t) :: st
    INTEGER(C_INT) :: ierr
    ierr = orctm_nh_solve(nh_handle, DIVG, PNH, st)
    IF (ierr /= ORCTM_NH_OK) PRINT *, 'orctm_nh_solve status ', ierr, ' rel.res ', st%rel_res
  END SUBROUTINE petsc_solve_3D

  ! The whole of OCVTOT_NH (ocvtot_nh.f90:1-259, single process) on the device: DIVG from the velocity
  ! tendencies (:11-62); with CG2D4NONHY the bottom values, the 2-D solve for PNH_bot (:64-147) and the
  ! corrections of DIVG (mo_petsc.f90:252-374), PNH_bot and DIVG_bot_surf copied back for the output
  ! (wrte_mean.f90:202-207); the solve (:150), the PNH halo, the U/V correction and WNO by continuity
  ! (:151-230); the copies of :232-249 stay here on the host.
  SUBROUTINE ocvtot_nh_device
    USE MO_PARAM1, ONLY: IE, JE, KE, KEP, I_start, J_start
    USE MO_COMMO1, ONLY: DLXP, DLYP, DDPO, WETO, AMSUO, DDUO, DTDXUO, DLYU, AMSUE, DDUE, DPYE, DLXV, &
                         DTDXPO, DTDYO, DT, DTI, UOO, UOO_1, VOE, VOE_1, WNO, WNO_1, WO, ZO, DIVG, PNH
#ifdef CG2D4NONHY
    USE MO_COMMO1, ONLY: PNH_bot, DIVG_bot_surf
    INTEGER(C_INT32_T) :: nit
    REAL(C_DOUBLE) :: rr, ms
#endif
    TYPE(orctm_nh_vgrid_t) :: vg
    TYPE(orctm_nh_stats_t) :: st
    INTEGER(C_INT) :: ierr
    IF (.NOT. nh_vgrid_attached) THEN
      vg%dlxp = C_LOC(DLXP); vg%dlyp = C_LOC(DLYP); vg%ddpo = C_LOC(DDPO); vg%weto = C_LOC(WETO)
      vg%amsuo = C_LOC(AMSUO); vg%dduo = C_LOC(DDUO); vg%dtdxuo = C_LOC(DTDXUO); vg%dlyu = C_LOC(DLYU)
      vg%amsue = C_LOC(AMSUE); vg%ddue = C_LOC(DDUE); vg%dpye = C_LOC(DPYE); vg%dlxv = C_LOC(DLXV)
      vg%dtdxpo = C_LOC(DTDXPO); vg%dtdyo = C_LOC(DTDYO)
      vg%dt = DT; vg%dti = DTI
      ierr = orctm_nh_attach_vgrid(nh_handle, vg)
      IF (ierr /= ORCTM_NH_OK) PRINT *, 'orctm_nh_attach_vgrid status ', ierr
      nh_vgrid_attached = .TRUE.
    ENDIF
    ierr = orctm_nh_divergence(nh_handle, UOO, UOO_1, VOE, VOE_1, WNO, WNO_1, ZO, DIVG)
    IF (ierr /= ORCTM_NH_OK) PRINT *, 'orctm_nh_divergence status ', ierr
#ifdef CG2D4NONHY
    ierr = orctm_nh_get_bottom(nh_handle, PNH_bot, DIVG_bot_surf, nit, rr, ms)
    IF (ierr /= ORCTM_NH_OK) PRINT *, 'orctm_nh_get_bottom status ', ierr, ' rel.res ', rr
#endif
    ierr = orctm_nh_solve_resident(nh_handle, st)
    IF (ierr /= ORCTM_NH_OK) PRINT *, 'orctm_nh_solve_resident status ', ierr, ' rel.res ', st%rel_res
    ierr = orctm_nh_project(nh_handle, PNH, UOO, VOE, WNO)
    IF (ierr /= ORCTM_NH_OK) PRINT *, 'orctm_nh_project status ', ierr
    UOO_1(:,:,:) = UOO(:,:,:)
    VOE_1(:,:,:) = VOE(:,:,:)
    WNO_1(:,:,:) = WNO(:,:,:)
    WO(:,:,:) = WNO(:,:,:)
  END SUBROUTINE ocvtot_nh_device

  ! TRONEU (troneu.f90:48-131, single process) on the device; the first call also does the work of
  ! ITPREP (itprep.f90:15-38) there. Replace `CALL TRONEU` (orctm.f90:1414) by `CALL troneu_device`.
  SUBROUTINE troneu_device
    USE MO_PARAM1, ONLY: IE, JE, I_start, J_start, ICYCLI_X, ICYCLI_Y
    USE MO_PARA2, ONLY: SORPAR
    USE MO_COMMO1, ONLY: DLXP, DLYP, DTDXPO, DTDYO, WETO, DLXU, DLYU, DEUTO, DLXV, DLYV, DEUTE, G, CONN, CONO, &
                         STABN, DT, U1O, UZO, V1E, VZE, B1O, Z1O
    TYPE(orctm_nh_sor_grid_t) :: sg
    INTEGER(C_INT) :: ierr
    IF (.NOT. C_ASSOCIATED(sor_handle)) THEN
      sg%ie = IE; sg%je = JE; sg%i_start = I_start; sg%j_start = J_start
      sg%cyclic_x = ICYCLI_X; sg%cyclic_y = ICYCLI_Y; sg%iterations = 250; sg%device = 0
      sg%half_sweeps = 0; sg%reserved = 0
      sg%dlxp = C_LOC(DLXP); sg%dlyp = C_LOC(DLYP); sg%dtdxpo = C_LOC(DTDXPO); sg%dtdyo = C_LOC(DTDYO)
      sg%weto = C_LOC(WETO); sg%dlxu = C_LOC(DLXU); sg%dlyu = C_LOC(DLYU); sg%deuto = C_LOC(DEUTO)
      sg%dlxv = C_LOC(DLXV); sg%dlyv = C_LOC(DLYV); sg%deute = C_LOC(DEUTE)
      sg%g = G; sg%conn = CONN; sg%cono = CONO; sg%stabn = STABN; sg%dt = DT; sg%sorpar = SORPAR
      ierr = orctm_nh_sor_create(sg, sor_handle)
      IF (ierr /= ORCTM_NH_OK) THEN
        PRINT *, 'orctm_nh_sor_create failed, status ', ierr
        STOP
      ENDIF
    ENDIF
    ierr = orctm_nh_troneu(sor_handle, U1O, UZO, V1E, VZE, B1O, Z1O, C_NULL_PTR)
    IF (ierr /= ORCTM_NH_OK) PRINT *, 'orctm_nh_troneu status ', ierr
  END SUBROUTINE troneu_device

  ! The resident pressure (the interior of PNH from the last solve, from which the next solve warm-starts)
  ! as two records on the restart unit IUNIT, sequential and unformatted like the restart files
  ! (aufw_fragment.f90:38-47): an EXTRA-style header IDATE, ICODE = NH_RESTART_CODE, KCODE = KE,
  ! IEDIM = IE*JE, then the (IE,JE,KE) field with a zero halo.
  SUBROUTINE aufw_nh(iunit, idate)
    USE MO_PARAM1, ONLY: IE, JE, KE
    INTEGER, INTENT(IN) :: iunit
    INTEGER(C_INT32_T), INTENT(IN) :: idate
    REAL(C_DOUBLE), ALLOCATABLE :: p(:,:,:)
    INTEGER(C_INT) :: ierr
    ALLOCATE(p(IE, JE, KE))
    p(:,:,:) = 0.0_C_DOUBLE
    ierr = orctm_nh_get_solution(nh_handle, p)
    IF (ierr /= ORCTM_NH_OK) PRINT *, 'orctm_nh_get_solution status ', ierr
    WRITE(iunit) idate, NH_RESTART_CODE, INT(KE, C_INT32_T), INT(IE * JE, C_INT32_T)
    WRITE(iunit) p
    DEALLOCATE(p)
  END SUBROUTINE aufw_nh

  ! Reads the two records aufw_nh wrote (after petsc_start) and makes their field the initial guess of
  ! the next solve, which then warm-starts exactly as it would have without the restart (mo_petsc.f90:232).
  SUBROUTINE aufr_nh(iunit)
    USE MO_PARAM1, ONLY: IE, JE, KE
    INTEGER, INTENT(IN) :: iunit
    INTEGER(C_INT32_T) :: hdr(4)
    REAL(C_DOUBLE), ALLOCATABLE :: p(:,:,:)
    INTEGER(C_INT) :: ierr
    READ(iunit) hdr
    IF (hdr(2) /= NH_RESTART_CODE .OR. hdr(3) /= KE .OR. hdr(4) /= IE * JE) THEN
      PRINT *, 'aufr_nh: not a resident-pressure record of this grid: ', hdr
      RETURN
    ENDIF
    ALLOCATE(p(IE, JE, KE))
    READ(iunit) p
    ierr = orctm_nh_set_guess(nh_handle, p)
    IF (ierr /= ORCTM_NH_OK) PRINT *, 'orctm_nh_set_guess status ', ierr
    DEALLOCATE(p)
  END SUBROUTINE aufr_nh

  ! mo_petsc.f90:131-143
  SUBROUTINE petsc_end
    INTEGER(C_INT) :: ierr
    IF (C_ASSOCIATED(nh_handle)) ierr = orctm_nh_destroy(nh_handle)
    nh_handle = C_NULL_PTR
    nh_vgrid_attached = .FALSE.
  END SUBROUTINE petsc_end

END MODULE orctm_nh_c
