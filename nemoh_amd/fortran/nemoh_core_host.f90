! Handle lifetime for the reference Solver: CORE_START after INITIALIZE has
! filled COM_VAR (geometry from PRE_PROC_MESH, tables from CREK), CORE_STOP
! before DEALLOCATE_DATA.  NEMOH_CORE_PRECISION=MIXED in the environment
! selects complex64 factors with FP64 refinement; the default is FP64.
MODULE NEMOH_CORE_HOST
  USE ISO_C_BINDING
  USE COM_VAR
  USE NEMOH_CORE_BIND
  IMPLICIT NONE
  ! only the two entry points are exported: the Solver's main program has
  ! locals of the same names as COM_VAR's NFA, NSYMY, T, XEFF, YEFF
  PRIVATE
  PUBLIC :: CORE_START, CORE_STOP
CONTAINS
  SUBROUTINE CORE_START()
    CHARACTER(LEN=16) :: mode
    INTEGER :: length, stat
    CALL CORE_CHECK(nemoh_core_create(CORE, NP, NFA, NSYMY, X, Y, Z, M1, M2, M3, M4, &
         XG, YG, ZG, XN, YN, ZN, AIRE, TDIS, ZER, G, RHO, Depth, XR, XZ, APD1X, APD1Z, APD2X, APD2Z))
    CALL GET_ENVIRONMENT_VARIABLE('NEMOH_CORE_PRECISION', mode, length, stat)
    IF (stat == 0 .AND. mode(1:5) == 'MIXED') THEN
      CALL CORE_CHECK(nemoh_core_set_precision(CORE, NEMOH_MIXED, 0.D0, 0))
    END IF
  END SUBROUTINE
  SUBROUTINE CORE_STOP()
    CALL nemoh_core_destroy(CORE)
    CORE = C_NULL_PTR
  END SUBROUTINE
END MODULE NEMOH_CORE_HOST
