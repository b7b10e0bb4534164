! ISO_C_BINDING interface to the solver core's C ABI (include/nemoh_core.h)
! for a Fortran host.  Built with the host's FP64 flags
! (-fdefault-real-8 -fdefault-double-8): REAL and COMPLEX are then C_DOUBLE
! and C_DOUBLE_COMPLEX, so the Solver's COM_VAR arrays pass straight through;
! INTEGER stays C_INT.
MODULE NEMOH_CORE_BIND
  USE ISO_C_BINDING
  IMPLICIT NONE
  TYPE(C_PTR), SAVE :: CORE = C_NULL_PTR
  INTEGER(C_INT), PARAMETER :: NEMOH_FP64 = 0, NEMOH_MIXED = 1
  INTERFACE
    ! nemoh_core_create: the COM_VAR state after INITIALIZE (INITIALIZATION.f90:28-69)
    INTEGER(C_INT) FUNCTION nemoh_core_create(ctx, np, nfa, nsym, x, y, z, m1, m2, m3, m4, &
        xg, yg, zg, xn, yn, zn, aire, tdis, zer, g, rho, depth, xr, xz, apd1x, apd1z, apd2x, apd2z) &
        BIND(C, NAME='nemoh_core_create')
      IMPORT
      TYPE(C_PTR) :: ctx
      INTEGER(C_INT), VALUE :: np, nfa, nsym
      REAL(C_DOUBLE) :: x(*), y(*), z(*)
      INTEGER(C_INT) :: m1(*), m2(*), m3(*), m4(*)
      REAL(C_DOUBLE) :: xg(*), yg(*), zg(*), xn(*), yn(*), zn(*), aire(*), tdis(*)
      REAL(C_DOUBLE), VALUE :: zer, g, rho, depth
      REAL(C_DOUBLE) :: xr(*), xz(*), apd1x(*), apd1z(*), apd2x(*), apd2z(*)
    END FUNCTION
    ! nemoh_core_solve: SOLVE_POTENTIAL_INFD_DIRECT / SOLVE_POTENTIAL_FD_DIRECT
    INTEGER(C_INT) FUNCTION nemoh_core_solve(ctx, period, amh, nexp, ambda, ar, nq, qq, cq, &
        nvel, zigb, zigs, zpb, zps) BIND(C, NAME='nemoh_core_solve')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: period, amh
      INTEGER(C_INT), VALUE :: nexp, nq
      REAL(C_DOUBLE) :: ambda(*), ar(*), qq(*), cq(*)
      COMPLEX(C_DOUBLE_COMPLEX) :: nvel(*), zigb(*), zigs(*), zpb(*), zps(*)
    END FUNCTION
    ! nemoh_core_kochin: COMPUTE_KOCHIN per angle (SOLVE_BEM.f90:98-103),
    ! for the sources of the last nemoh_core_solve
    INTEGER(C_INT) FUNCTION nemoh_core_kochin(ctx, xeff, yeff, ntheta, theta, hkochin) &
        BIND(C, NAME='nemoh_core_kochin')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: xeff, yeff
      INTEGER(C_INT), VALUE :: ntheta
      REAL(C_DOUBLE) :: theta(*)
      COMPLEX(C_DOUBLE_COMPLEX) :: hkochin(*)
    END FUNCTION
    ! nemoh_core_potential: COMPUTE_POTENTIAL_DOMAIN per field point
    ! (SOLVE_BEM.f90:104-110), xyz(3, npoints)
    INTEGER(C_INT) FUNCTION nemoh_core_potential(ctx, npoints, xyz, phi) BIND(C, NAME='nemoh_core_potential')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: npoints
      REAL(C_DOUBLE) :: xyz(*)
      COMPLEX(C_DOUBLE_COMPLEX) :: phi(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION nemoh_core_set_precision(ctx, mode, tol, max_iter) &
        BIND(C, NAME='nemoh_core_set_precision')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: mode, max_iter
      REAL(C_DOUBLE), VALUE :: tol
    END FUNCTION
    TYPE(C_PTR) FUNCTION nemoh_core_last_error(ctx) BIND(C, NAME='nemoh_core_last_error')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    SUBROUTINE nemoh_core_destroy(ctx) BIND(C, NAME='nemoh_core_destroy')
      IMPORT
      TYPE(C_PTR), VALUE :: ctx
    END SUBROUTINE
  END INTERFACE
CONTAINS
  ! Prints the core's message and stops, as the reference stops on its own
  ! errors (e.g. the pivot check of M_SOLVER.f90:49-52).
  SUBROUTINE CORE_CHECK(ierr)
    INTEGER(C_INT), INTENT(IN) :: ierr
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    INTEGER :: n
    IF (ierr == 0) RETURN
    n = 0
    IF (C_ASSOCIATED(CORE)) THEN
      CALL C_F_POINTER(nemoh_core_last_error(CORE), msg, [512])
      DO WHILE (n < 512)
        IF (msg(n + 1) == C_NULL_CHAR) EXIT
        n = n + 1
      END DO
      WRITE(*,*) 'nemoh_core error ', ierr, ': ', msg(1:n)
    ELSE
      WRITE(*,*) 'nemoh_core error ', ierr, ' (no handle: is a HIP device present?)'
    END IF
    STOP 1
  END SUBROUTINE
END MODULE NEMOH_CORE_BIND
