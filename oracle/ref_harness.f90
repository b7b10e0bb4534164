! Golden-value dump harness for the FP64 reference build.
!
! Test infrastructure only: it is linked against the reference NEMOH Solver
! modules (built by oracle/Makefile.ref into oracle/_ref/) and calls the
! reference routines on inputs read from stdin, printing their outputs with
! 17 significant digits.  tests/golden/make_golden.py drives it and stores the
! results as fixtures.  Nothing in the product links or calls this.
!
! Commands (one per stdin record):
!   gg     n  then n lines "re im"       -> GG(z, exp(z))           ELEMENTARY_FNS.f90:31
!   crek                                  -> XR, XZ, APD1X..APD2Z    NEMOH.f90:179
!   cinti  n  then n values of AKK        -> N, module NQ, QQ, CQ    COMPUTE_GREEN_INFD.f90:505
!   x0     n  then n values of AK         -> X0(AK)                  COMPUTE_GREEN_FD.f90:1635
!   wavenumber n then n lines "w depth g" -> the preProcessor's k   Common/Environment.f90:45-84
!   kochin kwave xeff yeff ntheta, ntheta angles, NFA lines "ZIGB ZIGS" (re im re im)
!          (after mesh)                   -> HKochin(theta)          COMPUTE_KOCHIN.f90:25-153
!   potential period npt, npt lines "x y z", NFA lines "ZIGB ZIGS" (after mesh)
!                                         -> PHI at each point        COMPUTE_POTENTIAL_DOMAIN.f90:24-82
!          (kwave, AMH, NEXP and LISC's AMBDA/AR set as SOLVE_BEM.f90:66-89 and
!          SOLVE_BEM_FD_DIRECT.f90:53-100 set them)
!   lisc   n  then n lines "akh amh"      -> NEXP, AMBDA, AR         COMPUTE_GREEN_FD.f90:1007
!   mesh   np nfa nsym depth, np lines "x y z", nfa lines "m1 m2 m3 m4"
!                                          -> sets COM_VAR, runs PRE_PROC_MESH, CREK
!   depth  h                              -> sets Depth only (geometry kept)
!   pairs  period  (after mesh)           -> every (ISP,IFP) pair of VAV*+VNS*
!   bench  period rows stride m (after mesh) -> wall seconds of rows x NFA pair
!                                            evaluations (field rows 1, 1+stride, ...;
!                                            VAV*+VNS*, same set-up as pairs) and of
!                                            GAUSSZ on an m x 2m system (M_SOLVER.f90:31)
!   end
PROGRAM ref_harness
  USE COM_VAR
  USE ELEMENTARY_FNS
  USE COMPUTE_GREEN_INFD, ONLY: VAVINFD, VNSINFD, CINT_INFD
  USE COMPUTE_GREEN_FD, ONLY: VAVFD, VNSFD, LISC, X0, CINT_FD, FF, HOUSRS, SPRBM
  USE PREPARE_MESH
  USE MMesh
  USE M_SOLVER, ONLY: GAUSSZ
  USE MEnvironment, ONLY: TEnvironment, Wavenumber
  USE MIDENTIFICATION, ONLY: TID
  IMPLICIT NONE
  CHARACTER(LEN=16) :: cmd
  INTEGER :: n, i, j, k, nq_out, nexp, isp, ifp, ios, iperiod
  REAL :: re, im, akk, ak, akh, amh, period, w, am0, gm, dij, bj
  COMPLEX :: zz, cexv, gv
  TYPE(TMesh) :: mesh
  INTEGER :: np_in, nfa_in, nsym_in
  REAL :: depth_in
  INTEGER :: nl, ncc, ic
  REAL :: hs(120,32), hc(31), hr(31), hi(31), hp(31)
  INTEGER :: nrows, mg, stride, r
  INTEGER(8) :: c0, c1, crate
  REAL :: tpairs, tgauss, acc
  COMPLEX, ALLOCATABLE :: zmat(:,:)
  LOGICAL :: fd
  TYPE(TEnvironment) :: env
  REAL :: kw, r1, i1, r2, i2
  REAL, ALLOCATABLE :: thetas(:)
  COMPLEX :: hk
  EXTERNAL :: COMPUTE_KOCHIN  ! external subroutine of COMPUTE_KOCHIN.f90
  EXTERNAL :: COMPUTE_POTENTIAL_DOMAIN
  TYPE(TID) :: idd
  REAL, ALLOCATABLE :: pts(:,:)
  COMPLEX :: phi

  DO
    READ(*,*,IOSTAT=ios) cmd
    IF (ios /= 0) EXIT
    SELECT CASE (TRIM(cmd))
    CASE ('gg')
      READ(*,*) n
      DO i = 1, n
        READ(*,*) re, im
        zz = CMPLX(re, im)
        cexv = EXP(zz)
        gv = GG(zz, cexv)
        WRITE(*,'(4ES26.17E3)') re, im, REAL(gv), AIMAG(gv)
      END DO
    CASE ('crek')
      CALL CREK
      WRITE(*,'(ES26.17E3)') (XR(i), i = 1, 328)
      WRITE(*,'(ES26.17E3)') (XZ(j), j = 1, 46)
      DO j = 1, 46
        DO i = 1, 328
          WRITE(*,'(4ES26.17E3)') APD1X(i,j), APD1Z(i,j), APD2X(i,j), APD2Z(i,j)
        END DO
      END DO
    CASE ('cinti')
      READ(*,*) n
      DO i = 1, n
        READ(*,*) akk
        CALL CINT_INFD(akk, nq_out)
        WRITE(*,'(ES26.17E3,2I6)') akk, nq_out, NQ
        WRITE(*,'(2ES26.17E3)') (QQ(k), CQ(k), k = 1, nq_out)
      END DO
    CASE ('x0')
      READ(*,*) n
      DO i = 1, n
        READ(*,*) ak
        WRITE(*,'(2ES26.17E3)') ak, X0(ak)
      END DO
    CASE ('wavenumber')
      READ(*,*) n
      DO i = 1, n
        READ(*,*) w, env%Depth, env%G
        WRITE(*,'(4ES26.17E3)') w, env%Depth, env%G, Wavenumber(w, env)
      END DO
    CASE ('kochin')
      READ(*,*) kw, XEFF, YEFF, n
      ALLOCATE(thetas(n))
      DO i = 1, n
        READ(*,*) thetas(i)
      END DO
      DO i = 1, NFA
        READ(*,*) r1, i1, r2, i2
        ZIGB(i) = CMPLX(r1, i1)
        ZIGS(i) = CMPLX(r2, i2)
      END DO
      DO i = 1, n
        CALL COMPUTE_KOCHIN(kw, thetas(i), hk)
        WRITE(*,'(3ES26.17E3)') thetas(i), REAL(hk), AIMAG(hk)
      END DO
      DEALLOCATE(thetas)
    CASE ('potential')
      READ(*,*) period, n
      ALLOCATE(pts(3, n))
      DO i = 1, n
        READ(*,*) pts(1,i), pts(2,i), pts(3,i)
      END DO
      DO i = 1, NFA
        READ(*,*) r1, i1, r2, i2
        ZIGB(i) = CMPLX(r1, i1)
        ZIGS(i) = CMPLX(r2, i2)
      END DO
      T = period
      w = 2.*4.*ATAN(1.)/T
      AMBDA = 0.; AR = 0.
      IF ((Depth == 0.) .OR. (w*w/G*Depth >= 20)) THEN
        kw = w*w/G
        amh = kw*Depth
        nexp = 0
        WRITE(*,'(A,2ES26.17E3,I6)') 'INFD', kw, amh, nexp
      ELSE
        akh = w*w*Depth/G
        amh = X0(akh)
        kw = amh/Depth
        CALL LISC(amh*TANH(amh), amh, nexp)
        WRITE(*,'(A,2ES26.17E3,I6)') 'FD', kw, amh, nexp
      END IF
      DO i = 1, n
        CALL COMPUTE_POTENTIAL_DOMAIN(idd, phi, pts(1,i), pts(2,i), pts(3,i), kw, amh, nexp)
        WRITE(*,'(5ES26.17E3)') pts(1,i), pts(2,i), pts(3,i), REAL(phi), AIMAG(phi)
      END DO
      DEALLOCATE(pts)
    CASE ('lisc')
      READ(*,*) n
      DO i = 1, n
        READ(*,*) akh, amh
        AMBDA = 0.
        AR = 0.
        CALL LISC(akh, amh, nexp)
        WRITE(*,'(2ES26.17E3,I6)') akh, amh, nexp
        WRITE(*,'(2ES26.17E3)') (AMBDA(k), AR(k), k = 1, nexp)
      END DO
    CASE ('mesh')
      READ(*,*) np_in, nfa_in, nsym_in, depth_in
      CALL CreateTMesh(mesh, np_in, nfa_in, 1)
      DO i = 1, np_in
        READ(*,*) mesh%X(1,i), mesh%X(2,i), mesh%X(3,i)
      END DO
      DO i = 1, nfa_in
        READ(*,*) mesh%P(1,i), mesh%P(2,i), mesh%P(3,i), mesh%P(4,i)
      END DO
      NFA = nfa_in
      NP = np_in
      NSYMY = nsym_in
      Depth = depth_in
      G = 9.81
      RHO = 1025.
      Indiq_solver = 0
      CALL ALLOCATE_DATA
      CALL PRE_PROC_MESH(mesh)
      CALL CREK
      WRITE(*,'(ES26.17E3)') ZER
      DO i = 1, NFA
        WRITE(*,'(8ES26.17E3)') XG(i), YG(i), ZG(i), XN(i), YN(i), ZN(i), AIRE(i), TDIS(i)
      END DO
    CASE ('pairs')
      ! Same per-frequency set-up as SOLVE_POTENTIAL_*_DIRECT
      ! (SOLVE_BEM_INFD_DIRECT.f90:47-80, SOLVE_BEM_FD_DIRECT.f90:48-100).
      READ(*,*) period
      T = period
      w = 2.*4.*ATAN(1.)/T
      VSXP = 0.; VSXM = 0.; VSYP = 0.; VSYM = 0.; VSZP = 0.; VSZM = 0.
      CQ = 0.; QQ = 0.; AMBDA = 0.; AR = 0.
      gm = 0.
      DO i = 1, NP - 1
        DO k = 1, NSYMY + 1
          bj = (-1.)**(k+1)
          DO j = i + 1, NP
            dij = SQRT((X(j)-X(i))**2 + (Y(i)-Y(j)*bj)**2)
            gm = MAX(dij, gm)
          END DO
        END DO
      END DO
      IF ((Depth == 0.) .OR. (w*w/G*Depth >= 20)) THEN
        am0 = w**2/G
        CALL CINT_INFD(am0*gm, n)
        WRITE(*,'(A,ES26.17E3,I6)') 'INFD', am0, NQ
        DO isp = 1, NFA
          DO ifp = 1, NFA
            CALL VAVINFD(1, XG(isp), YG(isp), ZG(isp), isp, ifp)
            CALL VNSINFD(1, isp, ifp, XG(isp), YG(isp), ZG(isp))
            CALL dump_pair(isp, ifp)
          END DO
        END DO
      ELSE
        akh = w*w*Depth/G
        amh = X0(akh)
        akh = amh*TANH(amh)
        am0 = amh/Depth
        CALL CINT_FD(am0*gm, n)
        NQ = n
        CALL LISC(akh, amh, nexp)
        WRITE(*,'(A,ES26.17E3,I6,I6)') 'FD', amh, NQ, nexp
        WRITE(*,'(2ES26.17E3)') (AMBDA(k), AR(k), k = 1, nexp)
        DO isp = 1, NFA
          DO ifp = 1, NFA
            CALL VAVFD(2, XG(isp), YG(isp), ZG(isp), isp, ifp)
            CALL VNSFD(am0, amh, nexp, isp, ifp, XG(isp), YG(isp), ZG(isp))
            CALL dump_pair(isp, ifp)
          END DO
        END DO
      END IF
    CASE ('bench')
      READ(*,*) period, nrows, stride, mg
      T = period
      w = 2.*4.*ATAN(1.)/T
      VSXP = 0.; VSXM = 0.; VSYP = 0.; VSYM = 0.; VSZP = 0.; VSZM = 0.
      CQ = 0.; QQ = 0.; AMBDA = 0.; AR = 0.
      gm = 0.
      DO i = 1, NP - 1
        DO k = 1, NSYMY + 1
          bj = (-1.)**(k+1)
          DO j = i + 1, NP
            dij = SQRT((X(j)-X(i))**2 + (Y(i)-Y(j)*bj)**2)
            gm = MAX(dij, gm)
          END DO
        END DO
      END DO
      fd = .NOT. ((Depth == 0.) .OR. (w*w/G*Depth >= 20))
      acc = 0.
      CALL SYSTEM_CLOCK(c0, crate)
      IF (.NOT. fd) THEN
        am0 = w**2/G
        CALL CINT_INFD(am0*gm, n)
        DO r = 0, nrows - 1
          isp = 1 + MOD(r * stride, NFA)
          DO ifp = 1, NFA
            CALL VAVINFD(1, XG(isp), YG(isp), ZG(isp), isp, ifp)
            CALL VNSINFD(1, isp, ifp, XG(isp), YG(isp), ZG(isp))
            acc = acc + VSXP1 + SP2
          END DO
        END DO
      ELSE
        akh = w*w*Depth/G
        amh = X0(akh)
        akh = amh*TANH(amh)
        am0 = amh/Depth
        CALL CINT_FD(am0*gm, n)
        NQ = n
        CALL LISC(akh, amh, nexp)
        DO r = 0, nrows - 1
          isp = 1 + MOD(r * stride, NFA)
          DO ifp = 1, NFA
            CALL VAVFD(2, XG(isp), YG(isp), ZG(isp), isp, ifp)
            CALL VNSFD(am0, amh, nexp, isp, ifp, XG(isp), YG(isp), ZG(isp))
            acc = acc + VSXP1 + SP2
          END DO
        END DO
      END IF
      CALL SYSTEM_CLOCK(c1)
      tpairs = REAL(c1 - c0) / REAL(crate)
      ALLOCATE(zmat(mg, 2*mg))
      DO j = 1, mg
        DO i = 1, mg
          zmat(i, j) = CMPLX(SIN(1.3*i + 0.7*j), COS(0.9*i - 1.1*j)) / mg
          zmat(i, j + mg) = (0., 0.)
        END DO
        zmat(j, j) = zmat(j, j) + 2.
        zmat(j, j + mg) = (1., 0.)
      END DO
      CALL SYSTEM_CLOCK(c0)
      CALL GAUSSZ(zmat, mg, mg, 2*mg)
      CALL SYSTEM_CLOCK(c1)
      tgauss = REAL(c1 - c0) / REAL(crate)
      acc = acc + REAL(zmat(1, mg + 1))
      DEALLOCATE(zmat)
      WRITE(*,'(L2,I8,ES26.17E3,I8,2ES26.17E3)') fd, nrows, tpairs, mg, tgauss, acc
    CASE ('ff')
      ! FF (COMPUTE_GREEN_FD.f90:1612) at n points: "x ak am"
      READ(*,*) n
      DO i = 1, n
        READ(*,*) re, ak, amh
        WRITE(*,'(ES26.17E3)') FF(re, ak, amh)
      END DO
    CASE ('housrs')
      ! HOUSRS (COMPUTE_GREEN_FD.f90:1454) on an nl x (ncc+1) matrix
      READ(*,*) nl, ncc
      hs = 0.
      DO i = 1, nl
        READ(*,*) (hs(i,j), j = 1, ncc + 1)
      END DO
      CALL HOUSRS(hs, 120, nl, ncc, 1, 1.E-20)
      WRITE(*,'(ES26.17E3)') (hs(i,ncc+1), i = 1, ncc)
    CASE ('sprbm')
      ! SPRBM (COMPUTE_GREEN_FD.f90:1182) on coefficients c(1..ic)
      READ(*,*) ic
      READ(*,*) (hc(i), i = 1, ic)
      CALL SPRBM(hc, ic, hr, hi, hp, isp, ifp)
      WRITE(*,'(I6,I6)') isp, ifp
      WRITE(*,'(2ES26.17E3)') (hr(i), hi(i), i = 1, ic)
    CASE ('depth')
      READ(*,*) Depth
    CASE ('end')
      EXIT
    END SELECT
  END DO

CONTAINS

  SUBROUTINE dump_pair(a, b)
    INTEGER, INTENT(IN) :: a, b
    WRITE(*,'(2I7,4ES26.17E3)') a, b, SP1, SM1, SP2, SM2
    WRITE(*,'(6ES26.17E3)') VSXP1, VSYP1, VSZP1, VSXM1, VSYM1, VSZM1
    WRITE(*,'(6ES26.17E3)') VSXP2, VSYP2, VSZP2, VSXM2, VSYM2, VSZM2
  END SUBROUTINE dump_pair

END PROGRAM ref_harness
