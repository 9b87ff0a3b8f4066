! ib3d_main.f90 -- the reference's driver (main.f90:2-68) on the MI355X.
!
! Same call sequence, same files in and out (inputdata, gridx/y/z.dat,
! geometries/sphere.msh; iter.dat, dragcoff.dat, liftcoff.dat, result,
! tag/, out/).  An optional first command-line argument caps the number of
! steps below itamax, for short checks.  With a nonzero amplitude a0 in
! inputdata the body oscillates (y_c = a0 sin(2 pi freq t)) and the geometry
! is rebuilt every step (shiftSurfaceNodesMovingbody, computeSurfaceNorm,
! selectiveRetagging, cellCount, computeNormDistance; the sequence of the
! reference's unwired moving-body routines, subSearch.f90:41-512).
PROGRAM ib3d_main
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CHARACTER(LEN=32) :: arg
  INTEGER(KIND=8) :: cap
  cap = -1
  IF (COMMAND_ARGUMENT_COUNT() >= 1) THEN
    CALL GET_COMMAND_ARGUMENT(1, arg)
    READ(arg, *) cap
  END IF
  CALL readInput
  CALL readSurfaceMeshGmsh
  CALL allocateArrays
  CALL shiftSurfaceNodesInitial
  CALL computeSurfaceNorm
  CALL tagging
  CALL writeTagging
  CALL cellCount
  CALL computeNormDistance
  IF (istart == 0) CALL initialConditions
  IF (istart == 1) CALL lastConditions
  CALL coefficientMatrix
  CALL non_uni_coeff
  IF (cap >= 0) itamax = MIN(itamax, cap)
  ita = 1
  ita1 = 1
  totime = 0.
  totime = totime + deltat
  IF (a0 /= 0.0d0) THEN
    CALL shiftSurfaceNodesMovingbody
    CALL computeSurfaceNorm
    CALL selectiveRetagging
    CALL cellCount
    CALL computeNormDistance
  END IF
  CALL nsMomentum
  CALL velocityBC
  CALL solidCellBC
  CALL velocityForcing1
  CALL velocityBC
  CALL poissonSolver
  CALL pressureForcing1
  CALL stressCal1
  ! main.f90:41-67 tests ita < itamax after the step (a post-tested loop)
  DO
    ita = ita + 1
    ita1 = ita1 + 1
    totime = totime + deltat
    IF (a0 /= 0.0d0) THEN
      CALL shiftSurfaceNodesMovingbody
      CALL computeSurfaceNorm
      CALL selectiveRetagging
      CALL cellCount
      CALL computeNormDistance
    END IF
    CALL nsMomentum
    CALL velocityBC
    CALL solidCellBC
    CALL velocityForcing1
    CALL velocityBC
    CALL poissonSolver
    CALL pressureForcing1
    CALL writeOutput
    IF (a0 /= 0.0d0) THEN
      CALL stressCal1
    ELSE
      CALL stressCal2
    END IF
    CALL writeResult
    IF (ita >= itamax) EXIT
  END DO
  CALL ib3d_check(ib3d_finalize(), 'ib3d_finalize')
END PROGRAM ib3d_main
