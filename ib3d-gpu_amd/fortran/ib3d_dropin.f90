! ib3d_dropin.f90 -- the reference's hot-path subroutines, argument-less and
! working on MODULE global, implemented on the MI355X through the C ABI.
!
! Each routine names the reference subroutine it replaces (file:line).  The
! numerical work happens in libib3d_hip on the device copy of the module;
! host arrays are refreshed where the reference itself updates the host
! (setup results, output cadence), never per step.

! ---------------------------------------------------------------- helpers
SUBROUTINE ib3d_update_host(name)
  ! replaces `!$acc update host (name)` for one module array
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CHARACTER(LEN=*), INTENT(IN) :: name
  TYPE(c_ptr) :: ptr
  INTEGER(c_int64_t) :: nb
  CALL array_address(name, ptr, nb)
  CALL ib3d_check(ib3d_download(TRIM(name)//c_null_char, ptr, nb), 'download '//TRIM(name))
END SUBROUTINE ib3d_update_host

SUBROUTINE ib3d_update_device(name)
  ! replaces `!$acc update device (name)` for one module array
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CHARACTER(LEN=*), INTENT(IN) :: name
  TYPE(c_ptr) :: ptr
  INTEGER(c_int64_t) :: nb
  CALL array_address(name, ptr, nb)
  CALL ib3d_check(ib3d_upload(TRIM(name)//c_null_char, ptr, nb), 'upload '//TRIM(name))
END SUBROUTINE ib3d_update_device

SUBROUTINE array_address(name, ptr, nb)
  USE global
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE
  CHARACTER(LEN=*), INTENT(IN) :: name
  TYPE(c_ptr), INTENT(OUT) :: ptr
  INTEGER(c_int64_t), INTENT(OUT) :: nb
  SELECT CASE (TRIM(name))
  CASE ('u'); ptr = c_loc(u); nb = 8_8 * SIZE(u, KIND=8)
  CASE ('v'); ptr = c_loc(v); nb = 8_8 * SIZE(v, KIND=8)
  CASE ('w'); ptr = c_loc(w); nb = 8_8 * SIZE(w, KIND=8)
  CASE ('ut'); ptr = c_loc(ut); nb = 8_8 * SIZE(ut, KIND=8)
  CASE ('vt'); ptr = c_loc(vt); nb = 8_8 * SIZE(vt, KIND=8)
  CASE ('wt'); ptr = c_loc(wt); nb = 8_8 * SIZE(wt, KIND=8)
  CASE ('p'); ptr = c_loc(p); nb = 8_8 * SIZE(p, KIND=8)
  CASE ('pc'); ptr = c_loc(pc); nb = 8_8 * SIZE(pc, KIND=8)
  CASE ('resi_u'); ptr = c_loc(resi_u); nb = 8_8 * SIZE(resi_u, KIND=8)
  CASE ('resi_v'); ptr = c_loc(resi_v); nb = 8_8 * SIZE(resi_v, KIND=8)
  CASE ('resi_w'); ptr = c_loc(resi_w); nb = 8_8 * SIZE(resi_w, KIND=8)
  CASE ('cell'); ptr = c_loc(cell); nb = 8_8 * SIZE(cell, KIND=8)
  CASE ('x1'); ptr = c_loc(x1); nb = 8_8 * SIZE(x1, KIND=8)
  CASE ('y1'); ptr = c_loc(y1); nb = 8_8 * SIZE(y1, KIND=8)
  CASE ('z1'); ptr = c_loc(z1); nb = 8_8 * SIZE(z1, KIND=8)
  CASE ('deltax'); ptr = c_loc(deltax); nb = 8_8 * SIZE(deltax, KIND=8)
  CASE ('deltay'); ptr = c_loc(deltay); nb = 8_8 * SIZE(deltay, KIND=8)
  CASE ('deltaz'); ptr = c_loc(deltaz); nb = 8_8 * SIZE(deltaz, KIND=8)
  CASE ('xu'); ptr = c_loc(xu); nb = 8_8 * SIZE(xu, KIND=8)
  CASE ('yu'); ptr = c_loc(yu); nb = 8_8 * SIZE(yu, KIND=8)
  CASE ('zu'); ptr = c_loc(zu); nb = 8_8 * SIZE(zu, KIND=8)
  CASE ('xv'); ptr = c_loc(xv); nb = 8_8 * SIZE(xv, KIND=8)
  CASE ('yv'); ptr = c_loc(yv); nb = 8_8 * SIZE(yv, KIND=8)
  CASE ('zv'); ptr = c_loc(zv); nb = 8_8 * SIZE(zv, KIND=8)
  CASE ('xw'); ptr = c_loc(xw); nb = 8_8 * SIZE(xw, KIND=8)
  CASE ('yw'); ptr = c_loc(yw); nb = 8_8 * SIZE(yw, KIND=8)
  CASE ('zw'); ptr = c_loc(zw); nb = 8_8 * SIZE(zw, KIND=8)
  CASE ('xp'); ptr = c_loc(xp); nb = 8_8 * SIZE(xp, KIND=8)
  CASE ('yp'); ptr = c_loc(yp); nb = 8_8 * SIZE(yp, KIND=8)
  CASE ('zp'); ptr = c_loc(zp); nb = 8_8 * SIZE(zp, KIND=8)
  CASE ('Acx'); ptr = c_loc(Acx); nb = 8_8 * SIZE(Acx, KIND=8)
  CASE ('Acy'); ptr = c_loc(Acy); nb = 8_8 * SIZE(Acy, KIND=8)
  CASE ('Acz'); ptr = c_loc(Acz); nb = 8_8 * SIZE(Acz, KIND=8)
  CASE ('interceptedIndexPtr'); ptr = c_loc(interceptedIndexPtr); nb = 8_8 * SIZE(interceptedIndexPtr, KIND=8)
  CASE ('fluidIndexPtr'); ptr = c_loc(fluidIndexPtr); nb = 8_8 * SIZE(fluidIndexPtr, KIND=8)
  CASE ('solidIndexPtr'); ptr = c_loc(solidIndexPtr); nb = 8_8 * SIZE(solidIndexPtr, KIND=8)
  CASE ('redCellIndexPtr'); ptr = c_loc(redCellIndexPtr); nb = 8_8 * SIZE(redCellIndexPtr, KIND=8)
  CASE ('blackCellIndexPtr'); ptr = c_loc(blackCellIndexPtr); nb = 8_8 * SIZE(blackCellIndexPtr, KIND=8)
  CASE ('xnode1'); ptr = c_loc(xnode1); nb = 8_8 * SIZE(xnode1, KIND=8)
  CASE ('ynode1'); ptr = c_loc(ynode1); nb = 8_8 * SIZE(ynode1, KIND=8)
  CASE ('znode1'); ptr = c_loc(znode1); nb = 8_8 * SIZE(znode1, KIND=8)
  CASE ('xcent'); ptr = c_loc(xcent); nb = 8_8 * SIZE(xcent, KIND=8)
  CASE ('ycent'); ptr = c_loc(ycent); nb = 8_8 * SIZE(ycent, KIND=8)
  CASE ('zcent'); ptr = c_loc(zcent); nb = 8_8 * SIZE(zcent, KIND=8)
  CASE ('cosAlpha'); ptr = c_loc(cosAlpha); nb = 8_8 * SIZE(cosAlpha, KIND=8)
  CASE ('cosBeta'); ptr = c_loc(cosBeta); nb = 8_8 * SIZE(cosBeta, KIND=8)
  CASE ('cosGamma'); ptr = c_loc(cosGamma); nb = 8_8 * SIZE(cosGamma, KIND=8)
  CASE ('nelp'); ptr = c_loc(nelp); nb = 8_8 * SIZE(nelp, KIND=8)
  CASE ('pNormDis'); ptr = c_loc(pNormDis); nb = 8_8 * SIZE(pNormDis, KIND=8)
  CASE ('u1NormDis'); ptr = c_loc(u1NormDis); nb = 8_8 * SIZE(u1NormDis, KIND=8)
  CASE ('u2NormDis'); ptr = c_loc(u2NormDis); nb = 8_8 * SIZE(u2NormDis, KIND=8)
  CASE ('v1NormDis'); ptr = c_loc(v1NormDis); nb = 8_8 * SIZE(v1NormDis, KIND=8)
  CASE ('v2NormDis'); ptr = c_loc(v2NormDis); nb = 8_8 * SIZE(v2NormDis, KIND=8)
  CASE ('w1NormDis'); ptr = c_loc(w1NormDis); nb = 8_8 * SIZE(w1NormDis, KIND=8)
  CASE ('w2NormDis'); ptr = c_loc(w2NormDis); nb = 8_8 * SIZE(w2NormDis, KIND=8)
  CASE DEFAULT
    WRITE(*, '(A,A)') 'ib3d: no host copy of array ', TRIM(name)
    ERROR STOP 1
  END SELECT
END SUBROUTINE array_address

! ---------------------------------------------------------------- input
! subReadInput.f90:2-133: 27 list-directed values, grid nodes, then the
! staggered coordinates.  The coordinates are built by the library from the
! node values (the same expressions) and copied back to the module.
SUBROUTINE readInput
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  REAL(KIND=8), ALLOCATABLE :: gx(:), gy(:), gz(:)
  INTEGER(KIND=8) :: i
  INTEGER :: unit
  OPEN(NEWUNIT=unit, FILE='inputdata', FORM='formatted', STATUS='old')
  READ(unit, *) nx, ny, nz, re, itamax, eps1, pcItaMax, u0, v0, w0, surGeoPoints, a0, freq, aoa, piv_pt, &
                xShift, yShift, zShift, istart, dt_order, inor, i_startSearch, i_endSearch, &
                j_startSearch, j_endSearch, k_startSearch, k_endSearch
  CLOSE(unit)
  rev = 1.0d0 / re
  pi = 4.0d0 * ATAN(1.0d0)
  deltat = dt_order * 1.0d-3
  ita = 0
  ita1 = 0
  totime = 0.0d0
  PRINT *, 'dt =', deltat, 'ita = ', ita, 'totime = ', totime
  PRINT *, 'total cells =', nx*ny*nz
  ALLOCATE(gx(nx+1), gy(ny+1), gz(nz+1))
  OPEN(NEWUNIT=unit, FILE='gridx.dat', FORM='formatted', STATUS='old')
  DO i = 1, nx+1
    READ(unit, *) gx(i)
  END DO
  CLOSE(unit)
  OPEN(NEWUNIT=unit, FILE='gridy.dat', FORM='formatted', STATUS='old')
  DO i = 1, ny+1
    READ(unit, *) gy(i)
  END DO
  CLOSE(unit)
  OPEN(NEWUNIT=unit, FILE='gridz.dat', FORM='formatted', STATUS='old')
  DO i = 1, nz+1
    READ(unit, *) gz(i)
  END DO
  CLOSE(unit)
  ALLOCATE(x1(nx+3), y1(ny+3), z1(nz+3), deltax(nx+2), deltay(ny+2), deltaz(nz+2))
  ALLOCATE(xu(nx+3), yu(ny+2), zu(nz+2), xv(nx+2), yv(ny+3), zv(nz+2), xw(nx+2), yw(ny+2), zw(nz+3), &
           xp(nx+2), yp(ny+2), zp(nz+2))
  CALL ib3d_check(ib3d_init(nx, ny, nz), 'ib3d_init')
  CALL ib3d_check(ib3d_set_grid(gx, gy, gz), 'set_grid')
  CALL ib3d_update_host('x1'); CALL ib3d_update_host('y1'); CALL ib3d_update_host('z1')
  CALL ib3d_update_host('deltax'); CALL ib3d_update_host('deltay'); CALL ib3d_update_host('deltaz')
  CALL ib3d_update_host('xu'); CALL ib3d_update_host('yu'); CALL ib3d_update_host('zu')
  CALL ib3d_update_host('xv'); CALL ib3d_update_host('yv'); CALL ib3d_update_host('zv')
  CALL ib3d_update_host('xw'); CALL ib3d_update_host('yw'); CALL ib3d_update_host('zw')
  CALL ib3d_update_host('xp'); CALL ib3d_update_host('yp'); CALL ib3d_update_host('zp')
END SUBROUTINE readInput

! subReadInput.f90:137-176: Gmsh 2.2 text with the reference's line accounting
SUBROUTINE readSurfaceMeshGmsh
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CHARACTER(LEN=128) :: word
  INTEGER(KIND=8) :: n, i1, i2, i3, i4, i5
  INTEGER :: unit
  OPEN(NEWUNIT=unit, FILE='geometries/sphere.msh', FORM='formatted', STATUS='old')
  DO n = 1, 4
    READ(unit, *) word
  END DO
  READ(unit, *) ibNodes
  ALLOCATE(xnode(ibNodes), ynode(ibNodes), znode(ibNodes))
  DO n = 1, ibNodes
    READ(unit, *) i1, xnode(n), ynode(n), znode(n)
  END DO
  DO n = 1, 2
    READ(unit, *) word
  END DO
  READ(unit, *) ibElems
  DO n = 1, surGeoPoints
    READ(unit, *) word
  END DO
  ibElems = ibElems - surGeoPoints
  ALLOCATE(ibElP1(ibElems), ibElP2(ibElems), ibElP3(ibElems))
  DO n = 1, ibElems
    READ(unit, *) i1, i2, i3, i4, i5, ibElP1(n), ibElP2(n), ibElP3(n)
  END DO
  CLOSE(unit)
  CALL ib3d_check(ib3d_set_mesh(ibNodes, xnode, ynode, znode, ibElems, ibElP1, ibElP2, ibElP3), 'set_mesh')
  PRINT *, 'SURFACE MESH READING COMPLETE'
  PRINT *, 'ibNodes =', ibNodes, 'ibElems =', ibElems
END SUBROUTINE readSurfaceMeshGmsh

! subAllocateArrays.f90:1-15 (the device copy was allocated by ib3d_init)
SUBROUTINE allocateArrays
  USE global
  IMPLICIT NONE
  ALLOCATE(u(nx+2,ny+2,nz+2), ut(nx+2,ny+2,nz+2), v(nx+2,ny+2,nz+2), vt(nx+2,ny+2,nz+2), &
           w(nx+2,ny+2,nz+2), wt(nx+2,ny+2,nz+2), p(nx+2,ny+2,nz+2), pc(nx+2,ny+2,nz+2), &
           resi_u(nx+2,ny+2,nz+2), resi_v(nx+2,ny+2,nz+2), resi_w(nx+2,ny+2,nz+2))
  ALLOCATE(cell(nx+2,ny+2,nz+2))
  ALLOCATE(Acx(nx,3), Acy(ny,3), Acz(nz,3))
END SUBROUTINE allocateArrays

! ---------------------------------------------------------------- geometry
SUBROUTINE shiftSurfaceNodesInitial          ! subSearch.f90:2-39
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  ALLOCATE(xnode1(ibNodes), ynode1(ibNodes), znode1(ibNodes))
  CALL ib3d_check(ib3d_shiftSurfaceNodesInitial(xShift, yShift, zShift), 'shiftSurfaceNodesInitial')
  CALL ib3d_update_host('xnode1'); CALL ib3d_update_host('ynode1'); CALL ib3d_update_host('znode1')
END SUBROUTINE shiftSurfaceNodesInitial

! subSearch.f90:41-76 with the free motion replaced by the prescribed
! transverse oscillation y_c(t) = a0 sin(2 pi freq t) (a0, freq from inputdata);
! the body's velocity and acceleration (finite difference of the velocity,
! the form of subForcing.f90:20) move the wall in the forcing and stress
SUBROUTINE shiftSurfaceNodesMovingbody
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  REAL(c_double) :: vel(3), acc(3)
  v_init = v_final
  xmove = 0.0d0
  ymove = a0 * SIN(2.0d0 * pi * freq * totime)
  zmove = 0.0d0
  v_final = a0 * (2.0d0 * pi * freq) * COS(2.0d0 * pi * freq * totime)
  vel = [0.0d0, v_final, 0.0d0]
  acc = [0.0d0, (v_final - v_init) / deltat, 0.0d0]
  CALL ib3d_check(ib3d_set_wall_motion(vel, acc), 'set_wall_motion')
  CALL ib3d_check(ib3d_shiftSurfaceNodesMovingbody(xShift, yShift, zShift, xmove, ymove, zmove), &
                  'shiftSurfaceNodesMovingbody')
END SUBROUTINE shiftSurfaceNodesMovingbody

SUBROUTINE selectiveRetagging                ! subSearch.f90:327-512
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  INTEGER(c_int32_t) :: counts(3)
  CALL ib3d_check(ib3d_selectiveRetagging(counts), 'selectiveRetagging')
  ibCellCount = counts(1)
  fluidCellCount = counts(2)
  solidCellCount = counts(3)
END SUBROUTINE selectiveRetagging

SUBROUTINE computeSurfaceNorm                ! subSearch.f90:80-131
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  IF (.NOT. ALLOCATED(xcent)) &
    ALLOCATE(xcent(ibElems), ycent(ibElems), zcent(ibElems), cosAlpha(ibElems), cosBeta(ibElems), cosGamma(ibElems))
  CALL ib3d_check(ib3d_computeSurfaceNorm(inor), 'computeSurfaceNorm')
  CALL ib3d_update_host('xcent'); CALL ib3d_update_host('ycent'); CALL ib3d_update_host('zcent')
  CALL ib3d_update_host('cosAlpha'); CALL ib3d_update_host('cosBeta'); CALL ib3d_update_host('cosGamma')
  PRINT *, 'SurfaceNorm done, inor =', inor
END SUBROUTINE computeSurfaceNorm

SUBROUTINE tagging                           ! subSearch.f90:134-321
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  INTEGER(c_int64_t) :: win(6)
  INTEGER(c_int32_t) :: counts(3)
  win = [i_startSearch, i_endSearch, j_startSearch, j_endSearch, k_startSearch, k_endSearch]
  CALL ib3d_check(ib3d_tagging(win, counts), 'tagging')
  ibCellCount = counts(1)
  fluidCellCount = counts(2)
  solidCellCount = counts(3)
  CALL ib3d_update_host('cell')
  PRINT *, 'search done'
  PRINT *, 'imms. cells=', ibCellCount
  PRINT *, 'fluid cells=', fluidCellCount
  PRINT *, 'solid cells=', solidCellCount
END SUBROUTINE tagging

SUBROUTINE cellCount                         ! subSearch.f90:515-596
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  INTEGER(c_int32_t) :: counts(5)
  CALL ib3d_check(ib3d_cellCount(counts), 'cellCount')
  ibCellCount = counts(1)
  fluidCellCount = counts(2)
  solidCellCount = counts(3)
  redCellCount = counts(4)
  blackCellCount = counts(5)
  IF (ALLOCATED(interceptedIndexPtr)) DEALLOCATE(interceptedIndexPtr, fluidIndexPtr, solidIndexPtr, &
                                                 redCellIndexPtr, blackCellIndexPtr)
  ALLOCATE(interceptedIndexPtr(ibCellCount,3), fluidIndexPtr(fluidCellCount,3), solidIndexPtr(solidCellCount,3), &
           redCellIndexPtr(redCellCount,3), blackCellIndexPtr(blackCellCount,3))
  IF (ibCellCount > 0) CALL ib3d_update_host('interceptedIndexPtr')
  IF (fluidCellCount > 0) CALL ib3d_update_host('fluidIndexPtr')
  IF (solidCellCount > 0) CALL ib3d_update_host('solidIndexPtr')
  IF (redCellCount > 0) CALL ib3d_update_host('redCellIndexPtr')
  IF (blackCellCount > 0) CALL ib3d_update_host('blackCellIndexPtr')
  PRINT *, fluidCellCount, redCellCount, blackCellCount
  PRINT *, 'cellCount done'
END SUBROUTINE cellCount

SUBROUTINE computeNormDistance               ! subSearch.f90:599-678
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  IF (ALLOCATED(nelp)) DEALLOCATE(nelp, pNormDis, u1NormDis, u2NormDis, v1NormDis, v2NormDis, w1NormDis, w2NormDis)
  ALLOCATE(nelp(ibCellCount), pNormDis(ibCellCount), u1NormDis(ibCellCount), u2NormDis(ibCellCount), &
           v1NormDis(ibCellCount), v2NormDis(ibCellCount), w1NormDis(ibCellCount), w2NormDis(ibCellCount))
  CALL ib3d_check(ib3d_computeNormDistance(), 'computeNormDistance')
  IF (ibCellCount > 0) THEN
    CALL ib3d_update_host('nelp'); CALL ib3d_update_host('pNormDis')
    CALL ib3d_update_host('u1NormDis'); CALL ib3d_update_host('u2NormDis')
    CALL ib3d_update_host('v1NormDis'); CALL ib3d_update_host('v2NormDis')
    CALL ib3d_update_host('w1NormDis'); CALL ib3d_update_host('w2NormDis')
  END IF
  PRINT *, 'computeNormDistance done'
END SUBROUTINE computeNormDistance

! ---------------------------------------------------------------- setup
SUBROUTINE initialConditions                 ! subInitialConditions.f90:2-48
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_initialConditions(), 'initialConditions')
  ita = 0
  ita1 = 0
  totime = 0.0d0
  PRINT *, 'initial'
END SUBROUTINE initialConditions

SUBROUTINE coefficientMatrix                 ! subCoefficientMatrix.f90:1-59
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_coefficientMatrix(), 'coefficientMatrix')
  CALL ib3d_update_host('Acx'); CALL ib3d_update_host('Acy'); CALL ib3d_update_host('Acz')
  PRINT *, 'Coefficient Matrix generated'
END SUBROUTINE coefficientMatrix

SUBROUTINE non_uni_coeff                     ! subNavierStokes.f90:3-619 (tables stay on the device)
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_non_uni_coeff(), 'non_uni_coeff')
END SUBROUTINE non_uni_coeff

! ---------------------------------------------------------------- per step
SUBROUTINE nsMomentum                        ! subNavierStokes.f90:622-1122
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_nsMomentum(rev, deltat), 'nsMomentum')
END SUBROUTINE nsMomentum

SUBROUTINE velocityBC                        ! subBoundaryConditions.f90:2-195
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_velocityBC(), 'velocityBC')
END SUBROUTINE velocityBC

SUBROUTINE solidCellBC                       ! subBoundaryConditions.f90:199-217
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_solidCellBC(), 'solidCellBC')
END SUBROUTINE solidCellBC

SUBROUTINE velocityForcing1                  ! subForcing.f90:107-652
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_velocityForcing1(), 'velocityForcing1')
END SUBROUTINE velocityForcing1

SUBROUTINE pressureForcing1                  ! subForcing.f90:2-103
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_pressureForcing1(), 'pressureForcing1')
END SUBROUTINE pressureForcing1

SUBROUTINE computeDiv                        ! subPcorVcor.f90:82-101
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_computeDiv(deltat, divmax), 'computeDiv')
END SUBROUTINE computeDiv

SUBROUTINE REDBLACKSOR(epsi, isum, derr, derr2)   ! subPcorVcor.f90:148-212
  USE ib3d_cabi
  IMPLICIT NONE
  REAL(KIND=8), INTENT(IN) :: epsi
  INTEGER(KIND=8), INTENT(OUT) :: isum
  REAL(KIND=8), INTENT(OUT) :: derr, derr2
  CALL ib3d_check(ib3d_REDBLACKSOR(epsi, isum, derr, derr2), 'REDBLACKSOR')
END SUBROUTINE REDBLACKSOR

SUBROUTINE correctPressure                   ! subPcorVcor.f90:105-120
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_correctPressure(), 'correctPressure')
END SUBROUTINE correctPressure

SUBROUTINE correctVelocity                   ! subPcorVcor.f90:125-142
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_correctVelocity(deltat), 'correctVelocity')
END SUBROUTINE correctVelocity

! subPcorVcor.f90:2-80, including the iter.dat line of 60-65
SUBROUTINE poissonSolver
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  REAL(KIND=8) :: derrStdSt, sors
  INTEGER :: unit
  CALL ib3d_check(ib3d_poissonSolver(deltat, eps1, nIterPcor, divmax, derrStdSt, sors), 'poissonSolver')
  dstart = 0.0d0
  dfinish = sors
  OPEN(NEWUNIT=unit, FILE='iter.dat', ACCESS='append', STATUS='unknown')
  WRITE(unit, 16) ita, nIterPcor, divmax, derrStdSt, dfinish - dstart
  CLOSE(unit)
  WRITE(*, 16) ita, nIterPcor, divmax, derrStdSt, dfinish - dstart
16 FORMAT(' ', I8, I10, 4E15.6)
END SUBROUTINE poissonSolver

! subStressCal1.f90:2-460 / subStressCal2.f90:2-381 with the files of
! subStressCal2.f90:374-379
SUBROUTINE stress_write(coef)
  USE global
  IMPLICIT NONE
  REAL(KIND=8), INTENT(IN) :: coef(4)
  INTEGER :: unit
  OPEN(NEWUNIT=unit, FILE='dragcoff.dat', ACCESS='append', STATUS='unknown')
  WRITE(unit, *) coef(1), coef(2), totime
  CLOSE(unit)
  OPEN(NEWUNIT=unit, FILE='liftcoff.dat', ACCESS='append', STATUS='unknown')
  WRITE(unit, *) coef(3), coef(4), totime
  CLOSE(unit)
END SUBROUTINE stress_write

SUBROUTINE stressCal1
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  REAL(c_double) :: coef(4)
  CALL ib3d_check(ib3d_stressCal1(re, coef), 'stressCal1')
  CALL stress_write(coef)
END SUBROUTINE stressCal1

SUBROUTINE stressCal2
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  REAL(c_double) :: coef(4)
  CALL ib3d_check(ib3d_stressCal2(re, coef), 'stressCal2')
  CALL stress_write(coef)
END SUBROUTINE stressCal2

! ---------------------------------------------------------------- output
! subWriteOutput.f90:2-26: field dump every 5000 steps (needs out/)
SUBROUTINE writeOutput
  USE global
  IMPLICIT NONE
  CHARACTER(LEN=70) :: fname
  INTEGER(KIND=8) :: i, j, k
  INTEGER :: unit
  IF (MOD(ita, 5000_8) /= 0) RETURN
  CALL ib3d_update_host('u'); CALL ib3d_update_host('v'); CALL ib3d_update_host('w')
  CALL ib3d_update_host('p'); CALL ib3d_update_host('cell')
  WRITE(fname, '(A,I9.9,A)') 'out/fielddata.', ita, '.dat'
  OPEN(NEWUNIT=unit, FILE=fname, STATUS='unknown')
  WRITE(unit, *) 'variables= "x","y","z","u","v","w","p","totime","cellid"'
  WRITE(unit, *) 'zone, ', 'i = ', nx, ' j = ', ny, ' k = ', nz
  DO k = 2, nz+1
    DO j = 2, ny+1
      DO i = 2, nx+1
        WRITE(unit, *) xp(i), yp(j), zp(k), 0.5d0*(u(i,j,k)+u(i-1,j,k)), 0.5d0*(v(i,j,k)+v(i,j-1,k)), &
                       0.5d0*(w(i,j,k)+w(i,j,k-1)), p(i,j,k), totime, cell(i,j,k)
      END DO
    END DO
  END DO
  CLOSE(unit)
END SUBROUTINE writeOutput

! subWriteOutput.f90:58-72: restart file every 1000 steps
SUBROUTINE writeResult
  USE global
  IMPLICIT NONE
  INTEGER(KIND=8) :: i, j, k
  INTEGER :: unit
  IF (MOD(ita, 1000_8) /= 0) RETURN
  CALL ib3d_update_host('u'); CALL ib3d_update_host('v'); CALL ib3d_update_host('w'); CALL ib3d_update_host('p')
  OPEN(NEWUNIT=unit, FILE='result', FORM='formatted', STATUS='unknown')
  DO k = 1, nz+2
    DO j = 1, ny+2
      DO i = 1, nx+2
        WRITE(unit, *) u(i,j,k), v(i,j,k), w(i,j,k), p(i,j,k), totime, ita, ita1
      END DO
    END DO
  END DO
  CLOSE(unit)
END SUBROUTINE writeResult

! subLastConditions.f90:2-42: restart from `result`, ut = u
SUBROUTINE lastConditions
  USE global
  IMPLICIT NONE
  INTEGER(KIND=8) :: i, j, k
  INTEGER :: unit
  OPEN(NEWUNIT=unit, FILE='result', FORM='formatted', STATUS='old')
  DO k = 1, nz+2
    DO j = 1, ny+2
      DO i = 1, nx+2
        READ(unit, *) u(i,j,k), v(i,j,k), w(i,j,k), p(i,j,k), totime, ita, ita1
      END DO
    END DO
  END DO
  CLOSE(unit)
  ut = u
  vt = v
  wt = w
  resi_u = 0.0d0
  resi_v = 0.0d0
  resi_w = 0.0d0
  CALL ib3d_update_device('u'); CALL ib3d_update_device('v'); CALL ib3d_update_device('w')
  CALL ib3d_update_device('ut'); CALL ib3d_update_device('vt'); CALL ib3d_update_device('wt')
  CALL ib3d_update_device('p')
  CALL ib3d_update_device('resi_u'); CALL ib3d_update_device('resi_v'); CALL ib3d_update_device('resi_w')
END SUBROUTINE lastConditions

! subWriteOutput.f90:76-94: cell tags after tagging (needs tag/)
SUBROUTINE writeTagging
  USE global
  IMPLICIT NONE
  CHARACTER(LEN=70) :: fname
  INTEGER(KIND=8) :: i, j, k
  INTEGER :: unit
  WRITE(fname, '(A,I2.2,A)') 'tag/tagdata.', ita, '.dat'
  OPEN(NEWUNIT=unit, FILE=fname, STATUS='unknown')
  WRITE(unit, *) 'variables= "x","y","z","cellid"'
  WRITE(unit, *) 'zone, ', 'i = ', nx, ' j = ', ny, ' k = ', nz
  DO k = 2, nz+1
    DO j = 2, ny+1
      DO i = 2, nx+1
        WRITE(unit, *) xp(i), yp(j), zp(k), cell(i,j,k)
      END DO
    END DO
  END DO
  CLOSE(unit)
END SUBROUTINE writeTagging
