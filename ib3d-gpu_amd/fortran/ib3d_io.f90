! ib3d_io.f90 -- the reference's input and output routines for the drop-in
! driver (bin/ib3d_main): the same files, record layouts and format strings
! (subReadInput.f90:2-176, subWriteOutput.f90:2-94).  Field outputs copy the
! device fields to the module at their own cadence instead of every step.
!
! A build that keeps the reference's own subReadInput.f90 and
! subWriteOutput.f90 leaves this file out: allocateArrays then hands the
! grid and the surface mesh the reference's readers left in the module to the
! device, and the host fields are refreshed with ib3d_update_host_fields.

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
  ib3d_grid_on_device = .TRUE.
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
  ib3d_mesh_on_device = .TRUE.
  PRINT *, 'SURFACE MESH READING COMPLETE'
  PRINT *, 'ibNodes =', ibNodes, 'ibElems =', ibElems
END SUBROUTINE readSurfaceMeshGmsh

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
