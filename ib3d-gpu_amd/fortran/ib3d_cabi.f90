! ib3d_cabi.f90 -- iso_c_binding interfaces to include/ib3d.h.
MODULE ib3d_cabi
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE
  ! set once the device module holds the grid / the surface mesh (by the
  ! drop-in driver's readers, or by allocateArrays from the module arrays
  ! that the reference's own readers filled)
  LOGICAL, SAVE :: ib3d_grid_on_device = .FALSE., ib3d_mesh_on_device = .FALSE.
  INTERFACE
    INTEGER(c_int) FUNCTION ib3d_init(nx, ny, nz) BIND(C, name='ib3d_init')
      IMPORT :: c_int, c_int64_t
      INTEGER(c_int64_t), VALUE :: nx, ny, nz
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_finalize() BIND(C, name='ib3d_finalize')
      IMPORT :: c_int
    END FUNCTION
    TYPE(c_ptr) FUNCTION ib3d_last_error() BIND(C, name='ib3d_last_error')
      IMPORT :: c_ptr
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_sync() BIND(C, name='ib3d_sync')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_upload(name, host, nbytes) BIND(C, name='ib3d_upload')
      IMPORT :: c_int, c_char, c_ptr, c_int64_t
      CHARACTER(kind=c_char), DIMENSION(*) :: name
      TYPE(c_ptr), VALUE :: host
      INTEGER(c_int64_t), VALUE :: nbytes
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_download(name, host, nbytes) BIND(C, name='ib3d_download')
      IMPORT :: c_int, c_char, c_ptr, c_int64_t
      CHARACTER(kind=c_char), DIMENSION(*) :: name
      TYPE(c_ptr), VALUE :: host
      INTEGER(c_int64_t), VALUE :: nbytes
    END FUNCTION
    INTEGER(c_int64_t) FUNCTION ib3d_array_count(name) BIND(C, name='ib3d_array_count')
      IMPORT :: c_char, c_int64_t
      CHARACTER(kind=c_char), DIMENSION(*) :: name
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_set_grid(gx, gy, gz) BIND(C, name='ib3d_set_grid')
      IMPORT :: c_int, c_double
      REAL(c_double), DIMENSION(*) :: gx, gy, gz
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_set_mesh(nn, xn, yn, zn, ne, p1, p2, p3) BIND(C, name='ib3d_set_mesh')
      IMPORT :: c_int, c_int64_t, c_double
      INTEGER(c_int64_t), VALUE :: nn, ne
      REAL(c_double), DIMENSION(*) :: xn, yn, zn
      INTEGER(c_int64_t), DIMENSION(*) :: p1, p2, p3
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_shiftSurfaceNodesInitial(xs, ys, zs) BIND(C, name='ib3d_shiftSurfaceNodesInitial')
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: xs, ys, zs
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_shiftSurfaceNodesMovingbody(xs, ys, zs, dx, dy, dz) &
        BIND(C, name='ib3d_shiftSurfaceNodesMovingbody')
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: xs, ys, zs, dx, dy, dz
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_selectiveRetagging(counts) BIND(C, name='ib3d_selectiveRetagging')
      IMPORT :: c_int, c_int32_t
      INTEGER(c_int32_t), DIMENSION(3) :: counts
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_set_wall_motion(vel, acc) BIND(C, name='ib3d_set_wall_motion')
      IMPORT :: c_int, c_double
      REAL(c_double), DIMENSION(3) :: vel, acc
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_computeSurfaceNorm(inor) BIND(C, name='ib3d_computeSurfaceNorm')
      IMPORT :: c_int, c_int64_t
      INTEGER(c_int64_t), VALUE :: inor
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_tagging(window, counts) BIND(C, name='ib3d_tagging')
      IMPORT :: c_int, c_int64_t, c_int32_t
      INTEGER(c_int64_t), DIMENSION(6) :: window
      INTEGER(c_int32_t), DIMENSION(3) :: counts
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_cellCount(counts) BIND(C, name='ib3d_cellCount')
      IMPORT :: c_int, c_int32_t
      INTEGER(c_int32_t), DIMENSION(5) :: counts
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_computeNormDistance() BIND(C, name='ib3d_computeNormDistance')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_initialConditions() BIND(C, name='ib3d_initialConditions')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_coefficientMatrix() BIND(C, name='ib3d_coefficientMatrix')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_non_uni_coeff() BIND(C, name='ib3d_non_uni_coeff')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_nsMomentum(rev, deltat) BIND(C, name='ib3d_nsMomentum')
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: rev, deltat
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_velocityBC() BIND(C, name='ib3d_velocityBC')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_solidCellBC() BIND(C, name='ib3d_solidCellBC')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_velocityForcing1() BIND(C, name='ib3d_velocityForcing1')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_pressureForcing1() BIND(C, name='ib3d_pressureForcing1')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_computeDiv(deltat, divmax) BIND(C, name='ib3d_computeDiv')
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: deltat
      REAL(c_double) :: divmax
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_REDBLACKSOR(epsi, isum, derr, derr2) BIND(C, name='ib3d_REDBLACKSOR')
      IMPORT :: c_int, c_double, c_int64_t
      REAL(c_double), VALUE :: epsi
      INTEGER(c_int64_t) :: isum
      REAL(c_double) :: derr, derr2
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_correctPressure() BIND(C, name='ib3d_correctPressure')
      IMPORT :: c_int
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_correctVelocity(deltat) BIND(C, name='ib3d_correctVelocity')
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: deltat
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_poissonSolver(deltat, eps1, niter, divmax, derrstd, sors) &
        BIND(C, name='ib3d_poissonSolver')
      IMPORT :: c_int, c_double, c_int64_t
      REAL(c_double), VALUE :: deltat, eps1
      INTEGER(c_int64_t) :: niter
      REAL(c_double) :: divmax, derrstd, sors
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_stressCal1(re, coef) BIND(C, name='ib3d_stressCal1')
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: re
      REAL(c_double), DIMENSION(4) :: coef
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_stressCal2(re, coef) BIND(C, name='ib3d_stressCal2')
      IMPORT :: c_int, c_double
      REAL(c_double), VALUE :: re
      REAL(c_double), DIMENSION(4) :: coef
    END FUNCTION
    INTEGER(c_int) FUNCTION ib3d_set_poisson_mode(mode, omega, tol, check, warm, maxsw) &
        BIND(C, name='ib3d_set_poisson_mode')
      IMPORT :: c_int, c_double, c_int64_t
      INTEGER(c_int), VALUE :: mode, check, warm
      REAL(c_double), VALUE :: omega, tol
      INTEGER(c_int64_t), VALUE :: maxsw
    END FUNCTION
  END INTERFACE
CONTAINS
  ! Abort with the library's message on a non-zero status (the reference has
  ! no status codes; its failures are Fortran runtime aborts).
  SUBROUTINE ib3d_check(rc, where)
    INTEGER(c_int), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: where
    CHARACTER(kind=c_char), DIMENSION(:), POINTER :: msg
    CHARACTER(LEN=512) :: text
    INTEGER :: k
    IF (rc == 0) RETURN
    CALL c_f_pointer(ib3d_last_error(), msg, [512])
    text = ' '
    DO k = 1, 512
      IF (msg(k) == c_null_char) EXIT
      text(k:k) = msg(k)
    END DO
    WRITE(*, '(A,A,A,A)') 'ib3d: ', where, ': ', TRIM(text)
    ERROR STOP 1
  END SUBROUTINE ib3d_check
END MODULE ib3d_cabi
