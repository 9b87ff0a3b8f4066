! ib3d_global.f90 -- host copy of the module the hot-path subroutines share.
!
! The reference keeps all state in MODULE global (modGlobal.f90:1-106); its
! subroutines take no arguments.  This module carries the same names and
! kinds for everything the hot path reads or writes, so code written against
! the reference's module keeps compiling.  The device copy (inside
! libib3d_hip) is authoritative during a run; ib3d_update_host(name) and
! ib3d_update_device(name) move one array between the two, which replaces the
! reference's `!$acc update host/device` statements (e.g. main.f90:165).
! Arrays are TARGET so their addresses can be handed to the C ABI.
MODULE global
  IMPLICIT NONE
  ! counts (default INTEGER) and sizes / loop scalars (INTEGER(8)), as in
  ! modGlobal.f90:4-10
  INTEGER :: istart, ibCellCount, fluidCellCount, solidCellCount, redCellCount, blackCellCount
  INTEGER(KIND=8) :: nx, ny, nz, itamax, pcItaMax, ita, nIterPcor, ita1, inor, &
                     i_startSearch, i_endSearch, j_startSearch, j_endSearch, &
                     k_startSearch, k_endSearch, surGeoPoints, ibElems, ibNodes
  REAL(KIND=8) :: re, rev, eps1, dt_order, deltat, totime, divmax, dstart, dfinish, pi, &
                  u0, v0, w0, a0, freq, aoa, piv_pt, xShift, yShift, zShift
  ! body motion (modGlobal.f90:23-24)
  REAL(KIND=8) :: u_init = 0.0d0, u_final = 0.0d0, v_init = 0.0d0, v_final = 0.0d0, &
                  w_init = 0.0d0, w_final = 0.0d0, xmove = 0.0d0, ymove = 0.0d0, zmove = 0.0d0
  ! grid (modGlobal.f90:29-30; shapes of subReadInput.f90:48-53)
  REAL(KIND=8), ALLOCATABLE, TARGET :: x1(:), y1(:), z1(:), deltax(:), deltay(:), deltaz(:), &
                                       xu(:), yu(:), zu(:), xv(:), yv(:), zv(:), &
                                       xw(:), yw(:), zw(:), xp(:), yp(:), zp(:)
  ! fields (subAllocateArrays.f90:4-12)
  REAL(KIND=8), ALLOCATABLE, TARGET :: u(:,:,:), v(:,:,:), w(:,:,:), ut(:,:,:), vt(:,:,:), wt(:,:,:), &
                                       p(:,:,:), pc(:,:,:), resi_u(:,:,:), resi_v(:,:,:), resi_w(:,:,:)
  INTEGER(KIND=8), ALLOCATABLE, TARGET :: cell(:,:,:)
  REAL(KIND=8), ALLOCATABLE, TARGET :: Acx(:,:), Acy(:,:), Acz(:,:)
  ! lists (count,3) (subSearch.f90:527-528, 568)
  INTEGER(KIND=8), ALLOCATABLE, TARGET :: interceptedIndexPtr(:,:), fluidIndexPtr(:,:), solidIndexPtr(:,:), &
                                          redCellIndexPtr(:,:), blackCellIndexPtr(:,:)
  ! surface mesh (subReadInput.f90:148, 160; subSearch.f90:8, 87-89)
  REAL(KIND=8), ALLOCATABLE, TARGET :: xnode(:), ynode(:), znode(:), xnode1(:), ynode1(:), znode1(:), &
                                       xcent(:), ycent(:), zcent(:), cosAlpha(:), cosBeta(:), cosGamma(:)
  INTEGER(KIND=8), ALLOCATABLE, TARGET :: ibElP1(:), ibElP2(:), ibElP3(:)
  ! intercepted-cell data (subSearch.f90:608-612)
  INTEGER(KIND=8), ALLOCATABLE, TARGET :: nelp(:)
  REAL(KIND=8), ALLOCATABLE, TARGET :: pNormDis(:), u1NormDis(:), u2NormDis(:), v1NormDis(:), v2NormDis(:), &
                                       w1NormDis(:), w2NormDis(:)
END MODULE global
