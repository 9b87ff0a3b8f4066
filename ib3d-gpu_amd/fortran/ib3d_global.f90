! ib3d_global.f90 -- host copy of the module the hot-path subroutines share.
!
! The reference keeps all state in MODULE global (modGlobal.f90:1-106) and its
! subroutines take no arguments.  This module declares every name of that
! module with its kind and rank, so every reference file that is not replaced
! (main.f90, subReadInput.f90, subWriteOutput.f90, subCalculateuc.f90) still
! compiles against it.  The device copy inside libib3d_hip is authoritative
! during a run; ib3d_update_host(name) and ib3d_update_device(name) move one
! array between the two and replace the reference's `!$acc update host/device`
! statements.  Arrays are TARGET so their addresses can be handed to the C
! ABI; the OpenACC `declare create` lines (modGlobal.f90:67-105) have no
! counterpart because the library owns the device copy.
MODULE global
  IMPLICIT NONE
  ! modGlobal.f90:3-24
  CHARACTER(LEN=128) :: line
  INTEGER :: istart, id1, ibCellCount, fluidCellCount, solidCellCount, redCellCount, blackCellCount
  INTEGER(KIND=8) :: nx, ny, nz, itamax, pcItaMax, ita, nIterPcor, nc, ita1, ital, sumIterPc, itaSola, &
                     totIterPc, inor, k_startSearch, k_endSearch, j_startSearch, j_endSearch, &
                     i_startSearch, i_endSearch
  REAL(KIND=8) :: lx, ly, dt_order, deltx2, delty2, deltz2, a0, freq, u0, v0, w0, p0, eps1, epsDiv, re, rev, &
                  divmax, fx, fy, alpha, pct, pct1, deltat, totime, dfinish, dstart, solverTime, pi, msTime, &
                  mindx, al, uc
  REAL(KIND=8) :: y11, y12, z11, z12, y21, y22, z21, z22, uc11, uc22
  REAL(KIND=8) :: u_init = 0.0d0, u_final = 0.0d0, v_init = 0.0d0, v_final = 0.0d0, w_init = 0.0d0, &
                  w_final = 0.0d0, xmove = 0.0d0, ymove = 0.0d0, zmove = 0.0d0, Total_Force_X = 0.0d0, &
                  Total_Force_Y = 0.0d0, Total_Force_Z = 0.0d0
  ! modGlobal.f90:25-28: tags, nearest elements, stress stencil cells
  INTEGER(KIND=8), ALLOCATABLE, TARGET :: cell(:,:,:), minElemcell(:,:,:)
  INTEGER(KIND=8), ALLOCATABLE, TARGET :: Elemcell(:,:), ucell(:,:), vcell(:,:), wcell(:,:), pcell(:,:)
  REAL(KIND=8), ALLOCATABLE, TARGET :: SUMWSS(:), SIGNWSS(:), areaElem(:)
  ! modGlobal.f90:29-49: grid (shapes of subReadInput.f90:48-53) and the
  ! convection tables of non_uni_coeff (n+2 entries each)
  REAL(KIND=8), ALLOCATABLE, TARGET :: deltax(:), deltay(:), deltaz(:), x1(:), y1(:), z1(:), &
                                       xu(:), yu(:), zu(:), xv(:), yv(:), zv(:), xw(:), yw(:), zw(:), &
                                       xp(:), yp(:), zp(:)
  REAL(KIND=8), ALLOCATABLE, TARGET :: &
    ca1_uu(:), ca2_uu(:), ca3_uu(:), ca4_uu(:), ca5_uu(:), ca6_uu(:), &
    ck1_uu(:), ck2_uu(:), ck3_uu(:), ck4_uu(:), ck5_uu(:), ck6_uu(:), &
    ca1_vv(:), ca2_vv(:), ca3_vv(:), ca4_vv(:), ca5_vv(:), ca6_vv(:), &
    ck1_vv(:), ck2_vv(:), ck3_vv(:), ck4_vv(:), ck5_vv(:), ck6_vv(:), &
    ca1_ww(:), ca2_ww(:), ca3_ww(:), ca4_ww(:), ca5_ww(:), ca6_ww(:), &
    ck1_ww(:), ck2_ww(:), ck3_ww(:), ck4_ww(:), ck5_ww(:), ck6_ww(:), &
    ca1_uv(:), ca2_uv(:), ca3_uv(:), ca4_uv(:), ca5_uv(:), ca6_uv(:), &
    ck1_uv(:), ck2_uv(:), ck3_uv(:), ck4_uv(:), ck5_uv(:), ck6_uv(:), &
    ca1_uw(:), ca2_uw(:), ca3_uw(:), ca4_uw(:), ca5_uw(:), ca6_uw(:), &
    ck1_uw(:), ck2_uw(:), ck3_uw(:), ck4_uw(:), ck5_uw(:), ck6_uw(:), &
    ca1_vu(:), ca2_vu(:), ca3_vu(:), ca4_vu(:), ca5_vu(:), ca6_vu(:), &
    ck1_vu(:), ck2_vu(:), ck3_vu(:), ck4_vu(:), ck5_vu(:), ck6_vu(:), &
    ca1_vw(:), ca2_vw(:), ca3_vw(:), ca4_vw(:), ca5_vw(:), ca6_vw(:), &
    ck1_vw(:), ck2_vw(:), ck3_vw(:), ck4_vw(:), ck5_vw(:), ck6_vw(:), &
    ca1_wu(:), ca2_wu(:), ca3_wu(:), ca4_wu(:), ca5_wu(:), ca6_wu(:), &
    ck1_wu(:), ck2_wu(:), ck3_wu(:), ck4_wu(:), ck5_wu(:), ck6_wu(:), &
    ca1_wv(:), ca2_wv(:), ca3_wv(:), ca4_wv(:), ca5_wv(:), ca6_wv(:), &
    ck1_wv(:), ck2_wv(:), ck3_wv(:), ck4_wv(:), ck5_wv(:), ck6_wv(:)
  ! modGlobal.f90:50-55: Laplacian rows and fields (subAllocateArrays.f90:4-12)
  REAL(KIND=8), ALLOCATABLE, TARGET :: Ac(:,:), Acx(:,:), Acy(:,:), Acz(:,:)
  REAL(KIND=8), ALLOCATABLE, TARGET :: b(:,:,:), u(:,:,:), ut(:,:,:), u_sum(:,:,:), u_avg(:,:,:), &
                                       v(:,:,:), vt(:,:,:), v_sum(:,:,:), v_avg(:,:,:), &
                                       w(:,:,:), wt(:,:,:), w_sum(:,:,:), w_avg(:,:,:), &
                                       p(:,:,:), p_sum(:,:,:), p_avg(:,:,:), pc(:,:,:), pco(:,:,:), &
                                       resi_u(:,:,:), resi_v(:,:,:), resi_w(:,:,:)
  ! modGlobal.f90:56-57: cell lists (count,3) (subSearch.f90:527-528, 568)
  INTEGER(KIND=8), ALLOCATABLE, TARGET :: interceptedIndexPtr(:,:), fluidIndexPtr(:,:), solidIndexPtr(:,:), &
                                          redCellIndexPtr(:,:), blackCellIndexPtr(:,:), nodeId(:,:), &
                                          fluidInterceptedIndexPtr(:,:)
  ! modGlobal.f90:58-65: surface mesh and intercepted-cell data
  INTEGER :: bcType
  REAL(KIND=8) :: xShift, yShift, zShift, aoa, piv_pt, var_surf, thetaDot, thetaDDot, piv_x, piv_y, piv_z, aoa1
  INTEGER(KIND=8) :: surGeoPoints, ibElems, ibNodes
  INTEGER(KIND=8), ALLOCATABLE, TARGET :: ibElP1(:), ibElP2(:), ibElP3(:), nelp(:), nelu1(:), nelu2(:), &
                                          nelv1(:), nelv2(:), nelw1(:), nelw2(:)
  REAL(KIND=8), ALLOCATABLE, TARGET :: xnode(:), ynode(:), znode(:), xnode1(:), ynode1(:), znode1(:), &
                                       xcent(:), ycent(:), zcent(:), cosAlpha(:), cosBeta(:), cosGamma(:), &
                                       alpha3(:), beta3(:), gamma3(:), bcSurf(:), &
                                       pNormDis(:), u1NormDis(:), u2NormDis(:), v1NormDis(:), v2NormDis(:), &
                                       w1NormDis(:), w2NormDis(:)
END MODULE global
