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
  CASE ('nelu1'); ptr = c_loc(nelu1); nb = 8_8 * SIZE(nelu1, KIND=8)
  CASE ('nelu2'); ptr = c_loc(nelu2); nb = 8_8 * SIZE(nelu2, KIND=8)
  CASE ('nelv1'); ptr = c_loc(nelv1); nb = 8_8 * SIZE(nelv1, KIND=8)
  CASE ('nelv2'); ptr = c_loc(nelv2); nb = 8_8 * SIZE(nelv2, KIND=8)
  CASE ('nelw1'); ptr = c_loc(nelw1); nb = 8_8 * SIZE(nelw1, KIND=8)
  CASE ('nelw2'); ptr = c_loc(nelw2); nb = 8_8 * SIZE(nelw2, KIND=8)
  CASE ('minElemcell'); ptr = c_loc(minElemcell); nb = 8_8 * SIZE(minElemcell, KIND=8)
  CASE ('xnode'); ptr = c_loc(xnode); nb = 8_8 * SIZE(xnode, KIND=8)
  CASE ('ynode'); ptr = c_loc(ynode); nb = 8_8 * SIZE(ynode, KIND=8)
  CASE ('znode'); ptr = c_loc(znode); nb = 8_8 * SIZE(znode, KIND=8)
  CASE ('alpha3'); ptr = c_loc(alpha3); nb = 8_8 * SIZE(alpha3, KIND=8)
  CASE ('beta3'); ptr = c_loc(beta3); nb = 8_8 * SIZE(beta3, KIND=8)
  CASE ('gamma3'); ptr = c_loc(gamma3); nb = 8_8 * SIZE(gamma3, KIND=8)
  CASE ('Elemcell'); ptr = c_loc(Elemcell); nb = 8_8 * SIZE(Elemcell, KIND=8)
  CASE ('ucell'); ptr = c_loc(ucell); nb = 8_8 * SIZE(ucell, KIND=8)
  CASE ('vcell'); ptr = c_loc(vcell); nb = 8_8 * SIZE(vcell, KIND=8)
  CASE ('wcell'); ptr = c_loc(wcell); nb = 8_8 * SIZE(wcell, KIND=8)
  CASE ('pcell'); ptr = c_loc(pcell); nb = 8_8 * SIZE(pcell, KIND=8)
  CASE ('ca1_uu'); ptr = c_loc(ca1_uu); nb = 8_8 * SIZE(ca1_uu, KIND=8)
  CASE ('ca2_uu'); ptr = c_loc(ca2_uu); nb = 8_8 * SIZE(ca2_uu, KIND=8)
  CASE ('ca3_uu'); ptr = c_loc(ca3_uu); nb = 8_8 * SIZE(ca3_uu, KIND=8)
  CASE ('ca4_uu'); ptr = c_loc(ca4_uu); nb = 8_8 * SIZE(ca4_uu, KIND=8)
  CASE ('ca5_uu'); ptr = c_loc(ca5_uu); nb = 8_8 * SIZE(ca5_uu, KIND=8)
  CASE ('ca6_uu'); ptr = c_loc(ca6_uu); nb = 8_8 * SIZE(ca6_uu, KIND=8)
  CASE ('ck1_uu'); ptr = c_loc(ck1_uu); nb = 8_8 * SIZE(ck1_uu, KIND=8)
  CASE ('ck2_uu'); ptr = c_loc(ck2_uu); nb = 8_8 * SIZE(ck2_uu, KIND=8)
  CASE ('ck3_uu'); ptr = c_loc(ck3_uu); nb = 8_8 * SIZE(ck3_uu, KIND=8)
  CASE ('ck4_uu'); ptr = c_loc(ck4_uu); nb = 8_8 * SIZE(ck4_uu, KIND=8)
  CASE ('ck5_uu'); ptr = c_loc(ck5_uu); nb = 8_8 * SIZE(ck5_uu, KIND=8)
  CASE ('ck6_uu'); ptr = c_loc(ck6_uu); nb = 8_8 * SIZE(ck6_uu, KIND=8)
  CASE ('ca1_vv'); ptr = c_loc(ca1_vv); nb = 8_8 * SIZE(ca1_vv, KIND=8)
  CASE ('ca2_vv'); ptr = c_loc(ca2_vv); nb = 8_8 * SIZE(ca2_vv, KIND=8)
  CASE ('ca3_vv'); ptr = c_loc(ca3_vv); nb = 8_8 * SIZE(ca3_vv, KIND=8)
  CASE ('ca4_vv'); ptr = c_loc(ca4_vv); nb = 8_8 * SIZE(ca4_vv, KIND=8)
  CASE ('ca5_vv'); ptr = c_loc(ca5_vv); nb = 8_8 * SIZE(ca5_vv, KIND=8)
  CASE ('ca6_vv'); ptr = c_loc(ca6_vv); nb = 8_8 * SIZE(ca6_vv, KIND=8)
  CASE ('ck1_vv'); ptr = c_loc(ck1_vv); nb = 8_8 * SIZE(ck1_vv, KIND=8)
  CASE ('ck2_vv'); ptr = c_loc(ck2_vv); nb = 8_8 * SIZE(ck2_vv, KIND=8)
  CASE ('ck3_vv'); ptr = c_loc(ck3_vv); nb = 8_8 * SIZE(ck3_vv, KIND=8)
  CASE ('ck4_vv'); ptr = c_loc(ck4_vv); nb = 8_8 * SIZE(ck4_vv, KIND=8)
  CASE ('ck5_vv'); ptr = c_loc(ck5_vv); nb = 8_8 * SIZE(ck5_vv, KIND=8)
  CASE ('ck6_vv'); ptr = c_loc(ck6_vv); nb = 8_8 * SIZE(ck6_vv, KIND=8)
  CASE ('ca1_ww'); ptr = c_loc(ca1_ww); nb = 8_8 * SIZE(ca1_ww, KIND=8)
  CASE ('ca2_ww'); ptr = c_loc(ca2_ww); nb = 8_8 * SIZE(ca2_ww, KIND=8)
  CASE ('ca3_ww'); ptr = c_loc(ca3_ww); nb = 8_8 * SIZE(ca3_ww, KIND=8)
  CASE ('ca4_ww'); ptr = c_loc(ca4_ww); nb = 8_8 * SIZE(ca4_ww, KIND=8)
  CASE ('ca5_ww'); ptr = c_loc(ca5_ww); nb = 8_8 * SIZE(ca5_ww, KIND=8)
  CASE ('ca6_ww'); ptr = c_loc(ca6_ww); nb = 8_8 * SIZE(ca6_ww, KIND=8)
  CASE ('ck1_ww'); ptr = c_loc(ck1_ww); nb = 8_8 * SIZE(ck1_ww, KIND=8)
  CASE ('ck2_ww'); ptr = c_loc(ck2_ww); nb = 8_8 * SIZE(ck2_ww, KIND=8)
  CASE ('ck3_ww'); ptr = c_loc(ck3_ww); nb = 8_8 * SIZE(ck3_ww, KIND=8)
  CASE ('ck4_ww'); ptr = c_loc(ck4_ww); nb = 8_8 * SIZE(ck4_ww, KIND=8)
  CASE ('ck5_ww'); ptr = c_loc(ck5_ww); nb = 8_8 * SIZE(ck5_ww, KIND=8)
  CASE ('ck6_ww'); ptr = c_loc(ck6_ww); nb = 8_8 * SIZE(ck6_ww, KIND=8)
  CASE ('ca1_uv'); ptr = c_loc(ca1_uv); nb = 8_8 * SIZE(ca1_uv, KIND=8)
  CASE ('ca2_uv'); ptr = c_loc(ca2_uv); nb = 8_8 * SIZE(ca2_uv, KIND=8)
  CASE ('ca3_uv'); ptr = c_loc(ca3_uv); nb = 8_8 * SIZE(ca3_uv, KIND=8)
  CASE ('ca4_uv'); ptr = c_loc(ca4_uv); nb = 8_8 * SIZE(ca4_uv, KIND=8)
  CASE ('ca5_uv'); ptr = c_loc(ca5_uv); nb = 8_8 * SIZE(ca5_uv, KIND=8)
  CASE ('ca6_uv'); ptr = c_loc(ca6_uv); nb = 8_8 * SIZE(ca6_uv, KIND=8)
  CASE ('ck1_uv'); ptr = c_loc(ck1_uv); nb = 8_8 * SIZE(ck1_uv, KIND=8)
  CASE ('ck2_uv'); ptr = c_loc(ck2_uv); nb = 8_8 * SIZE(ck2_uv, KIND=8)
  CASE ('ck3_uv'); ptr = c_loc(ck3_uv); nb = 8_8 * SIZE(ck3_uv, KIND=8)
  CASE ('ck4_uv'); ptr = c_loc(ck4_uv); nb = 8_8 * SIZE(ck4_uv, KIND=8)
  CASE ('ck5_uv'); ptr = c_loc(ck5_uv); nb = 8_8 * SIZE(ck5_uv, KIND=8)
  CASE ('ck6_uv'); ptr = c_loc(ck6_uv); nb = 8_8 * SIZE(ck6_uv, KIND=8)
  CASE ('ca1_uw'); ptr = c_loc(ca1_uw); nb = 8_8 * SIZE(ca1_uw, KIND=8)
  CASE ('ca2_uw'); ptr = c_loc(ca2_uw); nb = 8_8 * SIZE(ca2_uw, KIND=8)
  CASE ('ca3_uw'); ptr = c_loc(ca3_uw); nb = 8_8 * SIZE(ca3_uw, KIND=8)
  CASE ('ca4_uw'); ptr = c_loc(ca4_uw); nb = 8_8 * SIZE(ca4_uw, KIND=8)
  CASE ('ca5_uw'); ptr = c_loc(ca5_uw); nb = 8_8 * SIZE(ca5_uw, KIND=8)
  CASE ('ca6_uw'); ptr = c_loc(ca6_uw); nb = 8_8 * SIZE(ca6_uw, KIND=8)
  CASE ('ck1_uw'); ptr = c_loc(ck1_uw); nb = 8_8 * SIZE(ck1_uw, KIND=8)
  CASE ('ck2_uw'); ptr = c_loc(ck2_uw); nb = 8_8 * SIZE(ck2_uw, KIND=8)
  CASE ('ck3_uw'); ptr = c_loc(ck3_uw); nb = 8_8 * SIZE(ck3_uw, KIND=8)
  CASE ('ck4_uw'); ptr = c_loc(ck4_uw); nb = 8_8 * SIZE(ck4_uw, KIND=8)
  CASE ('ck5_uw'); ptr = c_loc(ck5_uw); nb = 8_8 * SIZE(ck5_uw, KIND=8)
  CASE ('ck6_uw'); ptr = c_loc(ck6_uw); nb = 8_8 * SIZE(ck6_uw, KIND=8)
  CASE ('ca1_vu'); ptr = c_loc(ca1_vu); nb = 8_8 * SIZE(ca1_vu, KIND=8)
  CASE ('ca2_vu'); ptr = c_loc(ca2_vu); nb = 8_8 * SIZE(ca2_vu, KIND=8)
  CASE ('ca3_vu'); ptr = c_loc(ca3_vu); nb = 8_8 * SIZE(ca3_vu, KIND=8)
  CASE ('ca4_vu'); ptr = c_loc(ca4_vu); nb = 8_8 * SIZE(ca4_vu, KIND=8)
  CASE ('ca5_vu'); ptr = c_loc(ca5_vu); nb = 8_8 * SIZE(ca5_vu, KIND=8)
  CASE ('ca6_vu'); ptr = c_loc(ca6_vu); nb = 8_8 * SIZE(ca6_vu, KIND=8)
  CASE ('ck1_vu'); ptr = c_loc(ck1_vu); nb = 8_8 * SIZE(ck1_vu, KIND=8)
  CASE ('ck2_vu'); ptr = c_loc(ck2_vu); nb = 8_8 * SIZE(ck2_vu, KIND=8)
  CASE ('ck3_vu'); ptr = c_loc(ck3_vu); nb = 8_8 * SIZE(ck3_vu, KIND=8)
  CASE ('ck4_vu'); ptr = c_loc(ck4_vu); nb = 8_8 * SIZE(ck4_vu, KIND=8)
  CASE ('ck5_vu'); ptr = c_loc(ck5_vu); nb = 8_8 * SIZE(ck5_vu, KIND=8)
  CASE ('ck6_vu'); ptr = c_loc(ck6_vu); nb = 8_8 * SIZE(ck6_vu, KIND=8)
  CASE ('ca1_vw'); ptr = c_loc(ca1_vw); nb = 8_8 * SIZE(ca1_vw, KIND=8)
  CASE ('ca2_vw'); ptr = c_loc(ca2_vw); nb = 8_8 * SIZE(ca2_vw, KIND=8)
  CASE ('ca3_vw'); ptr = c_loc(ca3_vw); nb = 8_8 * SIZE(ca3_vw, KIND=8)
  CASE ('ca4_vw'); ptr = c_loc(ca4_vw); nb = 8_8 * SIZE(ca4_vw, KIND=8)
  CASE ('ca5_vw'); ptr = c_loc(ca5_vw); nb = 8_8 * SIZE(ca5_vw, KIND=8)
  CASE ('ca6_vw'); ptr = c_loc(ca6_vw); nb = 8_8 * SIZE(ca6_vw, KIND=8)
  CASE ('ck1_vw'); ptr = c_loc(ck1_vw); nb = 8_8 * SIZE(ck1_vw, KIND=8)
  CASE ('ck2_vw'); ptr = c_loc(ck2_vw); nb = 8_8 * SIZE(ck2_vw, KIND=8)
  CASE ('ck3_vw'); ptr = c_loc(ck3_vw); nb = 8_8 * SIZE(ck3_vw, KIND=8)
  CASE ('ck4_vw'); ptr = c_loc(ck4_vw); nb = 8_8 * SIZE(ck4_vw, KIND=8)
  CASE ('ck5_vw'); ptr = c_loc(ck5_vw); nb = 8_8 * SIZE(ck5_vw, KIND=8)
  CASE ('ck6_vw'); ptr = c_loc(ck6_vw); nb = 8_8 * SIZE(ck6_vw, KIND=8)
  CASE ('ca1_wu'); ptr = c_loc(ca1_wu); nb = 8_8 * SIZE(ca1_wu, KIND=8)
  CASE ('ca2_wu'); ptr = c_loc(ca2_wu); nb = 8_8 * SIZE(ca2_wu, KIND=8)
  CASE ('ca3_wu'); ptr = c_loc(ca3_wu); nb = 8_8 * SIZE(ca3_wu, KIND=8)
  CASE ('ca4_wu'); ptr = c_loc(ca4_wu); nb = 8_8 * SIZE(ca4_wu, KIND=8)
  CASE ('ca5_wu'); ptr = c_loc(ca5_wu); nb = 8_8 * SIZE(ca5_wu, KIND=8)
  CASE ('ca6_wu'); ptr = c_loc(ca6_wu); nb = 8_8 * SIZE(ca6_wu, KIND=8)
  CASE ('ck1_wu'); ptr = c_loc(ck1_wu); nb = 8_8 * SIZE(ck1_wu, KIND=8)
  CASE ('ck2_wu'); ptr = c_loc(ck2_wu); nb = 8_8 * SIZE(ck2_wu, KIND=8)
  CASE ('ck3_wu'); ptr = c_loc(ck3_wu); nb = 8_8 * SIZE(ck3_wu, KIND=8)
  CASE ('ck4_wu'); ptr = c_loc(ck4_wu); nb = 8_8 * SIZE(ck4_wu, KIND=8)
  CASE ('ck5_wu'); ptr = c_loc(ck5_wu); nb = 8_8 * SIZE(ck5_wu, KIND=8)
  CASE ('ck6_wu'); ptr = c_loc(ck6_wu); nb = 8_8 * SIZE(ck6_wu, KIND=8)
  CASE ('ca1_wv'); ptr = c_loc(ca1_wv); nb = 8_8 * SIZE(ca1_wv, KIND=8)
  CASE ('ca2_wv'); ptr = c_loc(ca2_wv); nb = 8_8 * SIZE(ca2_wv, KIND=8)
  CASE ('ca3_wv'); ptr = c_loc(ca3_wv); nb = 8_8 * SIZE(ca3_wv, KIND=8)
  CASE ('ca4_wv'); ptr = c_loc(ca4_wv); nb = 8_8 * SIZE(ca4_wv, KIND=8)
  CASE ('ca5_wv'); ptr = c_loc(ca5_wv); nb = 8_8 * SIZE(ca5_wv, KIND=8)
  CASE ('ca6_wv'); ptr = c_loc(ca6_wv); nb = 8_8 * SIZE(ca6_wv, KIND=8)
  CASE ('ck1_wv'); ptr = c_loc(ck1_wv); nb = 8_8 * SIZE(ck1_wv, KIND=8)
  CASE ('ck2_wv'); ptr = c_loc(ck2_wv); nb = 8_8 * SIZE(ck2_wv, KIND=8)
  CASE ('ck3_wv'); ptr = c_loc(ck3_wv); nb = 8_8 * SIZE(ck3_wv, KIND=8)
  CASE ('ck4_wv'); ptr = c_loc(ck4_wv); nb = 8_8 * SIZE(ck4_wv, KIND=8)
  CASE ('ck5_wv'); ptr = c_loc(ck5_wv); nb = 8_8 * SIZE(ck5_wv, KIND=8)
  CASE ('ck6_wv'); ptr = c_loc(ck6_wv); nb = 8_8 * SIZE(ck6_wv, KIND=8)
  CASE DEFAULT
    WRITE(*, '(A,A)') 'ib3d: no host copy of array ', TRIM(name)
    ERROR STOP 1
  END SELECT
END SUBROUTINE array_address


! subAllocateArrays.f90:1-15: the module's field arrays, as the reference
! allocates them.  The device copy of the module is created here from the
! grid and the surface mesh unless the readers already handed them over (the
! drop-in driver's readers do; the reference's own readInput and
! readSurfaceMeshGmsh leave them in the module: x1(2:nx+2) holds the grid
! nodes, xnode/ibElP1.. the mesh).
SUBROUTINE allocateArrays
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  ALLOCATE(u(nx+2,ny+2,nz+2), ut(nx+2,ny+2,nz+2), v(nx+2,ny+2,nz+2), vt(nx+2,ny+2,nz+2), &
           w(nx+2,ny+2,nz+2), wt(nx+2,ny+2,nz+2), p(nx+2,ny+2,nz+2), &
           u_sum(nx+2,ny+2,nz+2), v_sum(nx+2,ny+2,nz+2), w_sum(nx+2,ny+2,nz+2), p_sum(nx+2,ny+2,nz+2), &
           u_avg(nx+2,ny+2,nz+2), v_avg(nx+2,ny+2,nz+2), w_avg(nx+2,ny+2,nz+2), p_avg(nx+2,ny+2,nz+2), &
           resi_u(nx+2,ny+2,nz+2), resi_v(nx+2,ny+2,nz+2), resi_w(nx+2,ny+2,nz+2))
  ALLOCATE(cell(nx+2,ny+2,nz+2))
  ALLOCATE(Acx(nx,3), Acy(ny,3), Acz(nz,3))
  ALLOCATE(pc(nx+2,ny+2,nz+2), pco(nx+2,ny+2,nz+2))
  ALLOCATE(SUMWSS(ibNodes), SIGNWSS(ibNodes))
  IF (.NOT. ib3d_grid_on_device) THEN
    CALL ib3d_check(ib3d_init(nx, ny, nz), 'ib3d_init')
    CALL ib3d_check(ib3d_set_grid(x1(2:nx+2), y1(2:ny+2), z1(2:nz+2)), 'set_grid')
    ib3d_grid_on_device = .TRUE.
  END IF
  IF (.NOT. ib3d_mesh_on_device) THEN
    CALL ib3d_check(ib3d_set_mesh(ibNodes, xnode, ynode, znode, ibElems, ibElP1, ibElP2, ibElP3), 'set_mesh')
    ib3d_mesh_on_device = .TRUE.
  END IF
END SUBROUTINE allocateArrays

! replaces main.f90:55, `!$acc update host (u, v, w, ut, vt, wt, p, cell)`:
! the same arrays, device -> module.  Call it where host output needs them
! (the reference's writeOutput every 5000 steps, writeResult every 1000).
SUBROUTINE ib3d_update_host_fields
  IMPLICIT NONE
  CALL ib3d_update_host('u'); CALL ib3d_update_host('v'); CALL ib3d_update_host('w')
  CALL ib3d_update_host('ut'); CALL ib3d_update_host('vt'); CALL ib3d_update_host('wt')
  CALL ib3d_update_host('p'); CALL ib3d_update_host('cell')
END SUBROUTINE ib3d_update_host_fields

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
    ALLOCATE(xcent(ibElems), ycent(ibElems), zcent(ibElems), cosAlpha(ibElems), cosBeta(ibElems), cosGamma(ibElems), &
             alpha3(ibElems), beta3(ibElems), gamma3(ibElems))
  CALL ib3d_check(ib3d_computeSurfaceNorm(inor), 'computeSurfaceNorm')
  ! subSearch.f90:128 copies the nine element arrays to the host
  CALL ib3d_update_host('xcent'); CALL ib3d_update_host('ycent'); CALL ib3d_update_host('zcent')
  CALL ib3d_update_host('cosAlpha'); CALL ib3d_update_host('cosBeta'); CALL ib3d_update_host('cosGamma')
  CALL ib3d_update_host('alpha3'); CALL ib3d_update_host('beta3'); CALL ib3d_update_host('gamma3')
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
  ! the reference sets the module's fields on the host and copies them to the
  ! device (subInitialConditions.f90:8-44); here the device sets them and the
  ! module receives the same values
  CALL ib3d_update_host('u'); CALL ib3d_update_host('v'); CALL ib3d_update_host('w')
  CALL ib3d_update_host('ut'); CALL ib3d_update_host('vt'); CALL ib3d_update_host('wt')
  CALL ib3d_update_host('p')
  CALL ib3d_update_host('resi_u'); CALL ib3d_update_host('resi_v'); CALL ib3d_update_host('resi_w')
  u_sum = 0.0d0; v_sum = 0.0d0; w_sum = 0.0d0; p_sum = 0.0d0
  u_avg = 0.0d0; v_avg = 0.0d0; w_avg = 0.0d0; p_avg = 0.0d0
  SUMWSS = 0.0d0; SIGNWSS = 0.0d0
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

SUBROUTINE non_uni_coeff                     ! subNavierStokes.f90:3-619
  USE global
  USE ib3d_cabi
  IMPLICIT NONE
  CALL ib3d_check(ib3d_non_uni_coeff(), 'non_uni_coeff')
  ! the reference builds the 108 tables on the host (subNavierStokes.f90:17-33)
  IF (.NOT. ALLOCATED(ca1_uu)) ALLOCATE( &
           ca1_uu(nx+2), ca2_uu(nx+2), ca3_uu(nx+2), ca4_uu(nx+2), ca5_uu(nx+2), ca6_uu(nx+2), &
           ck1_uu(nx+2), ck2_uu(nx+2), ck3_uu(nx+2), ck4_uu(nx+2), ck5_uu(nx+2), ck6_uu(nx+2), &
           ca1_vv(ny+2), ca2_vv(ny+2), ca3_vv(ny+2), ca4_vv(ny+2), ca5_vv(ny+2), ca6_vv(ny+2), &
           ck1_vv(ny+2), ck2_vv(ny+2), ck3_vv(ny+2), ck4_vv(ny+2), ck5_vv(ny+2), ck6_vv(ny+2), &
           ca1_ww(nz+2), ca2_ww(nz+2), ca3_ww(nz+2), ca4_ww(nz+2), ca5_ww(nz+2), ca6_ww(nz+2), &
           ck1_ww(nz+2), ck2_ww(nz+2), ck3_ww(nz+2), ck4_ww(nz+2), ck5_ww(nz+2), ck6_ww(nz+2), &
           ca1_uv(nx+2), ca2_uv(nx+2), ca3_uv(nx+2), ca4_uv(nx+2), ca5_uv(nx+2), ca6_uv(nx+2), &
           ck1_uv(nx+2), ck2_uv(nx+2), ck3_uv(nx+2), ck4_uv(nx+2), ck5_uv(nx+2), ck6_uv(nx+2), &
           ca1_uw(nx+2), ca2_uw(nx+2), ca3_uw(nx+2), ca4_uw(nx+2), ca5_uw(nx+2), ca6_uw(nx+2), &
           ck1_uw(nx+2), ck2_uw(nx+2), ck3_uw(nx+2), ck4_uw(nx+2), ck5_uw(nx+2), ck6_uw(nx+2), &
           ca1_vu(ny+2), ca2_vu(ny+2), ca3_vu(ny+2), ca4_vu(ny+2), ca5_vu(ny+2), ca6_vu(ny+2), &
           ck1_vu(ny+2), ck2_vu(ny+2), ck3_vu(ny+2), ck4_vu(ny+2), ck5_vu(ny+2), ck6_vu(ny+2), &
           ca1_vw(ny+2), ca2_vw(ny+2), ca3_vw(ny+2), ca4_vw(ny+2), ca5_vw(ny+2), ca6_vw(ny+2), &
           ck1_vw(ny+2), ck2_vw(ny+2), ck3_vw(ny+2), ck4_vw(ny+2), ck5_vw(ny+2), ck6_vw(ny+2), &
           ca1_wu(nz+2), ca2_wu(nz+2), ca3_wu(nz+2), ca4_wu(nz+2), ca5_wu(nz+2), ca6_wu(nz+2), &
           ck1_wu(nz+2), ck2_wu(nz+2), ck3_wu(nz+2), ck4_wu(nz+2), ck5_wu(nz+2), ck6_wu(nz+2), &
           ca1_wv(nz+2), ca2_wv(nz+2), ca3_wv(nz+2), ca4_wv(nz+2), ca5_wv(nz+2), ca6_wv(nz+2), &
           ck1_wv(nz+2), ck2_wv(nz+2), ck3_wv(nz+2), ck4_wv(nz+2), ck5_wv(nz+2), ck6_wv(nz+2))
  CALL ib3d_update_host('ca1_uu')
  CALL ib3d_update_host('ca2_uu')
  CALL ib3d_update_host('ca3_uu')
  CALL ib3d_update_host('ca4_uu')
  CALL ib3d_update_host('ca5_uu')
  CALL ib3d_update_host('ca6_uu')
  CALL ib3d_update_host('ck1_uu')
  CALL ib3d_update_host('ck2_uu')
  CALL ib3d_update_host('ck3_uu')
  CALL ib3d_update_host('ck4_uu')
  CALL ib3d_update_host('ck5_uu')
  CALL ib3d_update_host('ck6_uu')
  CALL ib3d_update_host('ca1_vv')
  CALL ib3d_update_host('ca2_vv')
  CALL ib3d_update_host('ca3_vv')
  CALL ib3d_update_host('ca4_vv')
  CALL ib3d_update_host('ca5_vv')
  CALL ib3d_update_host('ca6_vv')
  CALL ib3d_update_host('ck1_vv')
  CALL ib3d_update_host('ck2_vv')
  CALL ib3d_update_host('ck3_vv')
  CALL ib3d_update_host('ck4_vv')
  CALL ib3d_update_host('ck5_vv')
  CALL ib3d_update_host('ck6_vv')
  CALL ib3d_update_host('ca1_ww')
  CALL ib3d_update_host('ca2_ww')
  CALL ib3d_update_host('ca3_ww')
  CALL ib3d_update_host('ca4_ww')
  CALL ib3d_update_host('ca5_ww')
  CALL ib3d_update_host('ca6_ww')
  CALL ib3d_update_host('ck1_ww')
  CALL ib3d_update_host('ck2_ww')
  CALL ib3d_update_host('ck3_ww')
  CALL ib3d_update_host('ck4_ww')
  CALL ib3d_update_host('ck5_ww')
  CALL ib3d_update_host('ck6_ww')
  CALL ib3d_update_host('ca1_uv')
  CALL ib3d_update_host('ca2_uv')
  CALL ib3d_update_host('ca3_uv')
  CALL ib3d_update_host('ca4_uv')
  CALL ib3d_update_host('ca5_uv')
  CALL ib3d_update_host('ca6_uv')
  CALL ib3d_update_host('ck1_uv')
  CALL ib3d_update_host('ck2_uv')
  CALL ib3d_update_host('ck3_uv')
  CALL ib3d_update_host('ck4_uv')
  CALL ib3d_update_host('ck5_uv')
  CALL ib3d_update_host('ck6_uv')
  CALL ib3d_update_host('ca1_uw')
  CALL ib3d_update_host('ca2_uw')
  CALL ib3d_update_host('ca3_uw')
  CALL ib3d_update_host('ca4_uw')
  CALL ib3d_update_host('ca5_uw')
  CALL ib3d_update_host('ca6_uw')
  CALL ib3d_update_host('ck1_uw')
  CALL ib3d_update_host('ck2_uw')
  CALL ib3d_update_host('ck3_uw')
  CALL ib3d_update_host('ck4_uw')
  CALL ib3d_update_host('ck5_uw')
  CALL ib3d_update_host('ck6_uw')
  CALL ib3d_update_host('ca1_vu')
  CALL ib3d_update_host('ca2_vu')
  CALL ib3d_update_host('ca3_vu')
  CALL ib3d_update_host('ca4_vu')
  CALL ib3d_update_host('ca5_vu')
  CALL ib3d_update_host('ca6_vu')
  CALL ib3d_update_host('ck1_vu')
  CALL ib3d_update_host('ck2_vu')
  CALL ib3d_update_host('ck3_vu')
  CALL ib3d_update_host('ck4_vu')
  CALL ib3d_update_host('ck5_vu')
  CALL ib3d_update_host('ck6_vu')
  CALL ib3d_update_host('ca1_vw')
  CALL ib3d_update_host('ca2_vw')
  CALL ib3d_update_host('ca3_vw')
  CALL ib3d_update_host('ca4_vw')
  CALL ib3d_update_host('ca5_vw')
  CALL ib3d_update_host('ca6_vw')
  CALL ib3d_update_host('ck1_vw')
  CALL ib3d_update_host('ck2_vw')
  CALL ib3d_update_host('ck3_vw')
  CALL ib3d_update_host('ck4_vw')
  CALL ib3d_update_host('ck5_vw')
  CALL ib3d_update_host('ck6_vw')
  CALL ib3d_update_host('ca1_wu')
  CALL ib3d_update_host('ca2_wu')
  CALL ib3d_update_host('ca3_wu')
  CALL ib3d_update_host('ca4_wu')
  CALL ib3d_update_host('ca5_wu')
  CALL ib3d_update_host('ca6_wu')
  CALL ib3d_update_host('ck1_wu')
  CALL ib3d_update_host('ck2_wu')
  CALL ib3d_update_host('ck3_wu')
  CALL ib3d_update_host('ck4_wu')
  CALL ib3d_update_host('ck5_wu')
  CALL ib3d_update_host('ck6_wu')
  CALL ib3d_update_host('ca1_wv')
  CALL ib3d_update_host('ca2_wv')
  CALL ib3d_update_host('ca3_wv')
  CALL ib3d_update_host('ca4_wv')
  CALL ib3d_update_host('ca5_wv')
  CALL ib3d_update_host('ca6_wv')
  CALL ib3d_update_host('ck1_wv')
  CALL ib3d_update_host('ck2_wv')
  CALL ib3d_update_host('ck3_wv')
  CALL ib3d_update_host('ck4_wv')
  CALL ib3d_update_host('ck5_wv')
  CALL ib3d_update_host('ck6_wv')
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
  u_sum = 0.0d0; v_sum = 0.0d0; w_sum = 0.0d0; p_sum = 0.0d0
  u_avg = 0.0d0; v_avg = 0.0d0; w_avg = 0.0d0; p_avg = 0.0d0
  SUMWSS = 0.0d0; SIGNWSS = 0.0d0
  resi_u = 0.0d0
  resi_v = 0.0d0
  resi_w = 0.0d0
  CALL ib3d_update_device('u'); CALL ib3d_update_device('v'); CALL ib3d_update_device('w')
  CALL ib3d_update_device('ut'); CALL ib3d_update_device('vt'); CALL ib3d_update_device('wt')
  CALL ib3d_update_device('p')
  CALL ib3d_update_device('resi_u'); CALL ib3d_update_device('resi_v'); CALL ib3d_update_device('resi_w')
END SUBROUTINE lastConditions

