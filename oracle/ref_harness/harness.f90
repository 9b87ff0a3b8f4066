! Test harness around the reference subroutines (built from the reference's
! own source files, main.f90 excluded).  It replays the call order of the
! reference driver (main.f90:115-177) and dumps module arrays as raw
! little-endian binaries so the C restatement and the GPU path can be
! compared against the reference bit for bit.
!
! Usage (run inside a case directory holding inputdata, grid?.dat and
! geometries/sphere.msh):
!   ref_harness run    NSTEPS DUMPDIR [STEP ...]   full steps, stage dumps at STEPs
!   ref_harness sample NSWEEPS                     time the stages of step 1 and
!                                                  NSWEEPS single SOR iterations
!   ref_harness time   NSTEPS                      time NSTEPS whole steps (setup
!                                                  and each step into
!                                                  harness_time.txt)
!   ref_harness move   NSTEPS DUMPDIR DY           rigid body shifted by s*DY in y
!                                                  at step s, then the reference's
!                                                  computeSurfaceNorm,
!                                                  selectiveRetagging, cellCount and
!                                                  computeNormDistance; geometry
!                                                  dumps per step (no flow solve)
!
! This is test infrastructure: only the tests, the smoke check and the CPU
! baseline leg of the benchmark run it.
module harness_io
  implicit none
  character(len=512) :: dumpdir = '.'
contains
  subroutine dump_r8(name, a, n)
    character(len=*), intent(in) :: name
    integer(8), intent(in) :: n
    real(8), intent(in) :: a(n)
    integer :: u
    open(newunit=u, file=trim(dumpdir)//'/'//name//'.f64', access='stream', &
         form='unformatted', status='replace')
    write(u) a
    close(u)
  end subroutine
  subroutine dump_i8(name, a, n)
    character(len=*), intent(in) :: name
    integer(8), intent(in) :: n
    integer(8), intent(in) :: a(n)
    integer :: u
    open(newunit=u, file=trim(dumpdir)//'/'//name//'.i64', access='stream', &
         form='unformatted', status='replace')
    write(u) a
    close(u)
  end subroutine
end module harness_io

program ref_harness
  use global
  use harness_io
  implicit none
  character(len=512) :: arg, mode
  integer(8) :: nsteps, nsw, s, k
  integer :: nargs, ia, ndump
  integer(8) :: dsteps(64), inode
  real(8) :: move_dy, yc
  integer(8) :: c0, c1, crate
  real(8) :: t_setup, t_mom, t_bc, t_sol, t_frc, t_div, t_sor, t_corr, t_std, t_prf, t_str
  real(8) :: derr1, derr2, huge_eps
  integer :: ustep, utime
  logical :: dumping

  nargs = command_argument_count()
  if (nargs < 2) then
    print *, 'usage: ref_harness run NSTEPS DUMPDIR [STEP ...] | sample NSWEEPS | time NSTEPS'
    stop 2
  end if
  call get_command_argument(1, mode)
  call get_command_argument(2, arg)
  read(arg, *) nsteps
  ndump = 0
  if (trim(mode) == 'move') then
    call get_command_argument(3, dumpdir)
    call get_command_argument(4, arg)
    read(arg, *) move_dy
  end if
  if (trim(mode) == 'run') then
    call get_command_argument(3, dumpdir)
    do ia = 4, nargs
      call get_command_argument(ia, arg)
      ndump = ndump + 1
      read(arg, *) dsteps(ndump)
    end do
  end if
  call system_clock(c0, crate)

  call readInput
  call readSurfaceMeshGmsh
  call allocateArrays
  call shiftSurfaceNodesInitial
  call computeSurfaceNorm
  call tagging
  if (trim(mode) == 'run') call dump_setup_a()
  call cellCount
  call computeNormDistance
  call initialConditions
  call coefficientMatrix
  call non_uni_coeff
  if (trim(mode) == 'run') call dump_setup_b()
  call system_clock(c1)
  t_setup = real(c1 - c0, 8) / real(crate, 8)

  ita = 1
  ita1 = 1
  totime = 0.
  totime = totime + deltat

  if (trim(mode) == 'move') then
    open(newunit=ustep, file=trim(dumpdir)//'/move_steps.txt', status='replace')
    do s = 1, nsteps
      yc = real(s, 8) * move_dy
      do inode = 1, ibNodes
        xnode1(inode) = (xnode(inode) + xShift) + 0.0d0
        ynode1(inode) = (ynode(inode) + yShift) + yc
        znode1(inode) = (znode(inode) + zShift) + 0.0d0
      end do
      DEALLOCATE(xcent, ycent, zcent, cosAlpha, cosBeta, cosGamma, alpha3, beta3, gamma3)
      call computeSurfaceNorm
      call selectiveRetagging
      call dump_move(s, 'rt', .false.)
      DEALLOCATE(interceptedIndexPtr, fluidIndexPtr, solidIndexPtr, redCellIndexPtr, blackCellIndexPtr)
      call cellCount
      DEALLOCATE(pNormDis, nelp, nelu1, nelu2, nelv1, nelv2, nelw1, nelw2, &
                 u1NormDis, u2NormDis, v1NormDis, v2NormDis, w1NormDis, w2NormDis)
      call computeNormDistance
      call dump_move(s, 'nd', .true.)
      write(ustep, '(I8,ES25.17,5I12)') s, yc, ibCellCount, fluidCellCount, solidCellCount, &
                                        redCellCount, blackCellCount
      flush(ustep)
    end do
    close(ustep)
    stop
  end if

  if (trim(mode) == 'sample') then
    nsw = nsteps
    call tick(c0)
    call nsMomentum
    call tock(c0, t_mom)
    call velocityBC
    call tock(c0, t_bc)
    call solidCellBC
    call tock(c0, t_sol)
    call velocityForcing1
    call velocityBC
    call tock(c0, t_frc)
    ALLOCATE(b(nx+2, ny+2, nz+2))
    b = 0.
    pc = 0.
    pco = 0.
    call computeDiv
    call tock(c0, t_div)
    huge_eps = 1.0d300
    do k = 1, nsw
      call REDBLACKSOR(huge_eps, nIterPcor, derr1, derr2)
    end do
    call tock(c0, t_sor)
    call correctPressure
    call correctVelocity
    call velocityBC
    call tock(c0, t_corr)
    DEALLOCATE(b)
    call pressureForcing1
    call tock(c0, t_prf)
    call stressCal1
    call tock(c0, t_str)
    open(newunit=ustep, file='harness_sample.txt', status='replace')
    write(ustep, '(A,I12)') 'cells ', nx*ny*nz
    write(ustep, '(A,I12)') 'fluid ', fluidCellCount
    write(ustep, '(A,I12)') 'sweeps ', nsw
    write(ustep, '(A,ES25.17)') 'setup ', t_setup
    write(ustep, '(A,ES25.17)') 'nsMomentum ', t_mom
    write(ustep, '(A,ES25.17)') 'velocityBC ', t_bc
    write(ustep, '(A,ES25.17)') 'solidCellBC ', t_sol
    write(ustep, '(A,ES25.17)') 'velocityForcing1+BC ', t_frc
    write(ustep, '(A,ES25.17)') 'computeDiv ', t_div
    write(ustep, '(A,ES25.17)') 'sor_per_sweep ', t_sor / real(nsw, 8)
    write(ustep, '(A,ES25.17)') 'correct+BC ', t_corr
    write(ustep, '(A,ES25.17)') 'pressureForcing1 ', t_prf
    write(ustep, '(A,ES25.17)') 'stressCal1 ', t_str
    close(ustep)
    stop
  end if

  if (trim(mode) == 'time') then
    open(newunit=utime, file='harness_time.txt', status='replace')
    write(utime, '(A,I12)') 'cells ', nx*ny*nz
    write(utime, '(A,ES25.17)') 'setup ', t_setup
  end if
  open(newunit=ustep, file=trim(dumpdir)//'/steps.txt', status='replace')
  do s = 1, nsteps
    call system_clock(c0)
    if (s > 1) then
      ita = ita + 1
      ita1 = ita1 + 1
      totime = totime + deltat
    end if
    dumping = any(dsteps(1:ndump) == s)
    call nsMomentum
    if (dumping) call dump_stage(s, 'mom')
    call velocityBC
    if (dumping) call dump_stage(s, 'bc1')
    call solidCellBC
    if (dumping) call dump_stage(s, 'sol')
    call velocityForcing1
    if (dumping) call dump_stage(s, 'frc')
    call velocityBC
    if (dumping) call dump_stage(s, 'bc2')
    call poissonSolver
    if (dumping) call dump_stage(s, 'poi')
    call pressureForcing1
    if (dumping) call dump_stage(s, 'prf')
    if (s == 1) then
      call stressCal1
      call dump_cells()
    else
      call stressCal2
    end if
    write(ustep, '(I8,I12,2ES25.17)') ita, nIterPcor, divmax, totime
    flush(ustep)
    if (trim(mode) == 'time') then
      call system_clock(c1)
      write(utime, '(A,I8,I12,ES25.17)') 'step ', s, nIterPcor, real(c1 - c0, 8) / real(crate, 8)
      flush(utime)
    end if
  end do
  close(ustep)
  if (trim(mode) == 'time') close(utime)

contains

  subroutine tick(c)
    integer(8), intent(out) :: c
    call system_clock(c)
  end subroutine

  subroutine tock(c, t)
    integer(8), intent(inout) :: c
    real(8), intent(out) :: t
    integer(8) :: cn
    call system_clock(cn)
    t = real(cn - c, 8) / real(crate, 8)
    c = cn
  end subroutine

  subroutine dump_setup_a()
    integer(8) :: np
    np = (nx+2)*(ny+2)*(nz+2)
    call dump_i8('cell_tagging', cell, np)
    call dump_i8('minElemcell', minElemcell, np)
    call dump_r8('xnode1', xnode1, ibNodes)
    call dump_r8('ynode1', ynode1, ibNodes)
    call dump_r8('znode1', znode1, ibNodes)
    call dump_r8('xcent', xcent, ibElems)
    call dump_r8('ycent', ycent, ibElems)
    call dump_r8('zcent', zcent, ibElems)
    call dump_r8('cosAlpha', cosAlpha, ibElems)
    call dump_r8('cosBeta', cosBeta, ibElems)
    call dump_r8('cosGamma', cosGamma, ibElems)
    call dump_r8('alpha3', alpha3, ibElems)
    call dump_r8('beta3', beta3, ibElems)
    call dump_r8('gamma3', gamma3, ibElems)
    call dump_r8('x1', x1, nx+3)
    call dump_r8('y1', y1, ny+3)
    call dump_r8('z1', z1, nz+3)
    call dump_r8('deltax', deltax, nx+2)
    call dump_r8('deltay', deltay, ny+2)
    call dump_r8('deltaz', deltaz, nz+2)
    call dump_r8('xu', xu, nx+3)
    call dump_r8('yu', yu, ny+2)
    call dump_r8('zu', zu, nz+2)
    call dump_r8('xv', xv, nx+2)
    call dump_r8('yv', yv, ny+3)
    call dump_r8('zv', zv, nz+2)
    call dump_r8('xw', xw, nx+2)
    call dump_r8('yw', yw, ny+2)
    call dump_r8('zw', zw, nz+3)
    call dump_r8('xp', xp, nx+2)
    call dump_r8('yp', yp, ny+2)
    call dump_r8('zp', zp, nz+2)
  end subroutine

  subroutine dump_setup_b()
    integer :: u
    call dump_i8('cell', cell, (nx+2)*(ny+2)*(nz+2))
    call dump_i8('interceptedIndexPtr', interceptedIndexPtr, int(ibCellCount, 8)*3)
    call dump_i8('fluidIndexPtr', fluidIndexPtr, int(fluidCellCount, 8)*3)
    call dump_i8('solidIndexPtr', solidIndexPtr, int(solidCellCount, 8)*3)
    call dump_i8('redCellIndexPtr', redCellIndexPtr, int(redCellCount, 8)*3)
    call dump_i8('blackCellIndexPtr', blackCellIndexPtr, int(blackCellCount, 8)*3)
    call dump_i8('nelp', nelp, int(ibCellCount, 8))
    call dump_r8('pNormDis', pNormDis, int(ibCellCount, 8))
    call dump_r8('u1NormDis', u1NormDis, int(ibCellCount, 8))
    call dump_r8('u2NormDis', u2NormDis, int(ibCellCount, 8))
    call dump_r8('v1NormDis', v1NormDis, int(ibCellCount, 8))
    call dump_r8('v2NormDis', v2NormDis, int(ibCellCount, 8))
    call dump_r8('w1NormDis', w1NormDis, int(ibCellCount, 8))
    call dump_r8('w2NormDis', w2NormDis, int(ibCellCount, 8))
    call dump_r8('Acx', Acx, nx*3)
    call dump_r8('Acy', Acy, ny*3)
    call dump_r8('Acz', Acz, nz*3)
    call dump_r8('ca_uu', [ca1_uu(2:nx), ca2_uu(2:nx), ca3_uu(2:nx), ca4_uu(2:nx), ca5_uu(2:nx), ca6_uu(2:nx), &
                 ck1_uu(2:nx), ck2_uu(2:nx), ck3_uu(2:nx), ck4_uu(2:nx), ck5_uu(2:nx), ck6_uu(2:nx)], 12*(nx-1))
    call dump_r8('ca_uv', [ca1_uv(2:nx), ca2_uv(2:nx), ca3_uv(2:nx), ca4_uv(2:nx), ca5_uv(2:nx), ca6_uv(2:nx), &
                 ck1_uv(2:nx), ck2_uv(2:nx), ck3_uv(2:nx), ck4_uv(2:nx), ck5_uv(2:nx), ck6_uv(2:nx)], 12*(nx-1))
    call dump_r8('ca_vv', [ca1_vv(2:ny), ca2_vv(2:ny), ca3_vv(2:ny), ca4_vv(2:ny), ca5_vv(2:ny), ca6_vv(2:ny), &
                 ck1_vv(2:ny), ck2_vv(2:ny), ck3_vv(2:ny), ck4_vv(2:ny), ck5_vv(2:ny), ck6_vv(2:ny)], 12*(ny-1))
    call dump_r8('ca_vu', [ca1_vu(2:ny), ca2_vu(2:ny), ca3_vu(2:ny), ca4_vu(2:ny), ca5_vu(2:ny), ca6_vu(2:ny), &
                 ck1_vu(2:ny), ck2_vu(2:ny), ck3_vu(2:ny), ck4_vu(2:ny), ck5_vu(2:ny), ck6_vu(2:ny)], 12*(ny-1))
    call dump_r8('ca_ww', [ca1_ww(2:nz), ca2_ww(2:nz), ca3_ww(2:nz), ca4_ww(2:nz), ca5_ww(2:nz), ca6_ww(2:nz), &
                 ck1_ww(2:nz), ck2_ww(2:nz), ck3_ww(2:nz), ck4_ww(2:nz), ck5_ww(2:nz), ck6_ww(2:nz)], 12*(nz-1))
    call dump_r8('ca_wu', [ca1_wu(2:nz), ca2_wu(2:nz), ca3_wu(2:nz), ca4_wu(2:nz), ca5_wu(2:nz), ca6_wu(2:nz), &
                 ck1_wu(2:nz), ck2_wu(2:nz), ck3_wu(2:nz), ck4_wu(2:nz), ck5_wu(2:nz), ck6_wu(2:nz)], 12*(nz-1))
    open(newunit=u, file=trim(dumpdir)//'/counts.txt', status='replace')
    write(u, '(9I12)') nx, ny, nz, ibCellCount, fluidCellCount, solidCellCount, &
                       redCellCount, blackCellCount, ibElems
    write(u, '(2ES25.17)') deltat, rev
    close(u)
  end subroutine

  subroutine dump_cells()
    call dump_i8('Elemcell', Elemcell, ibElems*3)
    call dump_i8('ucell', ucell, ibElems*3)
    call dump_i8('vcell', vcell, ibElems*3)
    call dump_i8('wcell', wcell, ibElems*3)
    call dump_i8('pcell', pcell, ibElems*3)
  end subroutine

  subroutine dump_move(s, tag, lists)
    integer(8), intent(in) :: s
    character(len=*), intent(in) :: tag
    logical, intent(in) :: lists
    character(len=32) :: pre
    write(pre, '(A,I4.4,A,A,A)') 'm', s, '_', tag, '_'
    if (.not. lists) then
      call dump_i8(trim(pre)//'cell', cell, (nx+2)*(ny+2)*(nz+2))
      call dump_i8(trim(pre)//'minElemcell', minElemcell, (nx+2)*(ny+2)*(nz+2))
      call dump_r8(trim(pre)//'ycent', ycent, ibElems)
      call dump_r8(trim(pre)//'cosBeta', cosBeta, ibElems)
      return
    end if
    call dump_i8(trim(pre)//'interceptedIndexPtr', interceptedIndexPtr, int(ibCellCount, 8)*3)
    call dump_i8(trim(pre)//'fluidIndexPtr', fluidIndexPtr, int(fluidCellCount, 8)*3)
    call dump_i8(trim(pre)//'solidIndexPtr', solidIndexPtr, int(solidCellCount, 8)*3)
    call dump_i8(trim(pre)//'redCellIndexPtr', redCellIndexPtr, int(redCellCount, 8)*3)
    call dump_i8(trim(pre)//'blackCellIndexPtr', blackCellIndexPtr, int(blackCellCount, 8)*3)
    call dump_i8(trim(pre)//'nelp', nelp, int(ibCellCount, 8))
    call dump_r8(trim(pre)//'pNormDis', pNormDis, int(ibCellCount, 8))
    call dump_r8(trim(pre)//'u1NormDis', u1NormDis, int(ibCellCount, 8))
    call dump_r8(trim(pre)//'u2NormDis', u2NormDis, int(ibCellCount, 8))
    call dump_r8(trim(pre)//'v1NormDis', v1NormDis, int(ibCellCount, 8))
    call dump_r8(trim(pre)//'v2NormDis', v2NormDis, int(ibCellCount, 8))
    call dump_r8(trim(pre)//'w1NormDis', w1NormDis, int(ibCellCount, 8))
    call dump_r8(trim(pre)//'w2NormDis', w2NormDis, int(ibCellCount, 8))
  end subroutine

  subroutine dump_stage(s, tag)
    integer(8), intent(in) :: s
    character(len=*), intent(in) :: tag
    character(len=32) :: pre
    integer(8) :: np
    np = (nx+2)*(ny+2)*(nz+2)
    write(pre, '(A,I4.4,A,A,A)') 's', s, '_', tag, '_'
    call dump_r8(trim(pre)//'ut', ut, np)
    call dump_r8(trim(pre)//'vt', vt, np)
    call dump_r8(trim(pre)//'wt', wt, np)
    call dump_r8(trim(pre)//'p', p, np)
    if (tag == 'mom') then
      call dump_r8(trim(pre)//'resi_u', resi_u, np)
      call dump_r8(trim(pre)//'resi_v', resi_v, np)
      call dump_r8(trim(pre)//'resi_w', resi_w, np)
    end if
    if (tag == 'poi') then
      call dump_r8(trim(pre)//'u', u, np)
      call dump_r8(trim(pre)//'v', v, np)
      call dump_r8(trim(pre)//'w', w, np)
      call dump_r8(trim(pre)//'pc', pc, np)
    end if
  end subroutine

end program ref_harness
