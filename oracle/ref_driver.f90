!-----------------------------------------------------------------------
! Test-only driver around the flang-built COMCOT reference Fortran.
!
! This file is part of the ORACLE (test infrastructure): it is linked
! with objects that oracle/Makefile compiles from the reference sources
! where they lie, into oracle/_ref/libcomcot_ref.so.  Only tests/,
! tests/golden/make_golden.py and bench.py's cpu_baseline leg load it,
! as a checker and as the timed CPU baseline; the product never does.
!
! It replays the CPU path of the reference main loop (comcot.f90:322-440
! with the DEBUG_CORE body of CHANGE, comcot.f90:480-505) on layers it
! builds itself, so that the reference's own MASS/OPEN/ALL_GRID/MOMENT
! code computes every value:
!   setup      ALLOC (comcot.f90:516), PQ_DEPTH (initialization.f90:2501),
!              PARA (comcot.f90:694)
!   one step   MASS (mass.f90:3) -> OPEN (boundaries.f90:22)
!              -> ALL_GRID (all_grids.f90:4) -> MOMENT (moment.f90:3)
!              -> level-2 to level-1 copy of Z, M, N (comcot.f90:487-489)
! Single routines (JNQ, NEWQ, JNZ, UPDATE, ...) can be called one at a
! time for step-local fixtures, like the reference's DEBUG_CORE mode.
!-----------------------------------------------------------------------
module ref_state
  use layer_params
  use fault_params
  use wave_params
  implicit none
  type(layer), save :: lo
  type(layer), allocatable, save :: la(:)
  integer, save :: nla = 0
  ! the maximum that the GPU half keeps for the outermost layer and that
  ! MAX_AMP reads back with CUDA_GETZMAX: starts at the initial elevation
  ! (GPUcomcot.cu:79), raised after each root step by MAXAMP_LAUNCH
  real, allocatable, save :: zmaxr(:,:)
  ! BC_TYPE = 1: SPONGE_LAYER on the outermost layer after MOMENT
  logical, save :: sponge_on = .false.
contains


!.....resolve a layer id (1 = root, k+1 = la(k)); stops on a bad id
subroutine ref_pick(id, k)
  implicit none
  integer :: id, k
  if (id .eq. 1) then
     k = 0
  else if (id .ge. 2 .and. id-1 .le. nla) then
     k = id - 1
  else
     write(*,*) 'ref_driver: bad layer id', id
     stop 1
  end if
end subroutine ref_pick

subroutine ref_reset() bind(C, name="ref_reset")
  implicit none
  integer :: i
  integer :: ngrid
  real :: elmax, grav, pi, r_earth, gx, eps, zero, one, v_limit
  real :: rad_deg, rad_min
  integer :: num_flt
  common /cons/ elmax, grav, pi, r_earth, gx, eps, zero, one, ngrid,  &
                num_flt, v_limit, rad_deg, rad_min
  if (allocated(la)) deallocate(la)
  allocate(la(ngrid))
  do i = 1, ngrid
     la(i)%layswitch = 1
     la(i)%parent = -99
     la(i)%id = i + 1
     la(i)%num_child = 0
  end do
  nla = 0
  lo%num_child = 0
end subroutine ref_reset

!.....common per-layer fields; arrays Y/DEL_Y/X are only used by PARA
subroutine ref_fill(l, nx, ny, laycord, laygov, dx, dy, dt, y, dely,    &
                    fric_switch, fric_coef)
  implicit none
  type(layer) :: l
  integer :: nx, ny, laycord, laygov, fric_switch
  real :: dx, dy, dt, fric_coef
  real :: y(ny), dely(ny)
  l%nx = nx
  l%ny = ny
  l%laycord = laycord
  l%laygov = laygov
  l%layswitch = 0
  l%dx = dx
  l%dy = dy
  l%dt = dt
  l%ini_switch = 0
  l%bc_type = 0
  l%dim = 2
  l%fric_switch = fric_switch
  l%fric_coef = fric_coef
  l%h_limit = 0.0
  l%tide_level = 0.0
  l%sc_option = 0
  l%num_child = 0
  allocate(l%y(ny))
  allocate(l%del_y(ny))
  allocate(l%x(nx))
  l%y = y
  l%del_y = dely
  l%x = 0.0
end subroutine ref_fill

subroutine ref_init_root(nx, ny, laycord, laygov, dx, dy, dt, y, dely,  &
                         h, z0, fric_switch, fric_coef)               &
                         bind(C, name="ref_init_root")
  implicit none
  integer :: nx, ny, laycord, laygov, fric_switch
  real :: dx, dy, dt, fric_coef
  real :: y(ny), dely(ny), h(nx*ny), z0(nx*ny)
  call ref_reset()
  sponge_on = .false.
  lo%id = 1
  lo%parent = -1
  lo%level = 1
  lo%rel_size = 1
  lo%rel_time = 1
  call ref_fill(lo, nx, ny, laycord, laygov, dx, dy, dt, y, dely,       &
                fric_switch, fric_coef)
  call alloc(lo, 1)
  lo%h = reshape(h, (/nx, ny/))
  lo%z(:,:,1) = reshape(z0, (/nx, ny/))
  if (allocated(zmaxr)) deallocate(zmaxr)
  allocate(zmaxr(nx, ny))
  zmaxr = lo%z(:,:,1)
  call pq_depth(lo)
  if (laycord .eq. 0) call para(lo)
end subroutine ref_init_root

subroutine ref_init_child(parent, laycord, laygov, rel_size, rel_time,  &
                          corners, nx, ny, dx, dy, dt, y, dely, h, z0,  &
                          fric_switch, fric_coef)                       &
                          bind(C, name="ref_init_child")
  implicit none
  integer :: parent, laycord, laygov, rel_size, rel_time, nx, ny
  integer :: corners(4), fric_switch
  real :: dx, dy, dt, fric_coef
  real :: y(ny), dely(ny), h(nx*ny), z0(nx*ny)
  integer :: kp
  nla = nla + 1
  call ref_pick(parent, kp)
  la(nla)%id = nla + 1
  la(nla)%parent = parent
  if (kp .eq. 0) then
     la(nla)%level = lo%level + 1
     lo%num_child = lo%num_child + 1
  else
     la(nla)%level = la(kp)%level + 1
     la(kp)%num_child = la(kp)%num_child + 1
  end if
  la(nla)%rel_size = rel_size
  la(nla)%rel_time = rel_time
  call ref_fill(la(nla), nx, ny, laycord, laygov, dx, dy, dt, y, dely,  &
                fric_switch, fric_coef)
  call alloc(la(nla), 2)
  la(nla)%corners = corners
  la(nla)%h = reshape(h, (/nx, ny/))
  la(nla)%z(:,:,1) = reshape(z0, (/nx, ny/))
  call pq_depth(la(nla))
  if (laycord .eq. 0) call para(la(nla))
end subroutine ref_init_child

!.....nsteps root steps of the reference CPU time loop
subroutine ref_step(nsteps) bind(C, name="ref_step")
  implicit none
  integer :: nsteps
  integer :: k
  do k = 1, nsteps
     call mass(lo)
     call open(lo)
     if (lo%num_child .ge. 1) call all_grid(lo, la)
     call moment(lo)
     if (sponge_on) call sponge_layer(lo)
     lo%z(:,:,1) = lo%z(:,:,2)
     lo%m(:,:,1) = lo%m(:,:,2)
     lo%n(:,:,1) = lo%n(:,:,2)
  end do
end subroutine ref_step

!.....one reference routine on one layer (step-local fixtures)
!     op: 1 MASS, 2 OPEN, 3 MOMENT, 4 JNQ(parent,id), 5 NEWQ(parent,id,t)
!         6 JNZ(parent,id), 7 UPDATE(id), 8 ALL_GRID(root), 9 root copy
subroutine ref_call(op, id, t) bind(C, name="ref_call")
  implicit none
  integer :: op, id, t
  integer :: k, kp
  call ref_pick(id, k)
  select case (op)
  case (1)
     if (k .eq. 0) then
        call mass(lo)
     else
        call mass(la(k))
     end if
  case (2)
     call open(lo)
  case (3)
     if (k .eq. 0) then
        call moment(lo)
     else
        call moment(la(k))
     end if
  case (4, 5, 6)
     call ref_pick(la(k)%parent, kp)
     if (kp .eq. 0) then
        if (op .eq. 4) call jnq(lo, la(k))
        if (op .eq. 5) call newq(lo, la(k), t)
        if (op .eq. 6) call jnz(lo, la(k))
     else
        if (op .eq. 4) call jnq(la(kp), la(k))
        if (op .eq. 5) call newq(la(kp), la(k), t)
        if (op .eq. 6) call jnz(la(kp), la(k))
     end if
  case (7)
     call update(la(k))
  case (8)
     call all_grid(lo, la)
  case (9)
     lo%z(:,:,1) = lo%z(:,:,2)
     lo%m(:,:,1) = lo%m(:,:,2)
     lo%n(:,:,1) = lo%n(:,:,2)
  case default
     write(*,*) 'ref_driver: bad op', op
     stop 1
  end select
end subroutine ref_call

!.....copy one layer's state out (which: 0 get, 1 set)
!     planes: z(:,:,1:2), m(:,:,1:2), n(:,:,1:2), dz(:,:,1:2), each nx*ny
subroutine ref_state_io(id, which, z, m, n, dz) bind(C, name="ref_state_io")
  implicit none
  integer :: id, which
  real :: z(*), m(*), n(*), dz(*)
  integer :: k, nxy
  call ref_pick(id, k)
  if (k .eq. 0) then
     nxy = lo%nx*lo%ny
     if (which .eq. 0) then
        z(1:2*nxy) = reshape(lo%z, (/2*nxy/))
        m(1:2*nxy) = reshape(lo%m, (/2*nxy/))
        n(1:2*nxy) = reshape(lo%n, (/2*nxy/))
        dz(1:2*nxy) = reshape(lo%dz, (/2*nxy/))
     else
        lo%z = reshape(z(1:2*nxy), (/lo%nx, lo%ny, 2/))
        lo%m = reshape(m(1:2*nxy), (/lo%nx, lo%ny, 2/))
        lo%n = reshape(n(1:2*nxy), (/lo%nx, lo%ny, 2/))
        lo%dz = reshape(dz(1:2*nxy), (/lo%nx, lo%ny, 2/))
     end if
  else
     nxy = la(k)%nx*la(k)%ny
     if (which .eq. 0) then
        z(1:2*nxy) = reshape(la(k)%z, (/2*nxy/))
        m(1:2*nxy) = reshape(la(k)%m, (/2*nxy/))
        n(1:2*nxy) = reshape(la(k)%n, (/2*nxy/))
        dz(1:2*nxy) = reshape(la(k)%dz, (/2*nxy/))
     else
        la(k)%z = reshape(z(1:2*nxy), (/la(k)%nx, la(k)%ny, 2/))
        la(k)%m = reshape(m(1:2*nxy), (/la(k)%nx, la(k)%ny, 2/))
        la(k)%n = reshape(n(1:2*nxy), (/la(k)%nx, la(k)%ny, 2/))
        la(k)%dz = reshape(dz(1:2*nxy), (/la(k)%nx, la(k)%ny, 2/))
     end if
  end if
end subroutine ref_state_io

!.....setup products of one layer: HP, HQ (planes), the spherical
!     coefficient planes R0,R1,R11,R2,R21,R22,R3,R4,R5 (9 planes, zero for
!     Cartesian layers), R6 (ny), and RX, RY, GRX, GRY, DT (5 scalars)
subroutine ref_setup_io(id, hp, hq, rpl, r6, sc) bind(C, name="ref_setup_io")
  implicit none
  integer :: id
  real :: hp(*), hq(*), rpl(*), r6(*), sc(5)
  integer :: k
  call ref_pick(id, k)
  if (k .eq. 0) then
     call ref_setup_one(lo, hp, hq, rpl, r6, sc)
  else
     call ref_setup_one(la(k), hp, hq, rpl, r6, sc)
  end if
end subroutine ref_setup_io

subroutine ref_setup_one(l, hp, hq, rpl, r6, sc)
  implicit none
  type(layer) :: l
  real :: hp(*), hq(*), rpl(*), r6(*), sc(5)
  integer :: nxy
  nxy = l%nx*l%ny
  hp(1:nxy) = reshape(l%hp, (/nxy/))
  hq(1:nxy) = reshape(l%hq, (/nxy/))
  rpl(1:9*nxy) = 0.0
  r6(1:l%ny) = 0.0
  if (l%laycord .eq. 0) then
     rpl(1:nxy) = reshape(l%r0, (/nxy/))
     rpl(nxy+1:2*nxy) = reshape(l%r1, (/nxy/))
     rpl(2*nxy+1:3*nxy) = reshape(l%r11, (/nxy/))
     if (l%laygov .le. 1) rpl(3*nxy+1:4*nxy) = reshape(l%r2, (/nxy/))
     rpl(4*nxy+1:5*nxy) = reshape(l%r21, (/nxy/))
     rpl(5*nxy+1:6*nxy) = reshape(l%r22, (/nxy/))
     rpl(6*nxy+1:7*nxy) = reshape(l%r3, (/nxy/))
     if (l%laygov .le. 1) rpl(7*nxy+1:8*nxy) = reshape(l%r4, (/nxy/))
     rpl(8*nxy+1:9*nxy) = reshape(l%r5, (/nxy/))
     r6(1:l%ny) = l%r6
  end if
  sc(1) = l%rx
  sc(2) = l%ry
  sc(3) = l%grx
  sc(4) = l%gry
  sc(5) = l%dt
end subroutine ref_setup_one

!.....parent-side interface flux cache written by NEWQ (all_grids.f90:720,790)
subroutine ref_flux_io(id, xflux, yflux) bind(C, name="ref_flux_io")
  implicit none
  integer :: id
  real :: xflux(*), yflux(*)
  integer :: k
  call ref_pick(id, k)
  if (k .eq. 0) then
     xflux(1:2*lo%nx) = reshape(lo%xflux, (/2*lo%nx/))
     yflux(1:2*lo%ny) = reshape(lo%yflux, (/2*lo%ny/))
  else
     xflux(1:2*la(k)%nx) = reshape(la(k)%xflux, (/2*la(k)%nx/))
     yflux(1:2*la(k)%ny) = reshape(la(k)%yflux, (/2*la(k)%ny/))
  end if
end subroutine ref_flux_io

!.....time-step check of the reference (initialization.f90:1136-1307):
!     returns the adjusted root DT and each child's REL_TIME and DT
subroutine ref_cr_check(dt_root, rel_time, dt_child) bind(C, name="ref_cr_check")
  implicit none
  real :: dt_root, dt_child(*)
  integer :: rel_time(*)
  integer :: i
  call cr_check(lo, la)
  dt_root = lo%dt
  do i = 1, nla
     rel_time(i) = la(i)%rel_time
     dt_child(i) = la(i)%dt
  end do
end subroutine ref_cr_check
!.....grid coordinates X(NX), Y(NY) of a layer (what GET_TS searches)
subroutine ref_set_coords(id, x, y) bind(C, name="ref_set_coords")
  implicit none
  integer :: id
  real :: x(*), y(*)
  integer :: k
  call ref_pick(id, k)
  if (k .eq. 0) then
     lo%x(1:lo%nx) = x(1:lo%nx)
     lo%y(1:lo%ny) = y(1:lo%ny)
  else
     la(k)%x(1:la(k)%nx) = x(1:la(k)%nx)
     la(k)%y(1:la(k)%ny) = y(1:la(k)%ny)
  end if
end subroutine ref_set_coords

!.....the reference's gauge sample (output.f90:1035-1155) of the current
!     level-2 elevation of layer id at (x0, y0)
subroutine ref_get_ts(x0, y0, id, dat) bind(C, name="ref_get_ts")
  implicit none
  real :: x0, y0, dat
  integer :: id
  call get_ts(dat, x0, y0, lo, la, id)
end subroutine ref_get_ts

!.....MAXAMP_LAUNCH's recorder (GPUOutput.cu:46-53) on the level-2
!     elevation of the outermost layer: a value that is not below the
!     maximum leaves it, anything else (a NaN too) replaces it
subroutine ref_maxamp() bind(C, name="ref_maxamp")
  implicit none
  where (.not. (zmaxr .ge. lo%z(:,:,2))) zmaxr = lo%z(:,:,2)
end subroutine ref_maxamp

!.....the reference's snapshot writer DATA_PRT (output.f90:758-851) for
!     layer id at step k, into the working directory
subroutine ref_data_prt(k, id, fluxswitch, tide) bind(C, name="ref_data_prt")
  implicit none
  integer :: k, id, fluxswitch
  real :: tide
  integer :: j
  call ref_pick(id, j)
  if (j .eq. 0) then
     lo%fluxswitch = fluxswitch
     lo%tide_level = tide
     lo%sedi_switch = 1
     call data_prt(k, lo)
  else
     la(j)%fluxswitch = fluxswitch
     la(j)%tide_level = tide
     la(j)%sedi_switch = 1
     call data_prt(k, la(j))
  end if
end subroutine ref_data_prt

!.....the reference's maximum writer MAX_AMP (output.f90:855-1031) at model
!     time `time` of a run ending at `tend`, into the working directory
subroutine ref_max_amp(time, tend, tide) bind(C, name="ref_max_amp")
  implicit none
  real :: time, tend, tide
  lo%tide_level = tide
  call max_amp(lo, la, time, tend)
end subroutine ref_max_amp

!.....one FAULT record from 11 REALs (HH, L, W, D, TH, DL, RD, YO, XO, Y0,
!     X0) and the epicentre switch (0 centre, 9 top edge)
subroutine ref_fault_fill(f, params, fswitch)
  implicit none
  type(fault) :: f
  real :: params(11)
  integer :: fswitch
  f%hh = params(1)
  f%l = params(2)
  f%w = params(3)
  f%d = params(4)
  f%th = params(5)
  f%dl = params(6)
  f%rd = params(7)
  f%yo = params(8)
  f%xo = params(9)
  f%y0 = params(10)
  f%x0 = params(11)
  f%t0 = 0.0
  f%switch = fswitch
  f%num_flt = 1
  f%fs = 9          ! no WRITE_DEFORM file
end subroutine ref_fault_fill

!.....the reference's DEFORM_OKADA (deform.f90:74-191) on the outermost
!     layer alone: LO%DEFORM, nothing applied
subroutine ref_okada(params, fswitch) bind(C, name="ref_okada")
  implicit none
  real :: params(11)
  integer :: fswitch
  type(fault) :: f
  call ref_fault_fill(f, params, fswitch)
  call deform_okada(lo, f)
end subroutine ref_okada

!.....the reference's GET_FLOOR_DEFORM (deform.f90:2-71) for one fault
!     rupturing at T = 0: DEFORM_OKADA, ININTERP to every nested layer,
!     UPDATE_BATH, and the deformation added to every level-1 elevation
subroutine ref_floor_deform(params, fswitch) bind(C, name="ref_floor_deform")
  implicit none
  real :: params(11)
  integer :: fswitch
  type(fault), allocatable :: flt(:)
  real :: time
  integer :: nflt
  real :: elmax, grav, pi, r_earth, gx, eps, zero, one, v_limit, rad_deg, rad_min
  integer :: ngrid
  common /cons/ elmax, grav, pi, r_earth, gx, eps, zero, one, ngrid,  &
                nflt, v_limit, rad_deg, rad_min
  allocate(flt(nflt))
  call ref_fault_fill(flt(1), params, fswitch)
  time = 0.0
  call get_floor_deform(lo, la, flt, time)
  deallocate(flt)
end subroutine ref_floor_deform

!.....the reference's SPONGE_COEF (boundaries.f90:250-330) on layer 1;
!     on /= 0 also runs SPONGE_LAYER after every MOMENT of ref_step
!     (comcot.f90:403-406).  cx, cy receive COEFX, COEFY (NX*NY each).
subroutine ref_sponge(on, cx, cy) bind(C, name="ref_sponge")
  implicit none
  integer :: on
  real :: cx(*), cy(*)
  integer :: n
  n = lo%nx * lo%ny
  if (associated(lo%sponge_coefx)) deallocate(lo%sponge_coefx)
  if (associated(lo%sponge_coefy)) deallocate(lo%sponge_coefy)
  allocate(lo%sponge_coefx(lo%nx, lo%ny), lo%sponge_coefy(lo%nx, lo%ny))
  call sponge_coef(lo)
  cx(1:n) = reshape(lo%sponge_coefx, (/n/))
  cy(1:n) = reshape(lo%sponge_coefy, (/n/))
  sponge_on = on .ne. 0
end subroutine ref_sponge

!.....the reference's BC_WALL (boundaries.f90:369-401) on layer 1, without a
!     wave maker (INI_SWITCH /= 2)
subroutine ref_bc_wall() bind(C, name="ref_bc_wall")
  implicit none
  type(wave) :: w
  call bc_wall(lo, w)
end subroutine ref_bc_wall

!.....a Cartesian child of the spherical layer 1 laid out by the reference's
!     own SUBGRID_MATCHING (initialization.f90:882-1131) from its extent in
!     degrees (X_START, X_END, Y_START, Y_END), with SC_OPTION = 1 (the
!     coupling the reference selects for such a child, initialization.f90:
!     362-370): grid size and spacing, POS/CXY and the corners it sets
subroutine ref_sc_child(laygov, rel_size, rel_time, xyse, dt, fric_switch, fric_coef,  &
                        nx, ny, dx, dy, corners) bind(C, name="ref_sc_child")
  implicit none
  integer :: laygov, rel_size, rel_time, fric_switch, nx, ny, corners(4)
  real :: xyse(4), dt, fric_coef, dx, dy
  nla = nla + 1
  associate (c => la(nla))
  c%id = nla + 1
  c%parent = 1
  c%level = 2
  lo%num_child = lo%num_child + 1
  c%rel_size = rel_size
  c%rel_time = rel_time
  c%laycord = 1
  c%laygov = laygov
  c%layswitch = 0
  c%dt = dt
  c%ini_switch = 0
  c%bc_type = 0
  c%dim = 2
  c%fric_switch = fric_switch
  c%fric_coef = fric_coef
  c%h_limit = 0.0
  c%tide_level = 0.0
  c%num_child = 0
  c%upz = .true.
  c%sc_option = 1
  c%x_start = xyse(1)
  c%x_end = xyse(2)
  c%y_start = xyse(3)
  c%y_end = xyse(4)
  call subgrid_matching(lo, c)
  call alloc(c, 2)
  nx = c%nx
  ny = c%ny
  dx = c%dx
  dy = c%dy
  corners = c%corners
  end associate
end subroutine ref_sc_child

!.....POS (NX*NY*2), CXY (NX*NY*4), X (NX), Y (NY) of an SC_OPTION = 1 child
subroutine ref_sc_io(id, pos, cxy, x, y) bind(C, name="ref_sc_io")
  implicit none
  integer :: id, pos(*)
  real :: cxy(*), x(*), y(*)
  integer :: k, n
  call ref_pick(id, k)
  n = la(k)%nx * la(k)%ny
  pos(1:2*n) = reshape(la(k)%pos, (/2*n/))
  cxy(1:4*n) = reshape(la(k)%cxy, (/4*n/))
  x(1:la(k)%nx) = la(k)%x
  y(1:la(k)%ny) = la(k)%y
end subroutine ref_sc_io

!.....depth and initial elevation of a layer (then PQ_DEPTH)
subroutine ref_set_hz(id, h, z0) bind(C, name="ref_set_hz")
  implicit none
  integer :: id
  real :: h(*), z0(*)
  integer :: k
  call ref_pick(id, k)
  if (k .eq. 0) then
     lo%h = reshape(h(1:lo%nx*lo%ny), (/lo%nx, lo%ny/))
     lo%z(:,:,1) = reshape(z0(1:lo%nx*lo%ny), (/lo%nx, lo%ny/))
     call pq_depth(lo)
  else
     la(k)%h = reshape(h(1:la(k)%nx*la(k)%ny), (/la(k)%nx, la(k)%ny/))
     la(k)%z(:,:,1) = reshape(z0(1:la(k)%nx*la(k)%ny), (/la(k)%nx, la(k)%ny/))
     call pq_depth(la(k))
  end if
end subroutine ref_set_hz

!.....read DEFORM and H of layer id
subroutine ref_deform_io(id, deform, h) bind(C, name="ref_deform_io")
  implicit none
  integer :: id
  real :: deform(*), h(*)
  integer :: k, n
  call ref_pick(id, k)
  if (k .eq. 0) then
     n = lo%nx * lo%ny
     deform(1:n) = reshape(lo%deform, (/n/))
     h(1:n) = reshape(lo%h, (/n/))
  else
     n = la(k)%nx * la(k)%ny
     deform(1:n) = reshape(la(k)%deform, (/n/))
     h(1:n) = reshape(la(k)%h, (/n/))
  end if
end subroutine ref_deform_io
end module ref_state

!.....the CPU side of GCOMCOT_GET_Z (GPUOutput.cu:29-44) that GET_TS calls:
!     level-2 Z at (i,j), (i+1,j), (i,j+1), (i+1,j+1) of layer lid
subroutine ref_get_z(lid, l1, l2, l3, l4, i, j) bind(C, name="gcomcot_get_z_")
  use ref_state
  implicit none
  integer :: lid, i, j, k
  real :: l1, l2, l3, l4
  call ref_pick(lid, k)
  if (k .eq. 0) then
     l1 = lo%z(i, j, 2)
     l2 = lo%z(i + 1, j, 2)
     l3 = lo%z(i, j + 1, 2)
     l4 = lo%z(i + 1, j + 1, 2)
  else
     l1 = la(k)%z(i, j, 2)
     l2 = la(k)%z(i + 1, j, 2)
     l3 = la(k)%z(i, j + 1, 2)
     l4 = la(k)%z(i + 1, j + 1, 2)
  end if
end subroutine ref_get_z

!.....CUDA_GETZ / CUDA_GETMN / CUDA_GETZMAX (GPUOutput.cu:65-103) as DATA_PRT
!     and MAX_AMP call them: the level-2 elevation and fluxes of a layer
!     (what the GPU's Zout/MNout hold before CHANGE), and the recorder's
!     maximum of the outermost layer
subroutine ref_cuda_getz(z, lid) bind(C, name="cuda_getz_")
  use ref_state
  implicit none
  integer :: lid, k
  real :: z(*)
  call ref_pick(lid, k)
  if (k .eq. 0) then
     z(1:lo%nx*lo%ny) = reshape(lo%z(:,:,2), (/lo%nx*lo%ny/))
  else
     z(1:la(k)%nx*la(k)%ny) = reshape(la(k)%z(:,:,2), (/la(k)%nx*la(k)%ny/))
  end if
end subroutine ref_cuda_getz

subroutine ref_cuda_getmn(m, n, lid) bind(C, name="cuda_getmn_")
  use ref_state
  implicit none
  integer :: lid, k
  real :: m(*), n(*)
  call ref_pick(lid, k)
  if (k .eq. 0) then
     m(1:lo%nx*lo%ny) = reshape(lo%m(:,:,2), (/lo%nx*lo%ny/))
     n(1:lo%nx*lo%ny) = reshape(lo%n(:,:,2), (/lo%nx*lo%ny/))
  else
     m(1:la(k)%nx*la(k)%ny) = reshape(la(k)%m(:,:,2), (/la(k)%nx*la(k)%ny/))
     n(1:la(k)%nx*la(k)%ny) = reshape(la(k)%n(:,:,2), (/la(k)%nx*la(k)%ny/))
  end if
end subroutine ref_cuda_getmn

subroutine ref_cuda_getzmax(zmax, lid) bind(C, name="cuda_getzmax_")
  use ref_state
  implicit none
  integer :: lid
  real :: zmax(*)
  if (lid .ne. 1) then
     write(*,*) 'ref_driver: the recorder maximum is kept for layer 1 only'
     stop 1
  end if
  zmax(1:lo%nx*lo%ny) = reshape(zmaxr, (/lo%nx*lo%ny/))
end subroutine ref_cuda_getzmax

!.....HOT_START (hotstart.f90:3-25) on the driver's layers, in the current
!     working directory: START_TYPE 0 writes the snapshot files of every
!     layer, 1 loads layer 1 and 2 every layer
subroutine ref_hot_start(start_type, start_step) bind(C, name="ref_hot_start")
  use ref_state
  implicit none
  integer :: start_type, start_step
  call hot_start(lo, la, start_type, start_step)
end subroutine ref_hot_start

!.....Z_MAX and Z_MIN of a layer, in (which = 1) or out (which = 0)
subroutine ref_zmaxmin_io(id, which, zmax, zmin) bind(C, name="ref_zmaxmin_io")
  use ref_state
  implicit none
  integer :: id, which, k, nx, ny
  real :: zmax(*), zmin(*)
  call ref_pick(id, k)
  if (k .eq. 0) then
     nx = lo%nx
     ny = lo%ny
     if (which .eq. 1) then
        lo%z_max = reshape(zmax(1:nx*ny), (/nx, ny/))
        lo%z_min = reshape(zmin(1:nx*ny), (/nx, ny/))
     else
        zmax(1:nx*ny) = reshape(lo%z_max, (/nx*ny/))
        zmin(1:nx*ny) = reshape(lo%z_min, (/nx*ny/))
     end if
  else
     nx = la(k)%nx
     ny = la(k)%ny
     if (which .eq. 1) then
        la(k)%z_max = reshape(zmax(1:nx*ny), (/nx, ny/))
        la(k)%z_min = reshape(zmin(1:nx*ny), (/nx, ny/))
     else
        zmax(1:nx*ny) = reshape(la(k)%z_max, (/nx*ny/))
        zmin(1:nx*ny) = reshape(la(k)%z_min, (/nx*ny/))
     end if
  end if
end subroutine ref_zmaxmin_io
