! ref_driver.F90 -- C-callable harness around the reference's own solver
! module (m_fldsolver, src/algorithms/field_solver.F90), compiled from the
! reference sources where they lie by oracle/Makefile into oracle/_ref/.
!
! TEST INFRASTRUCTURE ONLY: it lets tests/ check the C oracle against the
! reference's own initRKstep / rk3Step / cleanEpar and its output
! interpolation (interpFromEdges / interpFromFaces), built once with default
! REAL as shipped (FP32) and once with -fdefault-real-8 -fdefault-double-8.
! It sets the module state that src/logistics/initialize.F90:121-149,185-240
! would set for a single meshblock and moves arrays in and out; it adds no
! arithmetic of its own.
module m_refdriver
  use iso_c_binding
  use m_globalnamespace
  use m_domain
  use m_fields
  use m_fldsolver
  use m_helpers
  implicit none
contains
  integer(c_int) function ref_real_size() bind(C, name='ref_real_size')
    real :: x
    ref_real_size = int(storage_size(x) / 8, c_int)
  end function ref_real_size

  subroutine alloc3(fld)
    real, allocatable, intent(inout) :: fld(:, :, :)
    if (allocated(fld)) deallocate (fld)
    allocate (fld(-NGHOST:this_meshblock%ptr%sx - 1 + NGHOST, &
                  -NGHOST:this_meshblock%ptr%sy - 1 + NGHOST, &
                  -NGHOST:this_meshblock%ptr%sz - 1 + NGHOST))
    fld = 0.0
  end subroutine alloc3

  subroutine ref_setup(nx, ny, nz, bndx, bndy, bndz, absorb, cc_in) bind(C, name='ref_setup')
    integer(c_int), value :: nx, ny, nz, bndx, bndy, bndz, absorb
    real(c_double), value :: cc_in
    global_mesh%x0 = 0; global_mesh%y0 = 0; global_mesh%z0 = 0
    global_mesh%sx = nx; global_mesh%sy = ny; global_mesh%sz = nz
    if (allocated(meshblocks)) deallocate (meshblocks)
    allocate (meshblocks(1))
    meshblocks(1)%rnk = 0
    meshblocks(1)%x0 = 0; meshblocks(1)%y0 = 0; meshblocks(1)%z0 = 0
    meshblocks(1)%sx = nx; meshblocks(1)%sy = ny; meshblocks(1)%sz = nz
    this_meshblock%ptr => meshblocks(1)
    boundary_x = bndx; boundary_y = bndy; boundary_z = bndz
    if ((bndx == 2) .or. (bndy == 2) .or. (bndz == 2)) then
      boundary_x = 2; boundary_y = 2; boundary_z = 2
    end if
    absorb_buff = absorb
    CC = cc_in
    CCINV = 1.0 / CC
    call alloc3(ex); call alloc3(ey); call alloc3(ez)
    call alloc3(bx); call alloc3(by); call alloc3(bz)
    call alloc3(b0x); call alloc3(b0y); call alloc3(b0z)
    call alloc3(dex); call alloc3(dey); call alloc3(dez)
    call alloc3(dbx); call alloc3(dby); call alloc3(dbz)
    call alloc3(enx); call alloc3(eny); call alloc3(enz)
    call alloc3(bnx); call alloc3(bny); call alloc3(bnz)
    call alloc3(rho)
  end subroutine ref_setup

  ! field ids follow the oracle's order: ex ey ez bx by bz b0x b0y b0z
  ! enx eny enz bnx bny bnz dex dey dez dbx dby dbz rho
  subroutine ref_xfer(which, p, to_ref) bind(C, name='ref_xfer')
    integer(c_int), value :: which, to_ref
    type(c_ptr), value :: p
    real, pointer :: a(:, :, :)
    integer :: n(3)
    n = [this_meshblock%ptr%sx + 2 * NGHOST, this_meshblock%ptr%sy + 2 * NGHOST, &
         this_meshblock%ptr%sz + 2 * NGHOST]
    call c_f_pointer(p, a, n)
    select case (which)
    case (0); call mv(ex)
    case (1); call mv(ey)
    case (2); call mv(ez)
    case (3); call mv(bx)
    case (4); call mv(by)
    case (5); call mv(bz)
    case (6); call mv(b0x)
    case (7); call mv(b0y)
    case (8); call mv(b0z)
    case (9); call mv(enx)
    case (10); call mv(eny)
    case (11); call mv(enz)
    case (12); call mv(bnx)
    case (13); call mv(bny)
    case (14); call mv(bnz)
    case (15); call mv(dex)
    case (16); call mv(dey)
    case (17); call mv(dez)
    case (18); call mv(dbx)
    case (19); call mv(dby)
    case (20); call mv(dbz)
    case (21); call mv(rho)
    end select
  contains
    subroutine mv(f)
      real, intent(inout) :: f(:, :, :)
      if (to_ref /= 0) then
        f = a
      else
        a = f
      end if
    end subroutine mv
  end subroutine ref_xfer

  subroutine ref_init_rk_step() bind(C, name='ref_init_rk_step')
    call initRKstep()
  end subroutine ref_init_rk_step

  subroutine ref_rk3_step(c1, c2, c3) bind(C, name='ref_rk3_step')
    real(c_double), value :: c1, c2, c3
    real :: r1, r2, r3
    r1 = c1; r2 = c2; r3 = c3
    call rk3Step(r1, r2, r3)
  end subroutine ref_rk3_step

  subroutine ref_clean_epar() bind(C, name='ref_clean_epar')
    call cleanEpar()
  end subroutine ref_clean_epar

  ! Harness ghost fill for one meshblock (the reference's MPI exchange,
  ! src/logistics/exchange_fields.F90, cannot be built here): a ghost cell
  ! whose out-of-range axes are all periodic takes the wrapped interior value.
  subroutine ref_exchange() bind(C, name='ref_exchange')
    integer :: i, j, k, si, sj, sk, nx, ny, nz
    logical :: ox, oy, oz
    nx = this_meshblock%ptr%sx; ny = this_meshblock%ptr%sy; nz = this_meshblock%ptr%sz
    do k = -NGHOST, nz - 1 + NGHOST
      oz = (k < 0) .or. (k >= nz)
      if (oz .and. boundary_z /= 1) cycle
      sk = modulo(k, nz)
      do j = -NGHOST, ny - 1 + NGHOST
        oy = (j < 0) .or. (j >= ny)
        if (oy .and. boundary_y /= 1) cycle
        sj = modulo(j, ny)
        do i = -NGHOST, nx - 1 + NGHOST
          ox = (i < 0) .or. (i >= nx)
          if (ox .and. boundary_x /= 1) cycle
          if (.not. (ox .or. oy .or. oz)) cycle
          si = modulo(i, nx)
          ex(i, j, k) = ex(si, sj, sk); ey(i, j, k) = ey(si, sj, sk); ez(i, j, k) = ez(si, sj, sk)
          bx(i, j, k) = bx(si, sj, sk); by(i, j, k) = by(si, sj, sk); bz(i, j, k) = bz(si, sj, sk)
        end do
      end do
    end do
  end subroutine ref_exchange

  ! The reference's interpFromEdges (E) and interpFromFaces (B) at offset
  ! (dx, dy, dz) from cell (i, j, k) of the module arrays
  ! (src/tools/helpers.F90:93-257): out(1:3) = E, out(4:6) = B, widened to
  ! double (exactly).
  subroutine ref_interp(dx, dy, dz, i, j, k, out) bind(C, name='ref_interp')
    real(c_double), value :: dx, dy, dz
    integer(c_int), value :: i, j, k
    real(c_double), intent(out) :: out(6)
    real :: rx, ry, rz, a, b, c
    rx = dx; ry = dy; rz = dz
    call interpFromEdges(rx, ry, rz, i, j, k, ex, ey, ez, a, b, c)
    out(1) = a; out(2) = b; out(3) = c
    call interpFromFaces(rx, ry, rz, i, j, k, bx, by, bz, a, b, c)
    out(4) = a; out(5) = b; out(6) = c
  end subroutine ref_interp

  ! nsteps timesteps in the order of src/tristanmainloop.F90:50-83 for a
  ! problem without a user boundary hook (used to time the reference).
  subroutine ref_run(nsteps) bind(C, name='ref_run')
    integer(c_int), value :: nsteps
    integer :: n
    do n = 1, nsteps
      call initRKstep()
      call rk3Step(1.0, 0.0, 1.0)
      call ref_exchange(); call cleanEpar(); call ref_exchange()
      call rk3Step(3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0)
      call ref_exchange(); call cleanEpar(); call ref_exchange()
      call rk3Step(1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0)
      call ref_exchange(); call cleanEpar(); call ref_exchange()
    end do
  end subroutine ref_run
end module m_refdriver
