! user_psr.F90 -- rotating dipole inside a conducting sphere
! (user/user_psr.F90:26-69); the sphere hook of user_psr.F90:72-332 runs on
! the device (TFF_BC_PSR).  Use absorbing boundaries (boundary_* = 2).
#define NGHOST 2
module m_userfile
  use m_globalnamespace
  use m_readinput
  use m_domain
  use m_fields
  use m_device
  implicit none
  real(dprec), parameter, private :: PSR_PI = 3.14159265d0  ! user_psr.F90:11
  real(dprec), private :: xc_g, yc_g, zc_g, psr_angle, psr_period, psr_omega, psr_radius
contains
  subroutine userInitialize()
    user_output_kind = TFF_USEROUT_RHO  ! userOutput, user_psr.F90:386-398
    call userReadInput()
    call userInitFields()
    call tffCheck(tff_set_bc(tff_h, TFF_BC_PSR, [psr_radius, psr_angle, psr_omega, xc_g, yc_g, &
                                                 zc_g], 6), 'tff_set_bc')
  end subroutine userInitialize

  subroutine userReadInput()
    call getInput('problem', 'psr_radius', psr_radius)
    call getInput('problem', 'psr_angle', psr_angle)
    psr_angle = psr_angle * PSR_PI / 180.d0
    call getInput('problem', 'psr_period', psr_period)
    psr_omega = 2.d0 * PSR_PI / psr_period
    xc_g = 0.5d0 * global_mesh%sx
    yc_g = 0.5d0 * global_mesh%sy
    zc_g = 0.5d0 * global_mesh%sz
  end subroutine userReadInput

  subroutine getDipole(x_g, y_g, z_g, obx, oby, obz)
    real(dprec), intent(in) :: x_g, y_g, z_g
    real(dprec), intent(out) :: obx, oby, obz
    real(dprec) :: nx, ny, nz, rr, mux, muy, muz, mdn
    nx = x_g - xc_g; ny = y_g - yc_g; nz = z_g - zc_g
    rr = 1.d0 / sqrt(nx**2 + ny**2 + nz**2)
    nx = nx * rr; ny = ny * rr; nz = nz * rr
    rr = rr**3
    mux = psr_radius**3 * sin(psr_angle)
    muy = 0.d0
    muz = psr_radius**3 * cos(psr_angle)
    mdn = mux * nx + muy * ny + muz * nz
    obx = (3.d0 * nx * mdn - mux) * rr
    oby = (3.d0 * ny * mdn - muy) * rr
    obz = (3.d0 * nz * mdn - muz) * rr
  end subroutine getDipole

  subroutine userInitFields()
    integer :: i, j, k
    real(dprec) :: b1, b2, b3
    do k = 0, this_meshblock%ptr%sz - 1
      do j = 0, this_meshblock%ptr%sy - 1
        do i = 0, this_meshblock%ptr%sx - 1
          call getDipole(real(i, dprec), j + 0.5d0, k + 0.5d0, b1, b2, b3); bx(i, j, k) = b1
          call getDipole(i + 0.5d0, real(j, dprec), k + 0.5d0, b1, b2, b3); by(i, j, k) = b2
          call getDipole(i + 0.5d0, j + 0.5d0, real(k, dprec), b1, b2, b3); bz(i, j, k) = b3
        end do
      end do
    end do
  end subroutine userInitFields

  subroutine userFieldBoundaryConditions(step)
    integer, optional, intent(in) :: step
    call tffCheck(tff_user_bc(tff_h, step), 'userFieldBoundaryConditions')
  end subroutine userFieldBoundaryConditions
end module m_userfile
