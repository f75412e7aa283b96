! user_loop.F90 -- two opposite magnetic sources below a conducting floor
! (user/user_loop.F90:27-93), x and y periodic, z open.  The floor/top
! hook runs on the device (TFF_BC_LOOP); <problem> twist = 1 adds the
! footpoint-1 rotation that the reference leaves commented out at
! user_loop.F90:122-150 (floor ghosts, r0 = r_L0 * L0, tanh(t/t0) ramp);
! twist = 2 applies its E on z <= 0 with a conducting top instead (the
! floor plane's bz then follows the update, which keeps div B there).
#define NGHOST 2
module m_userfile
  use m_globalnamespace
  use m_readinput
  use m_domain
  use m_fields
  use m_device
  implicit none
  real(dprec), private :: a0, L0, h0, r0, t0, beta_twist, xc, yc1, yc2, zc, bnorm
  integer, private :: twist
contains
  subroutine userInitialize()
    user_output_kind = TFF_USEROUT_RHO  ! userOutput, user_loop.F90:168-180
    call userReadInput()
    call userInitFields()
    call tffCheck(tff_set_bc(tff_h, TFF_BC_LOOP, [xc, yc1, yc2, zc, bnorm, r0, beta_twist, t0, &
                                                  real(twist, dprec)], 9), 'tff_set_bc')
  end subroutine userInitialize

  subroutine userReadInput()
    call getInput('problem', 'L0', L0)
    call getInput('problem', 'a_L0', a0)
    call getInput('problem', 'h_L0', h0)
    ! user_loop.F90:29-34: no defaults (the shipped inputs/input.loop lacks
    ! t0 and beta_twist, so the reference stops there, and so does this)
    call getInput('problem', 'r_L0', r0)
    call getInput('problem', 't0', t0)
    call getInput('problem', 'beta_twist', beta_twist)
    call getInput('problem', 'twist', twist, 0, register=.false.)
    xc = 0.5d0 * global_mesh%sx
    yc1 = 0.5d0 * global_mesh%sy - a0 * L0
    yc2 = 0.5d0 * global_mesh%sy + a0 * L0
    zc = -h0 * L0
    bnorm = 1000.d0 * h0**2 * L0**2
    r0 = r0 * L0
  end subroutine userReadInput

  subroutine userInitFields()
    integer :: i, j, k
    real(dprec) :: x, y, z, rx, ry1, ry2, rz, rr1, rr2
    do k = -NGHOST, this_meshblock%ptr%sz - 1 + NGHOST
      z = k
      do j = -NGHOST, this_meshblock%ptr%sy - 1 + NGHOST
        y = j
        do i = -NGHOST, this_meshblock%ptr%sx - 1 + NGHOST
          x = i
          rx = x - xc; ry1 = y + 0.5d0 - yc1; ry2 = y + 0.5d0 - yc2; rz = z + 0.5d0 - zc
          rr1 = sqrt(rx * rx + ry1 * ry1 + rz * rz); rr2 = sqrt(rx * rx + ry2 * ry2 + rz * rz)
          bx(i, j, k) = bnorm * (rx / rr1**3 - rx / rr2**3)
          rx = x + 0.5d0 - xc; ry1 = y - yc1; ry2 = y - yc2; rz = z + 0.5d0 - zc
          rr1 = sqrt(rx * rx + ry1 * ry1 + rz * rz); rr2 = sqrt(rx * rx + ry2 * ry2 + rz * rz)
          by(i, j, k) = bnorm * (ry1 / rr1**3 - ry2 / rr2**3)
          rx = x + 0.5d0 - xc; ry1 = y + 0.5d0 - yc1; ry2 = y + 0.5d0 - yc2; rz = z - zc
          rr1 = sqrt(rx * rx + ry1 * ry1 + rz * rz); rr2 = sqrt(rx * rx + ry2 * ry2 + rz * rz)
          bz(i, j, k) = bnorm * (rz / rr1**3 - rz / rr2**3)
        end do
      end do
    end do
    b0x = bx; b0y = by; b0z = bz
  end subroutine userInitFields

  subroutine userFieldBoundaryConditions(step)
    integer, optional, intent(in) :: step
    call tffCheck(tff_user_bc(tff_h, step), 'userFieldBoundaryConditions')
  end subroutine userFieldBoundaryConditions
end module m_userfile
