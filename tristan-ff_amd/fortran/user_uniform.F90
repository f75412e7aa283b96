! user_uniform.F90 -- vertical flux tube bz = 1/(1 + r/r0) with a rotating
! floor (user/user_uniform.F90:27-147); the floor hook runs on the device
! (TFF_BC_UNIFORM).  x and y periodic, z open.
#define NGHOST 2
module m_userfile
  use m_globalnamespace
  use m_readinput
  use m_domain
  use m_fields
  use m_device
  implicit none
  real(dprec), private :: L0, r0, beta_twist, xc, yc, bnorm, t0
contains
  subroutine userInitialize()
    user_output_kind = TFF_USEROUT_UNIFORM  ! userOutput, user_uniform.F90:151-173
    call userReadInput()
    call userInitFields()
    call tffCheck(tff_set_bc(tff_h, TFF_BC_UNIFORM, [xc, yc, r0, beta_twist, t0, bnorm], 6), &
                  'tff_set_bc')
  end subroutine userInitialize

  subroutine userReadInput()
    call getInput('problem', 'L0', L0)
    call getInput('problem', 'r_L0', r0)
    call getInput('problem', 't0', t0)
    call getInput('problem', 'beta_twist', beta_twist)
    xc = 0.5d0 * global_mesh%sx
    yc = 0.5d0 * global_mesh%sy
    r0 = r0 * L0
    bnorm = 1.d0
  end subroutine userReadInput

  subroutine userInitFields()
    integer :: i, j, k
    real(dprec) :: rx, ry, rr
    do k = -NGHOST, this_meshblock%ptr%sz - 1 + NGHOST
      do j = -NGHOST, this_meshblock%ptr%sy - 1 + NGHOST
        do i = -NGHOST, this_meshblock%ptr%sx - 1 + NGHOST
          rx = i + 0.5d0 - xc; ry = j + 0.5d0 - yc
          rr = sqrt(rx**2 + ry**2 + 1.d-7)
          bz(i, j, k) = bnorm / (1.d0 + (rr / r0))
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
