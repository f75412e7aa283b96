! user_emwave.F90 -- vacuum plane wave over a uniform background (the
! user/user_emwave.F90 form: E on cell edges, B on faces, interior cells
! only).  The wave vector is (kx, ky, 0) 2 pi / n with the discrete Yee
! eigenvector, so the mode travels along +k without dispersion error
! beyond the scheme's own; B0 = b0 z^.
#define NGHOST 2
module m_userfile
  use m_globalnamespace
  use m_readinput
  use m_domain
  use m_fields
  use m_device
  implicit none
  real(dprec), private :: amp, kxm, kym, b0
contains
  subroutine userInitialize()
    call userReadInput()
    call userInitFields()
    call tffCheck(tff_set_bc(tff_h, TFF_BC_NONE, [0.d0], 0), 'tff_set_bc')
  end subroutine userInitialize

  subroutine userReadInput()
    call getInput('problem', 'amplitude', amp, 1.d-2)
    call getInput('problem', 'kx', kxm, 1.d0)
    call getInput('problem', 'ky', kym, 1.d0)
    call getInput('problem', 'b0', b0, 1.d0)
  end subroutine userReadInput

  subroutine userInitFields()
    integer :: i, j, k
    real(dprec) :: kx, ky, sx, sy, s, al, be, pi
    pi = 4.d0 * atan(1.d0)
    kx = kxm * 2 * pi / global_mesh%sx
    ky = kym * 2 * pi / global_mesh%sy
    sx = sin(kx / 2); sy = sin(ky / 2); s = sqrt(sx**2 + sy**2)
    al = -sy / s; be = sx / s
    do k = 0, this_meshblock%ptr%sz - 1
      do j = 0, this_meshblock%ptr%sy - 1
        do i = 0, this_meshblock%ptr%sx - 1
          ex(i, j, k) = amp * al * cos(kx * (i + 0.5d0) + ky * j)
          ey(i, j, k) = amp * be * cos(kx * i + ky * (j + 0.5d0))
          bz(i, j, k) = b0 + amp * cos(kx * (i + 0.5d0) + ky * (j + 0.5d0))
        end do
      end do
    end do
    b0z = b0
  end subroutine userInitFields

  subroutine userFieldBoundaryConditions(step)
    integer, optional, intent(in) :: step
    ! no boundary hook (user/user_emwave.F90:75-78)
  end subroutine userFieldBoundaryConditions
end module m_userfile
