! tristan.F90 -- initialisation, main loop and program of the Fortran
! driver (src/logistics/initialize.F90, src/tristanmainloop.F90,
! src/tristan.F90).  The whole grid is one meshblock on one GPU: the
! <node_configuration> block is read but does not split the grid.

#define NGHOST 2
#define STR_MAX 280

module m_initialize
  use m_globalnamespace
  use m_readinput
  use m_domain
  use m_fields
  use m_device
  use m_userfile
  implicit none
contains
  subroutine initializeAll()
    integer :: sizex, sizey, sizez
    call readCommandlineArgs()
    ! initialize.F90:121-149
    call getInput('node_configuration', 'sizex', sizex, 1)
    call getInput('node_configuration', 'sizey', sizey, 1)
    call getInput('node_configuration', 'sizez', sizez, 1)
    call getInput('grid', 'mx0', global_mesh%sx)
    call getInput('grid', 'my0', global_mesh%sy)
    call getInput('grid', 'mz0', global_mesh%sz)
    call getInput('grid', 'boundary_x', boundary_x, 1)
    call getInput('grid', 'boundary_y', boundary_y, 1)
    call getInput('grid', 'boundary_z', boundary_z, 1)
    if ((boundary_x == 2) .or. (boundary_y == 2) .or. (boundary_z == 2)) then
      boundary_x = 2; boundary_y = 2; boundary_z = 2
      call getInput('grid', 'absorb_buff', absorb_buff, 20)
      call getInput('grid', 'absorb_fixed', absorb_fixed, 0)
    end if
    this_meshblock%ptr => global_mesh
    if ((global_mesh%sx < NGHOST) .or. (global_mesh%sy < NGHOST) .or. (global_mesh%sz < NGHOST)) &
      call throwError('ERROR: ghost zones overflow the domain size')
    ! initialize.F90:175-190
    call getInput('output', 'start', output_start, 0)
    call getInput('output', 'interval', output_interval, 10)
    call getInput('output', 'dump', output_dump, 0)
    call getInput('output', 'istep', output_istep, 0)
    call getInput('time', 'last', final_timestep, 1000)
    call getInput('algorithm', 'c', CC, 0.45d0)
    call getInput('algorithm', 'precision', precision_mode, 0)
    call getInput('algorithm', 'fused', fused_step, 1)
    CCINV = 1.d0 / CC
    call allocateFields()
    call createDevice()
    call userInitialize()
    call uploadFields()
    write (*, '(A,3I6,A,3I2,A,F6.3,A,I2)') 'tristan_ff: grid', global_mesh%sx, global_mesh%sy, &
      global_mesh%sz, '  boundary', boundary_x, boundary_y, boundary_z, '  c =', CC, &
      '  precision', precision_mode
  end subroutine initializeAll

  subroutine allocateFields()
    integer :: nx, ny, nz
    nx = this_meshblock%ptr%sx; ny = this_meshblock%ptr%sy; nz = this_meshblock%ptr%sz
    allocate (ex(-NGHOST:nx - 1 + NGHOST, -NGHOST:ny - 1 + NGHOST, -NGHOST:nz - 1 + NGHOST))
    allocate (ey, ez, bx, by, bz, b0x, b0y, b0z, mold=ex)
    ex = 0; ey = 0; ez = 0; bx = 0; by = 0; bz = 0; b0x = 0; b0y = 0; b0z = 0
  end subroutine allocateFields
end module m_initialize

module m_mainloop
  use m_globalnamespace
  use m_device
  use m_fldsolver
  use m_exchangefields
  use m_userfile
  implicit none
  integer :: timestep
contains
  ! src/tristanmainloop.F90:28-106
  subroutine mainloop()
    integer :: nxt, n
    integer(c_int64_t) :: c0, c1, crate
    real(dprec) :: dt_ms

    call exchangeFields(exchangeB0=.true.)
    call userFieldBoundaryConditions(0)
    call writeOutput(-1, 0.d0)

    timestep = 0
    do while (timestep <= final_timestep)
      ! run to the next output step (inclusive) in one go
      nxt = timestep
      if (output_interval > 0) then
        do while (.not. ((modulo(nxt, output_interval) == 0) .and. (nxt >= output_start)) &
                  .and. nxt < final_timestep)
          nxt = nxt + 1
        end do
      else
        nxt = final_timestep
      end if
      n = nxt - timestep + 1
      call system_clock(c0, crate)
      if (fused_step /= 0) then
        call tffCheck(tff_run(tff_h, timestep, n), 'tff_run')
      else
        call referenceSteps(timestep, n)
      end if
      call tffCheck(tff_sync(tff_h), 'tff_sync')
      call system_clock(c1)
      dt_ms = 1.d3 * real(c1 - c0, dprec) / real(crate, dprec) / n
      timestep = nxt
      if ((output_interval > 0) .and. (modulo(timestep, output_interval) == 0) .and. &
          (timestep >= output_start)) call writeOutput(timestep, dt_ms)
      timestep = timestep + 1
    end do
  end subroutine mainloop

  ! the stage sequence of tristanmainloop.F90:50-83, one procedure at a time
  subroutine referenceSteps(first, n)
    integer, intent(in) :: first, n
    integer :: t
    do t = first, first + n - 1
      call initRKstep()
      call rk3Step(1.d0, 0.d0, 1.d0)
      call userFieldBoundaryConditions(t)
      call exchangeFields()
      call cleanEpar()
      call exchangeFields()
      call rk3Step(3.d0 / 4.d0, 1.d0 / 4.d0, 1.d0 / 4.d0)
      call userFieldBoundaryConditions(t)
      call exchangeFields()
      call cleanEpar()
      call exchangeFields()
      call rk3Step(1.d0 / 3.d0, 2.d0 / 3.d0, 2.d0 / 3.d0)
      call userFieldBoundaryConditions(t)
      call exchangeFields()
      call cleanEpar()
      call exchangeFields()
    end do
  end subroutine referenceSteps

  ! Energy and div B history (FP64 reductions on the device); with
  ! <output> dump = 1, the fields with ghosts as raw FP64 (x fastest); with
  ! <output> istep >= 2, the six field datasets of writeFields_hdf5
  ! (write_output.F90:334-374: ex ey ez bx by bz node interpolants at every
  ! istep-th node) as raw FP32, one after the other, x fastest.
  subroutine writeOutput(step, ms)
    integer, intent(in) :: step
    real(dprec), intent(in) :: ms
    real(c_double) :: d(4)
    character(len=STR_MAX) :: fname
    integer :: u, f
    integer(c_int) :: dims(3)
    real(c_float), allocatable, target :: sm(:, :, :, :)
    type(c_ptr) :: bufs(6)
    call tffCheck(tff_reduce(tff_h, d), 'tff_reduce')
    if (step < 0) then
      write (*, '(A,ES24.16,A,ES24.16,A,ES10.3)') 'init          E2/2', d(1), '  B2/2', d(2), &
        '  max|divB|', d(3)
    else
      write (*, '(A,I8,A,ES24.16,A,ES24.16,A,ES10.3,A,F10.4)') 'step', step, '  E2/2', d(1), &
        '  B2/2', d(2), '  max|divB|', d(3), '  ms/step', ms
    end if
    if ((output_dump /= 0) .and. (step >= 0)) then
      call downloadFields()
      write (fname, '(A,A,I5.5,A)') trim(output_dir_name), '/flds.', step, '.bin'
      call execute_command_line('mkdir -p '//trim(output_dir_name))
      open (newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
      write (u) ex, ey, ez, bx, by, bz
      close (u)
    end if
    if ((output_istep >= 2) .and. (step >= 0)) then
      call tffCheck(tff_output_dims(tff_h, output_istep, dims), 'tff_output_dims')
      allocate (sm(dims(1), dims(2), dims(3), 6))
      do f = 1, 6
        bufs(f) = c_loc(sm(1, 1, 1, f))
      end do
      call tffCheck(tff_output_fields(tff_h, output_istep, bufs, TFF_HOST_F32), 'tff_output_fields')
      write (fname, '(A,A,I5.5,A)') trim(output_dir_name), '/flds.tot.', step, '.bin'
      call execute_command_line('mkdir -p '//trim(output_dir_name))
      open (newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
      write (u) sm
      close (u)
      deallocate (sm)
    end if
  end subroutine writeOutput
end module m_mainloop

program tristan
  use m_initialize
  use m_mainloop
  implicit none
  call initializeAll()
  call mainloop()
  call tff_destroy(tff_h)
end program tristan
