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
  ! initialize.F90:26-72, in the reference's order of reading (which is
  ! the order of the parameter registry): domain, simulation, output, then
  ! the user file.  Keys this driver adds (grid absorb_fixed, algorithm
  ! precision and fused, output dump and report) are read without entering
  ! the registry.
  subroutine initializeAll()
    integer :: sizex, sizey, sizez
    call readCommandlineArgs()
    ! initializeDomain, initialize.F90:121-149
    call getInput('node_configuration', 'sizex', sizex)
    call getInput('node_configuration', 'sizey', sizey)
    call getInput('node_configuration', 'sizez', sizez)
    call getInput('grid', 'mx0', global_mesh%sx)
    call getInput('grid', 'my0', global_mesh%sy)
    call getInput('grid', 'mz0', global_mesh%sz)
    if ((modulo(global_mesh%sx, sizex) /= 0) .or. (modulo(global_mesh%sy, sizey) /= 0) .or. &
        (modulo(global_mesh%sz, sizez) /= 0)) &
      call throwError('ERROR: grid size is not evenly divisible by the number of cores')
    call getInput('grid', 'boundary_x', boundary_x, 1)
    call getInput('grid', 'boundary_y', boundary_y, 1)
    call getInput('grid', 'boundary_z', boundary_z, 1)
    if ((boundary_x == 2) .or. (boundary_y == 2) .or. (boundary_z == 2)) then
      boundary_x = 2; boundary_y = 2; boundary_z = 2
      call getInput('grid', 'absorb_buff', absorb_buff, 20)
      call getInput('grid', 'absorb_fixed', absorb_fixed, 0, register=.false.)
    end if
    this_meshblock%ptr => global_mesh
    ! initializeSimulation, initialize.F90:185-190
    call getInput('time', 'last', final_timestep, 1000)
    call getInput('algorithm', 'c', CC, 0.45d0)
    CCINV = 1.d0 / CC
    call getInput('algorithm', 'precision', precision_mode, 0, register=.false.)
    call getInput('algorithm', 'fused', fused_step, 1, register=.false.)
    ! initializeOutput, initialize.F90:175-180
    call getInput('output', 'start', output_start, 0)
    call getInput('output', 'interval', output_interval, 10)
    call getInput('output', 'istep', output_istep, 4)
    call getInput('output', 'dump', output_dump, 0, register=.false.)
    call getInput('output', 'report', report_interval, 1, register=.false.)
    if (output_istep < 2) &
      call throwError('ERROR: <output> istep = 1 (the debug layout with ghost zones) is not supported')
    ! checkEverything, initialize.F90:256-264
    if ((global_mesh%sx < NGHOST) .or. (global_mesh%sy < NGHOST) .or. (global_mesh%sz < NGHOST)) &
      call throwError('ERROR: ghost zones overflow the domain size in 0')
    call allocateFields()
    call createDevice()
    call userInitialize()
    call uploadFields()
    call printParams()
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
  use m_readinput
  use m_device
  use m_fldsolver
  use m_exchangefields
  use m_userfile
  implicit none
  integer :: timestep
contains
  ! src/tristanmainloop.F90:28-106.  The steps between two reports or
  ! outputs run as one tff_run (graph replays); with <output> report = 1
  ! every step runs eagerly with phase timers for makeReport.
  subroutine mainloop()
    integer :: nxt, n
    integer(c_int64_t) :: c0, c1, crate
    real(dprec) :: wall_ms, out_ms, ph(3)
    logical :: out_now, rep_now

    call exchangeFields(exchangeB0=.true.)
    call userFieldBoundaryConditions(0)
    call diagnostics(-1, 0.d0)
    if (report_interval > 0) call tffCheck(tff_set_phase_timing(tff_h, 1), 'tff_set_phase_timing')

    timestep = 0
    do while (timestep <= final_timestep)
      ! run to the next step that writes output or a report (inclusive)
      nxt = timestep
      do while (nxt < final_timestep .and. .not. (isOutputStep(nxt) .or. isReportStep(nxt)))
        nxt = nxt + 1
      end do
      n = nxt - timestep + 1
      call system_clock(c0, crate)
      if (fused_step /= 0) then
        call tffCheck(tff_run(tff_h, timestep, n), 'tff_run')
      else
        call referenceSteps(timestep, n)
      end if
      call tffCheck(tff_sync(tff_h), 'tff_sync')
      call system_clock(c1)
      wall_ms = 1.d3 * real(c1 - c0, dprec) / real(crate, dprec)
      timestep = nxt
      out_now = isOutputStep(timestep)
      rep_now = isReportStep(timestep)
      out_ms = 0.d0
      if (out_now) then
        call system_clock(c0)
        call writeOutput(timestep)
        call system_clock(c1)
        out_ms = 1.d3 * real(c1 - c0, dprec) / real(crate, dprec)
      end if
      if (rep_now) then
        call tffCheck(tff_phase_times(tff_h, ph), 'tff_phase_times')
        ! per step: the mean of the steps run since the last report, and
        ! this step's output
        call makeReport(timestep, wall_ms / n + out_ms, (ph(1) + ph(3)) / n, ph(2) / n, ph(3) / n, out_ms)
      end if
      if (out_now .or. timestep == final_timestep) call diagnostics(timestep, wall_ms / n)
      timestep = timestep + 1
    end do
  end subroutine mainloop

  ! tristanmainloop.F90:90-91
  logical function isOutputStep(t)
    integer, intent(in) :: t
    isOutputStep = output_interval > 0
    if (isOutputStep) isOutputStep = (modulo(t, output_interval) == 0) .and. (t >= output_start)
  end function isOutputStep

  logical function isReportStep(t)
    integer, intent(in) :: t
    isReportStep = report_interval > 0
    if (isReportStep) isReportStep = modulo(t, report_interval) == 0
  end function isReportStep

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

  ! Energy and div B (FP64 reductions on the device; no reference
  ! counterpart) with the mean wall time per step since the last line.
  subroutine diagnostics(step, ms)
    integer, intent(in) :: step
    real(dprec), intent(in) :: ms
    real(c_double) :: d(4)
    call tffCheck(tff_reduce(tff_h, d), 'tff_reduce')
    if (step < 0) then
      write (*, '(A,ES24.16,A,ES24.16,A,ES10.3)') 'init          E2/2', d(1), '  B2/2', d(2), &
        '  max|divB|', d(3)
    else
      write (*, '(A,I8,A,ES24.16,A,ES24.16,A,ES10.3,A,F10.4)') 'step', step, '  E2/2', d(1), &
        '  B2/2', d(2), '  max|divB|', d(3), '  ms/step', ms
    end if
  end subroutine diagnostics

  ! writeOutput(time), write_output.F90:27-43: params.NNNNN, flds.tot.NNNNN
  ! (+ .xdmf) and domain.NNNNN through libtff_io.so; with <output> dump = 1
  ! also the raw fields with ghosts as FP64 (x fastest), flds.NNNNN.bin.
  subroutine writeOutput(step)
    integer, intent(in) :: step
    character(len=STR_MAX) :: fname
    character(kind=c_char, len=:), allocatable :: names
    integer(c_int) :: types(max(param_count, 1)), ivals(max(param_count, 1))
    real(c_float) :: rvals(max(param_count, 1))
    integer :: u, q
    names = ''
    do q = 1, param_count
      names = names//trim(param_group(q))//':'//trim(param_name(q))//c_null_char
      types(q) = param_type(q)
      ivals(q) = param_int(q)
      rvals(q) = param_real(q)
    end do
    if (tff_io_write_output(tff_h, trim(output_dir_name)//c_null_char, output_index, step, output_istep, &
                            user_output_kind, param_count, names//c_null_char, types, ivals, rvals) /= TFF_OK) &
      call throwError('ERROR in writeOutput: '//trim(c_text(tff_io_last_error())))
    output_index = output_index + 1
    if (output_dump /= 0) then
      call downloadFields()
      write (fname, '(A,A,I5.5,A)') trim(output_dir_name), '/flds.', step, '.bin'
      open (newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
      write (u, iostat=q) ex, ey, ez, bx, by, bz
      close (u)
    end if
  end subroutine writeOutput

  ! makeReport, tristanmainloop.F90:108-151 with auxiliary.F90:123-215: the
  ! per-step times of the phases (ms) with their min / max over the ranks
  ! (one here) and their share of the full step.  fld_solver is the GPU time
  ! of the stage kernels and the ghost exchange, usr_funcs of the boundary
  ! hooks; the exchange's part of fld_solver is listed on its own line.
  subroutine makeReport(tstep, full, solver, hooks, exchange, output)
    integer, intent(in) :: tstep
    real(dprec), intent(in) :: full, solver, hooks, exchange, output
    character(len=72) :: line
    character(len=16) :: num
    line = repeat('-', 71)
    write (*, '(1X,A)') line
    write (num, '(I0)') tstep
    line = repeat('.', 71)
    line(1:10 + len_trim(num)) = 'Timestep: '//trim(num)
    line(66:71) = '[DONE]'
    write (*, '(1X,A)') line
    write (*, '(1X,A)') '[ROUTINE]          [TIME, ms]      [MIN  /  MAX, ms]      [FRACTION, %]'
    call timeLine('Full_step: ', full, -1.d0)
    call timeLine('  fld_solver: ', solver, full)
    call timeLine('    exchange: ', exchange, full)
    call timeLine('  usr_funcs: ', hooks, full)
    call timeLine('  output_step: ', output, full)
    write (*, '(1X,A)') repeat('.', 71)
    write (*, '(A)') ''
  end subroutine makeReport

  ! printTime (auxiliary.F90:166-215): mean at column 20, min at 32, max at
  ! 43, percentage at 62; a phase below 1e-4 of the step prints as zero
  subroutine timeLine(msg, t, full)
    character(len=*), intent(in) :: msg
    real(dprec), intent(in) :: t, full
    character(len=72) :: line
    character(len=10) :: f
    real :: v
    v = real(t)
    if (full > 0.d0) then
      if (t / full < 1e-4) v = 0.0
    end if
    line = ''
    line(1:len(msg)) = msg
    f = fmtReal(v)
    line(20:29) = f
    line(32:41) = f
    line(43:52) = f
    if (full > 0.d0) line(62:71) = fmtReal(real(v * 100 / full))
    write (*, '(1X,A)') line
  end subroutine timeLine

  function fmtReal(v) result(f)
    real, intent(in) :: v
    character(len=10) :: f
    character(len=16) :: fmt
    fmt = '('//trim(realFormat(v))//')'
    write (f, fmt) v
  end function fmtReal
end module m_mainloop

program tristan
  use m_initialize
  use m_mainloop
  implicit none
  call initializeAll()
  call mainloop()
  call tff_destroy(tff_h)
end program tristan
