! m_host.F90 -- host-side modules of the Fortran driver.
!
! They keep the module and procedure names of the reference
! (src/global_namespace.F90, src/objects/domain.F90, src/objects/fields.F90,
! src/io/read_input.F90, src/algorithms/field_solver.F90,
! src/logistics/exchange_fields.F90) so that a user file and the main loop
! read as they do there; the field arrays live in HBM behind libtff.so and
! the host copies here exist for initial conditions and output only.

#define NGHOST 2
#define STR_MAX 280

module m_globalnamespace
  implicit none
  integer, parameter :: dprec = kind(1.0d0)
  real(dprec) :: CC, CCINV
  integer :: final_timestep
  integer :: output_start = 0, output_interval = 10, output_dump = 0
  ! <output> istep (initialize.F90:179): node stride of the flds.tot datasets;
  ! below 2 (this driver's default) they are not written
  integer :: output_istep = 0
  integer :: precision_mode = 0, fused_step = 1
  character(len=STR_MAX) :: input_file_name = 'input', output_dir_name = 'output'
end module m_globalnamespace

module m_errors
  implicit none
contains
  ! src/tools/error_handling.F90:8-13
  subroutine throwError(msg)
    character(len=*), intent(in) :: msg
    write (*, '(A)') trim(msg)
    stop 1
  end subroutine throwError
end module m_errors

module m_aux
  use m_globalnamespace
  implicit none
  real(dprec) :: dseed = 123457.d0
contains
  ! Park-Miller minimal standard generator, src/tools/auxiliary.F90:246-269
  real(dprec) function randomNum(s)
    real(dprec), intent(inout) :: s
    s = dmod(16807.d0 * s, 2147483647.d0)
    randomNum = s / 2147483648.d0
  end function randomNum

  real(dprec) function random(s)
    real(dprec), intent(inout) :: s
    random = 1.d0
    do while (random == 1.d0)
      random = randomNum(s)
    end do
  end function random
end module m_aux

module m_readinput
  use m_globalnamespace
  use m_errors
  implicit none
  interface getInput
    module procedure getIntInput
    module procedure getRealInput
  end interface getInput
  private :: getIntInput, getRealInput, parseInput
contains
  ! src/io/read_input.F90:32-54
  subroutine readCommandlineArgs()
    integer :: i
    character(len=STR_MAX) :: arg, arg1
    do i = 1, command_argument_count()
      call get_command_argument(i, arg)
      select case (arg)
      case ('-i', '--input')
        call get_command_argument(i + 1, arg1)
        input_file_name = trim(arg1)
      case ('-o', '--output')
        call get_command_argument(i + 1, arg1)
        output_dir_name = trim(arg1)
      end select
    end do
  end subroutine readCommandlineArgs

  ! Deck semantics of read_input.F90:56-102: find the line `<blockname>`,
  ! then, before the next line that starts with '<', the line whose first
  ! token is varname; the value is the text between '=' and '#'.
  function parseInput(blockname, varname, found) result(val)
    character(len=*), intent(in) :: blockname, varname
    logical, intent(out) :: found
    character(len=STR_MAX) :: val, line, tok
    integer :: u, ios, k1, k2, p
    logical :: inblock
    found = .false.; val = ''; inblock = .false.
    open (newunit=u, file=trim(input_file_name), action='read', iostat=ios)
    if (ios /= 0) call throwError('ERROR while reading input file: `'//trim(input_file_name)//'`')
    do
      read (u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      p = scan(line, ' =,/'//achar(9))
      if (p == 0) then
        tok = line
      else
        tok = line(1:p - 1)
      end if
      if (.not. inblock) then
        if (trim(tok) == '<'//trim(blockname)//'>') inblock = .true.
        cycle
      end if
      if (line(1:1) == '<') exit
      if (trim(tok) == trim(varname)) then
        k1 = scan(line, '=')
        k2 = scan(line, '#')
        if (k2 == 0) k2 = len(line) + 1
        val = adjustl(line(k1 + 1:k2 - 1))
        found = .true.
        exit
      end if
    end do
    close (u)
  end function parseInput

  subroutine getIntInput(blockname, varname, val, def_val)
    character(len=*), intent(in) :: blockname, varname
    integer, intent(out) :: val
    integer, optional, intent(in) :: def_val
    character(len=STR_MAX) :: s
    logical :: found
    real(dprec) :: r
    integer :: ios
    s = parseInput(blockname, varname, found)
    if (found) then
      read (s, *, iostat=ios) r  ! integers are read through real(8)
      if (ios == 0) then
        val = int(r)
        return
      end if
    end if
    if (present(def_val)) then
      val = def_val
    else
      call throwError('ERROR variable `'//trim(varname)//'` not found in <'//trim(blockname)// &
                      '> -- no default value provided')
    end if
  end subroutine getIntInput

  subroutine getRealInput(blockname, varname, val, def_val)
    character(len=*), intent(in) :: blockname, varname
    real(dprec), intent(out) :: val
    real(dprec), optional, intent(in) :: def_val
    character(len=STR_MAX) :: s
    logical :: found
    integer :: ios
    s = parseInput(blockname, varname, found)
    if (found) then
      read (s, *, iostat=ios) val
      if (ios == 0) return
    end if
    if (present(def_val)) then
      val = def_val
    else
      call throwError('ERROR variable `'//trim(varname)//'` not found in <'//trim(blockname)// &
                      '> -- no default value provided')
    end if
  end subroutine getRealInput
end module m_readinput

module m_domain
  implicit none
  type :: mesh
    integer :: x0 = 0, y0 = 0, z0 = 0
    integer :: sx = 0, sy = 0, sz = 0
  end type mesh
  type :: meshptr
    type(mesh), pointer :: ptr => null()
  end type meshptr
  type(mesh), target :: global_mesh
  type(meshptr) :: this_meshblock
  integer :: boundary_x = 1, boundary_y = 1, boundary_z = 1
  integer :: absorb_buff = 20, absorb_fixed = 0
end module m_domain

module m_fields
  use m_globalnamespace
  implicit none
  real(dprec), allocatable :: ex(:, :, :), ey(:, :, :), ez(:, :, :)
  real(dprec), allocatable :: bx(:, :, :), by(:, :, :), bz(:, :, :)
  real(dprec), allocatable :: b0x(:, :, :), b0y(:, :, :), b0z(:, :, :)
end module m_fields

! The device meshblock and the host<->device copies.
module m_device
  use iso_c_binding
  use m_tff
  use m_errors
  use m_fields
  use m_domain
  implicit none
  type(c_ptr) :: tff_h = c_null_ptr
contains
  subroutine tffCheck(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= TFF_OK) call throwError('ERROR in '//trim(what)//': '//trim(tff_error_text(tff_h)))
  end subroutine tffCheck

  subroutine createDevice()
    type(tff_config) :: cfg
    cfg%nx = this_meshblock%ptr%sx
    cfg%ny = this_meshblock%ptr%sy
    cfg%nz = this_meshblock%ptr%sz
    cfg%boundary = [boundary_x, boundary_y, boundary_z]
    cfg%absorb_buff = absorb_buff
    cfg%absorb_fixed = absorb_fixed
    cfg%precision = precision_mode
    cfg%device = 0
    cfg%cc = CC
    if (tff_create(cfg, tff_h) /= TFF_OK) call throwError('ERROR: tff_create failed')
  end subroutine createDevice

  subroutine moveField(id, a, up)
    integer(c_int), intent(in) :: id
    real(dprec), target, intent(inout) :: a(:, :, :)
    logical, intent(in) :: up
    if (up) then
      call tffCheck(tff_upload(tff_h, id, c_loc(a), TFF_HOST_F64), 'tff_upload')
    else
      call tffCheck(tff_download(tff_h, id, c_loc(a), TFF_HOST_F64), 'tff_download')
    end if
  end subroutine moveField

  subroutine uploadFields()
    call moveField(TFF_EX, ex, .true.); call moveField(TFF_EY, ey, .true.)
    call moveField(TFF_EZ, ez, .true.); call moveField(TFF_BX, bx, .true.)
    call moveField(TFF_BY, by, .true.); call moveField(TFF_BZ, bz, .true.)
    call moveField(TFF_B0X, b0x, .true.); call moveField(TFF_B0Y, b0y, .true.)
    call moveField(TFF_B0Z, b0z, .true.)
  end subroutine uploadFields

  subroutine downloadFields()
    call moveField(TFF_EX, ex, .false.); call moveField(TFF_EY, ey, .false.)
    call moveField(TFF_EZ, ez, .false.); call moveField(TFF_BX, bx, .false.)
    call moveField(TFF_BY, by, .false.); call moveField(TFF_BZ, bz, .false.)
  end subroutine downloadFields
end module m_device

! src/algorithms/field_solver.F90 procedure names on the device fields.
module m_fldsolver
  use m_device
  implicit none
contains
  subroutine initRKstep()
    call tffCheck(tff_init_rk_step(tff_h), 'initRKstep')
  end subroutine initRKstep

  subroutine rk3Step(rk_c1, rk_c2, rk_c3)
    real(dprec), intent(in) :: rk_c1, rk_c2, rk_c3
    call tffCheck(tff_rk3_step(tff_h, rk_c1, rk_c2, rk_c3), 'rk3Step')
  end subroutine rk3Step

  subroutine cleanEpar()
    call tffCheck(tff_clean_epar(tff_h), 'cleanEpar')
  end subroutine cleanEpar
end module m_fldsolver

! src/logistics/exchange_fields.F90 on one meshblock.
module m_exchangefields
  use m_device
  implicit none
contains
  subroutine exchangeFields(exchangeB0)
    logical, optional, intent(in) :: exchangeB0
    integer(c_int) :: b0
    b0 = 0
    if (present(exchangeB0)) then
      if (exchangeB0) b0 = 1
    end if
    call tffCheck(tff_exchange_fields(tff_h, b0), 'exchangeFields')
  end subroutine exchangeFields
end module m_exchangefields
