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
  ! <output> istep (initialize.F90:179, default 4): node stride of the
  ! flds.tot datasets
  integer :: output_istep = 4
  ! output files written so far (write_output.F90:41 output_index)
  integer :: output_index = 0
  ! <output> report: the timing report every this many steps (the
  ! reference's makeReport after every step: 1; 0 = none, steps replayed
  ! from captured graphs)
  integer :: report_interval = 1
  integer :: precision_mode = 0, fused_step = 1
  ! what var1..var3 of flds.tot hold (TFF_USEROUT_*): set by the user file
  integer :: user_output_kind = 0
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
  ! The registry of every deck parameter read (src/tools/auxiliary.F90:22-40,
  ! filled by read_input.F90's getInput): echoed by printParams and written
  ! to params.NNNNN.  Types 1 integer, 2 real (default REAL), 3 logical.
  integer, parameter :: MAX_PARAMS = 1000
  integer :: param_count = 0
  integer :: param_type(MAX_PARAMS)
  character(len=3) :: param_group(MAX_PARAMS)
  character(len=STR_MAX) :: param_name(MAX_PARAMS)
  integer :: param_int(MAX_PARAMS)
  real :: param_real(MAX_PARAMS)
  interface getInput
    module procedure getIntInput
    module procedure getRealInput
  end interface getInput
  private :: getIntInput, getRealInput, parseInput, registerParam, lookup
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

  ! Deck semantics of read_input.F90:56-102.  Records are read one at a
  ! time; a record's first token is what a list-directed read of a
  ! character variable gives: the text up to the first blank, comma or
  ! slash ('=' does not end it).  After the record `<blockname>`, the
  ! record whose token is varname gives the value (the text between '=' and
  ! '#', left-adjusted); the search ends at the next record starting with
  ! '<'.  Blank records are skipped, as a list-directed read skips them.
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
      p = scan(line, ' ,/'//achar(9))
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

  ! read_input.F90:104-120: the registry's group names
  function simplifyBlockname(blockname) result(bname)
    character(len=*), intent(in) :: blockname
    character(len=3) :: bname
    select case (trim(blockname))
    case ('node_configuration')
      bname = 'cpu'
    case ('grid')
      bname = 'grd'
    case ('restart')
      bname = 'rst'
    case ('problem')
      bname = 'prb'
    case default
      bname = blockname(1:min(3, len(blockname)))
    end select
  end function simplifyBlockname

  integer function lookup(blockname, varname)
    character(len=*), intent(in) :: blockname, varname
    integer :: n
    lookup = 0
    do n = 1, param_count
      if (param_group(n) == simplifyBlockname(blockname) .and. trim(param_name(n)) == trim(varname)) then
        lookup = n
        return
      end if
    end do
  end function lookup

  ! read_input.F90:122-136 and the registration of each getInput: the first
  ! reading of a (block, name) pair is kept
  subroutine registerParam(blockname, varname, t, ival, rval)
    character(len=*), intent(in) :: blockname, varname
    integer, intent(in) :: t, ival
    real, intent(in) :: rval
    if (lookup(blockname, varname) /= 0) return
    if (param_count >= MAX_PARAMS) call throwError('ERROR: too many input parameters')
    param_count = param_count + 1
    param_type(param_count) = t
    param_group(param_count) = simplifyBlockname(blockname)
    param_name(param_count) = varname
    param_int(param_count) = ival
    param_real(param_count) = rval
  end subroutine registerParam

  ! the value text of a key, or the reference's errors (read_input.F90:148-166)
  subroutine valueOf(blockname, varname, has_default, s, found)
    character(len=*), intent(in) :: blockname, varname
    logical, intent(in) :: has_default
    character(len=STR_MAX), intent(out) :: s
    logical, intent(out) :: found
    s = parseInput(blockname, varname, found)
    if (.not. found .and. .not. has_default) &
      call throwError('ERROR variable `'//trim(varname)//'` not found in <'//trim(blockname)//'>'// &
                      new_line('a')//'-- no default value provided')
  end subroutine valueOf

  subroutine conversionError(blockname, varname, s)
    character(len=*), intent(in) :: blockname, varname, s
    call throwError('ERROR variable `'//trim(varname)//'`='//trim(s)//' from <'//trim(blockname)// &
                    '> not converted'//new_line('a')//'-- no default value provided')
  end subroutine conversionError

  ! getInt4Input (read_input.F90:138-186): integers are read through
  ! real(8) and truncated; `register` = .false. reads a key of this driver
  ! that the reference does not have, kept out of the registry
  subroutine getIntInput(blockname, varname, val, def_val, register)
    character(len=*), intent(in) :: blockname, varname
    integer, intent(out) :: val
    integer, optional, intent(in) :: def_val
    logical, optional, intent(in) :: register
    character(len=STR_MAX) :: s
    logical :: found
    real(dprec) :: r
    integer :: ios
    call valueOf(blockname, varname, present(def_val), s, found)
    if (found) then
      read (s, *, iostat=ios) r
      if (ios == 0) then
        val = int(r)
      else if (present(def_val)) then
        val = def_val
      else
        call conversionError(blockname, varname, s)
      end if
    else
      val = def_val
    end if
    if (present(register)) then
      if (.not. register) return
    end if
    call registerParam(blockname, varname, 1, val, 0.0)
  end subroutine getIntInput

  ! getReal4Input (read_input.F90:250-298); the registry keeps default REAL
  subroutine getRealInput(blockname, varname, val, def_val, register)
    character(len=*), intent(in) :: blockname, varname
    real(dprec), intent(out) :: val
    real(dprec), optional, intent(in) :: def_val
    logical, optional, intent(in) :: register
    character(len=STR_MAX) :: s
    logical :: found
    integer :: ios
    call valueOf(blockname, varname, present(def_val), s, found)
    if (found) then
      read (s, *, iostat=ios) val
      if (ios /= 0) then
        if (present(def_val)) then
          val = def_val
        else
          call conversionError(blockname, varname, s)
        end if
      end if
    else
      val = def_val
    end if
    if (present(register)) then
      if (.not. register) return
    end if
    call registerParam(blockname, varname, 2, 0, real(val))
  end subroutine getRealInput

  ! src/tools/auxiliary.F90:72-89 getFMTForReal (width 10)
  function realFormat(v) result(fmt)
    real, intent(in) :: v
    character(len=16) :: fmt
    if ((abs(v) >= 100000) .or. ((abs(v) < 1e-4) .and. (v /= 0.0))) then
      fmt = 'ES10.2'
    else
      fmt = 'F10.2'
    end if
  end function realFormat

  ! initialize.F90:74-107: the registry, one line per parameter
  subroutine printParams()
    integer :: n
    character(len=64) :: fmt
    write (*, '(A)') '== Simulation parameters ==============================================='
    do n = 1, param_count
      select case (param_type(n))
      case (1)
        write (*, '(A30,A1,A20,A1,I10)') trim(param_group(n)), ':', trim(param_name(n)), ':', param_int(n)
      case (2)
        fmt = '(A30,A1,A20,A1,'//trim(realFormat(param_real(n)))//')'
        write (*, fmt) trim(param_group(n)), ':', trim(param_name(n)), ':', param_real(n)
      case default
        write (*, '(A30,A1,A20,A1,L10)') trim(param_group(n)), ':', trim(param_name(n)), ':', param_int(n) /= 0
      end select
    end do
    write (*, '(A)') '........................................................................'
  end subroutine printParams
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
