! m_tff.F90 -- ISO_C_BINDING interface to libtff.so (include/tff.h).
module m_tff
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TFF_OK = 0
  integer(c_int), parameter :: TFF_FP64 = 0, TFF_FP32 = 1, TFF_FP32_BULK = 2, TFF_FP32_COMP = 3
  integer(c_int), parameter :: TFF_OPEN = 0, TFF_PERIODIC = 1, TFF_ABSORB = 2
  integer(c_int), parameter :: TFF_EX = 0, TFF_EY = 1, TFF_EZ = 2, TFF_BX = 3, TFF_BY = 4, &
                               TFF_BZ = 5, TFF_B0X = 6, TFF_B0Y = 7, TFF_B0Z = 8
  integer(c_int), parameter :: TFF_HOST_F64 = 0, TFF_HOST_F32 = 1
  integer(c_int), parameter :: TFF_BC_NONE = 0, TFF_BC_LOOP = 1, TFF_BC_PSR = 2, TFF_BC_UNIFORM = 3

  type, bind(C) :: tff_config
    integer(c_int) :: nx, ny, nz
    integer(c_int) :: boundary(3)
    integer(c_int) :: absorb_buff
    integer(c_int) :: absorb_fixed
    integer(c_int) :: precision
    integer(c_int) :: device
    real(c_double) :: cc
  end type tff_config

  interface
    integer(c_int) function tff_create(cfg, h) bind(C, name='tff_create')
      import :: c_int, c_ptr, tff_config
      type(tff_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function tff_create

    subroutine tff_destroy(h) bind(C, name='tff_destroy')
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine tff_destroy

    type(c_ptr) function tff_last_error(h) bind(C, name='tff_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function tff_last_error

    integer(c_int) function tff_upload(h, field, host, dtype) bind(C, name='tff_upload')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, host
      integer(c_int), value :: field, dtype
    end function tff_upload

    integer(c_int) function tff_download(h, field, host, dtype) bind(C, name='tff_download')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, host
      integer(c_int), value :: field, dtype
    end function tff_download

    integer(c_int) function tff_set_bc(h, kind, params, n) bind(C, name='tff_set_bc')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: kind, n
      real(c_double), intent(in) :: params(*)
    end function tff_set_bc

    integer(c_int) function tff_init_rk_step(h) bind(C, name='tff_init_rk_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function tff_init_rk_step

    integer(c_int) function tff_rk3_step(h, c1, c2, c3) bind(C, name='tff_rk3_step')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: c1, c2, c3
    end function tff_rk3_step

    integer(c_int) function tff_clean_epar(h) bind(C, name='tff_clean_epar')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function tff_clean_epar

    integer(c_int) function tff_exchange_fields(h, b0) bind(C, name='tff_exchange_fields')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: b0
    end function tff_exchange_fields

    integer(c_int) function tff_user_bc(h, step) bind(C, name='tff_user_bc')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: step
    end function tff_user_bc

    integer(c_int) function tff_step(h, step) bind(C, name='tff_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: step
    end function tff_step

    integer(c_int) function tff_run(h, first, n) bind(C, name='tff_run')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: first, n
    end function tff_run

    integer(c_int) function tff_reduce(h, out) bind(C, name='tff_reduce')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(4)
    end function tff_reduce

    integer(c_int) function tff_output_dims(h, istep, dims) bind(C, name='tff_output_dims')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: istep
      integer(c_int), intent(out) :: dims(3)
    end function tff_output_dims

    integer(c_int) function tff_output_fields(h, istep, out, host_dtype) bind(C, name='tff_output_fields')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: istep, host_dtype
      type(c_ptr), intent(in) :: out(6)
    end function tff_output_fields

    integer(c_int) function tff_sync(h) bind(C, name='tff_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function tff_sync

    integer(c_int) function tff_set_phase_timing(h, on) bind(C, name='tff_set_phase_timing')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function tff_set_phase_timing

    integer(c_int) function tff_phase_times(h, ms) bind(C, name='tff_phase_times')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ms(3)
    end function tff_phase_times

    ! include/tff_io.h (libtff_io.so)
    integer(c_int) function tff_io_write_output(h, dir, idx, timestep, istep, user_output, n, names, types, &
                                                ivals, rvals) bind(C, name='tff_io_write_output')
      import :: c_int, c_ptr, c_char, c_float
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: dir(*)
      integer(c_int), value :: idx, timestep, istep, user_output, n
      character(kind=c_char), intent(in) :: names(*)
      integer(c_int), intent(in) :: types(*), ivals(*)
      real(c_float), intent(in) :: rvals(*)
    end function tff_io_write_output

    type(c_ptr) function tff_io_last_error() bind(C, name='tff_io_last_error')
      import :: c_ptr
    end function tff_io_last_error
  end interface

  integer(c_int), parameter :: TFF_PHASE_SOLVER = 0, TFF_PHASE_HOOK = 1, TFF_PHASE_EXCHANGE = 2
  integer(c_int), parameter :: TFF_USEROUT_NONE = 0, TFF_USEROUT_RHO = 1, TFF_USEROUT_UNIFORM = 2

contains
  ! text of a C string as a Fortran string
  function c_text(p) result(msg)
    type(c_ptr), intent(in) :: p
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function c_text

  ! text of the handle's last error as a Fortran string
  function tff_error_text(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ''
    p = tff_last_error(h)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function tff_error_text
end module m_tff
