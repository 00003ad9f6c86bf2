!=======================================================================
! cc_iface -- ISO_C_BINDING interfaces to the MI355X time step
! (include/cc_abi.h).  The derived type mirrors cc_params field by field.
!=======================================================================
module cc_iface
use iso_c_binding
implicit none

integer(c_int), parameter :: CC_FP64_STRICT = 0, CC_FP64 = 1, CC_MIXED = 2

type, bind(C) :: cc_params
	integer(c_int) :: n1, n2, n3, istat, precision, reserved
	real(c_double) :: Re, Mac, dt, PexVal, PeyVal, PezVal, Ri_b, slip_len
end type cc_params

! wall condition of a study (cc_wall); filled by cc_wall_preset
type, bind(C) :: cc_wall
	integer(c_int) :: vel, stripes, temp, reserved
	real(c_double) :: amp, kx, omega, phase_odd, disc_omega, grad, Tw, dTdz
end type cc_wall

interface
	integer(c_int) function cc_set_wall(h, wall, w, x1, x2) bind(C, name="cc_set_wall")
		import :: c_int, c_ptr, c_double, cc_wall
		type(c_ptr), value :: h
		integer(c_int), value :: wall
		type(cc_wall), intent(in) :: w
		real(c_double), intent(in) :: x1(*), x2(*)
	end function cc_set_wall

	integer(c_int) function cc_wall_preset(name, wall, w) bind(C, name="cc_wall_preset")
		import :: c_int, c_char, cc_wall
		character(kind=c_char), intent(in) :: name(*)
		integer(c_int), intent(out) :: wall
		type(cc_wall), intent(out) :: w
	end function cc_wall_preset

	! the reference's solution files through the native parallel writer/reader
	integer(c_int) function cc_write_solution(path, time, nstep, n1, n2, n3, x1, x2, x3, var, &
			pattern, nthreads) bind(C, name="cc_write_solution")
		import :: c_int, c_char, c_double, c_ptr
		character(kind=c_char), intent(in) :: path(*)
		real(c_double), value :: time
		integer(c_int), value :: nstep, n1, n2, n3, nthreads
		real(c_double), intent(in) :: x1(*), x2(*), x3(*), var(*)
		type(c_ptr), value :: pattern
	end function cc_write_solution

	integer(c_int) function cc_read_solution(path, n1, n2, n3, x1, x2, x3, var, time, nstep, &
			nthreads) bind(C, name="cc_read_solution")
		import :: c_int, c_char, c_double
		character(kind=c_char), intent(in) :: path(*)
		integer(c_int), value :: n1, n2, n3, nthreads
		real(c_double), intent(out) :: x1(*), x2(*), x3(*), var(*)
		real(c_double), intent(out) :: time
		integer(c_int), intent(out) :: nstep
	end function cc_read_solution

	type(c_ptr) function cc_io_last_error() bind(C, name="cc_io_last_error")
		import :: c_ptr
	end function cc_io_last_error

	integer(c_int) function cc_create(p, h) bind(C, name="cc_create")
		import :: c_int, c_ptr, cc_params
		type(cc_params), intent(in) :: p
		type(c_ptr), intent(out) :: h
	end function cc_create

	integer(c_int) function cc_upload(h, var, nstep, time) bind(C, name="cc_upload")
		import :: c_int, c_ptr, c_double
		type(c_ptr), value :: h
		real(c_double), intent(in) :: var(*)
		integer(c_int), value :: nstep
		real(c_double), value :: time
	end function cc_upload

	integer(c_int) function cc_step(h, nsteps) bind(C, name="cc_step")
		import :: c_int, c_ptr
		type(c_ptr), value :: h
		integer(c_int), value :: nsteps
	end function cc_step

	integer(c_int) function cc_stats(h, tau, stat1, stat2, corr) bind(C, name="cc_stats")
		import :: c_int, c_ptr, c_double
		type(c_ptr), value :: h
		real(c_double), intent(out) :: tau(2), stat1(*), stat2(*), corr(*)
	end function cc_stats

	integer(c_int) function cc_diag(h, x3f, x2e, out) bind(C, name="cc_diag")
		import :: c_int, c_ptr, c_double
		type(c_ptr), value :: h
		real(c_double), intent(in) :: x3f(*)
		real(c_double), value :: x2e
		real(c_double), intent(out) :: out(16)
	end function cc_diag

	integer(c_int) function cc_download(h, var, nstep, time) bind(C, name="cc_download")
		import :: c_int, c_ptr, c_double
		type(c_ptr), value :: h
		real(c_double), intent(out) :: var(*)
		integer(c_int), intent(out) :: nstep
		real(c_double), intent(out) :: time
	end function cc_download

	integer(c_int) function cc_blown(h, ijk) bind(C, name="cc_blown")
		import :: c_int, c_ptr
		type(c_ptr), value :: h
		integer(c_int), intent(out) :: ijk(3)
	end function cc_blown

	subroutine cc_destroy(h) bind(C, name="cc_destroy")
		import :: c_ptr
		type(c_ptr), value :: h
	end subroutine cc_destroy

	type(c_ptr) function cc_last_error() bind(C, name="cc_last_error")
		import :: c_ptr
	end function cc_last_error
end interface

contains

! message of the last failed call, as a Fortran string
function cc_errmsg() result(msg)
	character(len=:), allocatable :: msg
	type(c_ptr) :: p
	character(kind=c_char), pointer :: s(:)
	integer :: n
	p = cc_last_error()
	call c_f_pointer(p, s, [512])
	n = 0
	do while (n < 512)
		if (s(n+1) == c_null_char) exit
		n = n + 1
	enddo
	allocate(character(len=n) :: msg)
	msg = transfer(s(1:n), msg)
end function cc_errmsg

function cc_ioerrmsg() result(msg)
	character(len=:), allocatable :: msg
	character(kind=c_char), pointer :: cs(:)
	type(c_ptr) :: p
	integer :: n
	p = cc_io_last_error()
	call c_f_pointer(p, cs, [512])
	n = 0
	do while (n < 512)
		if (cs(n+1) == c_null_char) exit
		n = n + 1
	enddo
	allocate(character(len=n) :: msg)
	msg = transfer(cs(1:n), msg)
end function cc_ioerrmsg

end module cc_iface
