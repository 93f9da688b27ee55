! sphgrav_mod.f90 -- iso_c_binding interfaces to libsphgrav_hip.so.
!
! Existing StarSmasher call sites need nothing from this module: they call
! firsthalf_grav_forces / lasthalf_grav_forces / gpu_init_dev through
! Fortran's implicit interface (pbe/src/grav.f:25-26, balAV3.f:418,
! main.f:270), which the library exports verbatim.  This module gives new
! Fortran code explicit interfaces: the three legacy entry points (so the
! compiler checks argument types) and the status-returning sg_* API of
! include/sphgrav.h.
module sphgrav
  use iso_c_binding, only: c_int, c_double, c_ptr, c_char, c_null_ptr
  implicit none
  private

  integer(c_int), parameter, public :: SG_OK = 0, SG_ERR_ARG = 1, SG_ERR_NKERNEL = 2, &
       SG_ERR_MASS = 3, SG_ERR_RANGE2 = 4, SG_ERR_NONFINITE = 5, SG_ERR_HIP = 6, &
       SG_ERR_NO_DEVICE = 7, SG_ERR_STATE = 8, SG_ERR_CONFIG = 9
  integer(c_int), parameter, public :: SG_PREC_FP32C = 0, SG_PREC_FP64 = 1

  public :: firsthalf_grav_forces, lasthalf_grav_forces, gpu_init_dev
  public :: sg_init_device, sg_firsthalf, sg_lasthalf, sg_last_status, sg_device_count
  public :: sg_grav_psi
  public :: sg_status_message

  interface
     ! grav_force_direct.cu:650-656 (all arguments by reference)
     subroutine firsthalf_grav_forces(n, n_lower, my_length, px, py, pz, mass, range2, q, nkernel) &
          bind(C, name='firsthalf_grav_forces_')
       import :: c_int, c_double
       integer(c_int), intent(in) :: n, n_lower, my_length, q, nkernel
       real(c_double), intent(in) :: px(*), py(*), pz(*), mass(*), range2(*)
     end subroutine firsthalf_grav_forces

     ! grav_force_direct.cu:658-664
     subroutine lasthalf_grav_forces(n, ax, ay, az, pot, myrank) bind(C, name='lasthalf_grav_forces_')
       import :: c_int, c_double
       integer(c_int), intent(in) :: n, myrank
       real(c_double), intent(inout) :: ax(*), ay(*), az(*), pot(*)
     end subroutine lasthalf_grav_forces

     ! grav_force_direct.cu:672-675
     subroutine gpu_init_dev(device, theta) bind(C, name='gpu_init_dev_')
       import :: c_int, c_double
       integer(c_int), intent(in) :: device
       real(c_double), intent(in) :: theta
     end subroutine gpu_init_dev

     integer(c_int) function sg_init_device(device) bind(C, name='sg_init_device')
       import :: c_int
       integer(c_int), value :: device
     end function sg_init_device

     integer(c_int) function sg_firsthalf(n, n_lower, my_length, x, y, z, m, range2, nkernel) &
          bind(C, name='sg_firsthalf')
       import :: c_int, c_double
       integer(c_int), value :: n, n_lower, my_length, nkernel
       real(c_double), intent(in) :: x(*), y(*), z(*), m(*), range2(*)
     end function sg_firsthalf

     integer(c_int) function sg_lasthalf(ax, ay, az, pot) bind(C, name='sg_lasthalf')
       import :: c_int, c_double
       real(c_double), intent(inout) :: ax(*), ay(*), az(*), pot(*)
     end function sg_lasthalf

     ! grad-h softening term psi of particles n_lower..n_lower+my_length-1
     ! (advance.f90:716-737), written to psi(n_lower:...); blocking
     integer(c_int) function sg_grav_psi(n, n_lower, my_length, x, y, z, m, range2, nkernel, psi) &
          bind(C, name='sg_grav_psi')
       import :: c_int, c_double
       integer(c_int), value :: n, n_lower, my_length, nkernel
       real(c_double), intent(in) :: x(*), y(*), z(*), m(*), range2(*)
       real(c_double), intent(inout) :: psi(*)
     end function sg_grav_psi

     integer(c_int) function sg_last_status() bind(C, name='sg_last_status')
       import :: c_int
     end function sg_last_status

     integer(c_int) function sg_device_count(count) bind(C, name='sg_device_count')
       import :: c_int
       integer(c_int), intent(out) :: count
     end function sg_device_count

     type(c_ptr) function sg_status_string_c(status) bind(C, name='sg_status_string')
       import :: c_int, c_ptr
       integer(c_int), value :: status
     end function sg_status_string_c
  end interface

contains

  ! The C status string as a Fortran character variable.
  function sg_status_message(status) result(msg)
    use iso_c_binding, only: c_f_pointer
    integer(c_int), intent(in) :: status
    character(len=128) :: msg
    character(kind=c_char), pointer :: chars(:)
    type(c_ptr) :: p
    integer :: k
    msg = ' '
    p = sg_status_string_c(status)
    call c_f_pointer(p, chars, [128])
    do k = 1, 128
       if (chars(k) == char(0)) exit
       msg(k:k) = chars(k)
    end do
  end function sg_status_message

end module sphgrav
