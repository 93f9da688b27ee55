! sphgrav_check.f90 -- new-style Fortran caller through the sphgrav module:
! explicit interfaces and status codes instead of the legacy exit-on-error.
!
! usage: sphgrav_check <input> <output>   (same file formats as gravity_driver)
! The run also checks that a bad nkernel comes back as SG_ERR_NKERNEL and
! prints "status-check ok" when it does, and appends the grad-h term psi of
! the same i-range (sg_grav_psi) to the output as a fifth array.
program sphgrav_check
  use iso_c_binding, only: c_int, c_double
  use sphgrav
  implicit none
  integer(c_int) :: n, nlo, nhi, nkernel, st, ndev
  real(c_double), allocatable :: x(:), y(:), z(:), am(:), hp(:), u(:)
  real(c_double), allocatable :: range2(:), gx(:), gy(:), gz(:), grpot(:), psi(:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=fin, access='stream', form='unformatted', status='old')
  read(10) n, nlo, nhi, nkernel
  allocate(x(n), y(n), z(n), am(n), hp(n), u(n), range2(n), gx(n), gy(n), gz(n), grpot(n), psi(n))
  read(10) x, y, z, am, hp, u
  close(10)
  range2 = 4.0d0 * hp**2
  where (u == 0.0d0) range2 = -range2
  gx = -7777.0d0; gy = -7777.0d0; gz = -7777.0d0; grpot = -7777.0d0; psi = -7777.0d0

  st = sg_device_count(ndev)
  if (st /= SG_OK) stop 'no HIP device'
  st = sg_init_device(0_c_int)
  if (st /= SG_OK) stop 'sg_init_device failed'

  st = sg_firsthalf(n, nlo, nhi - nlo + 1, x, y, z, am, range2, 7_c_int)
  if (st /= SG_ERR_NKERNEL) then
     print *, 'unexpected status for nkernel=7: ', st, trim(sg_status_message(st))
     stop 1
  end if
  print '(a)', 'status-check ok'

  st = sg_firsthalf(n, nlo, nhi - nlo + 1, x, y, z, am, range2, nkernel)
  if (st /= SG_OK) then
     print *, 'sg_firsthalf: ', trim(sg_status_message(st))
     stop 2
  end if
  st = sg_lasthalf(gx, gy, gz, grpot)
  if (st /= SG_OK) then
     print *, 'sg_lasthalf: ', trim(sg_status_message(st))
     stop 3
  end if

  st = sg_grav_psi(n, nlo, nhi - nlo + 1, x, y, z, am, range2, nkernel, psi)
  if (st /= SG_OK) then
     print *, 'sg_grav_psi: ', trim(sg_status_message(st))
     stop 4
  end if

  open(11, file=fout, access='stream', form='unformatted', status='replace')
  write(11) gx, gy, gz, grpot, psi
  close(11)
end program sphgrav_check
