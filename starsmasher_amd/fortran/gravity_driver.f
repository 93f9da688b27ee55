c     gravity_driver.f -- a Fortran host calling the gravity library the
c     way StarSmasher does: implicit interfaces, arguments by reference,
c     range = 4 hp^2 negated for point particles (pbe/src/grav.f:15-18),
c     then firsthalf_grav_forces / lasthalf_grav_forces (grav.f:25-26,
c     balAV3.f:418) and gpu_init_dev (main.f:270).
c
c     usage: gravity_driver <input> <output>
c     input  (stream): int32 n, nlo, nhi, nkernel; then x, y, z, am, hp, u
c                      as n float64 each
c     output (stream): gx, gy, gz, grpot as n float64 each; entries outside
c                      nlo..nhi keep the sentinel -7777
      program gravity_driver
      implicit none
      integer n, nlo, nhi, nkernel, i, q, mylen, dev
      real*8 theta
      real*8, allocatable :: x(:), y(:), z(:), am(:), hp(:), u(:)
      real*8, allocatable :: range2(:), gx(:), gy(:), gz(:), grpot(:)
      character*512 fin, fout

      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
      open(10, file=fin, access='stream', form='unformatted',
     $     status='old')
      read(10) n, nlo, nhi, nkernel
      allocate(x(n), y(n), z(n), am(n), hp(n), u(n), range2(n))
      allocate(gx(n), gy(n), gz(n), grpot(n))
      read(10) x, y, z, am, hp, u
      close(10)

      do i = 1, n
         range2(i) = 4*hp(i)**2
         if (u(i) .eq. 0.d0) range2(i) = -range2(i)
         gx(i) = -7777.d0
         gy(i) = -7777.d0
         gz(i) = -7777.d0
         grpot(i) = -7777.d0
      enddo

      dev = 0
      theta = 0.3d0
      call gpu_init_dev(dev, theta)
      q = 1
      mylen = nhi - nlo + 1
      call firsthalf_grav_forces(n, nlo, mylen, x, y, z, am, range2,
     $     q, nkernel)
      call lasthalf_grav_forces(n, gx, gy, gz, grpot)

      open(11, file=fout, access='stream', form='unformatted',
     $     status='replace')
      write(11) gx, gy, gz, grpot
      close(11)
      end
