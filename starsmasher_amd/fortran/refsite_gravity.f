c     refsite_gravity.f -- drives the reference's own gravity call sites,
c     compiled unmodified from StarSmasher's source tree and linked against
c     libsphgrav_hip.so:
c       set_nusegpus             (pbe/src/gpu_grav.f:1-8)
c       gravforce                (pbe/src/grav.f:1-31: range = +-4 hp^2,
c                                 firsthalf_grav_forces over this rank's chunk)
c       lasthalf_grav_forces     (the call of uvdots, pbe/src/balAV3.f:418)
c     The particle data sit in the reference's common blocks
c     (pbe/src/starsmasher.h), filled here as main.f / init.f would fill
c     them for one gravity rank.
c
c     usage: refsite_gravity <input> <output>
c     input  (stream): int32 n, nlo, nhi, nkernel; then x, y, z, am, hp, u
c                      as n float64 each
c     output (stream): gx, gy, gz, grpot as n float64 each; entries outside
c                      nlo..nhi keep the sentinel -7777
      program refsite_gravity
      include 'starsmasher.h'
      integer i, nlo, nhi, nk, dev
      real*8 theta
      character*512 fin, fout

      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
      open(10, file=fin, access='stream', form='unformatted',
     $     status='old')
      read(10) ntot, nlo, nhi, nk
      if (ntot .gt. nmax) then
         print *, 'refsite_gravity: n exceeds nmax of starsmasher.h'
         stop 2
      endif
      read(10) (x(i), i=1,ntot)
      read(10) (y(i), i=1,ntot)
      read(10) (z(i), i=1,ntot)
      read(10) (am(i), i=1,ntot)
      read(10) (hp(i), i=1,ntot)
      read(10) (u(i), i=1,ntot)
      close(10)

      do i = 1, ntot
         gx(i) = -7777.d0
         gy(i) = -7777.d0
         gz(i) = -7777.d0
         grpot(i) = -7777.d0
      enddo

c     one gravity rank owning particles nlo..nhi (main.f:284-295 with
c     ngravprocs = 1 gives 1..n; a sub-range stands for one rank's chunk)
      n = ntot
      myrank = 0
      nprocs = 1
      ngravprocs = 1
      ngrav_lower = nlo
      ngrav_upper = nhi
      q = 1
      nkernel = nk
      call set_nusegpus

      dev = 0
      theta = 0.3d0
      call gpu_init_dev(dev, theta)
      call gravforce
      call lasthalf_grav_forces(ntot, gx, gy, gz, grpot)

      open(11, file=fout, access='stream', form='unformatted',
     $     status='replace')
      write(11) (gx(i), i=1,ntot), (gy(i), i=1,ntot),
     $     (gz(i), i=1,ntot), (grpot(i), i=1,ntot)
      close(11)
      end
