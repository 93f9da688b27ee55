c     ref_driver.f -- caller for the reference's own CPU gravity twin.
c
c     TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
c     with pbe/src/cpu_grav.f and pbe/src/starsmasher.h from the reference
c     checkout (untouched, where they lie) into oracle/_ref/libref_cpugrav.so.
c     The driver fills the common blocks that get_gravity_using_cpus reads
c     (starsmasher.h:23-29,55-56,69), calls it once, and copies gx, gy, gz and
c     grpot back.  It is used only to pin oracle/sg_oracle.c's restatement
c     and, optionally, as the timed CPU baseline.
      subroutine ref_cpu_gravity(npart, nlo, nhi, nk, nsg,
     $     xx, yy, zz, mm, hh, uu, ogx, ogy, ogz, opot, istat)
     $     bind(C, name='ref_cpu_gravity')
      use iso_c_binding
      include 'starsmasher.h'
      integer(c_int), value :: npart, nlo, nhi, nk, nsg
      real(c_double), intent(in) :: xx(npart), yy(npart), zz(npart)
      real(c_double), intent(in) :: mm(npart), hh(npart), uu(npart)
      real(c_double), intent(out) :: ogx(npart), ogy(npart), ogz(npart), opot(npart)
      integer(c_int), intent(out) :: istat
      integer i

      istat = 0
      if (npart .lt. 1 .or. npart .gt. nmax) then
         istat = 1
         return
      endif
      ntot = npart
      n = npart
      ngrav_lower = nlo
      ngrav_upper = nhi
      nkernel = nk
      nselfgravity = nsg
      do i = 1, npart
         x(i) = xx(i)
         y(i) = yy(i)
         z(i) = zz(i)
         am(i) = mm(i)
         hp(i) = hh(i)
         u(i) = uu(i)
      enddo
      call get_gravity_using_cpus
      do i = 1, npart
         ogx(i) = gx(i)
         ogy(i) = gy(i)
         ogz(i) = gz(i)
         opot(i) = grpot(i)
      enddo
      return
      end
