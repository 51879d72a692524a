!
! drive_solve_cg.F90 -- a Fortran host program that drives the drop-in
! solver exactly as the solve_cg_hip adapter does inside PIERNIK: the block's
! host arrays u(nvar, nxg, nyg, nzg), b(3, ...), psi(...) are registered with
! their uh/magh/psih companions, and every call of the solve_cg_sub seam
! (sweeps.F90:140-151, called at :245) is pk_upload -> pk_solve_cg ->
! pk_download, followed by the host's own periodic ghost refresh, as
! update_boundaries runs after each stage (sweeps.F90:271).
!
! Usage: drive_solve_cg in.bin out.bin
!   in.bin  (stream): int32 nx, ny, nz, nb, limiter, limiter_b;
!                     float64 gamma, smalld, smallei, dt, chspeed;
!                     then u, b, psi (Fortran order, ghosts included)
!   out.bin (stream): int32 return code of every call, in order (20 calls:
!                     init, register, three per stage), int32 dn_negative,
!                     ei_negative, float64 magic_mass(2), then u, b, psi
!
program drive_solve_cg

   use iso_c_binding, only: c_int32_t, c_int64_t, c_double, c_loc
   use piernik_hip,   only: pk_config, pk_stage_report, pk_init, pk_finalize, pk_register_block, pk_upload, &
        &                   pk_solve_cg, pk_download, pk_get_report, pk_error_message, PK_RK2_1, PK_RK2_2

   implicit none

   integer(c_int32_t) :: hdr(6), n(3), nb, ng(3), rc, rcs(20), ncall
   real(c_double) :: par(5), dl(3)
   real(c_double), allocatable, target :: u(:,:,:,:), uh(:,:,:,:), b(:,:,:,:), bh(:,:,:,:), psi(:,:,:), psih(:,:,:)
   type(pk_config) :: cfg
   type(pk_stage_report) :: rep
   integer(c_int64_t), parameter :: id = 42
   integer :: d, s, istep
   character(len=512) :: fin, fout

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(10, file=trim(fin), access='stream', form='unformatted', status='old')
   read(10) hdr, par
   n = hdr(1:3)
   nb = hdr(4)
   ng = n + 2 * nb
   allocate(u(5, ng(1), ng(2), ng(3)), uh(5, ng(1), ng(2), ng(3)), b(3, ng(1), ng(2), ng(3)), &
        &   bh(3, ng(1), ng(2), ng(3)), psi(ng(1), ng(2), ng(3)), psih(ng(1), ng(2), ng(3)))
   read(10) u, b, psi
   close(10)
   uh = 0.0_c_double
   bh = 0.0_c_double
   psih = 0.0_c_double
   dl = 1.0_c_double / real(n, c_double)

   cfg%nfluids = 1
   cfg%limiter = hdr(5)
   cfg%limiter_b = hdr(6)
   cfg%integration_order = 2
   cfg%fluid(1)%gamma = par(1)
   cfg%fluid(1)%magnetized = 1
   cfg%fluid(1)%reserved = 0
   cfg%fluid(2)%gamma = par(1)
   cfg%fluid(2)%magnetized = 0
   cfg%fluid(2)%reserved = 0
   cfg%smalld = par(2)
   cfg%smallei = par(3)
   cfg%use_smalld = 1
   cfg%use_smallei = 1
   cfg%cfl = 0.7_c_double
   cfg%cfl_glm = 0.7_c_double
   cfg%glm_alpha = 0.1_c_double
   cfg%has_dir = 1
   cfg%tile_x = 0
   cfg%tile_yz = 0
   cfg%fast_math = 0
   cfg%reserved = 0

   ncall = 0
   rcs = -1
   call check(pk_init(cfg))
   call check(pk_register_block(id, n, nb, dl, c_loc(u), c_loc(uh), c_loc(b), c_loc(bh), c_loc(psi), c_loc(psih)))
   do d = 1, 3
      do s = 1, 2
         istep = merge(PK_RK2_1, PK_RK2_2, s == 1)
         ! solve_cg_hip(cg, ddim, istep): upload, one stage, download
         call check(pk_upload(id))
         call check(pk_solve_cg(id, int(d - 1, c_int32_t), int(istep, c_int32_t), par(4), par(5)))
         call check(pk_download(id))
         call periodic_ghosts()
      enddo
   enddo
   rc = pk_get_report(id, rep, 0_c_int32_t)
   ! pk_finalize after the report; its return code is not among the 20

   open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(11) rcs, rep%dn_negative, rep%ei_negative, rep%magic_mass, u, b, psi
   close(11)
   rc = pk_finalize()

contains

   subroutine check(code)
      integer(c_int32_t), intent(in) :: code
      if (ncall < size(rcs)) then
         ncall = ncall + 1
         rcs(ncall) = code
      endif
      if (code /= 0) then
         print '(a,i4,1x,a)', 'ERROR', code, pk_error_message()
         stop 1
      endif
   end subroutine check

   ! every ghost cell takes the value of its periodic interior image
   subroutine periodic_ghosts()
      integer :: i, j, k, si, sj, sk
      do k = 1, ng(3)
         sk = modulo(k - 1 - nb, n(3)) + nb + 1
         do j = 1, ng(2)
            sj = modulo(j - 1 - nb, n(2)) + nb + 1
            do i = 1, ng(1)
               si = modulo(i - 1 - nb, n(1)) + nb + 1
               if (si == i .and. sj == j .and. sk == k) cycle
               u(:, i, j, k) = u(:, si, sj, sk)
               b(:, i, j, k) = b(:, si, sj, sk)
               psi(i, j, k) = psi(si, sj, sk)
            enddo
         enddo
      enddo
   end subroutine periodic_ghosts

end program drive_solve_cg
