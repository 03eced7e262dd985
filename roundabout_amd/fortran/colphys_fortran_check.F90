!> Fortran host check of the column-physics C ABI.
!>   colphys_fortran_check cpu  -- ABI reachable without a GPU: defaults, error
!>                                 ring, refused knobs, NULL-safe destroy
!>   colphys_fortran_check gpu  -- a 16 x 12 x 8 ocean stepped 3 times; prints
!>                                 the outputs' checksums and a parity line
!>                                 against the staged entry points, and a
!>                                 two-member ensemble whose members equal
!>                                 the single-grid step bitwise, and an ALE
!>                                 remap through device pointers
program colphys_fortran_check
   use, intrinsic :: iso_c_binding
   use rdb_colphys
   implicit none
   type(rdb_colphys_config_t) :: cfg
   type(c_ptr) :: h
   integer(c_int) :: st
   character(len=8) :: mode
   integer, parameter :: nx = 16, ny = 12, nz = 8
   real(c_double), target :: wet(nx, ny)
   real(c_double), target :: hl(nx, ny, nz), T(nx, ny, nz), S(nx, ny, nz)
   real(c_double) :: u(nx + 1, ny, nz), v(nx, ny + 1, nz), tx(nx + 1, ny, 1), ty(nx, ny + 1, 1)
   real(c_double) :: qh(nx, ny, 1), qs(nx, ny, 1), T1(nx, ny, nz), T2(nx, ny, nz), kt1(nx, ny, nz + 1)
   real(c_double) :: kt2(nx, ny, nz + 1), heat, salt
   integer :: i, j, k, m
   integer(c_int) :: n_mem, enx(2), eny(2), enz(2)
   type(c_ptr) :: ewet(2)
   type(rdb_colphys_momentum_config_t) :: mcfg
   real(c_double) :: u1(nx + 1, ny, nz), hf(nz), m0, m1, mabs
   real(c_double) :: hb(nx, ny, nz), tnow
   integer(c_int) :: nsteps
   character(len=*), parameter :: nml = "&ocean_vmix_nml use_kpp = .true., pp81_alpha = 5.0 /"
   real(c_double), target :: hn(nx, ny, nz), cr(nx, ny, nz), bud(nx, ny, nz)
   type(c_ptr) :: d_ho, d_hn, d_c, d_b
   integer(c_int) :: n_bad
   real(c_double) :: w_rel, w_neg, c0, c1
   ! the HIP runtime's allocation and copy, for device arrays of this host
   interface
      integer(c_int) function hipMalloc(p, nbytes) bind(c, name="hipMalloc")
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), intent(out) :: p
         integer(c_size_t), value :: nbytes
      end function
      integer(c_int) function hipFree(p) bind(c, name="hipFree")
         import :: c_int, c_ptr
         type(c_ptr), value :: p
      end function
      integer(c_int) function hipMemcpy(dst, src, nbytes, kind) bind(c, name="hipMemcpy")
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: dst, src
         integer(c_size_t), value :: nbytes
         integer(c_int), value :: kind
      end function
      integer(c_int) function hipDeviceSynchronize() bind(c, name="hipDeviceSynchronize")
         import :: c_int
      end function
   end interface

   call get_command_argument(1, mode)
   st = rdb_colphys_config_default(cfg)
   if (st /= 0 .or. cfg%rho0 /= 1035.0_c_double .or. cfg%pp81_nu0 /= 1.0e-2_c_double) stop 2
   if (rdb_colphys_working_precision() /= 8) stop 3
   h = c_null_ptr
   st = rdb_colphys_destroy(h)
   if (st /= 0) stop 4

   wet = 1.0_c_double
   wet(1, 1) = 0.0_c_double
   do k = 1, nz
      do j = 1, ny
         do i = 1, nx
            hl(i, j, k) = 25.0_c_double + real(mod(i + j, 3), c_double)
            T(i, j, k) = 4.0_c_double + 0.5_c_double*real(k, c_double) + 0.01_c_double*real(i - j, c_double)
            S(i, j, k) = 35.0_c_double - 0.01_c_double*real(k, c_double)
         end do
      end do
   end do
   hl(1, 1, :) = 1.5e-4_c_double
   T(1, 1, :) = 0.0_c_double
   S(1, 1, :) = 0.0_c_double
   do k = 1, nz
      do j = 1, ny
         do i = 1, nx + 1
            u(i, j, k) = 0.05_c_double*sin(real(i*k, c_double))
         end do
      end do
      do j = 1, ny + 1
         do i = 1, nx
            v(i, j, k) = 0.05_c_double*cos(real(j*k, c_double))
         end do
      end do
   end do
   tx = 0.08_c_double
   ty = 0.02_c_double
   qh = -60.0_c_double
   qs = 0.0_c_double

   if (trim(mode) == "cpu") then
      ! a configure-time check of a contributor reaches Fortran (the EPBL
      ! fields of the bind(c) mirror line up with the C struct)
      cfg%epbl_enable = 1
      cfg%epbl_translay_scale = 1.0_c_double
      st = rdb_colphys_create(cfg, nx, ny, nz, 0, wet, c_null_ptr, c_null_ptr, h)
      if (st /= RDB_COLPHYS_ERR_CONFIG_VALIDATE .or. c_associated(h)) stop 5
      if (index(colphys_last_error(), "translay_scale") == 0) stop 6
      cfg%epbl_enable = 0
      cfg%epbl_translay_scale = 0.1_c_double
      st = rdb_colphys_create(cfg, nx, ny, nz, 0, wet, c_null_ptr, c_null_ptr, h)
      if (st /= RDB_COLPHYS_ERR_SETUP .or. c_associated(h)) stop 7
      print '(a)', "fortran-abi-cpu: ok"
      stop
   end if

   ! gpu: fused step vs the staged entry points on two identical handles
   ! (the first one created from namelist text, with the budget slots kept)
   st = rdb_colphys_create_from_string(nml, len(nml, c_int), nx, ny, nz, 0, wet, c_null_ptr, c_null_ptr, h)
   if (st /= 0) then
      print '(a)', trim(colphys_last_error())
      stop 10
   end if
   call upload()
   if (rdb_colphys_set_budgets(h, 1_c_int) /= 0) stop 18
   st = rdb_colphys_step(h, 3600.0_c_double, 3)
   if (st /= 0) stop 11
   if (rdb_colphys_get_step_count(h, nsteps) /= 0 .or. nsteps /= 3) stop 19
   if (rdb_colphys_get_time(h, tnow) /= 0 .or. tnow /= 3.0_c_double*3600.0_c_double) stop 19
   ! the vdiff heat budget telescopes in every column
   if (colphys_get(h, "heat_budget_vdiff", hb) /= 0) stop 40
   do j = 1, ny
      do i = 1, nx
         if (abs(sum(hb(i, j, :))) > 1.0e-12_c_double*sum(abs(hl(i, j, :)*T(i, j, :)))) stop 41
      end do
   end do
   if (maxval(abs(hb)) <= 0.0_c_double) stop 42
   st = colphys_get(h, "temperature", T1)
   st = colphys_get(h, "kt", kt1)
   st = rdb_colphys_heat_salt_totals(h, heat, salt)
   ! momentum vertical viscosity with the closure's kv: the flux form keeps
   ! sum(hf * u) of every interior u-face column
   st = rdb_colphys_momentum_config_default(mcfg)
   if (st /= 0 .or. mcfg%use_kv_source /= 1 .or. mcfg%hbbl_visc /= 10.0_c_double) stop 13
   st = rdb_colphys_vdiff_momentum(h, 3600.0_c_double, mcfg)
   if (st /= 0) then
      print '(a)', trim(colphys_last_error())
      stop 14
   end if
   if (colphys_get(h, "u_face_x_layer", u1) /= 0) stop 15
   if (all(u1 == u)) stop 16
   do j = 1, ny
      do i = 2, nx
         hf = 0.5_c_double*(hl(i - 1, j, :) + hl(i, j, :))
         m0 = sum(hf*u(i, j, :))
         m1 = sum(hf*u1(i, j, :))
         mabs = sum(abs(hf*u(i, j, :)))
         if (abs(m1 - m0) > 1.0e-13_c_double*mabs) stop 17
      end do
   end do
   st = rdb_colphys_destroy(h)
   if (c_associated(h)) stop 12

   st = rdb_colphys_create(cfg, nx, ny, nz, 0, wet, c_null_ptr, c_null_ptr, h)
   call upload()
   do i = 1, 3
      st = rdb_colphys_eos(h)
      st = rdb_colphys_coefficients(h, 3600.0_c_double)
      st = rdb_colphys_surface_flux(h, 3600.0_c_double)
      st = rdb_colphys_diffuse_tracers(h, 3600.0_c_double)
      st = rdb_colphys_nonlocal(h, 3600.0_c_double)
   end do
   st = colphys_get(h, "temperature", T2)
   st = colphys_get(h, "kt", kt2)
   st = rdb_colphys_destroy(h)
   if (any(T1 /= T2) .or. any(kt1 /= kt2)) then
      print '(a)', "fortran-abi-gpu: fused and staged paths differ"
      stop 13
   end if

   ! the same ocean twice as a two-member ensemble, one launch per step
   enx = nx
   eny = ny
   enz = nz
   ewet = [c_loc(wet), c_loc(wet)]
   st = rdb_colphys_ensemble_create(cfg, 2, enx, eny, enz, 0, ewet, c_null_ptr, h)
   if (st /= 0) then
      print '(a)', trim(colphys_last_error())
      stop 14
   end if
   st = rdb_colphys_member_count(h, n_mem)
   if (st /= 0 .or. n_mem /= 2) stop 15
   if (colphys_set(h, "temperature", T) /= RDB_COLPHYS_ERR_BAD_HANDLE) stop 16
   do m = 0, 1
      call upload_member(m)
   end do
   st = rdb_colphys_step(h, 3600.0_c_double, 3)
   if (st /= 0) stop 17
   do m = 0, 1
      T2 = 0.0_c_double
      kt2 = 0.0_c_double
      if (colphys_member_get(h, m, "temperature", T2) /= 0) stop 18
      if (colphys_member_get(h, m, "kt", kt2) /= 0) stop 18
      if (any(T1 /= T2) .or. any(kt1 /= kt2)) then
         print '(a, i0, a)', "fortran-abi-gpu: ensemble member ", m, " differs from the single grid"
         stop 19
      end if
   end do
   st = rdb_colphys_destroy(h)

   ! ALE remap of T through device pointers onto a grid whose bed layer gave
   ! 5 m to the surface layer: content conserved per column (the bounding
   ! moves nothing on this smooth profile), the precondition scan clean
   hn = hl
   do j = 1, ny
      do i = 1, nx
         if (hl(i, j, 1) > 1.0_c_double) then
            hn(i, j, 1) = hl(i, j, 1) - 5.0_c_double
            hn(i, j, nz) = hl(i, j, nz) + 5.0_c_double
         end if
      end do
   end do
   cr = T
   bud = 0.0_c_double
   if (hipMalloc(d_ho, int(8*nx*ny*nz, c_size_t)) /= 0) stop 50
   if (hipMalloc(d_hn, int(8*nx*ny*nz, c_size_t)) /= 0) stop 50
   if (hipMalloc(d_c, int(8*nx*ny*nz, c_size_t)) /= 0) stop 50
   if (hipMalloc(d_b, int(8*nx*ny*nz, c_size_t)) /= 0) stop 50
   st = hipMemcpy(d_ho, c_loc(hl), int(8*nx*ny*nz, c_size_t), 1_c_int)
   st = hipMemcpy(d_hn, c_loc(hn), int(8*nx*ny*nz, c_size_t), 1_c_int)
   st = hipMemcpy(d_c, c_loc(cr), int(8*nx*ny*nz, c_size_t), 1_c_int)
   st = hipMemcpy(d_b, c_loc(bud), int(8*nx*ny*nz, c_size_t), 1_c_int)
   st = rdb_remap_scan_preconditions(nx, ny, nz, d_ho, d_hn, 1.0e-9_c_double, n_bad, w_rel, w_neg, c_null_ptr)
   if (st /= 0 .or. n_bad /= 0 .or. w_neg /= 0.0_c_double) stop 51
   st = rdb_remap_tracer_field(RDB_REMAP_PPM, nx, ny, nz, d_ho, d_hn, d_c, d_c, 0_c_int, 0_c_int, d_b, &
                               c_null_ptr, c_null_ptr)
   if (st /= 0) then
      print '(a)', trim(colphys_last_error())
      stop 52
   end if
   if (hipDeviceSynchronize() /= 0) stop 53
   st = hipMemcpy(c_loc(cr), d_c, int(8*nx*ny*nz, c_size_t), 2_c_int)
   st = hipMemcpy(c_loc(bud), d_b, int(8*nx*ny*nz, c_size_t), 2_c_int)
   do j = 1, ny
      do i = 1, nx
         c0 = sum(hl(i, j, :)*T(i, j, :))
         c1 = sum(hn(i, j, :)*cr(i, j, :))
         if (abs(c1 - c0) > 1.0e-12_c_double*abs(c0)) stop 54
         if (abs(sum(bud(i, j, :)) - (c1 - c0)) > 1.0e-12_c_double*abs(c0)) stop 55
      end do
   end do
   if (all(cr == T)) stop 56
   st = hipFree(d_ho)
   st = hipFree(d_hn)
   st = hipFree(d_c)
   st = hipFree(d_b)
   print '(a, es24.16, a, es24.16)', "fortran-abi-gpu: ok heat=", heat, " sum(T)=", sum(T1)

contains

   subroutine upload()
      real(c_double) :: w3(nx, ny, 1)
      w3(:, :, 1) = wet
      if (colphys_set(h, "h_layer", hl) /= 0) stop 20
      if (colphys_set(h, "temperature", T) /= 0) stop 21
      if (colphys_set(h, "salinity", S) /= 0) stop 22
      if (colphys_set(h, "u_face_x_layer", u) /= 0) stop 23
      if (colphys_set(h, "v_face_y_layer", v) /= 0) stop 24
      if (colphys_set(h, "tau_x", tx) /= 0) stop 25
      if (colphys_set(h, "tau_y", ty) /= 0) stop 26
      if (colphys_set(h, "Q_heat", qh) /= 0) stop 27
      if (colphys_set(h, "Q_salt", qs) /= 0) stop 28
   end subroutine upload

   subroutine upload_member(mm)
      integer, intent(in) :: mm
      if (colphys_member_set(h, mm, "h_layer", hl) /= 0) stop 30
      if (colphys_member_set(h, mm, "temperature", T) /= 0) stop 31
      if (colphys_member_set(h, mm, "salinity", S) /= 0) stop 32
      if (colphys_member_set(h, mm, "u_face_x_layer", u) /= 0) stop 33
      if (colphys_member_set(h, mm, "v_face_y_layer", v) /= 0) stop 34
      if (colphys_member_set(h, mm, "tau_x", tx) /= 0) stop 35
      if (colphys_member_set(h, mm, "tau_y", ty) /= 0) stop 36
      if (colphys_member_set(h, mm, "Q_heat", qh) /= 0) stop 37
      if (colphys_member_set(h, mm, "Q_salt", qs) /= 0) stop 38
   end subroutine upload_member

end program colphys_fortran_check
