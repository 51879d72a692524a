! piernik_gdf.F90 -- ISO_C_BINDING interface to libpiernik_gdf.so: GDF-v2
! data files of device-resident blocks and the gdf_distance comparison
! (include/piernik_gdf.h).  The derived types are interoperable with the C
! structs member for member.
module piernik_gdf
   use iso_c_binding, only: c_int32_t, c_int64_t, c_double, c_char, c_ptr, c_null_char, c_f_pointer, c_associated
   implicit none
   private
   public :: pk_gdf_cg, pk_gdf_meta, pk_gdf_meta_default, pk_gdf_write, pk_gdf_write_host, pk_gdf_read, &
        &    pk_gdf_distance, pk_gdf_error_message

   ! one block of the file: cg%l%id, cg%n_b, cg%my_se(:, LO), cg%fbnd, cg%dl
   type, bind(C) :: pk_gdf_cg
      integer(c_int64_t) :: cg_id
      integer(c_int32_t) :: level
      integer(c_int32_t) :: n_b(3)
      integer(c_int64_t) :: off(3)
      real(c_double)     :: left_edge(3), right_edge(3), dl(3)
   end type pk_gdf_cg

   ! the run: the C arrays edge[3][2] and bnd[3][2] are edge(LO:HI, xdim:zdim)
   type, bind(C) :: pk_gdf_meta
      real(c_double)         :: time, timestep
      integer(c_int32_t)     :: nstep, nres, nhdf, require_problem_IC
      real(c_double)         :: last_log_time, last_tsl_time, last_hdf_time, last_res_time
      real(c_double)         :: c_all_old, stepcfl, cfl_c
      integer(c_int32_t)     :: nfluids
      real(c_double)         :: magic_mass(2)
      integer(c_int32_t)     :: n_d(3)
      real(c_double)         :: edge(2, 3)
      integer(c_int32_t)     :: bnd(2, 3)
      integer(c_int32_t)     :: eff_dim, h5_64bit, gdf_strict
      character(kind=c_char) :: problem_name(64)
      character(kind=c_char) :: run_id(16)
      real(c_double)         :: cm, gram, sek, miu0
      real(c_double)         :: lmtvB(5)
      character(kind=c_char) :: s_lmtvB(16, 7)
   end type pk_gdf_meta

   interface
      subroutine pk_gdf_meta_default(meta) bind(C, name="pk_gdf_meta_default")
         import :: pk_gdf_meta
         type(pk_gdf_meta), intent(out) :: meta
      end subroutine pk_gdf_meta_default

      ! fields: an array of c_ptr to NUL-terminated short names (hdf_vars)
      integer(c_int32_t) function pk_gdf_write(path, meta, ncg, cgs, nfields, fields) bind(C, name="pk_gdf_write")
         import :: c_int32_t, c_char, c_ptr, pk_gdf_meta, pk_gdf_cg
         character(kind=c_char), dimension(*), intent(in) :: path
         type(pk_gdf_meta),                    intent(in) :: meta
         integer(c_int32_t), value                        :: ncg, nfields
         type(pk_gdf_cg),     dimension(*),    intent(in) :: cgs
         type(c_ptr),         dimension(*),    intent(in) :: fields
      end function pk_gdf_write

      ! data(c * nfields + f + 1): c_loc of field f of block c, shape n_b
      integer(c_int32_t) function pk_gdf_write_host(path, meta, ncg, cgs, nfields, fields, data) &
           &                                      bind(C, name="pk_gdf_write_host")
         import :: c_int32_t, c_char, c_ptr, pk_gdf_meta, pk_gdf_cg
         character(kind=c_char), dimension(*), intent(in) :: path
         type(pk_gdf_meta),                    intent(in) :: meta
         integer(c_int32_t), value                        :: ncg, nfields
         type(pk_gdf_cg),     dimension(*),    intent(in) :: cgs
         type(c_ptr),         dimension(*),    intent(in) :: fields, data
      end function pk_gdf_write_host

      integer(c_int32_t) function pk_gdf_read(path, grid, name, out, n) bind(C, name="pk_gdf_read")
         import :: c_int32_t, c_int64_t, c_char, c_double
         character(kind=c_char), dimension(*), intent(in)  :: path, name
         integer(c_int32_t), value                         :: grid
         real(c_double),         dimension(*), intent(out) :: out
         integer(c_int64_t), value                         :: n
      end function pk_gdf_read

      integer(c_int32_t) function pk_gdf_distance(path1, path2, lev1, lev2, cap, norms, levels, names, names_len, &
           &                                    total) bind(C, name="pk_gdf_distance")
         import :: c_int32_t, c_int64_t, c_char, c_double
         character(kind=c_char), dimension(*), intent(in)  :: path1, path2
         integer(c_int32_t), value                         :: lev1, lev2, cap
         real(c_double),         dimension(*), intent(out) :: norms
         integer(c_int32_t),     dimension(*), intent(out) :: levels
         character(kind=c_char), dimension(*), intent(out) :: names
         integer(c_int64_t), value                         :: names_len
         real(c_double),                       intent(out) :: total
      end function pk_gdf_distance

      type(c_ptr) function pk_gdf_last_error_c() bind(C, name="pk_gdf_last_error")
         import :: c_ptr
      end function pk_gdf_last_error_c
   end interface

contains

   function pk_gdf_error_message() result(msg)
      character(len=:), allocatable :: msg
      character(kind=c_char), dimension(:), pointer :: p
      type(c_ptr) :: cp
      integer :: i, n
      cp = pk_gdf_last_error_c()
      msg = ''
      if (.not. c_associated(cp)) return
      call c_f_pointer(cp, p, [1024])
      n = 0
      do i = 1, 1024
         if (p(i) == c_null_char) exit
         n = i
      enddo
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = p(i)
      enddo
   end function pk_gdf_error_message

end module piernik_gdf
