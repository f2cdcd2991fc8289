! trj_hip.f90 -- ISO_C_BINDING interface to the trj_hip C ABI (include/trj_hip.h).
!
! The derived types mirror the C structs field for field (bind(C) gives them
! the C layout); every function returns TRJ_OK (0) or a negative TRJ_E* code.
module trj_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TRJ_OK = 0
  integer(c_int32_t), parameter :: TRJ_F32 = 0, TRJ_F64 = 1
  integer(c_int32_t), parameter :: TRJ_AOS = 0, TRJ_SOA = 1
  integer(c_int32_t), parameter :: TRJ_WANT_RDF = 1, TRJ_WANT_SQ = 2, TRJ_WANT_FQT = 4
  integer(c_int32_t), parameter :: TRJ_QSET_REF = 0, TRJ_QSET_HALFSPACE = 1
  integer(c_int32_t), parameter :: TRJ_PREC_FP64 = 0, TRJ_PREC_FP32 = 1, &
                                   TRJ_PREC_FP16X3 = 2, TRJ_PREC_BF16X3 = 3
  integer(c_int32_t), parameter :: TRJ_SQ_NQ_CONSISTENT = 1, TRJ_SQ_CHARGE = 2
  integer(c_int32_t), parameter :: TRJ_RDF_CHARGE = 1, TRJ_RDF_FLOOR_BINS = 2
  integer(c_int), parameter :: TRJ_PIPE_SERIAL = 0, TRJ_PIPE_OVERLAP = 1
  integer(c_int), parameter :: TRJ_STATE_REPLACE = 0, TRJ_STATE_ADD = 1

  type, bind(C) :: trj_rdf_params
    integer(c_int32_t) :: ndim, nsp
    real(c_double) :: deltar, rcrdf
    integer(c_int32_t) :: periodic(3)
    integer(c_int32_t) :: flags = 0
  end type

  type, bind(C) :: trj_sq_params
    integer(c_int32_t) :: ndim, nsp
    real(c_double) :: qmax, qmin
    real(c_double) :: bsc(10)
    integer(c_int32_t) :: qset
    integer(c_int32_t) :: grid(3)
    integer(c_int32_t) :: precision
    integer(c_int32_t) :: flags
  end type

  type, bind(C) :: trj_sq_info
    real(c_double) :: dq, fk(3), qmin_eff
    integer(c_int32_t) :: nqmax, nqmin, lmax(3)
    integer(c_int64_t) :: nvec_grid, nvec_axis, box_ext(3)
  end type

  ! F(Q,t) over time origins (qw and bsc: c_loc of real(c_double) arrays)
  type, bind(C) :: trj_fqt_params
    integer(c_int32_t) :: ndim, nsp, nqw, nbuffer, jump, ndyn, itlimit, pad = 0
    type(c_ptr) :: qw, bsc
  end type

  ! NetCDF trajectory (classic formats); the handle is an opaque type(c_ptr)
  type, bind(C) :: trj_nc_file_info
    integer(c_int64_t) :: nframes, natoms
    integer(c_int32_t) :: version, coord_dtype
    integer(c_int32_t) :: has_type, has_cell_origin, has_cell_lengths, has_time
    integer(c_int32_t) :: has_charge
  end type

  interface
    integer(c_int) function trj_abi_version() bind(C, name="trj_abi_version")
      import :: c_int
    end function
    integer(c_int) function trj_create(ctx, device) bind(C, name="trj_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function
    subroutine trj_destroy(ctx) bind(C, name="trj_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function trj_last_error(ctx) bind(C, name="trj_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function trj_rdf_init(ctx, p, box, lsmax, nit) bind(C, name="trj_rdf_init")
      import :: c_int, c_ptr, c_double, c_int32_t, trj_rdf_params
      type(c_ptr), value :: ctx
      type(trj_rdf_params), intent(in) :: p
      real(c_double), intent(in) :: box(3)
      integer(c_int32_t), intent(out) :: lsmax, nit
    end function
    integer(c_int) function trj_sq_init(ctx, p, box, info) bind(C, name="trj_sq_init")
      import :: c_int, c_ptr, c_double, trj_sq_params, trj_sq_info
      type(c_ptr), value :: ctx
      type(trj_sq_params), intent(in) :: p
      real(c_double), intent(in) :: box(3)
      type(trj_sq_info), intent(out) :: info
    end function
    integer(c_int) function trj_frame(ctx, pos, dtype, layout, itype, n, box, what) bind(C, name="trj_frame")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      type(c_ptr), value :: pos
      integer(c_int32_t), value :: dtype, layout
      integer(c_int32_t), intent(in) :: itype(*)
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: box(3)
      integer(c_int32_t), value :: what
    end function
    ! trj_frame with per-atom charges charge(n) (TRJ_RDF_CHARGE / TRJ_SQ_CHARGE)
    integer(c_int) function trj_frame_q(ctx, pos, dtype, layout, itype, charge, n, box, what) &
        bind(C, name="trj_frame_q")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      type(c_ptr), value :: pos
      integer(c_int32_t), value :: dtype, layout
      integer(c_int32_t), intent(in) :: itype(*)
      real(c_double), intent(in) :: charge(*)
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: box(3)
      integer(c_int32_t), value :: what
    end function
    integer(c_int) function trj_rdf_qq(ctx, acc) bind(C, name="trj_rdf_qq")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: acc(*)
    end function
    integer(c_int) function trj_rdf_gqq(ctx, nconf, out) bind(C, name="trj_rdf_gqq")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nconf
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function trj_sq_qq(ctx, sqfq) bind(C, name="trj_sq_qq")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sqfq(*)
    end function
    integer(c_int) function trj_sq_qq_column(ctx, nconf, out) bind(C, name="trj_sq_qq_column")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nconf
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function trj_fqt_init(ctx, p, box, nvec, nw) bind(C, name="trj_fqt_init")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t, trj_fqt_params
      type(c_ptr), value :: ctx
      type(trj_fqt_params), intent(in) :: p
      real(c_double), intent(in) :: box(3)
      integer(c_int32_t), intent(out) :: nvec
      integer(c_int64_t), intent(out) :: nw(*)
    end function
    integer(c_int) function trj_fqt_results(ctx, fkt, fskt, count_org) bind(C, name="trj_fqt_results")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: fkt(*), fskt(*)
      integer(c_int64_t), intent(out) :: count_org
    end function
    ! number fluctuations (nrandom of /INPUT_RDF/), after trj_rdf_init
    integer(c_int) function trj_rdf_sampling(ctx, nrandom, seed) bind(C, name="trj_rdf_sampling")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nrandom
      integer(c_int64_t), value :: seed
    end function
    integer(c_int) function trj_rdf_sampling_points(ctx, pts) bind(C, name="trj_rdf_sampling_points")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: pts(*)
    end function
    integer(c_int) function trj_rdf_s2n_sums(ctx, gi, g2i) bind(C, name="trj_rdf_s2n_sums")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: gi(*), g2i(*)
    end function
    integer(c_int) function trj_rdf_s2n(ctx, nconf, out) bind(C, name="trj_rdf_s2n")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nconf
      real(c_double), intent(out) :: out(*)
    end function
    ! z-slices (twoDstruc_3D), before trj_rdf_init / trj_sq_init; zslice(nslice) FP32
    integer(c_int) function trj_set_slices(ctx, nslice, zslice, zgrid) bind(C, name="trj_set_slices")
      import :: c_int, c_ptr, c_int32_t, c_float
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nslice
      real(c_float), intent(in) :: zslice(*)
      real(c_float), value :: zgrid
    end function
    integer(c_int) function trj_slice_counts(ctx, count, countp) bind(C, name="trj_slice_counts")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: count(*), countp(*)
    end function
    integer(c_int) function trj_rdf_xy(ctx, hxy, counts) bind(C, name="trj_rdf_xy")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: hxy(*)
      integer(c_int64_t), intent(out) :: counts(*)
    end function
    integer(c_int) function trj_rdf_xy_gmix(ctx, nconf, out) bind(C, name="trj_rdf_xy_gmix")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nconf
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function trj_rdf_xy_qq(ctx, acc) bind(C, name="trj_rdf_xy_qq")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: acc(*)
    end function
    integer(c_int) function trj_rdf_xy_gcc(ctx, nconf, out) bind(C, name="trj_rdf_xy_gcc")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nconf
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function trj_sq_xy(ctx, sqf, sqfp, sqfq) bind(C, name="trj_sq_xy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sqf(*), sqfp(*), sqfq(*)
    end function
    integer(c_int) function trj_sq_xy_columns(ctx, nconf, sqxy, sqpxy, sqqxy) bind(C, name="trj_sq_xy_columns")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nconf
      real(c_double), intent(out) :: sqxy(*), sqpxy(*), sqqxy(*)
    end function
    integer(c_int) function trj_sync(ctx) bind(C, name="trj_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function trj_rdf_counts(ctx, counts) bind(C, name="trj_rdf_counts")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: counts(*)
    end function
    integer(c_int) function trj_rdf_gmix(ctx, nconf, out) bind(C, name="trj_rdf_gmix")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nconf
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function trj_rdf_first_minimum(ctx, rmin) bind(C, name="trj_rdf_first_minimum")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rmin
    end function
    integer(c_int) function trj_sq_shells(ctx, sqf, sqfp, nq) bind(C, name="trj_sq_shells")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sqf(*), sqfp(*)
      integer(c_int64_t), intent(out) :: nq(*)
    end function
    integer(c_int) function trj_sq_columns(ctx, nconf, cols, ncols, mix) bind(C, name="trj_sq_columns")
      import :: c_int, c_ptr, c_int64_t, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nconf
      real(c_double), intent(out) :: cols(*)
      integer(c_int32_t), intent(out) :: ncols
      real(c_double), intent(out) :: mix(*)
    end function
    integer(c_int) function trj_sq_plan(p, box, info, nq) bind(C, name="trj_sq_plan")
      import :: c_int, c_double, c_int64_t, trj_sq_params, trj_sq_info
      type(trj_sq_params), intent(in) :: p
      real(c_double), intent(in) :: box(3)
      type(trj_sq_info), intent(out) :: info
      integer(c_int64_t), intent(out) :: nq(*)
    end function
    integer(c_int) function trj_rdf_plan(p, box, lsmax, rc_eff) bind(C, name="trj_rdf_plan")
      import :: c_int, c_double, c_int32_t, trj_rdf_params
      type(trj_rdf_params), intent(in) :: p
      real(c_double), intent(in) :: box(3)
      integer(c_int32_t), intent(out) :: lsmax
      real(c_double), intent(out) :: rc_eff
    end function
    integer(c_int) function trj_reset(ctx) bind(C, name="trj_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function trj_set_pipeline(ctx, mode) bind(C, name="trj_set_pipeline")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    integer(c_int64_t) function trj_state_bytes(ctx) bind(C, name="trj_state_bytes")
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function trj_state_save(ctx, buf, bytes) bind(C, name="trj_state_save")
      import :: c_int, c_ptr, c_int64_t, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int8_t), intent(out) :: buf(*)
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function trj_state_load(ctx, buf, bytes, mode) bind(C, name="trj_state_load")
      import :: c_int, c_ptr, c_int64_t, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int8_t), intent(in) :: buf(*)
      integer(c_int64_t), value :: bytes
      integer(c_int), value :: mode
    end function
    ! path: a C string (trim(name)//c_null_char)
    integer(c_int) function trj_nc_open(h, path, errbuf, errlen) bind(C, name="trj_nc_open")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), intent(out) :: h
      character(kind=c_char), intent(in) :: path(*)
      character(kind=c_char), intent(out) :: errbuf(*)
      integer(c_int32_t), value :: errlen
    end function
    subroutine trj_nc_close(h) bind(C, name="trj_nc_close")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine
    type(c_ptr) function trj_nc_last_error(h) bind(C, name="trj_nc_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function trj_nc_info(h, info) bind(C, name="trj_nc_info")
      import :: c_int, c_ptr, trj_nc_file_info
      type(c_ptr), value :: h
      type(trj_nc_file_info), intent(out) :: info
    end function
    integer(c_int) function trj_nc_time(h, frame, t) bind(C, name="trj_nc_time")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), value :: frame
      real(c_double), intent(out) :: t
    end function
    ! frame is 0-based; pos(3*natoms) as stored (AoS); type(natoms)
    integer(c_int) function trj_nc_read(h, frame, pos, type, origin, lengths) bind(C, name="trj_nc_read")
      import :: c_int, c_ptr, c_int64_t, c_double, c_int32_t
      type(c_ptr), value :: h
      integer(c_int64_t), value :: frame
      real(c_double), intent(out) :: pos(*)
      integer(c_int32_t), intent(out) :: type(*)
      real(c_double), intent(out) :: origin(3), lengths(3)
    end function
    ! pos_soa(natoms*3): x(1:n) y(1:n) z(1:n); periodic in: /INPUT/ flags, out: updated
    integer(c_int) function trj_nc_frame_ref(h, frame, nsp, sp_types, pos_soa, itype, n, box, periodic) &
        bind(C, name="trj_nc_frame_ref")
      import :: c_int, c_ptr, c_int64_t, c_double, c_int32_t
      type(c_ptr), value :: h
      integer(c_int64_t), value :: frame
      integer(c_int32_t), value :: nsp
      integer(c_int32_t), intent(in) :: sp_types(*)
      real(c_double), intent(out) :: pos_soa(*)
      integer(c_int32_t), intent(out) :: itype(*)
      integer(c_int64_t), intent(out) :: n
      real(c_double), intent(out) :: box(3)
      integer(c_int32_t), intent(inout) :: periodic(3)
    end function
    ! keep /= 0: non-periodic coordinates handed on as read (needed by the z-slices)
    integer(c_int) function trj_nc_keep_nonperiodic(h, keep) bind(C, name="trj_nc_keep_nonperiodic")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: keep
    end function
    ! trj_nc_frame_ref plus charge(natoms): the q variable of the same atoms
    integer(c_int) function trj_nc_frame_ref_q(h, frame, nsp, sp_types, pos_soa, itype, charge, n, box, &
        periodic) bind(C, name="trj_nc_frame_ref_q")
      import :: c_int, c_ptr, c_int64_t, c_double, c_int32_t
      type(c_ptr), value :: h
      integer(c_int64_t), value :: frame
      integer(c_int32_t), value :: nsp
      integer(c_int32_t), intent(in) :: sp_types(*)
      real(c_double), intent(out) :: pos_soa(*)
      integer(c_int32_t), intent(out) :: itype(*)
      real(c_double), intent(out) :: charge(*)
      integer(c_int64_t), intent(out) :: n
      real(c_double), intent(out) :: box(3)
      integer(c_int32_t), intent(inout) :: periodic(3)
    end function
  end interface

contains

  ! Print the context's last error and stop, as the reference does on a failed
  ! kernel (printcudaerror + stop, rdf.cuf:241-242).
  subroutine trj_check(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (rc == TRJ_OK) return
    write (*, '(" *** trj_hip error ",i0," in ",a,": ")', advance='no') rc, what
    if (c_associated(ctx)) then
      call c_f_pointer(trj_last_error(ctx), msg, [1024])
      do k = 1, 1024
        if (msg(k) == c_null_char) exit
        write (*, '(a)', advance='no') msg(k)
      end do
    end if
    write (*, *)
    stop 1
  end subroutine
end module trj_hip
