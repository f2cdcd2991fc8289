! abi_check.f90 -- host-only check of the Fortran binding (no GPU): sizes of
! the bind(C) types, the ABI version, and the q-set plan of the reference's
! 3D/LJmix example (qmin = 6.4, qmax = 25, dq printed as 0.3740891).  With a
! NetCDF file as argument it also reads that file's frame 0 (types 1 and 2).
program abi_check
  use iso_c_binding
  use trj_hip
  implicit none
  type(trj_rdf_params) :: rp
  type(trj_sq_params) :: sp
  type(trj_sq_info) :: info
  integer(c_int64_t) :: nq(0:200)
  real(c_double) :: box(3), L, rc
  integer(c_int32_t) :: lsmax
  type(trj_nc_file_info) :: fi
  type(c_ptr) :: nc
  character(len=512) :: arg
  character(kind=c_char) :: err(256)
  real(c_double), allocatable :: pos(:)
  integer(c_int32_t), allocatable :: ity(:)
  integer(c_int32_t) :: sel(2), per(3)
  integer(c_int64_t) :: n
  real(c_double) :: t, ncbox(3)
  print '(a,4i6)', "sizes", c_sizeof(rp), c_sizeof(sp), c_sizeof(info), c_sizeof(fi)
  print '(a,i6)', "abi", trj_abi_version()
  L = 2.0d0 * acos(-1.0d0) / 0.3740891d0
  box = L
  sp%ndim = 3
  sp%nsp = 2
  sp%qmax = 25.0d0
  sp%qmin = 6.4d0
  sp%bsc = 1.0d0
  sp%qset = TRJ_QSET_REF
  sp%grid = 0
  sp%precision = TRJ_PREC_FP64
  sp%flags = 0
  if (trj_sq_plan(sp, box, info, nq) /= TRJ_OK) stop 2
  print '(a,3i6)', "plan", info%nqmax, info%nqmin, info%lmax(1)
  print '(a,16i6)', "nq", nq(1:16)
  rp%ndim = 3
  rp%nsp = 2
  rp%deltar = 0.25d0
  rp%rcrdf = 15.0d0
  rp%periodic = 1
  if (trj_rdf_plan(rp, box, lsmax, rc) /= TRJ_OK) stop 3
  print '(a,i6,f12.6)', "rdf", lsmax, rc
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    if (trj_nc_open(nc, trim(arg)//c_null_char, err, 256) /= TRJ_OK) stop 4
    if (trj_nc_info(nc, fi) /= TRJ_OK) stop 5
    allocate(pos(3 * fi%natoms), ity(fi%natoms))
    sel = [1, 2]
    per = 1
    if (trj_nc_frame_ref(nc, 0_c_int64_t, 2, sel, pos, ity, n, ncbox, per) /= TRJ_OK) stop 6
    if (trj_nc_time(nc, 1_c_int64_t, t) /= TRJ_OK) stop 7
    print '(a,i4,2i8,i4,3i3,f8.2)', "nc", fi%version, fi%nframes, n, ity(1), per, t
    print '(a,3f12.5)', "ncpos", pos(1), pos(n + 1), pos(2 * n + 1)
    call trj_nc_close(nc)
  end if
end program abi_check
