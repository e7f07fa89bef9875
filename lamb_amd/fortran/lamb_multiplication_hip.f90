! lamb_multiplication_hip -- the reference's multiply call over host sector
! types laid out like the reference's own, backed by the device library.
!
! A host program of the reference calls
!   call lamb_multiply(matrix_a, matrix_b, matrix_c, error)
! (src/lamb_multiplication.F90:116-119). This module keeps that call: its
! matrices hold sectors whose index and data stores have the fields, kinds
! and conventions of lamb_index_type / lamb_blk_info_type
! (src/lamb_index_types.F90:86-105), lamb_data_X_type
! (src/lamb_data_X_types.F90:77-83) and lamb_sector_X_type
! (src/lamb_sector_X_types.F90:79-87), kinds from src/lamb_kinds.F90:155-179:
! row_p int64 0-based exclusive offsets (n_rows + 1), col_i int16 1-based
! subset-local columns ascending per row, extents int64 1-based inclusive
! into data store 1, blocks column-major. Two fields stand for objects this
! path does not carry: a sector's subsets are the two subset numbers (the
! reference points at subset objects), and a matrix's distribution is the
! device distribution handle (lamb_hip_dist) with its context.
!
! lamb_multiply uploads A's and B's sectors, runs C = A*B on the device with
! the norm filter set by lamb_multiply_set_eps_filter (0 by default: the
! reference's call has no threshold), and replaces the index and data of
! every sector listed in matrix_c (their subsets and index%n_rows set by the
! caller, as the reference's sector set-up does) with C's.
module lamb_multiplication_hip
  use iso_c_binding
  use lamb_hip
  implicit none
  private

  integer, parameter, public :: dataptr_k = c_int64_t, datastore_k = c_int8_t
  integer, parameter, public :: blkn_k = c_int64_t, locatm_k = c_int16_t
  integer, parameter, public :: subset_k = c_int32_t

  type, public :: lamb_blk_info_type
    integer(kind=dataptr_k), dimension(2) :: extent = 0
    integer(kind=datastore_k) :: store = 1
    integer(kind=c_int8_t) :: bits = 0
  end type

  type, public :: lamb_index_type
    integer :: refcount = 1
    integer :: id = 0
    logical :: csr = .true.
    integer(kind=blkn_k) :: n_idx_blks = 0
    integer(kind=locatm_k) :: n_rows = 0
    integer(kind=blkn_k), dimension(:), allocatable :: row_p
    integer(kind=locatm_k), dimension(:), allocatable :: row_i
    integer(kind=locatm_k), dimension(:), allocatable :: col_i
    type(lamb_blk_info_type), dimension(:), allocatable :: blk_info
  end type

  type, public :: lamb_data_d_type
    double precision, dimension(:), pointer :: d => null()
    integer(kind=dataptr_k) :: data_len = 0
    logical :: mine = .true.
    integer :: refcount = 1
    integer :: id = 0
  end type

  type, public :: lamb_sector_d_type
    integer :: refcount = 1
    integer :: id = 0
    type(lamb_index_type) :: index
    type(lamb_data_d_type), dimension(:), allocatable :: data_stores
    integer(kind=subset_k), dimension(2) :: subsets = 0
    integer(kind=blkn_k) :: n_blks = 0
  end type

  type, public :: lamb_matrix_d_type
    integer :: refcount = 1
    integer :: id = 0
    type(lamb_sector_d_type), dimension(:), allocatable :: sectors
    type(c_ptr) :: ctx = c_null_ptr
    type(c_ptr) :: distribution = c_null_ptr
  end type

  ! level and type as in src/lamb_error.F90:116-159
  type, public :: lamb_error_type
    integer :: level = 0
    integer :: type = 0
    character(len=256) :: message = ''
  end type

  integer, parameter :: failure_level = 3, wrong_args = 100
  real(c_double), save :: eps_filter_setting = 0.0_c_double

  public :: lamb_multiply, lamb_multiply_set_eps_filter

contains

  subroutine lamb_multiply_set_eps_filter(eps_filter)
    real(c_double), intent(in) :: eps_filter
    eps_filter_setting = eps_filter
  end subroutine

  subroutine set_error(error, ctx, st, what)
    type(lamb_error_type), intent(inout) :: error
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: st
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(256)
    integer(c_int32_t) :: lvl, typ, ign
    integer :: i
    ign = lamb_hip_last_error(ctx, lvl, typ, buf, 256_c_int32_t)
    error%level = max(int(st), int(lvl))
    error%type = typ
    error%message = what // ': '
    do i = 1, 256
      if (buf(i) == c_null_char) exit
      if (len_trim(what) + 2 + i > len(error%message)) exit
      error%message(len_trim(what) + 2 + i:len_trim(what) + 2 + i) = buf(i)
    end do
  end subroutine

  ! every sector of m into the device matrix dm, then commit
  function upload(m, dm) result(st)
    type(lamb_matrix_d_type), intent(in) :: m
    type(c_ptr), intent(in) :: dm
    integer(c_int32_t) :: st
    integer :: i
    integer(c_int64_t) :: nb, b
    integer(c_int64_t), allocatable :: ext(:, :)
    real(c_double), target :: empty(1)
    st = lamb_hip_ok
    empty = 0
    do i = 1, size(m%sectors)
      associate (s => m%sectors(i))
        nb = s%n_blks
        allocate (ext(2, max(nb, 1_c_int64_t)))
        do b = 1, nb
          ext(:, b) = s%index%blk_info(b)%extent
        end do
        if (s%data_stores(1)%data_len > 0) then
          st = lamb_hip_sector_upload(dm, s%subsets(1), s%subsets(2), nb, s%index%row_p, &
                                      s%index%col_i, ext, s%data_stores(1)%d, &
                                      s%data_stores(1)%data_len)
        else
          st = lamb_hip_sector_upload(dm, s%subsets(1), s%subsets(2), nb, s%index%row_p, &
                                      s%index%col_i, ext, empty, 0_c_int64_t)
        end if
        deallocate (ext)
      end associate
      if (st /= lamb_hip_ok) return
    end do
    st = lamb_hip_matrix_commit(dm)
  end function

  ! C's sectors (as listed in m) from the device matrix dm
  function download(dm, m) result(st)
    type(c_ptr), intent(in) :: dm
    type(lamb_matrix_d_type), intent(inout) :: m
    integer(c_int32_t) :: st
    integer :: i
    integer(c_int64_t) :: nb, nd, b
    integer(c_int64_t), allocatable :: ext(:, :)
    st = lamb_hip_ok
    do i = 1, size(m%sectors)
      associate (s => m%sectors(i))
        st = lamb_hip_sector_query(dm, s%subsets(1), s%subsets(2), nb, nd)
        if (st /= lamb_hip_ok) return
        if (allocated(s%index%row_p)) deallocate (s%index%row_p)
        if (allocated(s%index%col_i)) deallocate (s%index%col_i)
        if (allocated(s%index%blk_info)) deallocate (s%index%blk_info)
        if (allocated(s%index%row_i)) deallocate (s%index%row_i)
        allocate (s%index%row_p(int(s%index%n_rows) + 1), s%index%col_i(max(nb, 1_c_int64_t)))
        allocate (s%index%blk_info(nb), ext(2, max(nb, 1_c_int64_t)))
        if (.not. allocated(s%data_stores)) allocate (s%data_stores(1))
        if (associated(s%data_stores(1)%d) .and. s%data_stores(1)%mine) &
          deallocate (s%data_stores(1)%d)
        allocate (s%data_stores(1)%d(max(nd, 1_c_int64_t)))
        st = lamb_hip_sector_download(dm, s%subsets(1), s%subsets(2), s%index%row_p, &
                                      s%index%col_i, ext, s%data_stores(1)%d)
        if (st /= lamb_hip_ok) return
        do b = 1, nb
          s%index%blk_info(b)%extent = ext(:, b)
          s%index%blk_info(b)%store = 1
          s%index%blk_info(b)%bits = 0
        end do
        deallocate (ext)
        s%index%csr = .true.
        s%index%n_idx_blks = nb
        s%n_blks = nb
        s%data_stores(1)%data_len = nd
        s%data_stores(1)%mine = .true.
      end associate
    end do
  end function

  ! C = A*B (src/lamb_multiplication.F90:116-204): A, B and C share one
  ! distribution; C's listed sectors are overwritten.
  subroutine lamb_multiply(matrix_a, matrix_b, matrix_c, error)
    type(lamb_matrix_d_type), intent(in) :: matrix_a, matrix_b
    type(lamb_matrix_d_type), intent(inout) :: matrix_c
    type(lamb_error_type), intent(inout) :: error
    type(c_ptr) :: da, db, dc, ctx
    type(lamb_hip_stats) :: stats
    integer(c_int32_t) :: st, ign
    ctx = matrix_c%ctx
    if (.not. c_associated(matrix_a%distribution, matrix_b%distribution) .or. &
        .not. c_associated(matrix_a%distribution, matrix_c%distribution)) then
      error%level = failure_level
      error%type = wrong_args
      error%message = 'lamb_multiply: A, B and C must share one distribution'
      return
    end if
    da = c_null_ptr
    db = c_null_ptr
    dc = c_null_ptr
    st = lamb_hip_matrix_create(matrix_a%distribution, da)
    if (st == lamb_hip_ok) st = upload(matrix_a, da)
    if (st == lamb_hip_ok) st = lamb_hip_matrix_create(matrix_b%distribution, db)
    if (st == lamb_hip_ok) st = upload(matrix_b, db)
    if (st == lamb_hip_ok) st = lamb_hip_matrix_create(matrix_c%distribution, dc)
    if (st == lamb_hip_ok) &
      st = lamb_hip_multiply(ctx, da, db, dc, 1.0_c_double, 0.0_c_double, eps_filter_setting, &
                             lamb_hip_fp64, stats)
    if (st == lamb_hip_ok) st = download(dc, matrix_c)
    if (st /= lamb_hip_ok) then
      call set_error(error, ctx, st, 'lamb_multiply')
    else
      error%level = 0
      error%type = 0
      error%message = ''
    end if
    if (c_associated(dc)) ign = lamb_hip_matrix_destroy(dc)
    if (c_associated(db)) ign = lamb_hip_matrix_destroy(db)
    if (c_associated(da)) ign = lamb_hip_matrix_destroy(da)
  end subroutine

end module lamb_multiplication_hip
