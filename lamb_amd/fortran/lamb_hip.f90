! lamb_hip -- Fortran 2003 binding of the C-ABI in include/lamb_hip.h.
!
! The raw bind(C) interfaces mirror the header one to one. The wrappers at the
! end keep the reference's calling convention for the multiply,
!   lamb_multiply(matrix_a, matrix_b, matrix_c, error)
! (src/lamb_multiplication.F90:116-119), and for block insertion,
!   internal_distribute(matrix, n, subset_idx, local_idx, datastore_i,
!                       blkdata_extents, data_areas, error)
! (src/lamb_matrix_operations.F90:88-97), over device handles instead of the
! host sector types. Integer kinds follow src/lamb_kinds.F90:155-179.
module lamb_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: lamb_hip_ok = 0
  integer(c_int32_t), parameter, public :: lamb_hip_fp64 = 0
  integer(c_int32_t), parameter, public :: lamb_hip_bf16_split = 1

  ! lamb_hip_stats, field for field
  type, bind(C), public :: lamb_hip_stats
    integer(c_int64_t) :: n_triples_kept = 0
    integer(c_int64_t) :: n_triples_filtered = 0
    integer(c_int64_t) :: n_c_blocks_struct = 0
    integer(c_int64_t) :: n_c_blocks_kept = 0
    integer(c_int64_t) :: n_c_elems_kept = 0
    integer(c_int64_t) :: n_tiles = 0
    integer(c_int64_t) :: n_tile_entries = 0
    real(c_double) :: flops = 0
    real(c_double) :: ms_symbolic = 0
    real(c_double) :: ms_numeric = 0
    real(c_double) :: ms_index = 0
    real(c_double) :: ms_total = 0
  end type

  ! error state of one call: level and type as in src/lamb_error.F90:116-159
  type, public :: lamb_hip_error
    integer(c_int32_t) :: level = 0
    integer(c_int32_t) :: type = 0
    character(len=256) :: message = ''
  end type

  type, public :: lamb_hip_matrix
    type(c_ptr) :: ctx = c_null_ptr
    type(c_ptr) :: mat = c_null_ptr
  end type

  ! lamb_hip_dbcsr_info, field for field
  type, bind(C), public :: lamb_hip_dbcsr_info
    character(kind=c_char) :: version(11)
    character(kind=c_char) :: name(81)
    character(kind=c_char) :: matrix_type
    integer(c_int32_t) :: n_chunks
    integer(c_int32_t) :: data_type
    integer(c_int32_t) :: n_block_rows, n_block_cols
    integer(c_int64_t) :: n_blocks
    integer(c_int64_t) :: n_values
  end type

  public :: lamb_hip_init, lamb_hip_finalize, lamb_hip_last_error
  public :: lamb_hip_dist_create, lamb_hip_dist_destroy
  public :: lamb_hip_matrix_create, lamb_hip_matrix_destroy, lamb_hip_matrix_copy
  public :: lamb_hip_sector_upload, lamb_hip_distribute, lamb_hip_matrix_commit
  public :: lamb_hip_sector_query, lamb_hip_sector_download, lamb_hip_matrix_info
  public :: lamb_hip_goods_size, lamb_hip_goods_pack, lamb_hip_goods_unpack
  public :: lamb_hip_multiply, lamb_hip_trace, lamb_hip_dbcsr_load, lamb_hip_partition_atoms
  public :: lamb_multiply_hip, internal_distribute_hip, lamb_load_dbcsr_hip
  public :: lamb_hip_multiply_test, lamb_multiply_test_hip, lamb_hip_dbcsr_read_info

  interface
    function lamb_hip_init(device, ctx) bind(C, name="lamb_hip_init") result(st)
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: device
      type(c_ptr), intent(out) :: ctx
      integer(c_int32_t) :: st
    end function

    function lamb_hip_finalize(ctx) bind(C, name="lamb_hip_finalize") result(st)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: st
    end function

    function lamb_hip_last_error(ctx, level, type, msg, msg_len) &
        bind(C, name="lamb_hip_last_error") result(st)
      import :: c_int32_t, c_ptr, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: level, type
      character(kind=c_char), intent(out) :: msg(*)
      integer(c_int32_t), value :: msg_len
      integer(c_int32_t) :: st
    end function

    ! coords and box may be c_null_ptr
    function lamb_hip_dist_create(ctx, n_atoms, blk_size, n_subsets, subset_map, &
        order_map, coords, box, dist) bind(C, name="lamb_hip_dist_create") result(st)
      import :: c_int32_t, c_int64_t, c_int8_t, c_int16_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n_atoms
      integer(c_int8_t), intent(in) :: blk_size(*)
      integer(c_int32_t), value :: n_subsets
      integer(c_int32_t), intent(in) :: subset_map(*)
      integer(c_int16_t), intent(in) :: order_map(*)
      type(c_ptr), value :: coords, box
      type(c_ptr), intent(out) :: dist
      integer(c_int32_t) :: st
    end function

    function lamb_hip_dist_destroy(dist) bind(C, name="lamb_hip_dist_destroy") result(st)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: dist
      integer(c_int32_t) :: st
    end function

    function lamb_hip_matrix_create(dist, mat) bind(C, name="lamb_hip_matrix_create") result(st)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: dist
      type(c_ptr), intent(out) :: mat
      integer(c_int32_t) :: st
    end function

    function lamb_hip_matrix_destroy(mat) bind(C, name="lamb_hip_matrix_destroy") result(st)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: mat
      integer(c_int32_t) :: st
    end function

    function lamb_hip_matrix_copy(src, dst) bind(C, name="lamb_hip_matrix_copy") result(st)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: src, dst
      integer(c_int32_t) :: st
    end function

    function lamb_hip_sector_upload(mat, subset_row, subset_col, n_blks, row_p, col_i, &
        extents, data, data_len) bind(C, name="lamb_hip_sector_upload") result(st)
      import :: c_int32_t, c_int64_t, c_int16_t, c_double, c_ptr
      type(c_ptr), value :: mat
      integer(c_int32_t), value :: subset_row, subset_col
      integer(c_int64_t), value :: n_blks
      integer(c_int64_t), intent(in) :: row_p(*)
      integer(c_int16_t), intent(in) :: col_i(*)
      integer(c_int64_t), intent(in) :: extents(*)
      real(c_double), intent(in) :: data(*)
      integer(c_int64_t), value :: data_len
      integer(c_int32_t) :: st
    end function

    function lamb_hip_distribute(mat, n, subset_idx, local_idx, extents, data, data_len) &
        bind(C, name="lamb_hip_distribute") result(st)
      import :: c_int32_t, c_int64_t, c_int16_t, c_double, c_ptr
      type(c_ptr), value :: mat
      integer(c_int64_t), value :: n
      integer(c_int32_t), intent(in) :: subset_idx(*)
      integer(c_int16_t), intent(in) :: local_idx(*)
      integer(c_int64_t), intent(in) :: extents(*)
      real(c_double), intent(in) :: data(*)
      integer(c_int64_t), value :: data_len
      integer(c_int32_t) :: st
    end function

    function lamb_hip_matrix_commit(mat) bind(C, name="lamb_hip_matrix_commit") result(st)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: mat
      integer(c_int32_t) :: st
    end function

    function lamb_hip_sector_query(mat, subset_row, subset_col, n_blks, data_len) &
        bind(C, name="lamb_hip_sector_query") result(st)
      import :: c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: mat
      integer(c_int32_t), value :: subset_row, subset_col
      integer(c_int64_t), intent(out) :: n_blks, data_len
      integer(c_int32_t) :: st
    end function

    function lamb_hip_sector_download(mat, subset_row, subset_col, row_p, col_i, extents, &
        data) bind(C, name="lamb_hip_sector_download") result(st)
      import :: c_int32_t, c_int64_t, c_int16_t, c_double, c_ptr
      type(c_ptr), value :: mat
      integer(c_int32_t), value :: subset_row, subset_col
      integer(c_int64_t), intent(out) :: row_p(*)
      integer(c_int16_t), intent(out) :: col_i(*)
      integer(c_int64_t), intent(out) :: extents(*)
      real(c_double), intent(out) :: data(*)
      integer(c_int32_t) :: st
    end function

    function lamb_hip_multiply_test(ctx, path, eps_filter, stats) &
        bind(C, name="lamb_hip_multiply_test") result(st)
      import :: c_int32_t, c_double, c_char, c_ptr, lamb_hip_stats
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), value :: eps_filter
      type(lamb_hip_stats), intent(out) :: stats
      integer(c_int32_t) :: st
    end function

    ! Sector wire format of lamb_goods_X_create / lamb_goods_X_unpack
    ! (src/lamb_goods_X_types.F90:119-268, src/lamb_goods_X_methods.F90:166-218):
    ! index holds goods_X%index, data goods_X%data.
    function lamb_hip_goods_size(mat, subset_row, subset_col, n_index_words, n_data) &
        bind(C, name="lamb_hip_goods_size") result(st)
      import :: c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: mat
      integer(c_int32_t), value :: subset_row, subset_col
      integer(c_int64_t), intent(out) :: n_index_words, n_data
      integer(c_int32_t) :: st
    end function

    function lamb_hip_goods_pack(mat, subset_row, subset_col, index, n_index_words, data, &
        n_data) bind(C, name="lamb_hip_goods_pack") result(st)
      import :: c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: mat
      integer(c_int32_t), value :: subset_row, subset_col
      integer(c_int64_t), intent(out) :: index(*)
      integer(c_int64_t), value :: n_index_words
      real(c_double), intent(out) :: data(*)
      integer(c_int64_t), value :: n_data
      integer(c_int32_t) :: st
    end function

    function lamb_hip_goods_unpack(mat, subset_row, subset_col, index, n_index_words, data, &
        n_data) bind(C, name="lamb_hip_goods_unpack") result(st)
      import :: c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: mat
      integer(c_int32_t), value :: subset_row, subset_col
      integer(c_int64_t), intent(in) :: index(*)
      integer(c_int64_t), value :: n_index_words
      real(c_double), intent(in) :: data(*)
      integer(c_int64_t), value :: n_data
      integer(c_int32_t) :: st
    end function

    function lamb_hip_matrix_info(mat, n_blks, n_elems) bind(C, name="lamb_hip_matrix_info") &
        result(st)
      import :: c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: mat
      integer(c_int64_t), intent(out) :: n_blks, n_elems
      integer(c_int32_t) :: st
    end function

    function lamb_hip_multiply(ctx, a, b, c, alpha, beta, eps_filter, flags, stats) &
        bind(C, name="lamb_hip_multiply") result(st)
      import :: c_int32_t, c_double, c_ptr, lamb_hip_stats
      type(c_ptr), value :: ctx, a, b, c
      real(c_double), value :: alpha, beta, eps_filter
      integer(c_int32_t), value :: flags
      type(lamb_hip_stats), intent(out) :: stats
      integer(c_int32_t) :: st
    end function

    function lamb_hip_trace(a, trace) bind(C, name="lamb_hip_trace") result(st)
      import :: c_int32_t, c_double, c_ptr
      type(c_ptr), value :: a
      real(c_double), intent(out) :: trace
      integer(c_int32_t) :: st
    end function

    function lamb_hip_partition_atoms(ctx, n_atoms, coords, box, dims, subset_map, order_map) &
        bind(C, name="lamb_hip_partition_atoms") result(st)
      import :: c_int16_t, c_int32_t, c_int64_t, c_float, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n_atoms
      real(c_float), intent(in) :: coords(3, *), box(3)
      integer(c_int32_t), intent(in) :: dims(3)
      integer(c_int32_t), intent(out) :: subset_map(*)
      integer(c_int16_t), intent(out) :: order_map(*)
      integer(c_int32_t) :: st
    end function

    function lamb_hip_dbcsr_read_info(path, info, row_blk_size, col_blk_size, msg, msg_len) &
        bind(C, name="lamb_hip_dbcsr_read_info") result(st)
      import :: c_int32_t, c_char, lamb_hip_dbcsr_info
      character(kind=c_char), intent(in) :: path(*)
      type(lamb_hip_dbcsr_info), intent(out) :: info
      integer(c_int32_t), intent(out) :: row_blk_size(*), col_blk_size(*)
      character(kind=c_char), intent(out) :: msg(*)
      integer(c_int32_t), value :: msg_len
      integer(c_int32_t) :: st
    end function

    function lamb_hip_dbcsr_load(mat, path) bind(C, name="lamb_hip_dbcsr_load") result(st)
      import :: c_int32_t, c_ptr, c_char
      type(c_ptr), value :: mat
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t) :: st
    end function
  end interface

contains

  subroutine fetch_error(ctx, st, error)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: st
    type(lamb_hip_error), intent(inout) :: error
    character(kind=c_char) :: buf(256)
    integer(c_int32_t) :: lvl, typ, ign
    integer :: i
    error%level = st
    if (st == lamb_hip_ok) return
    ign = lamb_hip_last_error(ctx, lvl, typ, buf, 256_c_int32_t)
    error%level = lvl
    error%type = typ
    error%message = ''
    do i = 1, 256
      if (buf(i) == c_null_char) exit
      error%message(i:i) = buf(i)
    end do
  end subroutine

  ! C = A*B (alpha = 1, beta = 0: C is overwritten) over device matrices, in
  ! lamb_multiply's argument order; eps_filter (optional, default 0: no
  ! filtering, as the reference's call has no threshold) sets the norm filter.
  subroutine lamb_multiply_hip(matrix_a, matrix_b, matrix_c, error, eps_filter, stats)
    type(lamb_hip_matrix), intent(in) :: matrix_a, matrix_b
    type(lamb_hip_matrix), intent(inout) :: matrix_c
    type(lamb_hip_error), intent(inout) :: error
    real(c_double), intent(in), optional :: eps_filter
    type(lamb_hip_stats), intent(out), optional :: stats
    type(lamb_hip_stats) :: st_local
    real(c_double) :: eps
    integer(c_int32_t) :: st
    eps = 0.0_c_double
    if (present(eps_filter)) eps = eps_filter
    st = lamb_hip_multiply(matrix_c%ctx, matrix_a%mat, matrix_b%mat, matrix_c%mat, &
                           1.0_c_double, 0.0_c_double, eps, lamb_hip_fp64, st_local)
    call fetch_error(matrix_c%ctx, st, error)
    if (present(stats)) stats = st_local
  end subroutine

  ! Blocks of one data store into the matrix, then commit. subset_idx(2,n),
  ! local_idx(2,n) and blkdata_extents(2,n) as in internal_distribute; one
  ! data area (the reference's datastore_i selects among several).
  subroutine internal_distribute_hip(matrix, n, subset_idx, local_idx, blkdata_extents, &
                                     data_area, error)
    type(lamb_hip_matrix), intent(inout) :: matrix
    integer(c_int64_t), intent(in) :: n
    integer(c_int32_t), intent(in) :: subset_idx(2, *)
    integer(c_int16_t), intent(in) :: local_idx(2, *)
    integer(c_int64_t), intent(in) :: blkdata_extents(2, *)
    real(c_double), intent(in) :: data_area(:)
    type(lamb_hip_error), intent(inout) :: error
    integer(c_int32_t) :: st
    st = lamb_hip_distribute(matrix%mat, n, subset_idx, local_idx, blkdata_extents, &
                             data_area, int(size(data_area), c_int64_t))
    if (st == lamb_hip_ok) st = lamb_hip_matrix_commit(matrix%mat)
    call fetch_error(matrix%ctx, st, error)
  end subroutine

  ! A DBCSR dump into the device matrix, in lamb_load_dbcsr's argument order
  ! (src/lamb_io_operations.F90:267-271) without the MPI environment: the
  ! matrix already exists over a distribution of the file's atoms.
  subroutine lamb_load_dbcsr_hip(matrix, filename, error)
    type(lamb_hip_matrix), intent(inout) :: matrix
    character(len=*), intent(in) :: filename
    type(lamb_hip_error), intent(inout) :: error
    integer(c_int32_t) :: st
    st = lamb_hip_dbcsr_load(matrix%mat, trim(filename)//c_null_char)
    call fetch_error(matrix%ctx, st, error)
  end subroutine

  ! lamb_multiply_test (src/lamb_multiplication.F90:85-114) without the MPI
  ! environment: the dump at filename as A, copied to B and C, C = A*B.
  subroutine lamb_multiply_test_hip(ctx, filename, error, eps_filter, stats)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: filename
    type(lamb_hip_error), intent(inout) :: error
    real(c_double), intent(in), optional :: eps_filter
    type(lamb_hip_stats), intent(out), optional :: stats
    type(lamb_hip_stats) :: st_local
    real(c_double) :: eps
    integer(c_int32_t) :: st
    eps = 0.0_c_double
    if (present(eps_filter)) eps = eps_filter
    st = lamb_hip_multiply_test(ctx, trim(filename)//c_null_char, eps, st_local)
    call fetch_error(ctx, st, error)
    if (present(stats)) stats = st_local
  end subroutine

end module lamb_hip
