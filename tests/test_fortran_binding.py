"""The Fortran 2003 bindings, compiled with AMD flang and run.

- lamb_hip.f90 (bind(C) interfaces) and lamb_multiplication_hip.f90 (the
  reference's lamb_multiply(matrix_a, matrix_b, matrix_c, error) call over
  sector / index / data types laid out like the reference's) compile, and a
  program referencing every entry point links against liblamb_hip.so.
- CPU: a Fortran program reads the header of the committed DBCSR fixture
  through lamb_hip_dbcsr_read_info; every field equals the fixture's record.
- GPU: a Fortran program fills the reference-layout types from seeded sectors,
  calls lamb_multiply, and writes C's sectors back; they equal the oracle's
  product (pattern identical, 1e-12). The same program runs
  lamb_multiply_test_hip on the fixture; its statistics equal the oracle's
  A*A of the fixture.
"""
import json
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
LIB = ROOT / "lamb_amd" / "liblamb_hip.so"
FIXTURE = ROOT / "tests" / "golden" / "two_chunks.dbcsr"
MODULES = [ROOT / "lamb_amd" / "fortran" / "lamb_hip.f90",
           ROOT / "lamb_amd" / "fortran" / "lamb_multiplication_hip.f90"]

needs_flang = pytest.mark.skipif(not pathlib.Path(FLANG).exists(), reason="flang not installed")
needs_lib = pytest.mark.skipif(not LIB.exists(), reason="liblamb_hip.so not built")


def build(tmp_path, name, source):
    objs = []
    for k, mod in enumerate(MODULES):
        o = tmp_path / f"m{k}.o"
        subprocess.run([FLANG, "-c", str(mod), "-J", str(tmp_path), "-I", str(tmp_path),
                        "-o", str(o)], check=True, capture_output=True, text=True)
        objs.append(str(o))
    prog = tmp_path / f"{name}.f90"
    prog.write_text(source)
    subprocess.run([FLANG, "-c", str(prog), "-I", str(tmp_path), "-o", str(tmp_path / "p.o")],
                   check=True, capture_output=True, text=True)
    exe = tmp_path / name
    r = subprocess.run([FLANG, str(tmp_path / "p.o"), *objs, "-o", str(exe),
                        f"-L{LIB.parent}", "-llamb_hip", f"-Wl,-rpath,{LIB.parent}",
                        "-Wl,--no-undefined"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


LINK_PROGRAM = """
program link_check
  use iso_c_binding
  use lamb_hip
  use lamb_multiplication_hip
  implicit none
  type(c_ptr) :: ctx, dist, a
  type(lamb_hip_stats) :: stats
  type(lamb_hip_matrix) :: ma, mb, mc
  type(lamb_hip_error) :: err
  type(lamb_matrix_d_type) :: xa, xb, xc
  type(lamb_error_type) :: error
  integer(c_int32_t) :: st, lvl, typ
  integer(c_int64_t) :: nb, ne
  integer(c_int8_t) :: blk(1)
  integer(c_int32_t) :: smap(1), sidx(2, 1)
  integer(c_int16_t) :: omap(1), col(1), lidx(2, 1)
  integer(c_int64_t) :: row_p(2), ext(2), ext2(2, 1)
  real(c_double) :: d(1), tr
  character(kind=c_char) :: msg(8)
  if (command_argument_count() > 100) then
    st = lamb_hip_init(0_c_int32_t, ctx)
    st = lamb_hip_last_error(ctx, lvl, typ, msg, 8_c_int32_t)
    st = lamb_hip_dist_create(ctx, 1_c_int64_t, blk, 1_c_int32_t, smap, omap, &
                              c_null_ptr, c_null_ptr, dist)
    st = lamb_hip_matrix_create(dist, a)
    st = lamb_hip_sector_upload(a, 1_c_int32_t, 1_c_int32_t, 1_c_int64_t, row_p, col, ext, d, 1_c_int64_t)
    st = lamb_hip_distribute(a, 1_c_int64_t, smap, omap, ext, d, 1_c_int64_t)
    st = lamb_hip_matrix_commit(a)
    st = lamb_hip_sector_query(a, 1_c_int32_t, 1_c_int32_t, nb, ne)
    st = lamb_hip_sector_download(a, 1_c_int32_t, 1_c_int32_t, row_p, col, ext, d)
    st = lamb_hip_matrix_info(a, nb, ne)
    st = lamb_hip_goods_size(a, 1_c_int32_t, 1_c_int32_t, nb, ne)
    st = lamb_hip_goods_pack(a, 1_c_int32_t, 1_c_int32_t, row_p, 2_c_int64_t, d, 1_c_int64_t)
    st = lamb_hip_goods_unpack(a, 1_c_int32_t, 1_c_int32_t, row_p, 2_c_int64_t, d, 1_c_int64_t)
    st = lamb_hip_matrix_copy(a, a)
    st = lamb_hip_multiply(ctx, a, a, a, 1.0_c_double, 0.0_c_double, 0.0_c_double, 0_c_int32_t, stats)
    st = lamb_hip_trace(a, tr)
    call lamb_multiply_hip(ma, mb, mc, err)
    call lamb_multiply_hip(ma, mb, mc, err, 1.0e-8_c_double)
    call lamb_multiply_set_eps_filter(1.0e-8_c_double)
    call lamb_multiply(xa, xb, xc, error)
    call lamb_multiply_test_hip(ctx, 'small.dbcsr', err)
    st = lamb_hip_multiply_test(ctx, 'small.dbcsr'//c_null_char, 0.0_c_double, stats)
    call internal_distribute_hip(ma, 1_c_int64_t, sidx, lidx, ext2, d, err)
    st = lamb_hip_matrix_destroy(a)
    st = lamb_hip_dist_destroy(dist)
    st = lamb_hip_finalize(ctx)
  end if
  print '(a)', 'linked'
end program link_check
"""


@needs_flang
@needs_lib
def test_fortran_bindings_compile_and_link(tmp_path):
    exe = build(tmp_path, "link_check", LINK_PROGRAM)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "linked"


HEADER_PROGRAM = """
program read_header
  use iso_c_binding
  use lamb_hip
  implicit none
  type(lamb_hip_dbcsr_info) :: info
  integer(c_int32_t) :: rows(64), cols(64), st, i
  character(kind=c_char) :: msg(256)
  character(len=4096) :: path
  call get_command_argument(1, path)
  st = lamb_hip_dbcsr_read_info(trim(path)//c_null_char, info, rows, cols, msg, 256_c_int32_t)
  print '(a,i0)', 'status ', st
  print '(a,81a)', 'name ', (info%name(i), i = 1, 81)
  print '(a,a)', 'matrix_type ', info%matrix_type
  print '(a,i0)', 'n_chunks ', info%n_chunks
  print '(a,i0)', 'data_type ', info%data_type
  print '(a,i0)', 'n_block_rows ', info%n_block_rows
  print '(a,i0)', 'n_block_cols ', info%n_block_cols
  print '(a,i0)', 'n_blocks ', info%n_blocks
  print '(a,i0)', 'n_values ', info%n_values
  print '(a,64(1x,i0))', 'row_blk_size', (rows(i), i = 1, info%n_block_rows)
end program read_header
"""


@needs_flang
@needs_lib
def test_fortran_reads_dbcsr_header(tmp_path):
    """Runs on the CPU: lamb_hip_dbcsr_read_info is host file reading."""
    exe = build(tmp_path, "read_header", HEADER_PROGRAM)
    out = subprocess.run([str(exe), str(FIXTURE)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    got = {}
    for line in out.stdout.splitlines():
        key, _, val = line.partition(" ")
        got[key] = val
    fx = json.loads((ROOT / "tests" / "golden" / "two_chunks.json").read_text())
    assert got["status"] == "0"
    assert got["name"].split("\x00")[0].strip() == fx["name"]
    assert got["matrix_type"].strip() == fx["matrix_type"]
    assert int(got["n_chunks"]) == fx["n_chunks"]
    assert int(got["data_type"]) == fx["data_type"]
    n = len(fx["sizes"])
    assert int(got["n_block_rows"]) == n and int(got["n_block_cols"]) == n
    assert int(got["n_blocks"]) == len(fx["blocks"])
    assert int(got["n_values"]) == sum(len(v) for _, _, v in fx["blocks"])
    assert [int(x) for x in got["row_blk_size"].split()] == fx["sizes"]


MULTIPLY_PROGRAM = """
program drop_in
  use iso_c_binding
  use lamb_hip
  use lamb_multiplication_hip
  implicit none
  type(c_ptr) :: ctx, dist
  type(lamb_matrix_d_type) :: a, b, c
  type(lamb_error_type) :: error
  type(lamb_hip_error) :: err2
  type(lamb_hip_stats) :: stats
  integer(c_int32_t) :: n_atoms, n_subsets, st, i
  real(c_double) :: eps
  integer(c_int8_t), allocatable :: blk(:)
  integer(c_int32_t), allocatable :: smap(:)
  integer(c_int16_t), allocatable :: omap(:)
  real(c_float), allocatable, target :: coords(:)
  real(c_float), target :: box(3)
  integer :: u
  character(len=4096) :: fin, fout, fdbcsr
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, fdbcsr)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) n_atoms, n_subsets, eps
  allocate (blk(n_atoms), smap(n_atoms), omap(n_atoms), coords(3 * n_atoms))
  read (u) blk, smap, omap, coords, box
  call read_matrix(u, a)
  call read_matrix(u, b)
  close (u)
  st = lamb_hip_init(0_c_int32_t, ctx)
  if (st /= 0) stop 2
  st = lamb_hip_dist_create(ctx, int(n_atoms, c_int64_t), blk, n_subsets, smap, omap, &
                            c_loc(coords), c_loc(box), dist)
  if (st /= 0) stop 3
  a%ctx = ctx
  b%ctx = ctx
  c%ctx = ctx
  a%distribution = dist
  b%distribution = dist
  c%distribution = dist
  ! C's sectors: the same (row subset, col subset) list as A's
  allocate (c%sectors(size(a%sectors)))
  do i = 1, size(a%sectors)
    c%sectors(i)%subsets = a%sectors(i)%subsets
    c%sectors(i)%index%n_rows = a%sectors(i)%index%n_rows
  end do
  call lamb_multiply_set_eps_filter(eps)
  call lamb_multiply(a, b, c, error)
  print '(a,i0,1x,i0,1x,a)', 'lamb_multiply ', error%level, error%type, trim(error%message)
  if (error%level /= 0) stop 4
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) int(size(c%sectors), c_int32_t)
  do i = 1, size(c%sectors)
    associate (s => c%sectors(i))
      write (u) s%subsets, s%n_blks, s%data_stores(1)%data_len
      write (u) s%index%row_p
      if (s%n_blks > 0) write (u) s%index%col_i(1:s%n_blks), &
        (s%index%blk_info(st)%extent, st = 1, int(s%n_blks))
      if (s%data_stores(1)%data_len > 0) write (u) s%data_stores(1)%d(1:s%data_stores(1)%data_len)
    end associate
  end do
  close (u)
  call lamb_multiply_test_hip(ctx, trim(fdbcsr), err2, stats=stats)
  print '(a,i0,4(1x,i0),1x,es24.16)', 'multiply_test ', err2%level, stats%n_triples_kept, &
    stats%n_triples_filtered, stats%n_c_blocks_kept, stats%n_c_elems_kept, stats%flops
  st = lamb_hip_dist_destroy(dist)
  st = lamb_hip_finalize(ctx)
contains
  subroutine read_matrix(u, m)
    integer, intent(in) :: u
    type(lamb_matrix_d_type), intent(inout) :: m
    integer(c_int32_t) :: ns, k, sr, sc, nr
    integer(c_int64_t) :: nb, nd, t
    integer(c_int64_t), allocatable :: ext(:, :)
    read (u) ns
    allocate (m%sectors(ns))
    do k = 1, ns
      read (u) sr, sc, nr, nb, nd
      associate (s => m%sectors(k))
        s%subsets = [sr, sc]
        s%n_blks = nb
        s%index%n_rows = int(nr, c_int16_t)
        s%index%n_idx_blks = nb
        allocate (s%index%row_p(nr + 1), s%index%col_i(max(nb, 1_c_int64_t)))
        allocate (s%index%blk_info(nb), ext(2, max(nb, 1_c_int64_t)))
        allocate (s%data_stores(1))
        allocate (s%data_stores(1)%d(max(nd, 1_c_int64_t)))
        s%data_stores(1)%data_len = nd
        read (u) s%index%row_p
        if (nb > 0) read (u) s%index%col_i(1:nb), ext(:, 1:nb)
        if (nd > 0) read (u) s%data_stores(1)%d(1:nd)
        do t = 1, nb
          s%index%blk_info(t)%extent = ext(:, t)
        end do
        deallocate (ext)
      end associate
    end do
  end subroutine
end program drop_in
"""


def _write_input(path, dist, la, lb, eps):
    with open(path, "wb") as f:
        np.array([dist.n_atoms, dist.n_subsets], np.int32).tofile(f)
        np.array([eps], np.float64).tofile(f)
        dist.blk.astype(np.int8).tofile(f)
        dist.subset_map.astype(np.int32).tofile(f)
        dist.order_map.astype(np.int16).tofile(f)
        dist.coords.astype(np.float32).reshape(-1).tofile(f)
        np.array([dist.box] * 3, np.float32).tofile(f)
        for m in (la, lb):
            np.array([len(m.sectors)], np.int32).tofile(f)
            for sec in m.sectors:
                r, s = sec.subsets
                np.array([r, s, len(sec.row_p) - 1], np.int32).tofile(f)
                np.array([sec.n_blks, len(sec.data)], np.int64).tofile(f)
                sec.row_p.astype(np.int64).tofile(f)
                sec.col_i.astype(np.int16).tofile(f)
                sec.extents.astype(np.int64).reshape(-1).tofile(f)
                sec.data.astype(np.float64).tofile(f)


def _read_output(path, dist):
    from lamb_amd.matrix import LambMatrix, Sector
    raw = pathlib.Path(path).read_bytes()
    pos = 0

    def take(dtype, n):
        nonlocal pos
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=pos)
        pos += a.nbytes
        return a.copy()

    ns = int(take(np.int32, 1)[0])
    sectors = []
    for _ in range(ns):
        r, s = (int(x) for x in take(np.int32, 2))
        nb, nd = (int(x) for x in take(np.int64, 2))
        n_rows = len(dist.subset_atoms[r - 1])
        row_p = take(np.int64, n_rows + 1)
        col_i = take(np.int16, nb)
        ext = take(np.int64, 2 * nb).reshape(nb, 2)
        data = take(np.float64, nd)
        sectors.append(Sector((r, s), row_p, col_i, ext, data))
    assert pos == len(raw)
    return LambMatrix(dist, sectors)


def _fixture_csr(S):
    """The committed fixture's matrix (its JSON record) as a BlockCSR."""
    fx = json.loads((ROOT / "tests" / "golden" / "two_chunks.json").read_text())
    blk = np.asarray(fx["sizes"], np.int32)
    blocks = sorted((int(r) - 1, int(c) - 1, v) for r, c, v in fx["blocks"])
    row_p = np.zeros(len(blk) + 1, np.int64)
    for r, _, _ in blocks:
        row_p[r + 1] += 1
    row_p = np.cumsum(row_p)
    data = np.concatenate([np.asarray(v, np.float64) for _, _, v in blocks])
    sizes = np.array([len(v) for _, _, v in blocks], np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    return S.BlockCSR(blk, row_p, np.array([c for _, c, _ in blocks], np.int32), off, data)


@pytest.mark.gpu
@needs_flang
@needs_lib
def test_fortran_lamb_multiply_drop_in(tmp_path):
    import oracle
    from __graft_entry__ import compare
    from lamb_amd import systems as S
    from lamb_amd.matrix import Distribution, LambMatrix
    atoms = S.water_system(240, seed=151)
    a = S.decay_matrix(atoms, 8.0, 0.4, 152)
    b = S.decay_matrix(atoms, 8.0, 0.4, 153)
    eps = 1e-7
    dist = Distribution(atoms.blk, atoms.pos, atoms.box)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, dist, LambMatrix.from_global(dist, a), LambMatrix.from_global(dist, b), eps)
    exe = build(tmp_path, "drop_in", MULTIPLY_PROGRAM)
    print(f"case: Fortran lamb_multiply on {atoms.n_atoms} atoms, {dist.n_subsets} subsets",
          flush=True)
    out = subprocess.run([str(exe), str(fin), str(fout), str(FIXTURE)], capture_output=True,
                         text=True, timeout=300)
    print(out.stdout, flush=True)
    assert out.returncode == 0, out.stderr
    got = _read_output(fout, dist).to_global()
    ref, _ = oracle.multiply(a, b, eps=eps)
    same, err = compare(got, ref)
    print(f"  pattern identical {same}, relative error {err:.3e}", flush=True)
    assert same and err <= 1e-12
    # lamb_multiply_test_hip on the fixture against the oracle's A*A
    fa = _fixture_csr(S)
    _, rst = oracle.multiply(fa, fa, eps=0.0)
    line = [x for x in out.stdout.splitlines() if x.startswith("multiply_test")][0].split()
    assert int(line[1]) == 0
    assert int(line[2]) == rst["n_triples_kept"]
    assert int(line[3]) == rst["n_triples_filtered"]
    assert float(line[6]) == rst["flops"]
