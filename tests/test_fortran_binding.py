"""The Fortran 2003 binding (lamb_amd/fortran/lamb_hip.f90) compiles with AMD
flang and links every bind(C) interface against liblamb_hip.so.

No GPU work: the program references each entry point behind a branch that is
only taken with command-line arguments, and is run without any.
"""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
LIB = ROOT / "lamb_amd" / "liblamb_hip.so"

PROGRAM = """
program link_check
  use iso_c_binding
  use lamb_hip
  implicit none
  type(c_ptr) :: ctx, dist, a
  type(lamb_hip_stats) :: stats
  type(lamb_hip_matrix) :: ma, mb, mc
  type(lamb_hip_error) :: err
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
    call lamb_multiply_hip(ma, mb, mc, err, 1.0e-8_c_double)
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


@pytest.mark.skipif(not pathlib.Path(FLANG).exists(), reason="flang not installed")
@pytest.mark.skipif(not LIB.exists(), reason="liblamb_hip.so not built")
def test_fortran_binding_compiles_and_links(tmp_path):
    mod = ROOT / "lamb_amd" / "fortran" / "lamb_hip.f90"
    prog = tmp_path / "link_check.f90"
    prog.write_text(PROGRAM)
    subprocess.run([FLANG, "-c", str(mod), "-J", str(tmp_path), "-o", str(tmp_path / "m.o")],
                   check=True, capture_output=True, text=True)
    subprocess.run([FLANG, "-c", str(prog), "-I", str(tmp_path), "-o", str(tmp_path / "p.o")],
                   check=True, capture_output=True, text=True)
    exe = tmp_path / "link_check"
    r = subprocess.run([FLANG, str(tmp_path / "p.o"), str(tmp_path / "m.o"), "-o", str(exe),
                        f"-L{LIB.parent}", "-llamb_hip", f"-Wl,-rpath,{LIB.parent}",
                        "-Wl,--no-undefined"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "linked"
