"""DBCSR dump files: reading through the C-ABI, loading into device matrices,
and writing.

The format is the one lamb_load_dbcsr reads (src/lamb_io_operations.F90:
296-330; chunk offsets :557-597). The readers and the loader run in the
HIP library (lamb_hip_dbcsr_read_info / _read_blocks / _load); the readers
need no GPU. ``write`` produces the same layout from a BlockCSR, split into
chunks by block row the way a DBCSR writer's processes each own some rows.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N
from .systems import BlockCSR

_MSG = 512


class DbcsrError(RuntimeError):
    pass


def _path(path) -> bytes:
    return os.fsencode(os.fspath(path))


def read_info(path):
    """(info dict, row block sizes, column block sizes) of a DBCSR file."""
    lib = N.load()
    info = N.DbcsrInfo()
    msg = C.create_string_buffer(_MSG)
    rc = lib.lamb_hip_dbcsr_read_info(_path(path), C.byref(info), None, None, msg, _MSG)
    if rc != 0:
        raise DbcsrError(msg.value.decode(errors="replace"))
    rows = np.zeros(info.n_block_rows, np.int32)
    cols = np.zeros(info.n_block_cols, np.int32)
    rc = lib.lamb_hip_dbcsr_read_info(_path(path), C.byref(info), N.ptr(rows, C.c_int32),
                                      N.ptr(cols, C.c_int32), msg, _MSG)
    if rc != 0:
        raise DbcsrError(msg.value.decode(errors="replace"))
    d = {
        "version": info.version.decode("latin-1"),
        "name": info.name.decode("latin-1"),
        "matrix_type": info.matrix_type.decode("latin-1"),
        "n_chunks": info.n_chunks,
        "data_type": info.data_type,
        "n_block_rows": info.n_block_rows,
        "n_block_cols": info.n_block_cols,
        "n_blocks": info.n_blocks,
        "n_values": info.n_values,
    }
    return d, rows, cols


def read_blocks(path):
    """(rows, cols, values, sizes): every block in file order, rows / cols
    1-based, values the blocks back to back (each column-major), sizes the
    (row, column) block sizes."""
    d, rs, cs = read_info(path)
    lib = N.load()
    rows = np.zeros(d["n_blocks"], np.int32)
    cols = np.zeros(d["n_blocks"], np.int32)
    vals = np.zeros(d["n_values"], np.float64)
    msg = C.create_string_buffer(_MSG)
    rc = lib.lamb_hip_dbcsr_read_blocks(_path(path), d["n_blocks"], N.ptr(rows, C.c_int32),
                                        N.ptr(cols, C.c_int32), d["n_values"],
                                        N.ptr(vals, C.c_double), msg, _MSG)
    if rc != 0:
        raise DbcsrError(msg.value.decode(errors="replace"))
    return rows, cols, vals, (rs, cs)


def read(path) -> BlockCSR:
    """A square DBCSR file as a BlockCSR (columns ascending in each row)."""
    rows, cols, vals, (rs, cs) = read_blocks(path)
    if len(rs) != len(cs) or not np.array_equal(rs, cs):
        raise DbcsrError("read: the matrix is not square with equal row and column blocks")
    n = len(rs)
    sizes = rs[rows - 1].astype(np.int64) * cs[cols - 1]
    starts = np.concatenate([[0], np.cumsum(sizes)])[:-1]
    order = np.lexsort((cols, rows))
    if len(order) > 1:
        key = (rows[order].astype(np.int64) << 32) | cols[order]
        if np.any(key[1:] == key[:-1]):
            raise DbcsrError("read: a block appears twice")
    r, c, sz, st = rows[order] - 1, cols[order] - 1, sizes[order], starts[order]
    row_p = np.zeros(n + 1, np.int64)
    np.add.at(row_p, r + 1, 1)
    row_p = np.cumsum(row_p)
    data = np.concatenate([vals[s:s + z] for s, z in zip(st, sz)]) if len(sz) else np.zeros(0)
    off = np.concatenate([[0], np.cumsum(sz)])[:-1].astype(np.int64)
    return BlockCSR(rs.astype(np.int32), row_p, c.astype(np.int32), off, data)


def write(path, m: BlockCSR, *, n_chunks: int = 1, name: str = "", matrix_type: str = "N",
          version: str = "DBCSRv_1.0", data_type: int = 3):
    """Write m as a DBCSR dump with n_chunks chunks. Block row r belongs to
    chunk r % n_chunks; each chunk stores its blocks back to back."""
    n = m.n_atoms
    blk = np.asarray(m.blk, np.int32)
    rows = m.rows()
    owner = rows % n_chunks
    out = bytearray()
    out += version.encode("latin-1")[:10].ljust(10)
    out += name.encode("latin-1")[:80].ljust(80)
    out += matrix_type.encode("latin-1")[:1]
    out += np.array([n_chunks, data_type, n, n], "<i4").tobytes()
    out += blk.astype("<i4").tobytes() * 2
    h3, data = [], []
    for c in range(n_chunks):
        sel = np.flatnonzero(owner == c)
        cnt = np.bincount(rows[sel], minlength=n)
        row_p = np.concatenate([[0], np.cumsum(cnt)]).astype("<i4")
        sizes = blk[rows[sel]].astype(np.int64) * blk[m.col[sel]]
        blk_p = (np.concatenate([[0], np.cumsum(sizes)])[:-1] + 1).astype("<i4")
        vals = [m.data[m.blk_off[b]:m.blk_off[b] + z] for b, z in zip(sel, sizes)]
        area = np.concatenate(vals) if vals else np.zeros(0)
        out += np.array([len(sel), len(area)], "<i4").tobytes()
        h3.append(row_p.tobytes() + (m.col[sel] + 1).astype("<i4").tobytes() + blk_p.tobytes())
        data.append(area.astype("<f8").tobytes())
    with open(path, "wb") as f:
        f.write(bytes(out))
        for h in h3:
            f.write(h)
        for d in data:
            f.write(d)
