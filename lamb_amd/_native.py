"""ctypes binding of liblamb_hip.so (declared in include/lamb_hip.h).

The library is the only compute path: there is no CPU fallback. Loading
fails with a RuntimeError when the shared object has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "liblamb_hip.so"

i8p = C.POINTER(C.c_int8)
i16p = C.POINTER(C.c_int16)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
vpp = C.POINTER(C.c_void_p)


class Stats(C.Structure):
    _fields_ = [
        ("n_triples_kept", C.c_int64),
        ("n_triples_filtered", C.c_int64),
        ("n_c_blocks_struct", C.c_int64),
        ("n_c_blocks_kept", C.c_int64),
        ("n_c_elems_kept", C.c_int64),
        ("n_tiles", C.c_int64),
        ("n_tile_entries", C.c_int64),
        ("flops", C.c_double),
        ("ms_symbolic", C.c_double),
        ("ms_numeric", C.c_double),
        ("ms_index", C.c_double),
        ("ms_total", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class DbcsrInfo(C.Structure):
    """lamb_hip_dbcsr_info."""
    _fields_ = [
        ("version", C.c_char * 11),
        ("name", C.c_char * 81),
        ("matrix_type", C.c_char),
        ("n_chunks", C.c_int32),
        ("data_type", C.c_int32),
        ("n_block_rows", C.c_int32),
        ("n_block_cols", C.c_int32),
        ("n_blocks", C.c_int64),
        ("n_values", C.c_int64),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "lamb_hip_init": (C.c_int32, [C.c_int32, vpp]),
    "lamb_hip_finalize": (C.c_int32, [C.c_void_p]),
    "lamb_hip_last_error": (C.c_int32, [C.c_void_p, i32p, i32p, C.c_char_p, C.c_int32]),
    "lamb_hip_dist_create": (C.c_int32, [C.c_void_p, C.c_int64, i8p, C.c_int32, i32p, i16p,
                                         f32p, f32p, vpp]),
    "lamb_hip_dist_destroy": (C.c_int32, [C.c_void_p]),
    "lamb_hip_matrix_create": (C.c_int32, [C.c_void_p, vpp]),
    "lamb_hip_matrix_destroy": (C.c_int32, [C.c_void_p]),
    "lamb_hip_matrix_copy": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lamb_hip_sector_upload": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int64, i64p,
                                           i16p, i64p, f64p, C.c_int64]),
    "lamb_hip_distribute": (C.c_int32, [C.c_void_p, C.c_int64, i32p, i16p, i64p, f64p,
                                        C.c_int64]),
    "lamb_hip_matrix_commit": (C.c_int32, [C.c_void_p]),
    "lamb_hip_sector_query": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, i64p, i64p]),
    "lamb_hip_sector_download": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, i64p, i16p,
                                             i64p, f64p]),
    "lamb_hip_matrix_info": (C.c_int32, [C.c_void_p, i64p, i64p]),
    "lamb_hip_goods_size": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, i64p, i64p]),
    "lamb_hip_goods_pack": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, i64p, C.c_int64,
                                        f64p, C.c_int64]),
    "lamb_hip_goods_unpack": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, i64p, C.c_int64,
                                          f64p, C.c_int64]),
    "lamb_hip_multiply": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_double, C.c_double, C.c_double, C.c_int32,
                                      C.POINTER(Stats)]),
    "lamb_hip_trace": (C.c_int32, [C.c_void_p, f64p]),
    "lamb_hip_trace_noalloc": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lamb_hip_sector_multiply": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                             C.c_int32, C.POINTER(Stats)]),
    "lamb_hip_matrix_filter": (C.c_int32, [C.c_void_p, C.c_double, i64p]),
    "lamb_hip_plexus_create": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, vpp]),
    "lamb_hip_plexus_destroy": (C.c_int32, [C.c_void_p]),
    "lamb_hip_plexus_info": (C.c_int32, [C.c_void_p, i32p, i32p, i32p, i32p, i64p]),
    "lamb_hip_plexus_maps": (C.c_int32, [C.c_void_p, i32p, i32p, i32p, i32p]),
    "lamb_hip_plexus_lists": (C.c_int32, [C.c_void_p, i32p, i32p, i32p]),
    "lamb_hip_plexus_sched": (C.c_int32, [C.c_void_p, i32p]),
    "lamb_hip_tuning_info": (C.c_int32, [C.c_void_p, i32p, i32p]),
    "lamb_hip_tuning_order": (C.c_int32, [C.c_void_p, i32p]),
    "lamb_hip_tuning_launch": (C.c_int32, [C.c_void_p, i32p, i32p, i32p]),
    "lamb_hip_tuning_mode": (C.c_int32, [C.c_void_p, C.c_int32]),
    "lamb_hip_tuning_save": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "lamb_hip_tuning_load": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "lamb_hip_dist_set_groups": (C.c_int32, [C.c_void_p, C.c_int32, i32p]),
    "lamb_hip_group_trace": (C.c_int32, [C.c_void_p, f64p]),
    "lamb_hip_group_norm": (C.c_int32, [C.c_void_p, f64p]),
    "lamb_hip_group_diff_norm": (C.c_int32, [C.c_void_p, C.c_void_p, f64p]),
    "lamb_hip_group_gershgorin": (C.c_int32, [C.c_void_p, f64p, f64p]),
    "lamb_hip_group_scale_shift": (C.c_int32, [C.c_void_p, f64p, f64p]),
    "lamb_hip_mcweeny": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                                     f64p, f64p, C.POINTER(Stats)]),
    "lamb_hip_mcweeny_capture": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                                             C.c_int32, vpp]),
    "lamb_hip_graph_replay": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, f64p, f64p,
                                          f64p]),
    "lamb_hip_graph_info": (C.c_int32, [C.c_void_p, i32p, i64p, f64p]),
    "lamb_hip_graph_destroy": (C.c_int32, [C.c_void_p]),
    "lamb_hip_multiply_adjoint": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_int32,
                                              C.POINTER(Stats)]),
    "lamb_hip_matrix_transpose": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lamb_hip_matrix_project": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "lamb_hip_matrix_get_values": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    "lamb_hip_matrix_set_values": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    "lamb_hip_matrix_release_values": (C.c_int32, [C.c_void_p]),
    "lamb_hip_matrix_block_atoms": (C.c_int32, [C.c_void_p, i32p, i32p]),
    "lamb_hip_dbcsr_read_info": (C.c_int32, [C.c_char_p, C.c_void_p, i32p, i32p, C.c_char_p,
                                             C.c_int32]),
    "lamb_hip_dbcsr_read_blocks": (C.c_int32, [C.c_char_p, C.c_int64, i32p, i32p, C.c_int64,
                                               f64p, C.c_char_p, C.c_int32]),
    "lamb_hip_dbcsr_load": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "lamb_hip_multiply_test": (C.c_int32, [C.c_void_p, C.c_char_p, C.c_double,
                                           C.POINTER(Stats)]),
    "lamb_hip_partition_atoms": (C.c_int32, [C.c_void_p, C.c_int64, f32p, f32p, i32p, i32p,
                                             i16p]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library, raising if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m lamb_amd.build` "
            "(the HIP library is the only compute path)")
    lib = C.CDLL(os.fspath(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def library_info() -> dict:
    """Which library this process uses: its path and hash, and whether it is
    the one the build record describes and was built from the sources in
    this tree (the record is written by lamb_amd.build next to the library)."""
    import hashlib
    import json
    info = {"path": os.fspath(LIB_PATH), "loaded": _lib is not None}
    if not LIB_PATH.exists():
        return info
    info["sha256"] = hashlib.sha256(LIB_PATH.read_bytes()).hexdigest()
    record = LIB_PATH.with_name("liblamb_hip.build.json")
    if record.exists():
        from lamb_amd import build as B
        rec = json.loads(record.read_text())
        info["built_utc"] = rec.get("built_utc")
        info["matches_build_record"] = rec.get("library_sha256") == info["sha256"]
        info["built_from_these_sources"] = rec.get("sources_sha256") == B.sources_sha256()
    return info


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
