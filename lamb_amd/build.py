"""Compile the HIP sources of the package into lamb_amd/liblamb_hip.so (gfx950)."""
from __future__ import annotations

import concurrent.futures as cf
import datetime
import hashlib
import json
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "liblamb_hip.so"
RECORD = HERE / "liblamb_hip.build.json"
BUILD = HERE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LAMB_HIP_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    *os.environ.get("LAMB_HIP_EXTRA_FLAGS", "").split(),
]


def _compile(src: pathlib.Path) -> pathlib.Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    RECORD.write_text(json.dumps({
        "library": OUT.name,
        "library_sha256": file_sha256(OUT),
        "sources_sha256": sources_sha256(),
        "built_utc": datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
        "hipcc": HIPCC,
        "flags": FLAGS,
    }, indent=1) + "\n")
    if verbose:
        print(f"built {OUT}")
    return OUT


def file_sha256(path: pathlib.Path) -> str:
    return hashlib.sha256(path.read_bytes()).hexdigest()


def sources_sha256() -> str:
    """One hash over the HIP sources and the public header, in name order."""
    h = hashlib.sha256()
    for f in sorted(CSRC.glob("*")) + [HERE.parent / "include" / "lamb_hip.h"]:
        if f.is_file():
            h.update(f.name.encode())
            h.update(f.read_bytes())
    return h.hexdigest()


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
