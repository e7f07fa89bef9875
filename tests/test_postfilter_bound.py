"""The tile epilogue decides a C block's post-filter bit from a cheap sum
(16 lane partials of one FMA per element, then DPP row prefix sums)
whenever that sum lies farther than (len + ceil(len/16) + 8) * 2^-52
(relative) from eps^2, and runs the contract's sequential FMA chain
otherwise. Every term is a square, so both sums are within gamma_k * S of
the exact sum; the margin is twice their first-order difference. This checks the rule on the CPU
against the sequential chain with eps^2 placed 1e-17..1e-9 (relative) from
the sum, and that a zero margin would be caught."""
import pathlib
import subprocess

import pytest

SRC = pathlib.Path(__file__).resolve().parent / "cpu_src" / "postfilter_bound.c"
MARGIN = "const double margin = sum * ((double)(len + (len + 15) / 16 + 8) * 0x1p-52);"


def build(tmp_path, text):
    src = tmp_path / "bound.c"
    src.write_text(text)
    exe = tmp_path / "bound"
    r = subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"],
                       capture_output=True, text=True)
    if r.returncode:
        pytest.skip(f"gcc failed: {r.stderr[-300:]}")
    return exe


def test_certain_decisions_match_the_sequential_chain(tmp_path):
    exe = build(tmp_path, SRC.read_text())
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and " wrong 0" in r.stdout


def test_a_zero_margin_is_caught(tmp_path):
    text = SRC.read_text()
    assert MARGIN in text
    exe = build(tmp_path, text.replace(MARGIN, "const double margin = 0.0;"))
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and " wrong 0" not in r.stdout
