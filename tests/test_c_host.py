"""The C host driver in INTEGRATION.md compiles against include/ggbtd.h with
warnings as errors, links against the built library and runs (no GPU needed:
it checks the ABI version, the footprint query and that a mismatched ABI
version is refused before any device call)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ggbtddft_amd", "libggbtd.so")


def _c_block():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text[text.index("## Equivalent C host"):]
    m = re.search(r"```c\n(.*?)```", sec, re.S)
    assert m, "no C block in the C host section"
    return m.group(1)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_c_host_compiles_links_and_runs(tmp_path):
    src = tmp_path / "drive.c"
    src.write_text(_c_block())
    obj = tmp_path / "drive.o"
    cc = ["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include")]
    r = subprocess.run(cc + ["-c", str(src), "-o", str(obj)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    if not os.path.exists(LIB):
        pytest.skip("libggbtd.so not built")
    exe = tmp_path / "drive"
    r = subprocess.run(["gcc", str(obj), "-L", os.path.dirname(LIB), "-lggbtd",
                        "-Wl,-rpath," + os.path.dirname(LIB), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert re.match(r"abi \d+ footprint \d+ refused: ggb_create: abi_version mismatch", r.stdout)
