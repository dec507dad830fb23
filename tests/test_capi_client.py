"""include/semx.h from a compiled caller (CPU, no GPU calls): a C program,
compiled as C and as C++ against the header and linked against
libsemx.so, builds a mesh, hands its numbering back through
semx_mesh_create_numbered, and sees an unfit numbering and a bad argument
refused with their status codes; it switches a mesh to cylindrical
coordinates and builds a system for the coupled field v~ on it."""
import os
import shutil
import subprocess

import pytest

import semtex_amd as sx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "semtex_amd")


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_compiled_client(tmp_path, lang):
    sx.load()   # the library exists (built in-tree)
    cc = "gcc" if lang == "c" else "g++"
    if shutil.which(cc) is None:
        pytest.skip(f"{cc} not found")
    exe = str(tmp_path / f"client_{lang}")
    src = os.path.join(ROOT, "tests", "capi", "client.c")
    cmd = [cc, "-x", lang, src, "-x", "none", "-I", os.path.join(ROOT, "include"), "-o", exe,
           "-L", LIBDIR, "-lsemx", f"-Wl,-rpath,{LIBDIR}", "-Wall", "-Werror"]
    if lang == "c":
        cmd.insert(1, "-std=c99")
    subprocess.check_call(cmd)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = dict(line.split(" ", 1) for line in out.stdout.strip().splitlines())
    info = [int(v) for v in lines["info"].split()]
    # np nel next nint nglobal nsolve nband nbound nPlane planeSize
    assert info[:4] == [4, 2, 12, 4]
    assert info[4] == 2 * 12 - 4 and info[7] == 6 and info[8] == 32 and info[9] == 32
    assert lines["numbered"] == "0"
    assert lines["refused"].startswith("7 ")
    assert lines["bad-arg"] == "1"
    assert lines["cyl"] == "0" and lines["field-v"] == "0"
    assert lines["hbb00-positive"] == "1"
    assert lines["cyl-after-system"] == "1" and lines["field-x"] == "1"
