"""The C ABI from a compiled C and C++ caller (CPU): tests/c_abi/abi_check.c is compiled against
include/cravagpu.h with the C and the C++ compiler (warnings as errors) and linked with
libcravagpu.so; it checks the host-only entry points and the argument checks that run before
any device work, so a header that disagrees with the library fails to compile, link or run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "crava_amd")
SRC = os.path.join(HERE, "c_abi", "abi_check.c")


@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "libcravagpu.so")), reason="library not built")
@pytest.mark.parametrize("compiler,lang", [("gcc", ["-std=c99"]), ("g++", ["-x", "c++", "-std=c++17"])])
def test_compiled_caller(tmp_path, compiler, lang):
    if shutil.which(compiler) is None:
        pytest.skip("%s not available" % compiler)
    exe = str(tmp_path / ("abi_check_" + compiler.replace("+", "p")))
    cmd = ([compiler] + lang + ["-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-x", "none", "-L", LIBDIR, "-lcravagpu", "-Wl,-rpath," + LIBDIR, "-lm", "-o", exe])
    subprocess.run(cmd, check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "abi_check: ok" in r.stdout
