"""The gpu_bitvector shim of the integration notes (the first C++ block
under "Bitvector operators", which a FastBit maintainer would add)
compiles against the reference's own headers (ibis::bitvector,
ibis::array_t) and links against the reference library and
libfastbit_hip.so.  CPU only; needs the reference sources, so it is
skipped where they are absent."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
CLANG = "/opt/rocm/llvm/bin/clang++"


def test_shim_compiles_and_links(tmp_path):
    if not os.path.isdir(REF) or not os.path.exists(CLANG):
        pytest.skip("reference sources or clang not present")
    refso = os.path.join(ROOT, "oracle", "_ref", "libfbref.so")
    hipso = os.path.join(ROOT, "libfastlib_amd", "lib", "libfastbit_hip.so")
    if not (os.path.exists(refso) and os.path.exists(hipso)):
        pytest.skip("libraries not built")
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    i = doc.index("## Bitvector operators")
    code = re.search(r"```cpp\n(.*?)```", doc[i:], re.S).group(1)
    src = tmp_path / "shim.cpp"
    src.write_text(code + "\nint main() {\n    ibis::bitvector b;\n    b.set(1, 100);\n"
                   "    gpu_bitvector g(b);\n    g &= g;\n    ibis::bitvector o;\n    g.to(o);\n"
                   "    return (int)g.cnt() - (int)o.cnt();\n}\n")
    # the reference headers need the compiler switch and defines its own
    # build uses (oracle/Makefile)
    cmd = [CLANG, "-std=gnu++11", "-w", "-Xclang", "-fdebugger-support", "-DWITHOUT_FASTBIT_CONFIG_H",
           "-I" + os.path.join(ROOT, "include"), "-I" + REF, "-I/root/reference/win", str(src),
           "-o", str(tmp_path / "shim"), refso, hipso, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
