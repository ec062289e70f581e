"""The C++ interface header and the drop-in example compile against the C ABI
(no device work)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_interface_compiles(tmp_path):
    src = os.path.join(ROOT, "qmcbeaver_amd", "examples", "dropin_vmc.cpp")
    out = tmp_path / "dropin.o"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-c", "-I",
                        os.path.join(ROOT, "include"), src, "-o", str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_c_header_is_plain_c(tmp_path):
    c = tmp_path / "t.c"
    c.write_text('#include "qmcbeaver_amd.h"\nint main(void){return qmcb_abi_version()==1?0:1;}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-c", "-I",
                        os.path.join(ROOT, "include"), str(c), "-o", str(tmp_path / "t.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
