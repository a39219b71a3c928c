"""The C and C++ headers compile on their own (CPU only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_headers_compile(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text('#include "xdg_hip.hpp"\nint main() { return xdg_hip::ID_NONE + 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                    "-I", os.path.join(ROOT, "include"), str(src)], check=True)
    csrc = tmp_path / "h.c"
    csrc.write_text('#include "xdg_hip.h"\nint main(void) { return XDG_HIP_OK; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                    "-I", os.path.join(ROOT, "include"), str(csrc)], check=True)
