"""The RayTracer subclass of the integration guide (its C++ block that
starts with `// file: include/xdg/hip/ray_tracer.h`) compiles against the
reference's own headers, with the guide's constants.h edit (a HIP library
id) applied to a scratch copy (CPU only; skipped where the reference tree is
absent)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


def _fmt_include():
    try:
        import torch
        inc = os.path.join(os.path.dirname(torch.__file__), "include")
        if os.path.exists(os.path.join(inc, "fmt", "format.h")):
            return inc
    except Exception:
        pass
    return None


def test_binding_compiles_against_reference_headers(tmp_path):
    if not os.path.isdir(os.path.join(REF, "include", "xdg")):
        pytest.skip("reference headers not present")
    fmt = _fmt_include()
    if fmt is None:
        pytest.skip("no fmt headers")
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    binding = [b for b in blocks if b.startswith("// file: include/xdg/hip/ray_tracer.h")]
    assert len(binding) == 1
    inc = tmp_path / "include"
    shutil.copytree(os.path.join(REF, "include"), inc, symlinks=False,
                    ignore=shutil.ignore_patterns("linalg*"))
    const = inc / "xdg" / "constants.h"
    src = const.read_text()
    src2 = src.replace("  EMBREE,\n  GPRT\n};", "  EMBREE,\n  GPRT,\n  HIP\n};")
    src2 = src2.replace('  {RTLibrary::GPRT, "GPRT"}\n};', '  {RTLibrary::GPRT, "GPRT"},\n  {RTLibrary::HIP, "HIP"}\n};')
    assert src2 != src and "HIP" in src2
    const.write_text(src2)
    (inc / "xdg" / "hip").mkdir()
    (inc / "xdg" / "hip" / "ray_tracer.h").write_text(binding[0])
    unit = tmp_path / "ray_tracer.cpp"
    unit.write_text('#include "xdg/hip/ray_tracer.h"\n'
                    'std::shared_ptr<xdg::RayTracer> make() { return std::make_shared<xdg::HIPRayTracer>(); }\n')
    # ray_tracing_interface.h includes primitive_ref.h only for Embree's
    # callback declarations (nothing of it is used by the interface);
    # Embree is not installed here, so that header's guard is pre-defined.
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-DFMT_HEADER_ONLY",
                        "-D_XDG_TRIANGLE_REF_H",
                        "-I", str(inc), "-I", os.path.join(ROOT, "include"), "-I", fmt, str(unit)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
