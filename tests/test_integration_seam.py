"""The reference-side binding compiles against the reference's own headers.

integration/mem3d_amd.cpp and integration/mem2d_amd.cpp give
astraCUDA3d::FP/BP/FDK (include/astra/cuda/3d/mem3d.h:107-113) and
astraCUDA::FP/BP (include/astra/cuda/2d/mem2d.h:77-83) new bodies over the C
ABI.  They are checked with `hipcc -fsyntax-only` in the reference's HIP
header configuration (ASTRA_BUILDING_HIP maps the cuda* names of
include/astra/cuda/gpu_runtime_wrapper.h to HIP).  Skipped where the
reference sources are absent.  The product never includes these files.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

needs_ref = pytest.mark.skipif(
    not os.path.isfile(os.path.join(REF, "include", "astra", "cuda", "3d", "mem3d_internal.h"))
    or not (os.path.exists(HIPCC) or shutil.which("hipcc")),
    reason="reference headers or hipcc not available")


def _syntax(path):
    cmd = [HIPCC, "-fsyntax-only", "-std=c++17", "-DASTRA_BUILDING_HIP", "-DASTRA_CUDA",
           "-I" + os.path.join(REF, "include"), "-I" + os.path.join(REF, "lib", "include"),
           "-I" + os.path.join(ROOT, "include"), path]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300)


@needs_ref
@pytest.mark.parametrize("name", ["mem3d_amd.cpp", "mem2d_amd.cpp"])
def test_binding_compiles_against_reference_headers(name):
    r = _syntax(os.path.join(ROOT, "integration", name))
    errors = [l for l in r.stderr.splitlines() if "error" in l]
    assert r.returncode == 0 and not errors, r.stderr


@needs_ref
def test_binding_check_detects_errors(tmp_path):
    """The check is not vacuous: the binding with a wrong accessor (the
    non-existent astraCUDA3d::CDataGPU::getHandle) fails to compile."""
    src = open(os.path.join(ROOT, "integration", "mem3d_amd.cpp")).read()
    bad = src.replace("out = s->getPtr();", "out = s->getHandle();", 1)
    assert bad != src
    p = tmp_path / "bad.cpp"
    p.write_text(bad)
    r = _syntax(str(p))
    assert r.returncode != 0


def test_binding_declares_only_abi_symbols():
    """Every astra_amd_* function the binding calls is declared in the C ABI
    header (the binding uses nothing the library does not export)."""
    import re

    hdr = open(os.path.join(ROOT, "include", "astra_amd.h")).read()
    declared = set(re.findall(r"\b(astra_amd_\w+)\s*\(", hdr))
    for name in ("mem3d_amd.cpp", "mem2d_amd.cpp"):
        src = open(os.path.join(ROOT, "integration", name)).read()
        used = set(re.findall(r"\b(astra_amd_\w+)\s*\(", src))
        assert used, name
        assert used <= declared, (name, used - declared)
