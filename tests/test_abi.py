"""The C-ABI library loads and exports every entry point include/nuc34.h
declares (CPU: no device calls)."""
import os
import re
import subprocess

from conftest import BUILD, ROOT


def declared():
    text = open(os.path.join(ROOT, "include", "nuc34.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nuc34_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "nuc34_decompose" in names and "nuc34_load_csr" in names and len(names) >= 20


def test_library_exports_all():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(BUILD, "libnuc34.so")], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_python_mirror_binds_all():
    import nuc34
    lib = nuc34.lib()
    for n in declared():
        assert getattr(lib, n).restype is not None or n.endswith(("_free", "_destroy"))


def test_no_cpu_fallback_in_mirror():
    src = open(os.path.join(ROOT, "gpu-nucleus34_amd", "nuc34", "__init__.py")).read()
    assert "oracle" not in src.replace("oracle/", "")
