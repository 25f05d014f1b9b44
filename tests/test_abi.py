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


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of the ABI structs have the C layout (sizes and offsets)."""
    import ctypes
    import subprocess
    import nuc34
    fields = {"nuc34_serve_config": (nuc34.ServeConfig, ["hbm_budget", "large_fraction", "variant", "batch_window_ms", "large_workers", "large_wedges", "small_workers"]),
              "nuc34_job_info": (nuc34.JobInfo, ["n", "t_submit", "t_done", "error"]),
              "nuc34_serve_stats": (nuc34.ServeStats, ["splits", "busy_s"]),
              "nuc34_stats": (nuc34.Stats, ["k_max", "ms_total", "n_kernels", "kernels", "debits", "ms_tri_order",
                                             "ms_peel_init"]),
              "nuc34_tuning": (nuc34.Tuning, ["peel_tile_cap", "k4_list_grid_cap", "tri_tile_edges",
                                                "peel_finish_grid", "batch_peel", "graph_peel_max_tris",
                                                "tri_schedule"]),
              "nuc34_device_rates": (nuc34.DeviceRates, ["atomic_gops", "load_hbm_gops"])}
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "nuc34.h"', "int main(void) {"]
    for st, (_, fs) in fields.items():
        src.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            src.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-I", inc, str(c), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                                check=True).stdout.split("\n") if line)
    for st, (cls, fs) in fields.items():
        assert int(out[st]) == ctypes.sizeof(cls), st
        for f in fs:
            assert int(out[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)


def test_failed_context_creation_names_its_cause():
    """Without a usable device nuc34_ctx_create fails with a status, and
    nuc34_last_error(NULL) gives that failure's message (no context holds it)."""
    import ctypes
    import nuc34
    lib = nuc34.lib()
    ctx = ctypes.c_void_p()
    rc = lib.nuc34_ctx_create(0, ctypes.c_uint64(0), ctypes.byref(ctx))
    if rc == 0:  # a device is present: nothing to check
        lib.nuc34_ctx_destroy(ctx)
        return
    assert not ctx.value
    msg = lib.nuc34_last_error(None)
    msg = msg.decode() if isinstance(msg, bytes) else ctypes.cast(msg, ctypes.c_char_p).value.decode()
    assert msg and msg != "null context", msg
