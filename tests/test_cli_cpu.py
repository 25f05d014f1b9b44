"""CLI surface that needs no device: usage and dataset errors."""
import json
import os
import subprocess

from conftest import BUILD, GOLDEN


def run(*args):
    return subprocess.run([os.path.join(BUILD, args[0]), *args[1:]], capture_output=True, text=True)


def test_usage():
    for exe in ("nucleus34", "nucleus34_ondemand"):
        r = run(exe)
        assert r.returncode == 1 and "Usage:" in r.stderr


def test_unknown_extension_error_json():
    for exe in ("nucleus34", "nucleus34_ondemand"):
        r = run(exe, os.path.join(GOLDEN, "unknown.xyz"))
        assert r.returncode == 0
        rec = json.loads(r.stdout)
        assert rec["dataset"]["error"] == "Unknown dataset format: .xyz"
        assert rec["dataset"]["errorstep"].startswith("0.2")
        assert rec["environment"]["app"].startswith("hip-nucleus34")


def test_bad_mtx_header_error_json():
    r = run("nucleus34", os.path.join(GOLDEN, "badheader.mtx"))
    rec = json.loads(r.stdout)
    assert rec["dataset"]["error"] == "Wrong header for mtx format!"
