"""Sanitizer leg (CPU): the dataset loaders (gpu-nucleus34_amd/host/graph_io.cpp)
and the CLI front end built with AddressSanitizer and UndefinedBehaviorSanitizer
(host code only, g++), fed the committed fixtures, randomised texts and
malformed inputs (truncated or inconsistent .grh files, oversized numbers,
binary bytes, very long lines). No sanitizer report may appear, and every
result the sanitized loader gives equals the product library's."""
import json
import os
import random
import shutil
import struct
import subprocess

import numpy as np
import pytest

import nuc34
from conftest import BUILD, GOLDEN, PKG, ROOT
from test_loader import _random_text

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
REPORTS = ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    loader = d / "loader_check"
    subprocess.run(["g++", "-std=c++17", *SAN, "-fopenmp", "-pthread", os.path.join(PKG, "host", "graph_io.cpp"),
                    os.path.join(ROOT, "tests", "sanitize", "loader_check.cpp"), "-o", str(loader)], check=True)
    cli = d / "nucleus34_san"
    subprocess.run(["g++", "-std=c++17", *SAN, "-DNUC34_CLI_DIRECT", os.path.join(PKG, "host", "cli.cpp"), "-o",
                    str(cli), f"-L{BUILD}", "-lnuc34", f"-Wl,-rpath,{BUILD}"], check=True)
    return d, loader, cli


def malformed(d):
    """Inputs a careless parser would mishandle."""
    files = {}

    def grh(name, sv, nv, ne, body):
        p = d / name
        p.write_bytes(struct.pack("<QQQ", sv, nv, ne) + body)
        files[name] = p

    grh("trunc_header.grh", 4, 0, 0, b"")
    (d / "trunc_header.grh").write_bytes(struct.pack("<Q", 4) + b"\x01\x02")
    grh("trunc_body.grh", 4, 100, 50, b"\x00" * 40)
    grh("huge_header.grh", 4, 0xFFFFFFFF, 0xFFFFFFFF, b"\x00" * 64)
    grh("wide_ids.grh", 8, 3, 2, b"\x00" * 64)
    grh("wrapping_degrees.grh", 4, 2, 1, struct.pack("<II", 0xFFFFFFFF, 2) + struct.pack("<I", 1))
    grh("degree_mismatch.grh", 4, 3, 2, struct.pack("<III", 1, 0, 0) + struct.pack("<II", 1, 2))
    grh("good.grh", 4, 3, 2, struct.pack("<III", 1, 1, 0) + struct.pack("<II", 1, 2))
    texts = {
        "empty.edges": b"",
        "comments_only.edges": b"# a\n% b\n#\n",
        "header_only.mtx": b"%%MatrixMarket\n3 3 2\n",
        "bad_header.mtx": b"1 2 3 4\n1 2\n",
        "big_numbers.edges": b"0 0\n4294967295 1\n4294967296 2\n99999999999999999999 3\n1 2\n",
        "negative.edges": b"0 0\n-1 2\n3 -4\n5 6\n",
        "binary.edges": bytes(random.Random(1).randrange(256) for _ in range(4096)),
        "long_line.edges": b"0 0\n" + b"7" * (1 << 20) + b" 1\n2 3\n",
        "no_newline.txt": b"0 0\n1 2\n2 3\n3 1",
        "crlf_mix.edges": b"0 0\r\n1 2\r\n2 3\n3 1\r\n\r\n",
        "nul_bytes.edges": b"0 0\n1\x002\n2 3\n",
        "tabs_commas.txt": b"x y\n1,2\n3\t4\t5\n 6  7 \n",
    }
    for name, data in texts.items():
        (d / name).write_bytes(data)
        files[name] = d / name
    return files


def fnv(a: np.ndarray) -> int:
    h = 1469598103934665603
    for b in np.ascontiguousarray(a, dtype="<u4").tobytes():
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def product_line(path) -> str:
    try:
        g = nuc34.load_dataset(str(path))
    except nuc34.Nuc34Error as e:
        return "ERR " + str(e).split("] ", 1)[1]
    return f"OK {g.n} {g.m} {int(g.oriented)} {fnv(g.O)} {fnv(g.E)}"


def test_sanitized_loader_matches_product(built, tmp_path):
    d, loader, _ = built
    rng = random.Random(7)
    paths = [os.path.join(GOLDEN, f) for f in sorted(os.listdir(GOLDEN)) if not f.endswith((".json", ".py"))]
    for i in range(40):
        p = tmp_path / f"r{i}{rng.choice(['.edges', '.txt', '.mtx'])}"
        p.write_text(_random_text(rng, rng.randrange(0, 80)), newline="")
        paths.append(str(p))
    paths += [str(p) for p in malformed(tmp_path).values()]
    r = subprocess.run([str(loader), str(d), *paths], capture_output=True, text=True, env=ENV, timeout=300)
    assert not any(s in r.stderr for s in REPORTS), r.stderr[-4000:]
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    lines = r.stdout.splitlines()
    assert len(lines) == len(paths)
    for path, line in zip(paths, lines):
        assert line == product_line(path), path
    errs = {os.path.basename(p): line for p, line in zip(paths, lines) if line.startswith("ERR")}
    assert errs["huge_header.grh"] == errs["trunc_body.grh"] == "ERR Truncated .grh body."
    assert "does not match" in errs["wrapping_degrees.grh"] and "does not match" in errs["degree_mismatch.grh"]
    assert errs["wide_ids.grh"] == "ERR Incompatible types in dataset."


def test_sanitized_cli_front_end(built, tmp_path):
    """The CLI's argument handling, JSON writer and error paths under the
    sanitizers: usage, unknown format, bad header, missing file, and a valid
    dataset (on a host without a device the record ends in the device step's
    error; with one it runs through)."""
    _, _, cli = built
    cases = [[], [os.path.join(GOLDEN, "unknown.xyz")], [os.path.join(GOLDEN, "badheader.mtx"), "2"],
             [str(tmp_path / "missing.edges")], [os.path.join(GOLDEN, "rmat8.edges"), "0", str(tmp_path / "o.k")]]
    for args in cases:
        r = subprocess.run([str(cli), *args], capture_output=True, text=True, env=ENV, timeout=300)
        assert not any(s in r.stderr for s in REPORTS), (args, r.stderr[-4000:])
        if not args:
            assert r.returncode == 1 and "Usage:" in r.stderr
            continue
        assert r.returncode == 0, (args, r.stderr[-2000:])
        rec = json.loads(r.stdout)
        assert rec["dataset"]["file"] == os.path.basename(args[0])
