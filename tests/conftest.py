import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu-nucleus34_amd")
BUILD = os.path.join(PKG, "_build")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_built():
    if not os.path.exists(os.path.join(BUILD, "libnuc34.so")) or not os.path.exists(os.path.join(BUILD, "nuc34_gen")):
        subprocess.run(["make", "-s", "-j8", "-C", PKG], check=True)
    if not os.path.exists(os.path.join(ORACLE, "_build", "liboracle34.so")):
        subprocess.run(["make", "-s", "-C", ORACLE], check=True)


_ensure_built()


def gen(kind, *args, out):
    subprocess.run([os.path.join(BUILD, "nuc34_gen"), kind, *map(str, args), str(out)], check=True,
                   stdout=subprocess.DEVNULL)
    return str(out)


@pytest.fixture(scope="session")
def gen_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("graphs")
