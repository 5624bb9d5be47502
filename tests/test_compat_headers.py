"""The drop-in headers (include/spws/compat/) compile as C++ with g++ and
the reference-style drivers over them link against libspws.  Running the
drivers needs a GPU (tests/test_gpu_compat.py)."""
import glob
import os
import subprocess

import pytest

from helpers import build_compat_driver

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("header", sorted(glob.glob(os.path.join(REPO, "include", "spws", "compat", "*.h"))))
def test_header_compiles_alone(header, tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(f'#include "{header}"\nint main() {{ return 0; }}\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"),
                    str(src)], check=True)


@pytest.mark.parametrize("name", ["drv_csr", "drv_coord", "drv_cscT"])
def test_drivers_link(name, tmp_path):
    from sparsews_amd import build as B
    if not os.path.exists(B.LIB):
        pytest.skip("libspws.so is not built")
    assert os.access(build_compat_driver(name, tmp_path), os.X_OK)
