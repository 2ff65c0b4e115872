"""The C++ API client (tests/cpp/goofit_client.cpp) on the CPU: it is built against
include/ and linked to this repository's libgoofit.so, and without a GPU it fails
loudly with the API's GeneralError instead of computing anything elsewhere."""
import os
import subprocess

import pytest

from goofit_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
CLIENT = os.path.join(HERE, "cpp", "goofit_client")
LIBDIR = os.path.realpath(os.path.join(HERE, "..", "goofit_amd", "lib"))


def test_client_is_built_and_links_libgoofit():
    assert os.path.exists(CLIENT), "tests/cpp/goofit_client is not built (run the build first)"
    ldd = subprocess.run(["ldd", CLIENT], capture_output=True, text=True, check=True).stdout
    libs = {line.split()[0]: os.path.realpath(line.split("=>")[1].split()[0])
            for line in ldd.splitlines() if "=>" in line}
    assert libs.get("libgoofit.so") == os.path.join(LIBDIR, "libgoofit.so")
    assert libs.get("libgoofit_hip.so") == os.path.join(LIBDIR, "libgoofit_hip.so")


def test_client_fails_loudly_without_gpu(tmp_path):
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible")
    ev = tmp_path / "ev.txt"
    ev.write_text("0.1\n0.2\n")
    r = subprocess.run([CLIENT, "add1d", str(ev)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "GeneralError" in r.stderr and "GPU" in r.stderr
    assert "nll" not in r.stdout
