"""A C client of the C ABI: tests/c_client/ppmg_c_check.c, compiled with gcc
against libppmg.so by __graft_entry__.build().  Without a device it checks
the error paths; on the GPU it runs the registry entry points on host arrays
and checks mass conservation of the periodic fold, the duplicate face nodes,
M2P of a constant field (both kernels), and a masked ghost_get of an integer
field against its closed form."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "c_client", "ppmg_c_check")


def _run(mode):
    assert os.path.exists(EXE), f"{EXE} is missing: run __graft_entry__.build()"
    print(f"ppmg_c_check {mode}", flush=True)
    return subprocess.run([EXE, mode], capture_output=True, text=True, timeout=120)


def test_c_client_error_paths():
    r = _run("cpu")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "code=903" in r.stdout and "code=503" in r.stdout
    assert r.stdout.strip().endswith("ok")


def test_c_client_fails_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    r = _run("gpu")
    assert r.returncode != 0
    assert "code=9100" in r.stdout


@pytest.mark.gpu
def test_c_client_registry_on_host_arrays():
    r = _run("gpu")
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("duplicate face nodes equal node 1") == 2
    assert "integer ghost_get with a mask" in r.stdout
    assert r.stdout.strip().endswith("ok")
