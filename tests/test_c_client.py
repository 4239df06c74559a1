"""The C ABI from a C program (tests/abi_client.c), the way the nmesh-side
shim calls it: built with gcc -std=c99 -pedantic -Werror against
include/nmx.h alone and linked to nmesh_amd/libnmx.so.

- host mode (no device): a built, refined, mixed-order shell's tables
  handed to a second context with nmx_mesh_import come back identical; a
  swapped neighbour pair and elements out of order are refused;
- device mode: the same import into a device context; RHS, 3 RK4 steps,
  time and vars CRC bitwise equal between the built and the imported mesh,
  the state moved in through per-element arrays and read back through them.
"""
import os
import shutil
import subprocess

import pytest

import nmesh_amd as nm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_client(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc on this host")
    libdir = os.path.dirname(nm.LIB_PATH)
    exe = str(tmp_path / "abi_client")
    cmd = [gcc, "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O1",
           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "abi_client.c"),
           "-o", exe, "-L", libdir, "-lnmx", "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def run_client(exe, mode, timeout):
    print("abi_client", mode, flush=True)
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("abi_client %s ok" % mode), r.stdout


def test_c_client_host_import(tmp_path):
    run_client(build_client(tmp_path), "host", 120)


@pytest.mark.gpu
def test_c_client_device_transfers_and_steps(tmp_path):
    run_client(build_client(tmp_path), "device", 110)
