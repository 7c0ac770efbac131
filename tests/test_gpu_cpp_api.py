"""Runs the C++ drop-in API test program (tests/cpp/test_mesh_api.cpp) on
the GPU: fcl::collide / fcl::distance / BVHModel / collideBatch used the way
FCL's own mesh tests use them, checked against the oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_mesh_api():
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "test_mesh_api")
    assert os.path.exists(exe), "build the C++ tests first (__graft_entry__.build())"
    print("case: C++ drop-in API on env.obj / rob.obj", flush=True)
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden")], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0 and "PASS" in r.stdout


def test_cpp_shape_api():
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "test_shape_api")
    assert os.path.exists(exe), "build the C++ tests first (__graft_entry__.build())"
    print("case: C++ drop-in API on primitive pairs", flush=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=240)
    print(r.stdout)
    assert r.returncode == 0 and "PASS" in r.stdout


def test_cpp_broadphase():
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "test_broadphase")
    assert os.path.exists(exe), "build the C++ tests first (__graft_entry__.build())"
    print("case: C++ broadphase manager over the device sweep", flush=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=240)
    print(r.stdout)
    assert r.returncode == 0 and "PASS" in r.stdout
