"""The C++ drop-in's math types (include/fcl/fcl.h) against the Eigen
idioms FCL callers build poses with -- AngleAxis, Quaternion, Translation3,
Transform3's translate/rotate forms, normalisation -- checked against closed
forms by tests/cpp/test_types.cpp (host only: no GPU, no library)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_math_types(tmp_path):
    exe = tmp_path / "test_types"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-O1", "-I",
                    os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_types.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stdout
    assert "types ok" in r.stdout
