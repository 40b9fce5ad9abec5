"""CPU tier: the C++ facade (include/orb_gpu.hpp) compiles and links against
liborb_gpu.so, and its device-free entry points run."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_facade_builds_and_runs(orb, tmp_path):
    exe = str(tmp_path / "facade_example")
    pkg = os.path.dirname(orb.library_path)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "facade_example.cpp"), "-L", pkg, "-lorb_gpu",
                    f"-Wl,-rpath,{pkg}", "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert "facade ok" in out.stdout
