"""CPU check of the FP32 kernel's update-list schedule (csrc/schedule.hpp):
a C++ harness emulates the kernel's register / LDS-stack / HBM sources and
compares every root value with direct recursion (tests/sched_harness.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_schedule_emulation(tmp_path):
    exe = tmp_path / "sched_harness"
    obj = tmp_path / "oracle.o"
    subprocess.run(["gcc", "-O1", "-std=c11", "-c", os.path.join(ROOT, "oracle", "b3_oracle.c"),
                    "-o", str(obj)], check=True)
    subprocess.run(["g++", "-O1", "-std=c++17", os.path.join(ROOT, "tests", "sched_harness.cpp"),
                    str(obj), "-o", str(exe), "-lpthread", "-lm"], check=True)
    out = subprocess.run([str(exe), "300"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK"), out.stdout
