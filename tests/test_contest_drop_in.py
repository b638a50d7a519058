"""The reference's C++ entry points (Contest::build_context / execute / destroy_context,
include/plan.h:398-405) over the C ABI through include/hjoin_contest.hpp, run as a compiled C++
program against types of the reference's plan.h shape; results are compared with a nested-loop
join inside the program."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hash-join-optimizer_amd", "_build", "contest_drop_in")


def test_adapter_header_compiles(tmp_path):
    """the header-only adapter compiles against plan.h-shaped types (no GPU needed)"""
    out = tmp_path / "a.o"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "contest_drop_in.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_contest_execute_drop_in():
    assert os.path.exists(BIN), "build the harness first: python __graft_entry__.py build"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "contest ok" in r.stdout
