"""CPU checks of the search kernels' 4-ary heap and of the workspace arena.

The Heap4 struct is copied out of point_geometric_features_amd/csrc/
pgeof_search.hip into tests/cpp/heap4_check.cpp, compiled with the host
compiler (with address and undefined-behaviour checks) and run: for every k
up to 130, with and without the global tail for k > 21 (laid out per lane
of the launch or per hardware wave slot), partially and fully
filled heaps over poisoned memory must come out as the k smallest keys in
order (see the .cpp file for the details).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "point_geometric_features_amd", "csrc", "pgeof_search.hip")
HARNESS = os.path.join(HERE, "cpp", "heap4_check.cpp")


def _heap4_struct() -> str:
    s = open(SRC).read()
    start = s.index("template <bool TAIL>\nstruct Heap4\n{")
    end = s.index("\n};\n", start) + 4
    return s[start:end]


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_heap4_on_cpu(tmp_path):
    code = open(HARNESS).read()
    assert "// @HEAP4@" in code
    cpp = tmp_path / "heap4_check.cpp"
    cpp.write_text(code.replace("// @HEAP4@", _heap4_struct()))
    exe = tmp_path / "heap4_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-o", str(exe), str(cpp)], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout


ARENA_HARNESS = os.path.join(HERE, "cpp", "arena_check.cpp")
HEADER = os.path.join(HERE, "..", "point_geometric_features_amd", "csrc", "pgeof_internal.hpp")


def _arena_struct() -> str:
    s = open(HEADER).read()
    start = s.index("struct Arena\n{")
    end = s.index("\n};\n", start) + 4
    return s[start:end]


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_arena_regions_aligned_on_skewed_bases(tmp_path):
    """Workspace regions start on 256-byte address boundaries whatever the
    base's alignment (a 64-bit atomic on an address 4 mod 8 faults the GPU:
    the host searches put the workspace after outputs of odd length)."""
    code = open(ARENA_HARNESS).read()
    assert "// @ARENA@" in code
    cpp = tmp_path / "arena_check.cpp"
    cpp.write_text(code.replace("// @ARENA@", _arena_struct()))
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-o", str(exe), str(cpp)], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ", 0 failures" in r.stdout
