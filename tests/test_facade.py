"""The C++ facade (include/amethyst.hpp) builds with g++ and its iterators
match the reference's bit_iterator semantics (golden vectors from the
reference's own src/bit_iterator.cuh). Host only."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_facade_iterators(tmp_path):
    exe = tmp_path / "facade"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "facade_iterators.cpp"), "-o", str(exe),
                           "-L", os.path.join(ROOT, "amethyst_amd"), "-lamethyst",
                           "-Wl,-rpath," + os.path.join(ROOT, "amethyst_amd") + ":/opt/rocm/lib"])
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))["bit_iterator"]
    out = subprocess.run([str(exe)] + ["%02x" % b for b in g["bytes"]], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    bits = out.stdout.splitlines()[0]
    assert [int(c) for c in bits] == g["index"]
    assert out.stdout.strip().endswith("OK")


@pytest.mark.gpu
def test_facade_wrappers_on_gpu():
    """Every facade wrapper, called from C++ on the GPU, against the oracle
    (tests/cpp/facade_gpu.cpp, prebuilt by __graft_entry__.build())."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    exe = os.path.join(ROOT, "tests", "cpp", "facade_gpu")
    assert os.path.exists(exe), "tests/cpp/facade_gpu is not built (make -C tests/cpp)"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ALL OK")
    assert out.stdout.count("\nok ") + out.stdout.startswith("ok ") >= 16
