"""The C++20 host interface (diplodocus_amd/host: AccelerationStructure-shaped GpuBvh,
factory, path tracer, Whitted renderer, OBJ scene loader, JSON configuration, stats
exporters) driven by its C++ test program tests/cpp/host_api_test.cc."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "diplodocus_amd", "host", "_build", "host_api_test")
SCENES = ["--scenes", os.path.join(ROOT, "tests", "golden", "obj"), "--configs",
          os.path.join(ROOT, "tests", "golden", "config")]


def run(args, timeout):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: build with __graft_entry__.build()")
    p = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=timeout)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0 and "host_api_test: OK" in p.stdout, p.stdout + p.stderr


def test_host_api_cpu_checks():
    run(["--cpu"] + SCENES, 60)


@pytest.mark.gpu
def test_host_api_gpu():
    run(SCENES, 240)
