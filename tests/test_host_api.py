"""The C++ host interface (include/krabgpu.hpp): Field2D, State, Agent,
Schedule, simulate and the flockers State with the reference's names.

tests/cpp/test_host_api.cpp holds the cases: Schedule::step order and
update timing (src/engine/schedule.rs:347-413) and simulate!'s loop
(src/lib.rs:1158-1173) on the host; on the GPU, the reference's own Field2D
tests (tests/engine/field_2d.rs:30-247) and flockers runs through Schedule
and through Flocker::run, bit for bit against the oracle."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
BIN = os.path.join(CPP, "test_host_api")


def _binary(build):
    if build or not os.path.exists(BIN):
        lib = os.path.join(os.path.dirname(HERE), "krabmaga_amd", "libkrabgpu.so")
        assert os.path.exists(lib), "build with __graft_entry__.build()"
        subprocess.run(["make", "-s", "-C", os.path.join(os.path.dirname(HERE), "oracle")],
                       check=True)
        subprocess.run(["make", "-s", "-C", CPP], check=True)
    return BIN


def test_host_interface_builds_and_schedules():
    r = subprocess.run([_binary(True), "--cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout


@pytest.mark.gpu
def test_host_interface_on_the_device():
    r = subprocess.run([_binary(False)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
