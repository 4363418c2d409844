"""The device kernels' phase code, executed on the CPU.

krabmaga_amd/csrc/kg_phases.hpp and kg_rt.hpp hold the bodies of the
binning, flock and tile-resident kernels as __host__ __device__ functions.
tests/emu drives them block by block on the host; emu_check (built with
AddressSanitizer and UBSan) compares 35 scenarios bit for bit with the
oracle: both toroidal modes, padding cells, clustered and dense layouts,
stacked birds and positions on cell corners,
tiny LDS capacities (global paths), overflowing tile capacities (inbox and
spill), jumps longer than a cell, shuffled block and atomic orders.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")


@pytest.fixture(scope="module")
def emu_built():
    subprocess.run(["make", "-s", "-C", EMU], check=True)
    return EMU


def test_kernel_phases_match_oracle_under_sanitizers(emu_built):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(emu_built, "emu_check")], capture_output=True, text=True,
                       timeout=900, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "0 scenario(s) failed" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr
