"""The C++ call surface (include/mnrt_host.h: KDTreeTri, KDTreePoint, PhotonMap,
RTUpdateKernelData / KernelRTTraceRays, KernelRTSampleAreaLight / KernelRTTraceShadowRays*,
PMSetPhotonMap / KernelPMPhotonGather) runs on the
GPU and agrees with the C ABI it wraps; see tests/host/test_host_classes.cpp."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "test_host_classes")


def test_host_classes():
    assert os.path.exists(BIN), f"{BIN} is missing: build it with __graft_entry__.build()"
    print("running the host-class test program", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=180)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
