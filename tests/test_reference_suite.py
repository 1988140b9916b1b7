"""The reference's own device integration tests on the MI355X core.

oracle/build_ref.sh compiles the reference's tests/cuda_tests.cpp and
tests/lighting_tests.cpp (with RAY_TRACER_HAS_CUDA) unchanged against this
repository's drop-in binding of src/cuda/cuda_renderer.h
(cuda-path-tracer_amd/binding/cuda_renderer_hip.cpp) and libptrace.so.  Their
own bars apply: device images within 5% relative L2 of the portable CPU
renderer, MIS within 1e-4 with identical ray counts, RGBA within +-1, exact
tie and ownership checks.  Exit code 77 would mean "no device", which is a
failure here.

preview_session_check (tests/native/preview_session_check.cpp, built by the
same script) instantiates the reference viewer's preview_session<cuda_renderer>
(src/display/preview_session.h) on the binding: after idle frames, camera
moves, focus loss, resizes and every debug mode, the session's RGBA equals a
fresh renderer's byte for byte and the reference host renderer's within 1,
and it reports rays and kernel time per frame.
"""
import os
import subprocess

import pytest

from conftest import ROOT

REF = os.path.join(ROOT, "oracle", "_ref")


@pytest.mark.gpu
@pytest.mark.parametrize("binary", ["cuda_tests", "cuda_lighting_tests", "preview_session_check"])
def test_reference_device_suite_passes(binary):
    path = os.path.join(REF, binary)
    if not os.path.exists(path):
        pytest.skip(f"{binary} was not built (needs the reference tree at build time)")
    proc = subprocess.run([path], capture_output=True, text=True, timeout=600)
    print(proc.stdout[-3000:], proc.stderr[-3000:])
    assert proc.returncode == 0, (proc.returncode, proc.stdout[-2000:], proc.stderr[-2000:])
    assert "0 failures" in proc.stdout
