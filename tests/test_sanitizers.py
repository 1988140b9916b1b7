"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, and
fuzzing of the input decoders (CPU only).

- The scene library (BVH builders, OBJ/MTL import, PNG/PNM/JPEG decoding),
  the CPU oracle and the CPU model of the traversal, built with both
  sanitizers (tests/native/build_sanitized.py), run the workloads of
  tests/sanitized_workload.py in a child process that preloads the sanitizer
  runtime: no report, and outputs identical to the normal builds'.
- libFuzzer targets over the texture decoders and the OBJ/MTL import
  (tests/native/fuzz_image.cpp, fuzz_obj.cpp), seeded with the fixture images
  and models and with the inputs earlier fuzzing found (tests/golden/fuzz:
  a JPEG whose coefficients overflow the integer IDCT, a PNG header that
  declares a 4 GB image), run for a fixed time: no memory error, undefined
  behaviour, crash, failure without a message, invalid imported scene or
  allocation above 2 GB.
"""
import glob
import json
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, NATIVE)
import build_sanitized  # noqa: E402
import sanitized_workload  # noqa: E402

ARTIFACTS = ["libptscene_asan.so", "liboracle_asan.so", "libwidemodel_asan.so", "fuzz_image", "fuzz_obj"]
FUZZ_SECONDS = 20


@pytest.fixture(scope="module", autouse=True)
def sanitized_builds():
    if not all(os.path.exists(os.path.join(NATIVE, a)) for a in ARTIFACTS):
        build_sanitized.build_all()


def _env():
    env = dict(os.environ)
    env.update(build_sanitized.RUN_ENV)
    # Python itself leaks at exit; leaks are looked for in the fuzzers only.
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["LD_PRELOAD"] = build_sanitized.runtime_library()
    return env


@pytest.mark.parametrize("workload", ["scene", "oracle", "model"])
def test_sanitized_build_matches_normal_build(workload):
    proc = subprocess.run([sys.executable, os.path.join(HERE, "sanitized_workload.py"), workload], env=_env(),
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600)
    assert proc.returncode == 0, proc.stderr[-4000:]
    assert "runtime error" not in proc.stderr and "AddressSanitizer" not in proc.stderr, proc.stderr[-4000:]
    got = json.loads(proc.stdout.strip().splitlines()[-1])
    want = sanitized_workload.run(workload, sanitized=False)
    assert got == want


def _obj_seeds(dest):
    """OBJ fixtures packed as the OBJ target reads them: OBJ text, NUL, the
    model's MTL (its texture maps renamed to t.png), NUL, one of its images."""
    for i, obj in enumerate(sorted(glob.glob(os.path.join(GOLDEN, "assets", "models", "*", "*.obj")))):
        d = os.path.dirname(obj)
        mtls = sorted(glob.glob(os.path.join(d, "*.mtl")))
        imgs = sorted(glob.glob(os.path.join(d, "**", "*.png"), recursive=True) +
                      glob.glob(os.path.join(d, "**", "*.jpg"), recursive=True))[:1]
        text = open(obj, "rb").read().replace(b"\0", b" ")
        mtl = open(mtls[0], "rb").read().replace(b"\0", b" ") if mtls else b""
        mtl = mtl.replace(b"map_Kd ", b"map_Kd t.png #")
        img = open(imgs[0], "rb").read() if imgs else b""
        with open(os.path.join(dest, f"seed{i}"), "wb") as f:
            f.write(text + b"\0" + mtl + b"\0" + img)


def _image_seeds(dest):
    for path in sanitized_workload.IMAGES:
        shutil.copy(path, os.path.join(dest, os.path.basename(path)))


@pytest.mark.parametrize("target", ["fuzz_image", "fuzz_obj"])
def test_fuzz_decoders(target, tmp_path):
    corpus, work = tmp_path / "corpus", tmp_path / "work"
    corpus.mkdir()
    work.mkdir()
    (_image_seeds if target == "fuzz_image" else _obj_seeds)(str(corpus))
    env = dict(os.environ)
    env.update(build_sanitized.RUN_ENV)
    env["PT_FUZZ_DIR"] = str(work)
    # A small file may declare a large image that takes seconds to decode
    # under the sanitizers; stb_image accepts such files too, so the per-input
    # limit is generous and the allocation limit is what flags a decoder that
    # allocates for a declared size before checking it.
    cmd = [os.path.join(NATIVE, target), str(corpus), f"-max_total_time={FUZZ_SECONDS}", "-seed=1",
           "-timeout=60", "-rss_limit_mb=4096", "-malloc_limit_mb=2048", "-max_len=65536",
           f"-artifact_prefix={tmp_path}/", "-print_final_stats=1"]
    proc = subprocess.run(cmd, env=env, cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=FUZZ_SECONDS + 120)
    tail = "\n".join(proc.stdout.splitlines()[-40:])
    assert proc.returncode == 0, tail
    assert "Done" in proc.stdout, tail
