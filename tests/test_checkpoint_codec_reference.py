"""The reference-compatible checkpoint codec against the reference's own bytes
(CPU only, no device).

tests/golden/<preset>.structure.w<N>.swfx holds serializeStructureCheckpoint's
bytes of the reference's Structure before its first step and after 20 steps
(tiny wings: the bytes; small wings: SHA-256 digests). tests/native/
checkpoint_recode.cpp, compiled here against the built library, decodes them
with the build's codec and writes them again: verbatim, and with the
suspension section restated from the decoded state the way a checkpoint
taken after a step is written (its topology and state fingerprints,
finishSubstep's segment, payload and suspension diagnostics). Both must
reproduce the reference's bytes exactly.
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "tests" / "golden"))
import swfx  # noqa: E402

DT = 1.0 / 120.0  # the fixtures' substep (one substep per 1/120 s step)


@pytest.fixture(scope="module")
def recode(tmp_path_factory):
    library = ROOT / "simwing_amd" / "libsimwing_xpbd.so"
    if not library.exists():
        pytest.skip("the library is not built")
    compiler = shutil.which("g++")
    if compiler is None or not Path("/opt/rocm/include").exists():
        pytest.skip("g++ or the ROCm headers are missing")
    exe = tmp_path_factory.mktemp("recode") / "checkpoint_recode"
    cmd = [compiler, "-std=c++20", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{ROOT / 'simwing_amd' / 'csrc'}", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "native" /
           "checkpoint_recode.cpp"), "-o", str(exe), f"-L{library.parent}", "-lsimwing_xpbd",
           f"-Wl,-rpath,{library.parent}", "-L/opt/rocm/lib", "-lamdhip64"]
    subprocess.run(cmd, check=True, capture_output=True)

    def run(preset: str, data: bytes) -> str:
        path = exe.parent / f"{preset}.bytes"
        path.write_bytes(data)
        out = subprocess.run([str(exe), preset, str(path), repr(DT)], capture_output=True, text=True,
                             env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
        return out.stdout.strip()
    return run


CASES = [("tiny", "w1"), ("tiny", "w0"), ("tiny-tethered", "w1"), ("tiny-free", "w1")]


@pytest.mark.parametrize("preset,tag", CASES)
@pytest.mark.parametrize("key", ["checkpoint_bytes_initial", "checkpoint_bytes"])
def test_reference_bytes_are_rewritten_exactly(golden, recode, preset, tag, key):
    data = golden(f"{preset}.structure.{tag}.swfx")[key].tobytes()
    assert recode(preset, data) == "verbatim 1 restated 1"

