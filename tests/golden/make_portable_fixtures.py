"""Writes the export-scale fixture of the oracle run with portable
transcendentals (the functions the device kernels use).

The oracle is the CPU restatement of the reference substep; with
``use_portable_math`` it calls the portable double-double atan2/sin/cos/acos
instead of the C library's, and is otherwise the same code that reproduces
the reference's own states bit for bit. Its export-wing trajectory is the
bit-exact target of the device path (the C library's last-bit rounding
differences grow chaotically over 240 steps, so the reference's own
trajectory can only be matched in distribution at that horizon).

Output ``tests/golden/export.portable.swfx``: for each snapshot step, the
SHA-256 digest (4 uint64 words) and the stride-101 samples of every state
array, plus per-step state hashes.

Usage: python tests/golden/make_portable_fixtures.py [steps]
"""
from __future__ import annotations

import hashlib
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
for path in (ROOT, HERE, ROOT / "tests", ROOT / "oracle"):
    sys.path.insert(0, str(path))

import oracle as orc  # noqa: E402
import simwing_amd as sw  # noqa: E402
import swfx  # noqa: E402
from helpers import flight_case  # noqa: E402

STATE_KEYS = ("position", "previous", "velocity", "membrane_lambda", "membrane_resultant", "membrane_residual",
              "constraint_lambda", "dihedral_lambda", "segment_lambda", "payload")
STRIDE = 101
SNAPSHOTS = (1, 24, 240)


def digest(array: np.ndarray) -> np.ndarray:
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(array).tobytes()).digest(), dtype=np.uint64).copy()


def main() -> None:
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 240
    definition, pressure, loads = flight_case("export")
    body = orc.OracleBody(definition)
    body.use_portable_math(True)
    body.set_face_pressure(pressure)
    settings = sw.flight_step_settings()
    out: dict[str, np.ndarray] = {}
    start = time.time()
    for step in range(1, steps + 1):
        body.set_external_forces(loads)
        body.step(settings)
        if step in SNAPSHOTS:
            state = body.state()
            for key in STATE_KEYS:
                out[f"step{step}/{key}/digest"] = digest(state[key])
                out[f"step{step}/{key}/samples"] = state[key][::STRIDE].copy()
            print(f"step {step}: {time.time() - start:.0f} s", flush=True)
    swfx.write(str(HERE / "export.portable.swfx"), out)


if __name__ == "__main__":
    main()
