"""Regenerates the golden fixtures from the reference's own CPU solver.

Runs oracle/_ref/ref_driver (the reference's src/softwing and src/fsi
sources compiled by ``make -C oracle ref``; it needs /root/reference and so
runs only where the reference is present) on the build's seeded synthetic
wings, and writes into tests/golden/:

* ``<preset>.assembly.swfx``  -- the reference's assembleSceneStructure
  tables (tiny, small), plus the build's hinge colour order;
* ``<preset>.colouring.swfx`` -- SoftBody::constraintColouringView();
* ``<preset>.states.swfx``    -- full solver state after the listed steps;
* ``<preset>.frame.swfx``     -- ``ref_driver --frame``: the reference's own
  simwing::fsi::Structure stepped with the flight loads (no face pressure,
  loads pending again after the last step) and the diagnostic-frame fields
  simwing::viewer::buildStructureFrame makes of it (tiny: with the frame's
  serialized bytes; small: their SHA-256 digest and size);
* ``<preset>.scene.bin``      -- ``ref_driver --scene``: the reference's
  writeScene stream of the seeded wing (tiny: the stream itself; small:
  ``small.scene.digest.swfx`` with its SHA-256 digest and size);
* ``<preset>.structure.w<N>.swfx`` -- ``ref_driver --structure``: the
  reference's simwing::fsi::Structure stepped 20 times with the flight loads
  (no face pressure) with workerThreads N (0: the serial sweep, 1: the
  coloured sweep): definitionFingerprint(), suspensionState() before the
  first step and after the last, the checkpoint's public fields,
  serializeStructureCheckpoint's bytes after the last step and before the
  first, and the canonical text of the suspension definition the Structure
  builds (tiny: the bytes; small: their SHA-256 digests and sizes);
  (also, with workerThreads 1 only, for the tethered and free wings, whose
  checkpoints have no suspension section);
* ``<preset>.structure.damped.swfx`` -- the same with workerThreads 1 and
  every suspension segment's axial damping set to LINE_DAMPING;
* ``<preset>.structure.compressed.swfx`` -- the same with workerThreads 1 and
  every membrane's compressionStiffnessRatio set to COMPRESSION_RATIO;
* ``<preset>.digest.swfx``    -- for the ~70k-triangle wing (export) and the
  tethered and free small wings (small-tethered, small-free): SHA-256 digests
  of every assembly and colouring array, and strided samples plus digests of
  the state after 1 and 240 steps (the full arrays are too large to commit;
  dihedral arrays are in the reference's hinge insertion order, which is the
  build's colour-major order).

Usage: python tests/golden/make_fixtures.py [preset ...]
"""
from __future__ import annotations

import hashlib
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(HERE))
import swfx  # noqa: E402

DRIVER = ROOT / "oracle" / "_ref" / "ref_driver"
WORK = Path(tempfile.mkdtemp(prefix="simwing-fixtures-"))

PLAN = {
    "tiny": {"steps": 20, "snapshots": [1, 5, 20], "workers": 12},
    "small": {"steps": 120, "snapshots": [1, 120], "workers": 12},
    "export": {"steps": 240, "snapshots": [1, 240], "workers": 8},
    # The cfg0 wing with stiff hinges: it stays near rest under the flight
    # loads, so it carries the 240-step comparison with the reference.
    "small-stiff": {"steps": 240, "snapshots": [1, 120, 240], "workers": 12},
    # The same wings tethered at two fixed riser junctions: no pilot, the
    # lines are load-path Cable constraints swept in pairs.
    "tiny-tethered": {"steps": 20, "snapshots": [1, 5, 20], "workers": 12},
    "small-tethered": {"steps": 20, "snapshots": [1, 20], "workers": 12},
    # ... at two dynamic riser junctions: every node is dynamic and there is
    # no suspension, so the reference restores the free membrane's momentum.
    "tiny-free": {"steps": 20, "snapshots": [1, 5, 20], "workers": 12},
    "small-free": {"steps": 20, "snapshots": [1, 20], "workers": 12},
}
STRIDE = 101
# Written as digests and strided samples (the full arrays are too large).
DIGESTED = {"export", "small-stiff", "small-tethered", "small-free"}
FRAME_STEPS = {"tiny": 20, "small": 40}
# Wings without a suspension whose Structure checkpoints (SWST around SWBC
# alone) are also kept.
STRUCTURE_ONLY = {"tiny-tethered", "tiny-free", "small-tethered"}


def digest(array: np.ndarray) -> np.ndarray:
    """SHA-256 of the array's little-endian bytes as 4 uint64 words."""
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(array).tobytes()).digest(), dtype=np.uint64).copy()


def run(preset: str) -> None:
    plan = PLAN[preset]
    WORK.mkdir(parents=True, exist_ok=True)
    prefix = WORK / preset
    cmd = [str(DRIVER), preset, str(prefix), str(plan["steps"]), str(plan["workers"]),
           *map(str, plan["snapshots"])]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    assembly = swfx.read(f"{prefix}.assembly.swfx")
    colouring = swfx.read(f"{prefix}.colouring.swfx")
    states = swfx.read(f"{prefix}.states.swfx")
    if preset not in DIGESTED:
        swfx.write(str(HERE / f"{preset}.assembly.swfx"), assembly)
        swfx.write(str(HERE / f"{preset}.colouring.swfx"), colouring)
        keep = {k: v for k, v in states.items() if k != "state_hash"}
        swfx.write(str(HERE / f"{preset}.states.swfx"), keep)
        return
    out = {}
    for name, value in {**assembly, **colouring}.items():
        out[f"digest/{name}"] = digest(value)
        out[f"size/{name}"] = np.array([value.size], np.uint64)
    for step in plan["snapshots"]:
        tag = f"step{step}"
        for key in ("position", "previous", "velocity", "membrane_residual", "membrane_lambda", "dihedral_lambda",
                    "constraint_lambda"):
            value = states[f"{tag}/{key}"]
            out[f"{tag}/{key}/digest"] = digest(value)
            width = 3 if key in ("position", "previous", "velocity", "membrane_lambda") else 1
            rows = value.reshape(-1, width)
            picked = np.arange(0, rows.shape[0], STRIDE)
            out[f"{tag}/{key}/rows"] = picked.astype(np.uint64)
            out[f"{tag}/{key}/samples"] = rows[picked].reshape(-1).copy()
            out[f"{tag}/{key}/maxabs"] = np.array([np.abs(value).max()])
        for key in ("segment_lambda", "payload", "suspension_residual"):
            if f"{tag}/{key}" in states:  # absent without a pilot
                out[f"{tag}/{key}"] = states[f"{tag}/{key}"]
    out["dihedral_colour_order_digest"] = digest(assembly["dihedral_colour_order"])
    swfx.write(str(HERE / f"{preset}.digest.swfx"), out)


def run_frame(preset: str) -> None:
    WORK.mkdir(parents=True, exist_ok=True)
    path = WORK / f"{preset}.frame.swfx"
    cmd = [str(DRIVER), "--frame", preset, str(path), str(FRAME_STEPS[preset])]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    frame = swfx.read(str(path))
    n = frame["position"].size
    for key, field in (("position", "vertex_position"), ("velocity", "vertex_velocity"),
                       ("pending_force", "vertex_pending_force")):
        # The frame's first N vertices are the committed nodes themselves.
        assert np.array_equal(frame[key].view(np.uint64), frame[field][:n].view(np.uint64)), key
        del frame[key]
    frame["node_count"] = np.array([n // 3], np.uint64)
    if preset != "tiny":
        data = frame.pop("frame_bytes")
        frame["frame_bytes_digest"] = digest(data)
        frame["frame_bytes_size"] = np.array([data.size], np.uint64)
    swfx.write(str(HERE / f"{preset}.frame.swfx"), frame)


def run_scene(preset: str) -> None:
    WORK.mkdir(parents=True, exist_ok=True)
    path = WORK / f"{preset}.scene.bin"
    cmd = [str(DRIVER), "--scene", preset, str(path)]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    data = np.frombuffer(path.read_bytes(), np.uint8)
    if preset == "tiny":
        (HERE / "tiny.scene.bin").write_bytes(data.tobytes())
    else:
        swfx.write(str(HERE / f"{preset}.scene.digest.swfx"),
                   {"digest": digest(data), "size": np.array([data.size], np.uint64)})


STRUCTURE_STEPS = 20
# Axial damping (N s/m) given to every suspension segment for the damped run,
# and the compressionStiffnessRatio given to every membrane for the
# compressed run.
LINE_DAMPING = 40.0
COMPRESSION_RATIO = 0.25


def run_structure(preset: str, workers: int, variant: str = "") -> None:
    tag = variant or f"w{workers}"
    path = WORK / f"{preset}.structure.{tag}.swfx"
    cmd = [str(DRIVER), "--structure", preset, str(path), str(STRUCTURE_STEPS), str(workers)]
    if variant == "damped":
        cmd.append(repr(LINE_DAMPING))
    elif variant == "compressed":
        cmd += ["0", repr(COMPRESSION_RATIO)]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    fixture = swfx.read(str(path))
    if not preset.startswith("tiny"):
        for key in ("checkpoint_bytes", "checkpoint_bytes_initial", "suspension_text"):
            if key not in fixture:
                continue
            data = fixture.pop(key)
            fixture[f"{key}_digest"] = digest(data)
            fixture[f"{key}_size"] = np.array([data.size], np.uint64)
    swfx.write(str(HERE / f"{preset}.structure.{tag}.swfx"), fixture)


def main(argv: list[str]) -> int:
    if not DRIVER.exists():
        print(f"{DRIVER} is missing: run make -C oracle ref (needs /root/reference)", file=sys.stderr)
        return 2
    for preset in argv or list(PLAN):
        run(preset)
        if preset in FRAME_STEPS:
            run_frame(preset)
            run_scene(preset)
            for workers in (0, 1):
                run_structure(preset, workers)
            run_structure(preset, 1, "damped")
            run_structure(preset, 1, "compressed")
        if preset in STRUCTURE_ONLY:
            run_structure(preset, 1)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
