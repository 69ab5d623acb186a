"""Copies the reference's own input files used as test fixtures into this
folder, and checks them against the SHA-256 digests recorded here.

  3d_200k.h5  build/geometries_3d/3d_200k.h5: the raw point cloud that
              SnapshotManager::LoadRawPoints reads (snapshotmanager.cpp:20-113),
              202,329 Poisson-disk points with grain ids and a volume attribute
  p1.json     input_data/p1.json: a run configuration that names that cloud
              (SimParams3D::ParseFile, parameters_sim_3d.cpp:88-160)

These are input data of the reference, not code: the tests read them with
this repository's own HDF5 reader and JSON parameter parser.
Run: python tests/golden/fetch_reference_inputs.py [reference root]
"""
from __future__ import annotations

import hashlib
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
FILES = {
    "3d_200k.h5": "build/geometries_3d/3d_200k.h5",
    "p1.json": "input_data/p1.json",
}
SHA256 = {
    "3d_200k.h5": "947f9eb07e022ef38d33cb960979a47fb3c3f1c8603d20f40a432cbd49deba29",
    "p1.json": "7579c4c9bf1e66034e1c64eea82a64260ab894fe90451d5ba8cd74b8187c25a9",
}


def digest(path: Path) -> str:
    return hashlib.sha256(path.read_bytes()).hexdigest()


def main(ref_root: str = "/root/reference") -> None:
    for name, rel in FILES.items():
        src = Path(ref_root) / rel
        dst = HERE / name
        shutil.copyfile(src, dst)
        print(name, digest(dst))


def verify() -> None:
    for name, want in SHA256.items():
        got = digest(HERE / name)
        if got != want:
            raise SystemExit(f"{name}: digest {got} != recorded {want}")


if __name__ == "__main__":
    main(*sys.argv[1:])
    verify()
