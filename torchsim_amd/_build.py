"""Build the gfx950 HIP library in-tree.

Every HIP translation unit is compiled by ``hipcc --offload-arch=gfx950`` into
``torchsim_amd/csrc/_build/`` and linked into ``torchsim_amd/libtorchsim_hip.so``,
which is what :mod:`torchsim_amd._native` loads. The real-subspace kernel is
compiled once per state tile so the tiles build in parallel.
"""

from __future__ import annotations

import concurrent.futures
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
BUILD = CSRC / "_build"
LIBRARY = PACKAGE / "libtorchsim_hip.so"
ARCH = os.environ.get("TSIM_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are kept out of every kernel on purpose.
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _hipcc() -> str:
    found = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(found).exists():
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return found


def _tiles() -> list[tuple[int, int, int]]:
    text = (CSRC / "epg_real_impl.h").read_text()
    return [
        (int(i), int(sl), int(lanes))
        for i, sl, lanes in re.findall(r"X\((\d+), (\d+), (\d+)\)", text)
    ]


def _units() -> list[tuple[Path, list[str], Path]]:
    units = []
    for source in sorted(CSRC.glob("*.hip")):
        if source.name == "epg_real_tile.hip":
            for ident, sl, lanes in _tiles():
                defines = [
                    f"-DTSIM_TILE_ID={ident}",
                    f"-DTSIM_TILE_SL={sl}",
                    f"-DTSIM_TILE_L={lanes}",
                ]
                units.append((source, defines, BUILD / f"epg_real_tile_{ident}.o"))
        elif source.name == "epg_complex_tile.hip":
            for lanes in (4, 8, 10, 12, 16, 20, 32, 64):
                units.append((source, [f"-DTSIM_COMPLEX_L={lanes}"], BUILD / f"epg_complex_tile_{lanes}.o"))
        elif source.name == "epg_complex_adjoint_tile.hip":
            for lanes in (4, 8, 10, 12, 16, 20, 32, 64):
                units.append((source, [f"-DTSIM_COMPLEX_ADJOINT_L={lanes}"],
                              BUILD / f"epg_complex_adjoint_tile_{lanes}.o"))
        elif source.name == "epg_semisolid_tile.hip":
            for lanes in (4, 8, 10, 12, 16, 20, 32, 64):
                units.append((source, [f"-DTSIM_SEMISOLID_L={lanes}"], BUILD / f"epg_semisolid_tile_{lanes}.o"))
        elif source.name == "epg_complex_fovr_tile.hip":
            for lanes in (4, 8, 10, 12, 16, 20, 32, 64):
                units.append((source, [f"-DTSIM_COMPLEX_FOVR_L={lanes}"], BUILD / f"epg_complex_fovr_tile_{lanes}.o"))
        else:
            units.append((source, [], BUILD / f"{source.stem}.o"))
    return units


def _compile(hipcc: str, source: Path, defines: list[str], target: Path) -> None:
    command = [hipcc, *FLAGS, *defines, "-c", str(source), "-o", str(target)]
    done = subprocess.run(command, capture_output=True, text=True)
    if done.returncode != 0:
        raise RuntimeError(
            f"hipcc failed on {source.name} {' '.join(defines)}:\n{done.stderr[-4000:]}"
        )


_INCLUDE = re.compile(r'^\s*#\s*include\s+"([^"]+)"', re.MULTILINE)


def _dependencies(source: Path) -> set[Path]:
    """The source and every local header it includes, followed recursively."""
    seen: set[Path] = set()
    pending = [source]
    while pending:
        path = pending.pop()
        if path in seen or not path.exists():
            continue
        seen.add(path)
        pending += [(path.parent / name).resolve() for name in _INCLUDE.findall(path.read_text())]
    return seen


def _stale(source: Path, target: Path) -> bool:
    """Whether `target` is older than its source or a header it includes."""
    if not target.exists():
        return True
    return target.stat().st_mtime < max(path.stat().st_mtime for path in _dependencies(source))


def build(jobs: int | None = None, verbose: bool = False, incremental: bool = False) -> Path:
    """Compile every HIP source for gfx950 and link the shared library
    (with `incremental`, only the units older than their sources)."""
    hipcc = _hipcc()
    BUILD.mkdir(parents=True, exist_ok=True)
    units = _units()
    todo = [unit for unit in units if not incremental or _stale(unit[0], unit[2])]
    jobs = jobs or min(16, max(1, len(os.sched_getaffinity(0))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        futures = [pool.submit(_compile, hipcc, *unit) for unit in todo]
        for future in futures:
            future.result()
    objects = [str(target) for _, _, target in units]
    partial = LIBRARY.with_suffix(".so.tmp")
    command = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objects, "-o", str(partial)]
    done = subprocess.run(command, capture_output=True, text=True)
    if done.returncode != 0:
        raise RuntimeError(f"link failed:\n{done.stderr[-4000:]}")
    os.replace(partial, LIBRARY)
    if verbose:
        print(f"built {LIBRARY} from {len(units)} units", file=sys.stderr)
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
