"""Builds libgoqu_hip.so (HIP kernels + C ABI, gfx950) in-tree, and the CPU oracle.

`python -m goqu_amd.build` compiles every source under goqu_amd/csrc with hipcc for
gfx950 and links goqu_amd/libgoqu_hip.so.  The built library travels with the source tree;
nothing is JIT-compiled at import time.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libgoqu_hip.so"
ORACLE_DIR = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GOQU_OFFLOAD_ARCH", "gfx950")
# packed-FP32 instructions stay disabled in every HIP object.  Uniform branches are not
# structurized: the stage kernel's instruction dispatch is all uniform, and structurizing it
# makes the compiler copy the 16 register-resident amplitudes at every join of the dispatch.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


# sv_stage.hip holds twelve large kernel instantiations: each is compiled in an object of its own
# (GOQU_STAGE_KERNEL = 1..7), next to the object with the file's host code, so they build in
# parallel.
STAGE_SPLIT = "sv_stage.hip"
STAGE_KERNELS = 12


def _jobs() -> list[tuple[Path, list[str], str]]:
    jobs = []
    for src in sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp"))):
        jobs.append((src, [], src.name + ".o"))
        if src.name == STAGE_SPLIT:
            for k in range(1, STAGE_KERNELS + 1):
                jobs.append((src, [f"-DGOQU_STAGE_KERNEL={k}"], f"{src.name}.k{k}.o"))
    # the longest compiles first
    return sorted(jobs, key=lambda j: (j[0].name != STAGE_SPLIT or not j[1], j[0].name))


# kernels whose speed depends on their waves per SIMD (k_stage, k_stage_dm: four, at most 128
# VGPRs; the lean paired kernels: three, at most 168; the QAOA tile kernels: forward five, reverse
# three); the build reports a drop in their occupancy
OCCUPANCY_WATCH = {"_ZN4goqu7k_stageIdE": 4, "_ZN4goqu7k_stageIfE": 4, "_ZN4goqu12k_stage_pairIdE": 3,
                   "_ZN4goqu12k_stage_pairIfE": 3, "_ZN4goqu15k_stage_pair_dmIdE": 3, "_ZN4goqu10k_stage_dmIdE": 4,
                   "_ZN4goqu11k_qaoa_passILb0EE": 5, "_ZN4goqu11k_qaoa_passILb1EE": 3}
WATCHED_SOURCES = {"qaoa.hip"}


def _check_occupancy(src: Path, remarks: str) -> None:
    current = None
    for line in remarks.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            current = m.group(1)
            continue
        m = re.search(r"remark:\s+Occupancy \[waves/SIMD\]: (\d+)", line)
        if m and current:
            for prefix, want in OCCUPANCY_WATCH.items():
                if current.startswith(prefix) and int(m.group(1)) < want:
                    print(f"warning: {src.name}: {current} runs {m.group(1)} waves/SIMD (expected {want})",
                          file=sys.stderr)


def _compile(job: tuple[Path, list[str], str]) -> Path:
    src, defs, name = job
    obj = OBJ / name
    extra = ["-Rpass-analysis=kernel-resource-usage"] if defs or src.name in WATCHED_SOURCES else []
    cmd = [HIPCC, *HIP_FLAGS, *defs, *extra, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    if extra:
        _check_occupancy(src, r.stderr)
    return obj


def _stale(job: tuple[Path, list[str], str]) -> bool:
    """An object is stale when it is missing or older than its source or any header in csrc/ or
    include/."""
    obj = OBJ / job[2]
    if not obj.exists():
        return True
    deps = [job[0], *CSRC.glob("*.h"), *(ROOT / "include").glob("*.h")]
    return any(d.stat().st_mtime > obj.stat().st_mtime for d in deps)


def build_library(jobs: int | None = None, incremental: bool = False) -> Path:
    """Compiles every source (incremental=True: only objects older than their inputs) and links
    the library."""
    OBJ.mkdir(exist_ok=True)
    work = _jobs()
    if incremental:
        todo = [j for j in work if _stale(j)]
        done = [OBJ / j[2] for j in work if not _stale(j)]
    else:
        todo, done = work, []
    jobs = jobs or max(1, min(len(todo), os.cpu_count() or 4, 8))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, todo)) + done
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return ORACLE_DIR / "_build" / "liboracle.so"


def main() -> int:
    lib = build_library(incremental="--incremental" in sys.argv)
    orc = build_oracle()
    print(f"built {lib}\nbuilt {orc}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
