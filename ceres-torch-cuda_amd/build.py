"""Build recipe for libceres_mi355x.so (hipcc, gfx950).

Compiles every ``csrc/*.hip`` translation unit in parallel and links them into
``ceres_mi355x/libceres_mi355x.so`` next to the Python package, which loads it
with ctypes.  Run ``python ceres-torch-cuda_amd/build.py`` or call ``build()``.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT_DIR = HERE / "ceres_mi355x"
LIB = OUT_DIR / "libceres_mi355x.so"
OBJ_DIR = HERE / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CT_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-Wno-unknown-cuda-version",
]


# Kernels may not spill: no VGPR spills, no AGPRs used as spill space, no
# scratch.  A register-allocator spill copy placed inside divergent control
# flow only saves the lanes active there, so a value that is live in other
# lanes comes back as garbage when restored under a wider exec mask; kernels
# are kept small enough that the allocator never spills.  The kernels that
# accumulate on the matrix cores may hold their accumulators in AGPRs: each is
# allowed exactly that many (4 blocks of v_mfma_f64_16x16x4f64 results, 4
# doubles each = 32 registers) and no more, so any spill into AGPRs still fails.
MFMA_ACCUMULATOR_AGPRS = {"k_chol_tiles": 32, "k_gram_partials": 32}
RESOURCE_RE = re.compile(r"remark: .*?(Function Name|VGPRs|AGPRs|ScratchSize \[bytes/lane\]|VGPRs Spill|SGPRs Spill): (\S+)")


def check_resources(stderr: str, src: str) -> list[str]:
    problems, name = [], None
    for key, val in RESOURCE_RE.findall(stderr):
        if key == "Function Name":
            name = val
        elif key == "AGPRs" and int(val) > 0:
            allowed = next((v for k, v in MFMA_ACCUMULATOR_AGPRS.items() if name and k in name), 0)
            if int(val) > allowed:
                problems.append(f"{src}:{name} uses {val} AGPRs (allowed {allowed})")
        elif key == "VGPRs Spill" and int(val) > 0:
            problems.append(f"{src}:{name} spills {val} VGPRs")
        elif key.startswith("ScratchSize") and int(val) > 0:
            problems.append(f"{src}:{name} uses {val} bytes of scratch")
    return problems


# Per-file flags.  The batched bundle-adjustment kernel runs a whole LM solve
# in one loop; machine-level loop-invariant code motion would hoist every
# phase's FP64 constants and addresses to the top of that loop, where they
# stay live across all phases and push the kernel into spills.
FILE_FLAGS = {"ba_batch.hip": ["-mllvm", "-disable-machine-licm"]}


def _compile(src: Path, verbose: bool) -> Path:
    obj = OBJ_DIR / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, *FILE_FLAGS.get(src.name, []), "-Rpass-analysis=kernel-resource-usage", "-c", str(src),
           "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    problems = check_resources(res.stderr, src.name)
    if problems:
        raise RuntimeError("kernel resource check failed:\n  " + "\n  ".join(problems))
    return obj


def build(verbose: bool = False) -> Path:
    OBJ_DIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    jobs = min(len(sources), int(os.environ.get("MAX_JOBS", "8")) or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
