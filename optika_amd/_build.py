"""Build liboptk.so (gfx950) in-tree with hipcc.

The library links against the HIP runtime that torch ships (its
``libamdhip64.so`` carries the soname ``libamdhip64.so.7``), so a process
that imports torch first shares one HIP runtime with it.
"""

from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
LIB = HERE / "liboptk.so"

ARCH = os.environ.get("OPTK_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-mcode-object-version=5",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    f"-I{ROOT / 'include'}",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _torch_lib() -> str:
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or spec.origin is None:
        raise RuntimeError("torch is required to locate its HIP runtime")
    return str(pathlib.Path(spec.origin).parent / "lib")


# per-source flags: the float64 trace keeps the reference's rounding sequence
# (no FMA contraction); the float32 form contracts and uses the hardware's
# float32 division and square root, with float32 denormals flushed (its
# coordinates, directions and intensities never come near 1e-38, and the
# flush lets division and square root skip their denormal rescaling)
_SOURCES = {
    "optk_lib.hip": ["-ffp-contract=off"],
    "optk_fast.hip": ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
                      "-fgpu-flush-denormals-to-zero"],
    "optk_batch.hip": ["-ffp-contract=off"],
    "optk_mlp.hip": ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
}


def sources() -> list[pathlib.Path]:
    return [CSRC / name for name in _SOURCES]


def build(verbose: bool = True) -> pathlib.Path:
    """Compile every HIP source into ``optika_amd/liboptk.so``."""
    tl = _torch_lib()
    objs = []
    tmpdir = HERE / "_build"
    tmpdir.mkdir(exist_ok=True)
    procs = []
    for src in sources():
        obj = tmpdir / (src.stem + ".o")
        cmd = [_hipcc(), *COMMON, *_SOURCES[src.name], "-c", str(src), "-o", str(obj)]
        cmd = [c for c in cmd if c != "-shared"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(str(obj))
    for cmd, proc in procs:
        if proc.wait() != 0:
            raise subprocess.CalledProcessError(proc.returncode, cmd)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp),
           f"-L{tl}", f"-Wl,-rpath,{tl}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
