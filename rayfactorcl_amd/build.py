"""Build recipe of the native library (gfx950) and the CLI.

``build()`` compiles every HIP / C++ source of the package into
``rayfactorcl_amd/librayfactor_amd.so`` (in-tree, so it travels with the
repository) and the drop-in command-line driver ``rayfactorcl_amd/cli/rayfactor``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "librayfactor_amd.so"
CLI = PKG / "cli" / "rayfactor"
BUILD_DIR = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    # no floating-point contraction: every fused multiply-add is explicit
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    # packed FP32 (v_pk_*_f32) stays off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
# kernels: MFMA results written straight to VGPRs (no accumulator-register copies)
KERNEL_FLAGS = ["-mllvm", "-amdgpu-mfma-vgpr-form"]

SOURCES = ["rf_kernels.hip", "rf_runtime.cpp", "rf_xml.cpp", "rf_report.cpp"]


def _run(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; that note is harmless
        lines = [l for l in text.splitlines() if "unknown target feature" not in l and l.strip()]
        if p.returncode != 0:
            failed.append((c, "\n".join(lines)))
    if failed:
        msg = "\n\n".join(f"$ {' '.join(c)}\n{t}" for c, t in failed)
        raise RuntimeError(f"native build failed:\n{msg}")


def build(verbose: bool = False) -> Path:
    """Compile the library and the CLI; returns the library path."""
    BUILD_DIR.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = BUILD_DIR / (src + ".o")
        objs.append(obj)
        extra = KERNEL_FLAGS if src.endswith(".hip") else []
        cmd = [HIPCC, f"--offload-arch={ARCH}", *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]
        cmds.append(cmd)
    if verbose:
        for c in cmds:
            print(" ".join(c), file=sys.stderr)
    _run(cmds)
    tmp = LIB.with_suffix(".so.tmp")
    _run([[HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]])
    os.replace(tmp, LIB)
    CLI.parent.mkdir(exist_ok=True)
    _run([[HIPCC, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(PKG / "cli" / "rayfactor_main.cpp"),
           "-o", str(CLI), f"-L{PKG}", "-lrayfactor_amd", "-Wl,-rpath,$ORIGIN/.."]])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
