"""Build ``libpic_hip.so`` in-tree with hipcc for gfx950.

Flags: ``-ffp-contract=off`` keeps every product and sum separately rounded
(the reference's Numba kernels are compiled without fast-math, and this is what
makes an FP64 sweep bit-identical to it); ``-packed-fp32-ops`` is disabled so
no packed FP32 instructions are emitted.
"""

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent.parent
SOURCES = [PKG.parent / "csrc" / "pic_hip.hip"]
DEPENDS = SOURCES + [PKG.parent / "csrc" / "pic_kernels.hpp", PKG.parent / "csrc" / "pic_adjoint.hpp",
                     PKG.parent / "csrc" / "pic_sweep2.hpp",
                     ROOT / "include" / "pic_hip.h"]
OUTPUT = PKG / "libpic_hip.so"
ARCH = os.environ.get("PIC_HIP_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libpic_hip.so)")


def command(output: Path = OUTPUT):
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-result", "-Wno-unused-value",
        f"-I{ROOT / 'include'}", f"-I{PKG.parent / 'csrc'}",
        "-o", str(output), *map(str, SOURCES),
    ]


PROBE_SOURCE = PKG.parent / "csrc" / "hbm_copy_probe.hip"
PROBE_OUTPUT = PKG / "hbm_copy_probe"


def build_probe(verbose: bool = False) -> Path:
    """The copy-calibration executable behind the HBM traffic figures (a
    measuring instrument, not part of the library): built only on request,
    ``python -m pulsed_ion_chamber_hip._build probe``."""
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-o", str(PROBE_OUTPUT), str(PROBE_SOURCE)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stderr}")
    return PROBE_OUTPUT


def build(verbose: bool = False) -> Path:
    """Compile unconditionally (a few seconds to a minute)."""
    tmp = OUTPUT.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{err}")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    if sys.argv[1:] == ["probe"]:
        print(build_probe(verbose=True))
    else:
        print(build(verbose=True))
    sys.exit(0)
