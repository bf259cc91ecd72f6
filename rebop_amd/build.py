"""Builds the gfx950 HIP library ``rebop_amd/librebop_gpu.so`` in-tree.

Usage: ``python -m rebop_amd.build`` (or ``build()`` from Python).  hipcc
cross-compiles for gfx950 without a GPU present.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "librebop_gpu.so"
SG_LIB = HERE / "librebop_surrogate.so"
SG_SOURCES = ["surrogate.hip"]
ARCH = "gfx950"

SOURCES = ["ssa_kernels.hip", "ssa_lds.hip", "rebop_gpu.cpp", "jit.cpp", "noalloc.hip"]
# headers whose text is embedded in the library for run-time compilation
EMBEDDED_HEADERS = ["rebop_rng.h", "rebop_expr.h", "ssa_layout.h", "ssa_lane.h"]
DELIM = "REBOPSRC"

# -ffp-contract=off: the SSA's FP64 arithmetic must round exactly as the
# reference's separate multiplies and adds do (no fused multiply-add).
# -packed-fp32-ops is disabled for every HIP compile of this project.
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fPIC",
    "-Wall",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the rebop_amd HIP library cannot be built")


def embedded_headers_cpp() -> str:
    """C++ source defining the text of the headers the run-time compiler needs."""
    names, bodies = [], []
    for h in EMBEDDED_HEADERS:
        text = (CSRC / h).read_text()
        if f"){DELIM}\"" in text:
            raise RuntimeError(f"{h} contains the raw-string delimiter")
        names.append(f'"{h}"')
        bodies.append(f'R"{DELIM}({text}){DELIM}"')
    sources = ",\n".join(bodies)
    return (
        "namespace rebop {\n"
        f"extern const int kEmbeddedHeaderCount = {len(names)};\n"
        f"extern const char *const kEmbeddedHeaderNames[] = {{{', '.join(names)}}};\n"
        f"extern const char *const kEmbeddedHeaderSources[] = {{\n{sources}\n}};\n"
        "}\n"
    )


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    """Compile every HIP source into ``librebop_gpu.so`` (always recompiles).

    Each source is compiled to an object in its own hipcc process, several at
    once, then the objects are linked; the surrogate's library is built
    beside them."""
    from concurrent.futures import ThreadPoolExecutor

    tmp = LIB.with_suffix(".so.tmp")
    sg_tmp = SG_LIB.with_suffix(".so.tmp")
    # the surrogate's kernels (dense layers on the matrix cores; contraction is
    # allowed there: they are not part of the bit-exact SSA path)
    sg_flags = [f for f in FLAGS if f != "-ffp-contract=off"]
    jobs = min(8, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as td:
        emb = Path(td) / "embedded_headers.cpp"
        emb.write_text(embedded_headers_cpp())
        srcs = [CSRC / s for s in SOURCES] + [emb]
        objs = [Path(td) / f"{i}_{src.stem}.o" for i, src in enumerate(srcs)]
        cmds = [[hipcc(), *FLAGS, "-c", "-o", str(o), str(src)] for src, o in zip(srcs, objs)]
        cmds.append([hipcc(), *sg_flags, "-shared", "-o", str(sg_tmp),
                     *[str(CSRC / s) for s in SG_SOURCES]])
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            for f in [ex.submit(_run, c, verbose) for c in cmds]:
                f.result()
        _run([hipcc(), *FLAGS, "-shared", "-o", str(tmp), *[str(o) for o in objs], "-lhiprtc"],
             verbose)
    os.replace(tmp, LIB)
    os.replace(sg_tmp, SG_LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
