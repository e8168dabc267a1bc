"""Build libcupymag_hip.so for gfx950 in-tree (hipcc for the HIP sources,
g++ for the host-only assembler), so the library travels with the source
tree.  Usage: ``python -m cupymag_amd.build``."""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "lib" / "libcupymag_hip.so"
OBJ = ROOT / "_build" / "obj"
ARCH = os.environ.get("CPM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: keep the compiler from forming v_pk_*_f32 instructions
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics", f"-I{ROOT / 'include'}", f"-I{CSRC}",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-pthread", f"-I{ROOT / 'include'}", f"-I{CSRC}",
             "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    OBJ.mkdir(parents=True, exist_ok=True)
    OUT.parent.mkdir(parents=True, exist_ok=True)
    jobs = []
    for src in sorted(CSRC.glob("*.hip")):
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(OBJ / (src.stem + ".o"))])
    for src in sorted(CSRC.glob("*.cpp")):
        jobs.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(OBJ / (src.stem + ".o"))])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    objs = [str(OBJ / (Path(j[-3]).stem + ".o")) for j in jobs]
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp), "-pthread"])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
