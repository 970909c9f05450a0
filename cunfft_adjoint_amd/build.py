"""Build libcunaf.so (float) and libcunad.so (double) for gfx950 with hipcc.

    python -m cunfft_adjoint_amd.build [--jobs N] [--only f|d] [--incremental]

Both libraries come from the same HIP sources; the double build adds
-DDOUBLE_PRECISION, as the reference's Makefile does for libcunad.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(HERE, "lib")
BUILDDIR = os.path.join(HERE, "_build")

SOURCES = ["cuna_kernels.hip", "cuna_api.hip", "cuna_tune.hip", "cuna_chirp.hip"]
ARCH = os.environ.get("CUNA_OFFLOAD_ARCH", "gfx950")

# packed-fp32 instructions (v_pk_*_f32) are disabled for all device code: some
# of their forms have given wrong results on this hardware without an error
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

LIBS = {"f": ("libcunaf.so", []), "d": ("libcunad.so", ["-DDOUBLE_PRECISION"])}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _newest_input(src: str) -> float:
    """Latest mtime of a source and every header it may include."""
    heads = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    heads += [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h")]
    return max(os.path.getmtime(p) for p in [src, __file__, *heads])


def build(jobs: int = 4, only: str | None = None, verbose: bool = False,
          incremental: bool = False) -> list[str]:
    """Compile both libraries.  incremental=True skips objects newer than
    their source and every header (a development shortcut; the default
    always recompiles)."""
    cc = hipcc()
    os.makedirs(LIBDIR, exist_ok=True)
    tasks = []
    for tag, (libname, defs) in LIBS.items():
        if only and tag != only:
            continue
        bdir = os.path.join(BUILDDIR, tag)
        os.makedirs(bdir, exist_ok=True)
        for src in SOURCES:
            obj = os.path.join(bdir, src.replace(".hip", ".o"))
            spath = os.path.join(CSRC, src)
            if (incremental and os.path.exists(obj)
                    and os.path.getmtime(obj) > _newest_input(spath)):
                continue
            cmd = [cc, *COMMON, *defs, "-c", spath, "-o", obj]
            tasks.append((tag, obj, cmd))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = {ex.submit(_run, cmd): (tag, obj, cmd) for tag, obj, cmd in tasks}
        for f in cf.as_completed(futs):
            out = f.result()
            if verbose and out.strip():
                print(out)
    built = []
    for tag, (libname, defs) in LIBS.items():
        if only and tag != only:
            continue
        objs = [os.path.join(BUILDDIR, tag, s.replace(".hip", ".o")) for s in SOURCES]
        dst = os.path.join(LIBDIR, libname)
        tmp = dst + ".tmp"
        _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
        os.replace(tmp, dst)
        built.append(dst)
    return built


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--only", choices=["f", "d"])
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--incremental", action="store_true",
                    help="skip objects newer than their inputs")
    a = ap.parse_args(argv)
    for p in build(a.jobs, a.only, a.v, a.incremental):
        print(p)


if __name__ == "__main__":
    sys.exit(main())
