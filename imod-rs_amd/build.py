"""Build libtiltgpu.so (gfx950 HIP kernels + C ABI + host utilities) and the
`tilt` drop-in executable, in-tree under imod-rs_amd/_build/.

    python imod-rs_amd/build.py [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_SOURCES = ["csrc/kernels_bp.hip", "csrc/kernels_filter.hip", "csrc/kernels_fp.hip",
               "csrc/kernels_misc.hip", "csrc/session.hip", "csrc/server.hip", "csrc/ctf.hip",
               "csrc/noalloc.hip"]
HOST_SOURCES = ["host/textio.cpp", "host/synth.cpp", "host/mrc.cpp", "host/localalign.cpp",
                "csrc/tuning.cpp"]
# drop-in executables: name -> sources (linked against libtiltgpu)
CLIS = {"tilt": ["host/tilt_main.cpp"], "ctfphaseflip": ["host/ctfphaseflip_main.cpp"]}

# -ffp-contract=off: coordinate arithmetic must round like the reference (which
# never fuses); fused multiply-adds are written explicitly where wanted.
# -packed-fp32-ops off: no v_pk_*_f32 instructions are generated.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
             "-I" + os.path.join(REPO, "include")]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
              "-I" + os.path.join(REPO, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def lib_path() -> str:
    return os.path.join(OUT, "libtiltgpu.so")


def cli_path(name: str = "tilt") -> str:
    return os.path.join(OUT, name)


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    objs = []
    steps = []
    for src in HIP_SOURCES:
        path = os.path.join(HERE, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        steps.append([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        path = os.path.join(HERE, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        steps.append([CXX, *HOST_FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for r in ex.map(_run, steps):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    tmp = lib_path() + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp,
          "-lpthread"])
    os.replace(tmp, lib_path())
    cli_steps = []
    for name, srcs in CLIS.items():
        paths = [os.path.join(HERE, s) for s in srcs if os.path.exists(os.path.join(HERE, s))]
        if paths:
            cli_steps.append((name, [CXX, *HOST_FLAGS, *paths, "-o", cli_path(name) + ".tmp",
                                     "-L" + OUT, "-ltiltgpu", "-Wl,-rpath,$ORIGIN", "-lpthread"]))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, [cmd for _, cmd in cli_steps]))
    for name, _ in cli_steps:
        os.replace(cli_path(name) + ".tmp", cli_path(name))
    return lib_path()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))


if __name__ == "__main__":
    main()
