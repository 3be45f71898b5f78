"""Build libtonto_gpu.so for gfx950 in-tree.

Every angular-momentum class pair listed in csrc/gen/classes.inc is compiled as its own
translation unit (csrc/jk_class.hip with -DTGPU_LP.. defines) so that the classes build
in parallel; csrc/tgpu_host.hip holds the C ABI and the small kernels.  Objects go to
tonto_amd/_build/, the library to tonto_amd/libtonto_gpu.so.

    python -m tonto_amd.build [--jobs N] [--incremental]
"""
import argparse
import os
import re
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
# another copy of the kernel sources (variant builds)
CSRC = os.environ.get("TONTO_GPU_CSRC") or CSRC
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtonto_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
         # packed FP32 instructions stay off ("toolchain rules")
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def class_list():
    txt = open(os.path.join(CSRC, "gen", "classes.inc")).read()
    block = txt.split("#define TGPU_FOR_EACH_CLASS(M)")[1].split("/*")[0]
    return [tuple(int(x) for x in m) for m in re.findall(r"M\((\d), (\d), (\d), (\d), (\d)\)", block)]


def _deps(extra):
    return [os.path.join(CSRC, f) for f in ("tgpu_device.h", "rys_table.inc")] + \
        [os.path.join(CSRC, "gen", "classes.inc"), os.path.join(HERE, "..", "include", "tonto_gpu.h")] + extra


def _stale(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src, obj, defines, deps, incremental, log):
    if incremental and not _stale(obj, deps):
        return 0.0
    t0 = time.time()
    cmd = [HIPCC] + FLAGS + [f"-D{d}" for d in defines] + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {obj}:\n{r.stderr[-4000:]}")
    dt = time.time() - t0
    log(f"  {os.path.basename(obj)} {dt:.1f}s")
    return dt


def build(jobs=None, incremental=False, verbose=True, defines=(), lib=None, build_dir=None):
    """Compile every translation unit for gfx950 and link the library (default
    tonto_amd/libtonto_gpu.so).  defines: extra -D macros for all units (variant builds
    go to their own lib / build_dir)."""
    log = print if verbose else (lambda *a: None)
    global BUILD
    lib = lib or LIB
    if build_dir:
        BUILD = build_dir
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 8)
    tasks = []
    for (lp, lq, lr, ls, heavy) in class_list():
        name = f"{lp}{lq}{lr}{ls}"
        gen = os.path.join(CSRC, "gen", f"eri_{name}.inc")
        tasks.append((os.path.join(CSRC, "jk_class.hip"), os.path.join(BUILD, f"jk_{name}.o"),
                      [f"TGPU_LP={lp}", f"TGPU_LQ={lq}", f"TGPU_LR={lr}", f"TGPU_LS={ls}",
                       f"TGPU_HEAVY={heavy}"], _deps([gen, os.path.join(CSRC, "jk_class.hip")])))
    tasks.append((os.path.join(CSRC, "tgpu_host.hip"), os.path.join(BUILD, "host.o"), [],
                  _deps([os.path.join(CSRC, "tgpu_host.hip")])))
    tasks.append((os.path.join(CSRC, "mlp.hip"), os.path.join(BUILD, "mlp.o"), [],
                  [os.path.join(CSRC, "mlp.hip"), os.path.join(HERE, "..", "include", "tonto_mlp.h")]))
    # largest classes first so the tail is short
    tasks.sort(key=lambda t: -os.path.getsize(next((d for d in t[3] if "eri_" in d), t[0])))
    t0 = time.time()
    tasks = [(src, obj, list(dfs) + list(defines), deps) for (src, obj, dfs, deps) in tasks]
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda t: _compile(*t, incremental, log), tasks))
    objs = [t[1] for t in tasks]
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", lib] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    log(f"built {lib} in {time.time() - t0:.1f}s")
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--define", action="append", default=[], help="extra NAME=VALUE for every unit")
    ap.add_argument("--lib", default=None, help="output library (default tonto_amd/libtonto_gpu.so)")
    ap.add_argument("--build-dir", default=None)
    a = ap.parse_args()
    try:
        build(a.jobs, a.incremental, defines=a.define, lib=a.lib, build_dir=a.build_dir)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
