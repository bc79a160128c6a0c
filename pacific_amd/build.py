"""Build recipe of the in-tree HIP library ``pacific_amd/libgrains_gpu.so``.

Compiles every source under ``pacific_amd/csrc`` for gfx950 with hipcc.  The
FP64 path must evaluate expressions in source order (``-ffp-contract=off``);
packed FP32 instructions are disabled for all HIP code.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgrains_gpu.so")
CLI = os.path.join(HERE, "grains_gpu")
EXAMPLE = os.path.join(HERE, "fs_plugin_example")

SOURCES = ["dem_kernels.hip", "dem_forces_sph64.hip", "dem_forces_sph32.hip", "dem_forces_shp64.hip",
           "dem_forces_shp32.hip", "dem_narrow.hip", "dem_noalloc.hip", "dem_batch.hip", "dem_adjoint.hip", "dem_system.cpp", "dem_shapes.cpp",
           "grains_xml.cpp", "grains_basilisk.cpp", "grains_output.cpp", "grains_reload.cpp",
           "grains_insert.cpp"]
ARCH = os.environ.get("PACIFIC_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    # one compiler process per source (the kernel files take minutes each),
    # objects in a scratch folder outside the tree, then one link
    with tempfile.TemporaryDirectory(prefix="grains_gpu_build_") as bdir:
        objs = [os.path.join(bdir, os.path.basename(src) + ".o") for src in srcs]
        jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            outs = list(pool.map(lambda so: _run([hipcc()] + HIPCC_FLAGS + ["-c", "-o", so[1], so[0]]),
                                 zip(srcs, objs)))
        out = "\n".join(o for o in outs if o.strip())
        out += _run([hipcc()] + HIPCC_FLAGS + ["-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "grains_cli.cpp")
    if os.path.exists(cli_src):
        tmpc = CLI + ".tmp"
        out += _run([hipcc()] + HIPCC_FLAGS + ["-o", tmpc, cli_src, "-L" + HERE, "-lgrains_gpu",
                                               "-Wl,-rpath,$ORIGIN"])
        os.replace(tmpc, CLI)
    ex_src = os.path.join(CSRC, "fs_plugin_example.cpp")
    if os.path.exists(ex_src):
        tmpe = EXAMPLE + ".tmp"
        out += _run([hipcc()] + HIPCC_FLAGS + ["-o", tmpe, ex_src, "-L" + HERE, "-lgrains_gpu",
                                               "-Wl,-rpath,$ORIGIN"])
        os.replace(tmpe, EXAMPLE)
    if verbose and out.strip():
        print(out)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
    sys.exit(0)
