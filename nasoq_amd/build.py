"""Build the engine's shared library nasoq_amd/liblblg.so.

Host C++ (symbolic phase, C-ABI, drivers) is compiled with hipcc's host
compiler; device code (*.hip) is compiled for gfx950 only.  The library is
built in-tree so that it travels with the repository snapshot.

Usage:  python -m nasoq_amd.build   (or  python nasoq_amd/build.py)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "liblblg.so")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are disabled for all device code on this target.
# MFMA accumulators are kept in VGPRs (the gfx90a+ form): with AGPR
# accumulators the 128 x 128 tile kernels copy all 128 accumulator registers
# out before their epilogues and need 360 registers (one wave per SIMD); in
# VGPR form they fit in 256 (two waves per SIMD), which measured 20% faster.
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-munsafe-fp-atomics", "-mllvm", "-amdgpu-mfma-vgpr-form"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          "-Wno-unused-result"]


def _sources():
    host, dev = [], []
    for d, _, files in os.walk(CSRC):
        for f in sorted(files):
            p = os.path.join(d, f)
            if f.endswith(".cpp"):
                host.append(p)
            elif f.endswith(".hip"):
                dev.append(p)
    return sorted(host), sorted(dev)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    host, dev = _sources()
    objs, cmds = [], []
    for src in host + dev:
        rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
        obj = os.path.join(OBJ, rel + ".o")
        objs.append(obj)
        if src.endswith(".hip"):
            cmd = [HIPCC, "-c", src, "-o", obj] + COMMON + DEVICE_FLAGS
        else:
            cmd = [HIPCC, "-x", "c++", "-c", src, "-o", obj] + COMMON + ["-fopenmp", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
        cmds.append(cmd)
    # compile in parallel
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    link = [HIPCC, "-shared", "-o", OUT + ".tmp"] + objs + DEVICE_FLAGS + ["-fopenmp", "-lpthread"]
    _run(link)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
