"""Build the HIP library of the package in-tree with hipcc (gfx950 only).

    python -m <package>.build      or      __graft_entry__.build()

Produces libvesicle_ls.so next to this file.  -ffp-contract=off: every fused
multiply-add in the kernels is explicit, so results are one fixed IEEE
sequence (the CPU oracle states the same one).  The flag
`-Xclang -target-feature -Xclang -packed-fp32-ops` keeps the compiler from
emitting packed FP32 instructions; the host pass warns that it does not know
the feature, which is harmless.
"""
import concurrent.futures
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libvesicle_ls.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ["vls_kernels.hip", "vls_batch.hip", "vls_adjoint.hip", "vls_variants.hip", "vls_relax.hip", "vls_api.hip",
           "vls_probe.hip"]
CXX_SOURCES = ["vls_matio.cpp"]  # host-only C++ (the MAT v5 reader/writer)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    jobs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        jobs.append((obj, [HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + FLAGS))
    for src in CXX_SOURCES:
        obj = os.path.join(CSRC, src.replace(".cpp", ".o"))
        jobs.append((obj, ["g++", "-c", os.path.join(CSRC, src), "-o", obj, "-O2", "-std=c++17", "-fPIC", "-Wall"]))
    if verbose:
        for _, cmd in jobs:
            print(" ".join(cmd), flush=True)
    # the translation units are independent: compile them side by side
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(subprocess.run, cmd, check=True) for _, cmd in jobs]:
            f.result()
    objs = [obj for obj, _ in jobs]
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=gfx950", "-fPIC", "-Wl,-rpath,/opt/rocm/lib",
                                                  "-L/opt/rocm/lib", "-lrocfft", "-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    for o in objs:
        os.remove(o)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
