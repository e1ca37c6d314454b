"""Compile the HIP sources of cubismnova_amd into libcubismnova_amd.so (gfx950).

Every .hip file under csrc/ is compiled to an object with hipcc and linked
into one shared library next to this file.  Run as a module
(``python -m cubismnova_amd.build``) or call :func:`build`.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcubismnova_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    "--offload-arch=" + ARCH,
    # operation order of the reference must survive: no FMA contraction
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def scratch_kernels(remarks):
    """Kernels whose per-lane scratch (private) memory is not zero, from the
    compiler's kernel-resource-usage remarks: [(kernel, bytes per lane)]."""
    import re
    out, name = [], None
    for line in remarks.splitlines():
        m = re.search(r"remark:\s+Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"remark:\s+ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name and int(m.group(1)) > 0:
            out.append((name, int(m.group(1))))
    return out


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + COMMON + ["-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    # no kernel of the library spills registers to scratch memory: a tile
    # shape that would is excluded by its shape rule, and the build says so
    # if one ever does
    spills = scratch_kernels(r.stderr)
    if spills:
        raise RuntimeError("%s: kernels using scratch memory (bytes per lane): %s"
                           % (src, ", ".join("%s %d" % k for k in spills)))
    err = "\n".join(l for l in r.stderr.splitlines() if "remark:" not in l)
    return obj, err


def build(verbose=False, jobs=None):
    """Compile all sources and link the shared library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for o, err in results:
            warn = [l for l in err.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
            if warn:
                print("\n".join(warn), file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
