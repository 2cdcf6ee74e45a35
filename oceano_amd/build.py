"""Build liboceano_dg.so in-tree with hipcc for gfx950.

The library holds the host set-up (mesh builder, test-case functions, operator set-up) and
the hand-written HIP kernels; everything is compiled by hipcc into one shared object next
to this file, so it travels with the repository snapshot.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liboceano_dg.so")

# (source, extra flags, object name): the per-degree stage source is compiled once per degree
SOURCES = [("mesh.cpp", [], "mesh"), ("icbc.cpp", [], "icbc"),
           ("dg_operator.hip", [], "dg_operator"), ("dg_rk.hip", [], "dg_rk"),
           ("mlp.hip", [], "mlp"), ("dg_stage_multi.hip", [], "dg_stage_multi")] + \
          [("dg_stage_degree.hip", ["-DOC_DEGREE=%d" % p], "dg_stage_p%d" % p)
           for p in (1, 2, 3, 4)]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    "--offload-arch=gfx950",
    # packed FP32 ops are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP32 division and square root as the hardware's v_rcp_f32 / v_sqrt_f32 (1 ulp)
    # instead of the correctly rounded multi-instruction sequences: only the FP32 arithmetic
    # of the mixed-precision path and the closure network is affected (FP64 code is not)
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
    "-Wno-unused-command-line-argument",
]


def _headers():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    hs.append(os.path.join(ROOT, "include", "oceano_dg.h"))
    hs.append(os.path.join(ROOT, "include", "oceano_mlp.h"))
    return max(os.path.getmtime(h) for h in hs)


def build(verbose=False, extra=(), incremental=False):
    """Compile every source (incremental=True skips objects newer than their inputs)."""
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    procs = []
    newest_header = _headers()
    for src, flags, name in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(HERE, "_build", name + ".o")
        if (incremental and os.path.exists(obj)
                and os.path.getmtime(obj) > max(os.path.getmtime(path), newest_header)):
            objs.append(obj)
            continue
        cmd = [HIPCC, *COMMON, *flags, *extra, "-x", "hip", "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_cpp_test(verbose)
    return LIB


CPP_TEST = os.path.join(ROOT, "tests", "cpp", "test_adapter")


def build_cpp_test(verbose=False):
    """The C++ drop-in test links the library like a reference-side caller would."""
    src = CPP_TEST + ".cpp"
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-I" + os.path.join(ROOT, "include"),
           src, "-L" + HERE, "-loceano_dg", "-Wl,-rpath,$ORIGIN/../../oceano_amd", "-o", CPP_TEST]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv, incremental="-i" in sys.argv)
    print(LIB)
