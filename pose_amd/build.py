"""Build libsgm_hip.so for gfx950 with hipcc (in-tree, no JIT cache)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/sgm_api.hip", "csrc/sgm_path_first.hip", "csrc/sgm_path_mid.hip",
           "csrc/sgm_path_last.hip", "csrc/sgm_path_pair.hip", "csrc/sgm_batch.hip",
           "csrc/sgm_serve.hip"]
OUTPUT = os.path.join(HERE, "libsgm_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # bit-exact float semantics (no FMA contraction of the reference's double/float steps)
    "-ffp-contract=off",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + err)
    return err


def build(verbose=False):
    """Compile every translation unit (in parallel) and link libsgm_hip.so."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    with ThreadPoolExecutor(len(srcs)) as ex:
        errs = list(ex.map(lambda so: _run([hipcc] + compile_flags + ["-c", so[0], "-o", so[1]]),
                           zip(srcs, objs)))
    errs.append(_run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUTPUT + ".tmp"]
                     + objs + ["-lpthread"]))
    if verbose:
        for e in errs:
            if e:
                print(e, file=sys.stderr)
    os.replace(OUTPUT + ".tmp", OUTPUT)
    return OUTPUT


SYNTH_OUTPUT = os.path.join(HERE, "libsgm_synth.so")


def build_synth():
    """The synthetic-input generator (plain C, OpenMP)."""
    cc = os.environ.get("CC", "gcc")
    _run([cc, "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c11", "-o", SYNTH_OUTPUT + ".tmp",
          os.path.join(HERE, "csrc", "synth.c"), "-lm"])
    os.replace(SYNTH_OUTPUT + ".tmp", SYNTH_OUTPUT)
    return SYNTH_OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_synth())
