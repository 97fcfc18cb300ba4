"""Builds libmbs_hip.so (gfx950) in place with hipcc.

Called by __graft_entry__.build(); also runnable as a script.  Every HIP
translation unit is compiled with packed-FP32 instruction generation
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmbs_hip.so")
ARCH = os.environ.get("MBS_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast", "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False, out=None, defines=()):
    src = os.path.join(CSRC, "mbs_engine.hip")
    out = out or LIB
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-D" + d for d in defines] + ["-I", CSRC, src, "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout)
        raise RuntimeError("hipcc failed for %s" % src)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
