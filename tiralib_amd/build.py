"""Build recipe for libtira_hip.so (gfx950) and the oracle.

    python -m tiralib_amd.build            # library + oracle (+ oracle/_ref when the TIRA tree is present)
    python -m tiralib_amd.build --lib-only

Each HIP/C++ source under tiralib_amd/csrc is compiled to an object in
parallel with hipcc, then linked into tiralib_amd/libtira_hip.so in-tree.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtira_hip.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
REFERENCE_TREE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 anywhere in device code.
# -amdgpu-mfma-vgpr-form: MFMA results in VGPRs, so the BF16 network's
# activations feed the next layer without v_accvgpr_read copies.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", f"-I{INCLUDE}", f"-I{CSRC}",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas",
          "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the device-only target feature: drop that warning
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{err}")
    return err


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE_FLAGS, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *COMMON, "-x", "hip", *DEVICE_FLAGS, "-c", src, "-o", obj]
    return obj, _run(cmd)


def build_lib(verbose=False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    if not srcs:
        raise RuntimeError("no sources under tiralib_amd/csrc")
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for obj, warn in ex.map(_compile, srcs):
            objs.append(obj)
            if verbose and warn:
                print(warn, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle(with_ref=True):
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    if with_ref and os.path.isdir(os.path.join(REFERENCE_TREE, "tira")):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "ref", f"REF={REFERENCE_TREE}"], check=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise SystemExit(f"hipcc not found at {HIPCC}")
    print(build_lib(a.verbose))
    if not a.lib_only:
        build_oracle()
        print(os.path.join(ORACLE_DIR, "libtira_oracle.so"))


if __name__ == "__main__":
    main()
