"""Build libglobe_amd.so and the drop-in CLIs (runPCGsolver, runAIREML,
runAIREML_reducedModel) for gfx950.

Every HIP source is compiled with hipcc for gfx950 only, with packed FP32
instructions disabled; objects go to globe_amd/_build/, the shared library and
executables to globe_amd/ and globe_amd/bin/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
BIN = os.path.join(HERE, "bin")
LIB = os.path.join(HERE, "libglobe_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

LIB_SOURCES = ["k_record.hip", "k_geno.hip", "k_grm.hip", "k_grm_i8.hip", "k_mfma.hip", "k_probe.hip", "api.cpp", "aireml.cpp",
               "serve.cpp", "ingest.cpp", "loglik.cpp", "noalloc.hip"]
CLI_COMMON = ["io.cpp"]
CLIS = {"runPCGsolver": "cli_pcgsolver.cpp", "runAIREML": "cli_aireml.cpp",
        "runAIREML_reducedModel": "cli_aireml_reduced.cpp", "makeGRM": "cli_make_grm.cpp"}

DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1] if cmd else ''}")
    return r


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def _compile(src):
    path = os.path.join(CSRC, src)
    out = _obj(src)
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC, *DEVICE_FLAGS, *lang, "-c", path, "-o", out])
    return out


def build(verbose: bool = True, jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    srcs = LIB_SOURCES + [s for s in CLI_COMMON + list(CLIS.values())
                          if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = dict(zip(srcs, ex.map(_compile, srcs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
          *[objs[s] for s in LIB_SOURCES]])
    os.replace(tmp, LIB)
    # host replay of the contraction's launch plans (a checker; no device code)
    rep = os.path.join(HERE, "libglobe_plan_replay.so")
    _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", "-o", rep + ".tmp",
          os.path.join(CSRC, "plan_replay.cpp")])
    os.replace(rep + ".tmp", rep)
    for exe, src in CLIS.items():
        if src not in objs:
            continue
        common = [objs[s] for s in CLI_COMMON if s in objs]
        out = os.path.join(BIN, exe)
        _run([HIPCC, f"--offload-arch={ARCH}", "-o", out + ".tmp", objs[src], *common, LIB,
              f"-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{HERE}"])
        os.replace(out + ".tmp", out)
    if verbose:
        print(f"[globe_amd.build] {LIB}")
    return LIB


if __name__ == "__main__":
    build()
