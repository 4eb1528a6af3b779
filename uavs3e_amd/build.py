"""Build recipe of the HIP library (gfx950) and the host helpers.

Produces, in uavs3e_amd/lib/:
  libuavs3e_me_hip8.so / libuavs3e_me_hip10.so  -- the C-ABI library, one per
      sample bit depth (the reference fixes `pel` at compile time,
      inc/com_api.h:46-52);
  libme_host8.so / libme_host10.so             -- host-side C: the per-call ME
      driver (calls through the dispatch table) and the synthetic generator.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fvisibility=hidden", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def hip_lib(bd):
    return os.path.join(LIB, "libuavs3e_me_hip%d.so" % bd)


def host_lib(bd):
    return os.path.join(LIB, "libme_host%d.so" % bd)


def _build_hip(bd, extra=()):
    srcs = [os.path.join(CSRC, f) for f in ("percall.hip", "frame.hip")]
    out = hip_lib(bd)
    tmp = out + ".tmp%d" % os.getpid()
    _run([HIPCC] + HIP_FLAGS + list(extra) + ["-DBIT_DEPTH=%d" % bd] + srcs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def _build_host(bd):
    srcs = [os.path.join(CSRC, f) for f in ("me_host.c", "synth.c")]
    srcs = [s for s in srcs if os.path.exists(s)]
    out = host_lib(bd)
    tmp = out + ".tmp%d" % os.getpid()
    _run(["gcc", "-O2", "-std=gnu99", "-fPIC", "-shared", "-ffp-contract=off", "-DBIT_DEPTH=%d" % bd,
          "-I" + os.path.join(ROOT, "include")] + srcs + ["-o", tmp, "-lm", "-lpthread"])
    os.replace(tmp, out)
    return out


def build(jobs=4):
    os.makedirs(LIB, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_build_hip, 8), ex.submit(_build_hip, 10), ex.submit(_build_host, 8),
                ex.submit(_build_host, 10)]
        for f in futs:
            f.result()


if __name__ == "__main__":
    build()
    print("built:", ", ".join(sorted(os.listdir(LIB))))
    sys.exit(0)
