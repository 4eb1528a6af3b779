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


# Translation units of the HIP library: (source, extra defines).  The
# integer-search and sub-pel kernels are split by instance so that the
# expensive device compiles run in parallel.
HIP_UNITS = [("percall.hip", []), ("refs.hip", []), ("frame.hip", []), ("server.hip", []), ("mc.hip", []), ("affine_me.hip", []), ("tz.hip", ["-DME_TZ_PART=0"]), ("tz.hip", ["-DME_TZ_PART=1"]), ("tz.hip", ["-DME_TZ_PART=2"]),
             ("intsearch.hip", ["-DME_IS_WAVES=1"]), ("intsearch.hip", ["-DME_IS_WAVES=2"]),
             ("intsearch.hip", ["-DME_IS_WAVES=4"]), ("intsearch.hip", ["-DME_IS_WAVES=8"]),
             ("subpel.hip", [])]
BUILD_DIR = os.path.join(ROOT, "_build")
# Per-source compiler options.  subpel.hip: the matrix-core sub-pel kernel's
# products stay in ordinary VGPRs (read straight by the VALU sums; in AGPRs
# every coefficient costs a v_accvgpr_read, and the kernel needs 220
# registers instead of 152).
UNIT_FLAGS = {"subpel.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _obj_name(bd, src, defs):
    tag = "".join("_" + d.split("=")[-1] for d in defs)
    return os.path.join(BUILD_DIR, "bd%d_%s%s.o" % (bd, os.path.splitext(src)[0], tag))


def _compile_unit(bd, src, defs, extra=()):
    obj = _obj_name(bd, src, defs)
    flags = [f for f in HIP_FLAGS if f != "-shared"]
    _run([HIPCC] + flags + list(extra) + defs + ["-DBIT_DEPTH=%d" % bd, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def _link_hip(bd, objs):
    out = hip_lib(bd)
    tmp = out + ".tmp%d" % os.getpid()
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950"] + objs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def _build_host(bd):
    """Host C: the per-call ME driver, the synthetic generator and the
    serving load (which calls the HIP library's C ABI, so it links it)."""
    srcs = [os.path.join(CSRC, f) for f in ("me_host.c", "synth.c", "serve_load.c")]
    out = host_lib(bd)
    tmp = out + ".tmp%d" % os.getpid()
    _run(["gcc", "-O2", "-std=gnu99", "-fPIC", "-shared", "-ffp-contract=off", "-DBIT_DEPTH=%d" % bd,
          "-I" + os.path.join(ROOT, "include")] + srcs +
         ["-o", tmp, "-L" + LIB, "-luavs3e_me_hip%d" % bd, "-Wl,-rpath,$ORIGIN", "-lm", "-lpthread"])
    os.replace(tmp, out)
    return out


def peak_lib():
    return os.path.join(LIB, "libme_peak.so")


def _build_peak():
    """The VALU issue-rate microbenchmarks bench.py --valu-peak runs (a
    library of its own: not part of the ME C ABI)."""
    out = peak_lib()
    tmp = out + ".tmp%d" % os.getpid()
    _run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, "peak.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build(jobs=None, bit_depths=(8, 10)):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = jobs or max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        units = {bd: [ex.submit(_compile_unit, bd, src, defs, UNIT_FLAGS.get(src, ())) for src, defs in HIP_UNITS]
                 for bd in bit_depths}
        links = [ex.submit(_link_hip, bd, [f.result() for f in units[bd]]) for bd in bit_depths]
        for f in links:
            f.result()
        peak = ex.submit(_build_peak)
        hosts = [ex.submit(_build_host, bd) for bd in bit_depths]   # links the HIP library
        for f in hosts:
            f.result()
        peak.result()


if __name__ == "__main__":
    build()
    print("built:", ", ".join(sorted(os.listdir(LIB))))
    sys.exit(0)
