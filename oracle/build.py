"""Build recipe for the test-side checkers (TEST INFRASTRUCTURE ONLY).

* ``build_oracle()`` compiles the C restatement ``oracle/me_oracle.c`` once per
  bit depth into ``oracle/lib/liboracle{8,10}.so``.
* ``build_ref()`` compiles the reference uavs3e from its own sources where they
  lie under ``/root/reference`` (all of ``src/*.c`` except ``uavs3e.c``,
  ``me.c`` and ``inter.c``, plus ``src/sse`` and ``src/avx2``, with the reference's own flags
  from src/CMakeLists.txt:13-15) together with ``oracle/ref_harness.c`` into
  ``oracle/_ref/libuavs3e_ref{8,10}.so``.  ``uavs3e.c`` is left out because it
  needs the generated ``version.h``; ``me.c`` and ``inter.c`` are compiled
  inside the harness.
  The reference selects its sample type with a hard-coded ``#define BIT_DEPTH 8``
  (inc/com_api.h:46) and documents editing that line to switch (README.md:15-16),
  so the 10-bit library is compiled from a scratch copy of ``src/`` and ``inc/``
  under /tmp with that single line changed; the copy is removed afterwards and
  nothing of the reference's sources is written into this repository.
  The reference's own build system (CMake, version.sh) is not run.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.environ.get("UAVS3E_REF_DIR", "/root/reference")
LIB_DIR = os.path.join(HERE, "lib")
REF_OUT = os.path.join(HERE, "_ref")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build_oracle():
    os.makedirs(LIB_DIR, exist_ok=True)
    src = os.path.join(HERE, "me_oracle.c")
    for bd in (8, 10):
        out = os.path.join(LIB_DIR, "liboracle%d.so" % bd)
        _run(["gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-ffp-contract=off",
              "-DBIT_DEPTH=%d" % bd, src, "-o", out, "-lm"])


def _ref_sources(root):
    src = os.path.join(root, "src")
    core = sorted(os.path.join(src, f) for f in os.listdir(src)
                  if f.endswith(".c") and f not in ("uavs3e.c", "me.c", "inter.c"))
    sse = sorted(os.path.join(src, "sse", f) for f in os.listdir(os.path.join(src, "sse")) if f.endswith(".c"))
    avx2 = sorted(os.path.join(src, "avx2", f) for f in os.listdir(os.path.join(src, "avx2")) if f.endswith(".c"))
    return core, sse, avx2


def _build_ref_lib(root, bd, jobs):
    objdir = os.path.join(REF_OUT, "obj%d" % bd)
    os.makedirs(objdir, exist_ok=True)
    core, sse, avx2 = _ref_sources(root)
    inc = ["-I" + os.path.join(root, "inc"), "-I" + os.path.join(root, "src")]
    base = ["gcc", "-c", "-fPIC", "-std=c99", "-O3", "-w"] + inc
    cmds = []
    objs = []
    for group, extra, tag in ((core, [], "c"), (sse, ["-msse4.2"], "sse"), (avx2, ["-mavx2"], "avx2")):
        for f in group:
            o = os.path.join(objdir, "%s_%s.o" % (tag, os.path.basename(f)[:-2]))
            cmds.append(base + extra + [f, "-o", o])
            objs.append(o)
    harness_o = os.path.join(objdir, "ref_harness.o")
    cmds.append(["gcc", "-c", "-fPIC", "-std=gnu99", "-O3", "-w"] + inc +
                [os.path.join(HERE, "ref_harness.c"), "-o", harness_o])
    objs.append(harness_o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(REF_OUT, "libuavs3e_ref%d.so" % bd)
    _run(["gcc", "-shared", "-o", out] + objs + ["-lm", "-lpthread"])
    shutil.rmtree(objdir, ignore_errors=True)
    return out


def build_ref(jobs=8):
    """Compile the reference into oracle/_ref/ when its tree is present."""
    if not os.path.isdir(os.path.join(REF_DIR, "src")):
        return False
    os.makedirs(REF_OUT, exist_ok=True)
    _build_ref_lib(REF_DIR, 8, jobs)
    scratch = "/tmp/uavs3e_ref10_%d" % os.getpid()
    try:
        shutil.rmtree(scratch, ignore_errors=True)
        shutil.copytree(os.path.join(REF_DIR, "src"), os.path.join(scratch, "src"))
        shutil.copytree(os.path.join(REF_DIR, "inc"), os.path.join(scratch, "inc"))
        api = os.path.join(scratch, "inc", "com_api.h")
        with open(api) as f:
            text = f.read()
        if "#define BIT_DEPTH 8" not in text:
            raise RuntimeError("unexpected com_api.h: BIT_DEPTH line not found")
        with open(api, "w") as f:
            f.write(text.replace("#define BIT_DEPTH 8", "#define BIT_DEPTH 10", 1))
        _build_ref_lib(scratch, 10, jobs)
    finally:
        shutil.rmtree(scratch, ignore_errors=True)
    return True


if __name__ == "__main__":
    build_oracle()
    ok = build_ref()
    print("oracle built; reference %s" % ("built" if ok else "absent (skipped)"))
    sys.exit(0)
