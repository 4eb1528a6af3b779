"""The per-call adapter of INTEGRATION.md section 1 (src/hip_me.c, which
installs the HIP entries into uavs3e's own funs_handle_t) compiles against
the reference's headers with incompatible pointer types as errors, at both
bit depths: every slot assignment matches the reference's signature without
a cast.  Compile only (no GPU, nothing is run); skipped where the reference
tree is absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("UAVS3E_REF_DIR", "/root/reference")


def _adapter_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", text, flags=re.S)
    src = next(b for b in blocks if "void uavs3e_funs_init_hip(void)" in b)
    return src


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "inc")), reason="reference tree absent")
@pytest.mark.parametrize("bd", [8, 10])
def test_adapter_compiles_against_reference_headers(bd, tmp_path):
    inc = os.path.join(REF, "inc")
    if bd == 10:
        # the reference selects its sample type in com_api.h (README.md:15-16):
        # a scratch copy of the headers with that one line changed
        inc = str(tmp_path / "inc")
        shutil.copytree(os.path.join(REF, "inc"), inc)
        api = os.path.join(inc, "com_api.h")
        text = open(api).read()
        assert "#define BIT_DEPTH 8" in text
        open(api, "w").write(text.replace("#define BIT_DEPTH 8", "#define BIT_DEPTH 10", 1))
    src = tmp_path / "hip_me.c"
    src.write_text(_adapter_source())
    r = subprocess.run(["gcc", "-std=gnu99", "-c", "-Werror=incompatible-pointer-types", "-Wall", "-Wno-unused",
                        "-I" + inc, "-I" + os.path.join(REF, "src"), "-I" + os.path.join(ROOT, "include"),
                        "-DBIT_DEPTH=%d" % bd, str(src), "-o", str(tmp_path / "hip_me.o")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    errors = [l for l in r.stdout.splitlines() if "error" in l]
    assert not errors, r.stdout
