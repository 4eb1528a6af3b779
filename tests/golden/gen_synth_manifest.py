"""Records SHA-256 digests of the synthetic workload frames.

    python tests/golden/gen_synth_manifest.py

The generator (uavs3e_amd/csrc/synth.c) uses only IEEE +,-,*,/ and integer
hashing, so these digests must reproduce on any host; tests/test_lib_cpu.py
checks them.
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from uavs3e_amd import workloads as wl  # noqa: E402

SIZES = {"hd_16x16": None, "hd_allpu": None, "uhd10_allpu": (640, 360), "fuhd10_bi": (640, 360)}


def digests():
    out = {}
    for name, size in SIZES.items():
        w = wl.ALL[name]()
        if size:
            w = w.scaled(*size)
        cur, refs = w.frames()
        out[w.name] = {"cur": hashlib.sha256(cur.tobytes()).hexdigest(),
                       "refs": [hashlib.sha256(r.tobytes()).hexdigest() for r in refs],
                       "shape": list(cur.shape), "dtype": str(cur.dtype)}
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "synth_manifest.json")
    with open(path, "w") as f:
        json.dump(digests(), f, indent=1, sort_keys=True)
    print("wrote", path)
