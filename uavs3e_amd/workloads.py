"""The benchmark workloads: frame sizes, search settings and synthetic frames.

Each workload is generated from its seed by the host generator
(csrc/synth.c); the same seed gives the same samples on any machine.

  hd_16x16    1920x1080 8-bit, 1 reference, 16x16 PUs, full search +-32, QP 32
  hd_allpu    1920x1080 8-bit, 2 references, all 29 PU shapes, +-64      (bench line)
  uhd10_allpu 3840x2160 10-bit, 4 references (POC distance 1,2,4,8), 29 shapes, +-128
  fuhd10_bi   7680x4320 10-bit, 2 L0 + 2 L1 references, 29 shapes, +-64

Motion: the current frame samples the field at (x, y) + global + region
offset; the global motion is (+13, -6) quarter-pel per frame and every
64x64 region adds an integer quarter-pel offset in [-16, 16].
"""
import math
from dataclasses import dataclass, field
from typing import List

from . import MASK_16x16, MASK_ALL, MeHost

GLOBAL_QPEL = (13, -6)


def lambda_mv(qp):
    """floor(65536 * sqrt(1.43631 * 2^((qp-16)/4))) (uavs3e.c:834-851)."""
    return int(math.floor(65536.0 * math.sqrt(1.43631 * math.pow(2.0, (qp - 16.0) / 4.0))))


@dataclass
class Workload:
    name: str
    w: int
    h: int
    bd: int
    R: int
    mask: int
    qp: int = 32
    ref_shifts_qpel: List[tuple] = field(default_factory=list)  # field shift of each reference
    cur_shift_qpel: tuple = (0, 0)
    sigma_ref: float = 4.0
    sigma_cur: float = 3.0
    seed: int = 1
    drift: bool = False
    l1: int = 0

    @property
    def num_refs(self):
        return len(self.ref_shifts_qpel)

    @property
    def num_refs_l1(self):
        """References of list 1 (the last ones): the bi-directional workload
        has 2 L0 (past) and 2 L1 (future) references."""
        return self.l1

    @property
    def lam(self):
        return lambda_mv(self.qp)

    @property
    def n_ctu(self):
        return ((self.w + 127) // 128) * ((self.h + 127) // 128)

    def scaled(self, w, h, R=None):
        """The same generation rules on a smaller picture (parity tests)."""
        import dataclasses
        return dataclasses.replace(self, name=self.name + "_%dx%d" % (w, h), w=w, h=h,
                                   R=self.R if R is None else R)

    def frames(self, host=None):
        """(cur, [refs]) as numpy arrays."""
        host = host or MeHost(self.bd)
        refs = []
        for i, (qx, qy) in enumerate(self.ref_shifts_qpel):
            off = 0.0
            if self.drift:
                off = round(16 * host.uniform(host.key(self.seed, 0, 10 + i, 6), 0)) - 8
            refs.append(host.synth(self.seed, 0, 10 + i, self.w, self.h, qx / 4.0, qy / 4.0, 0, self.sigma_ref, 3, off))
        off = 0.0
        if self.drift:
            off = round(16 * host.uniform(host.key(self.seed, 0, 1, 6), 0)) - 8
        cx, cy = self.cur_shift_qpel
        cur = host.synth(self.seed, 0, 1, self.w, self.h, cx / 4.0, cy / 4.0, 16, self.sigma_cur, 4, off)
        return cur, refs

    def ref_chroma(self, host=None):
        """4:2:0 chroma (u, v) of each reference: two more fields of the same
        kind (streams 1 and 2 of the seed) at half resolution, moved by half
        the luma shift."""
        host = host or MeHost(self.bd)
        out = []
        for i, (qx, qy) in enumerate(self.ref_shifts_qpel):
            uv = []
            for c in (1, 2):
                uv.append(host.synth(self.seed, c, 10 + i, self.w // 2, self.h // 2, qx / 8.0, qy / 8.0, 0,
                                     self.sigma_ref, 3, 0.0))
            out.append(tuple(uv))
        return out


def _second_ref_shift(seed):
    host = MeHost(8)
    k = host.key(seed, 0, 0, 5)
    return (int(host.uniform(k, 0) * 17) - 8, int(host.uniform(k, 1) * 17) - 8)


def hd_16x16():
    return Workload("1080p8_16x16_1ref_r32", 1920, 1080, 8, 32, MASK_16x16, ref_shifts_qpel=[(0, 0)],
                    cur_shift_qpel=GLOBAL_QPEL, seed=1)


def hd_allpu():
    s = _second_ref_shift(2)
    return Workload("1080p8_allpu_2ref_r64", 1920, 1080, 8, 64, MASK_ALL,
                    ref_shifts_qpel=[(0, 0), (-s[0], -s[1])], cur_shift_qpel=GLOBAL_QPEL, seed=2)


def uhd10_allpu():
    gx, gy = GLOBAL_QPEL
    return Workload("2160p10_allpu_4ref_r128", 3840, 2160, 10, 128, MASK_ALL,
                    ref_shifts_qpel=[((8 - d) * gx, (8 - d) * gy) for d in (1, 2, 4, 8)],
                    cur_shift_qpel=(8 * gx, 8 * gy), sigma_ref=12.0, sigma_cur=12.0, seed=3)


def fuhd10_bi():
    gx, gy = GLOBAL_QPEL
    return Workload("4320p10_allpu_l0l1_r64", 7680, 4320, 10, 64, MASK_ALL,
                    ref_shifts_qpel=[(o * gx, o * gy) for o in (-1, -2, 1, 2)],
                    cur_shift_qpel=(0, 0), sigma_ref=12.0, sigma_cur=12.0, seed=5, drift=True, l1=2)


ALL = {"hd_16x16": hd_16x16, "hd_allpu": hd_allpu, "uhd10_allpu": uhd10_allpu, "fuhd10_bi": fuhd10_bi}
