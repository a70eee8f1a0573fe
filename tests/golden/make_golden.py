"""Writes tests/golden/cfg0_fp64.npz: the cfg0 dam break (2-D, FP64, 1250
fluid) as generated, and the CPU oracle's state after 1 and 10 steps.

The reference ships no test vectors; this fixture pins the scene generator,
the oracle and the HIP path against silent change.  Run from the repo root:
    python -m tests.golden.make_golden
"""
import os

import numpy as np

from mps_cuda_amd import configs
from oracle import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "cfg0_fp64.npz")


def make():
    pyoracle.load()
    sc, p = configs.cfg0("fp64")
    o = pyoracle.Oracle(p, sc)
    out = dict(typ=sc.typ, pos0=sc.pos[:2], normal0=sc.normal[:2],
               range_min=np.array(sc.range_min), range_max=np.array(sc.range_max),
               l0=p.l0, dt=p.dt, nu=p.nu, c=p.c, rho=p.rho, pmax_mode=p.pmax_mode)
    done = 0
    for steps in (1, 10):
        o.step(steps - done)
        done = steps
        st = o.state()
        cs, ids = o.buckets(0)
        out.update({f"pos_{steps}": st["pos"][:2], f"vel_{steps}": st["vel"][:2], f"prs_{steps}": st["prs"],
                    f"nd_{steps}": st["nd"], f"cs_{steps}": cs, f"ids_{steps}": ids})
    np.savez_compressed(PATH, **out)
    return PATH


if __name__ == "__main__":
    print(make(), os.path.getsize(PATH), "bytes")
