"""The committed cfg0 fixture (tests/golden/cfg0_fp64.npz, written by
tests/golden/make_golden.py): the scene generator, the CPU oracle and the HIP
path must reproduce it bit for bit."""
import os

import numpy as np
import pytest

from mps_cuda_amd import configs, ffi

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg0_fp64.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(PATH, allow_pickle=False)


def _check(h, gold, where):
    done = 0
    for steps in (1, 10):
        h.step(steps - done)
        done = steps
        st = h.state()
        for f in ("pos", "vel"):
            assert np.array_equal(st[f][:2], gold[f"{f}_{steps}"]), f"{where}: {f} after {steps}"
        for f in ("prs", "nd"):
            assert np.array_equal(st[f], gold[f"{f}_{steps}"]), f"{where}: {f} after {steps}"
        cs, ids = h.buckets(ffi.BKT_FLUID)
        assert np.array_equal(cs, gold[f"cs_{steps}"]) and np.array_equal(ids, gold[f"ids_{steps}"])


def test_generator_reproduces_the_fixture_scene(gold):
    sc, p = configs.cfg0("fp64")
    assert np.array_equal(sc.typ, gold["typ"])
    assert np.array_equal(sc.pos[:2], gold["pos0"]) and np.array_equal(sc.normal[:2], gold["normal0"])
    assert np.array_equal(np.array(sc.range_min), gold["range_min"])
    assert (p.l0, p.dt, p.nu, p.c, p.rho) == (gold["l0"], gold["dt"], gold["nu"], gold["c"], gold["rho"])


def test_oracle_reproduces_the_fixture(oracle_mod, gold):
    sc, p = configs.cfg0("fp64")
    _check(oracle_mod.Oracle(p, sc), gold, "oracle")


@pytest.mark.gpu
def test_hip_path_reproduces_the_fixture(gold):
    sc, p = configs.cfg0("fp64")
    print("[case] golden cfg0 FP64 on the GPU", flush=True)
    _check(ffi.Handle(p, sc), gold, "HIP")
