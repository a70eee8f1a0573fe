"""The C ABI library on the CPU: symbols, host-side setup, scene generator.

No device is needed: these tests call only the host helpers and check that
device entry points fail loudly (not silently) without a GPU.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mps_cuda_amd import configs, ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mps.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mps_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ffi.load_library()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


@pytest.mark.parametrize("make,prec", [(configs.cfg0, "fp64"), (configs.cfg0, "fp32"),
                                       (configs.small3d, "fp64"), (configs.small3d, "fp32")])
def test_host_constants_equal_the_oracle(oracle_mod, make, prec):
    scene, params = make(prec)
    mine = ffi.compute_constants(params, scene)
    ref = oracle_mod.Oracle(params, scene).constants()
    for k, v in ref.items():
        assert mine[k] == v, f"{k}: {mine[k]!r} != {v!r}"


@pytest.mark.parametrize("make", [configs.cfg0, configs.small3d])
def test_host_edge_flags_equal_the_oracle(oracle_mod, make):
    scene, params = make("fp64")
    e = ffi.compute_edge(params, scene)
    assert np.array_equal(e, oracle_mod.Oracle(params, scene).edge())
    walls = scene.typ == ffi.WLL
    assert e[walls].sum() > 0 and e[~walls].sum() == 0


def test_dam_break_generator():
    sc, p = configs.cfg0()
    assert (sc.typ == ffi.FLD).sum() == 1250 and (sc.typ == ffi.WLL).sum() == 678
    assert np.all(sc.typ[:1250] == ffi.FLD)  # fluid first
    n = np.linalg.norm(sc.normal[:, sc.typ == ffi.WLL], axis=0)
    assert np.allclose(n, 1.0)
    # innermost wall layer first (the reference reflects across the first wall met)
    w = np.where(sc.typ == ffi.WLL)[0]
    depth = np.maximum.reduce([np.where(sc.pos[0, w] < 0, -sc.pos[0, w], 0),
                               np.where(sc.pos[0, w] > 1.0, sc.pos[0, w] - 1.0, 0),
                               np.where(sc.pos[1, w] < 0, -sc.pos[1, w], 0)]) / 0.01
    assert np.all(np.diff(np.floor(depth)) >= 0)
    c = ffi.compute_constants(p, sc)
    assert c["nb"][:2] == [39, 38] and c["n_fluid"] == 1250


def test_jitter_is_bounded_and_quantized():
    sc = ffi.dambreak(dim=3, l0=0.005, tank=(0.1, 0.1, 0.05), column=(0.05, 0.05, 0.05),
                      jitter=0.01, seed=1)
    f = sc.typ == ffi.FLD
    lat = (np.floor(sc.pos[:, f] / 0.005) + 0.5) * 0.005
    j = sc.pos[:, f] - lat
    q = 2.0 ** -19
    assert np.all(np.abs(j) <= 0.01 * 0.005 + q / 2)
    assert np.all(np.abs(j) >= q / 2 * (1 - 1e-9))  # never on the lattice plane
    assert np.allclose(np.round(j / q - 0.5), j / q - 0.5, atol=1e-6)
    sc2 = ffi.dambreak(dim=3, l0=0.005, tank=(0.1, 0.1, 0.05), column=(0.05, 0.05, 0.05),
                       jitter=0.01, seed=1)
    assert np.array_equal(sc.pos, sc2.pos)  # seeded


def test_cfg1_workload_size():
    sc, p = configs.cfg1()
    assert (sc.typ == ffi.FLD).sum() == 300000
    c = ffi.compute_constants(p, sc)
    assert c["n_fluid"] == 300000 and c["n_wall"] == (sc.typ == ffi.WLL).sum()


def test_invalid_scenes_are_rejected():
    sc, p = configs.small2d()
    bad = ffi.SceneArrays(typ=sc.typ.copy(), pos=sc.pos.copy(), normal=sc.normal,
                          range_min=sc.range_min, range_max=sc.range_max)
    bad.typ[3] = 7
    with pytest.raises(ffi.MPSError, match="bad type"):
        ffi.compute_constants(p, bad)
    out = ffi.SceneArrays(typ=sc.typ, pos=sc.pos.copy(), normal=sc.normal,
                          range_min=sc.range_min, range_max=sc.range_max)
    out.pos[0, 0] = 50.0
    with pytest.raises(ffi.MPSError, match="outside the domain"):
        ffi.compute_constants(p, out)
    p2 = configs.small2d()[1]
    p2.dim = 4
    with pytest.raises(ffi.MPSError, match="dim"):
        ffi.compute_constants(p2, sc)


def test_device_entry_points_fail_loudly_without_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    sc, p = configs.small2d()
    with pytest.raises(ffi.MPSError, match="no HIP device|HIP"):
        ffi.Handle(p, sc)


def test_batch_checks_every_scene_before_touching_the_device():
    s2, p2 = configs.small2d("fp32")
    s3, p3 = configs.small3d("fp32")
    with pytest.raises(ffi.MPSError, match="same dim, precision"):
        ffi.Handle([p2, p3], [s2, s3])
    bad = ffi.SceneArrays(typ=s2.typ.copy(), pos=s2.pos, normal=s2.normal,
                          range_min=s2.range_min, range_max=s2.range_max)
    bad.typ[0] = 9
    with pytest.raises(ffi.MPSError, match="scene 1: .*bad type"):
        ffi.Handle([p2, p2], [s2, bad])


def test_cfg3_batch_workload():
    scenes, params = configs.cfg3()
    assert len(scenes) == 512
    nf = np.array([(s.typ == ffi.FLD).sum() for s in scenes])
    assert 1800 <= nf.min() and nf.max() <= 8400
    nu = np.array([p.nu for p in params])
    c = np.array([p.c for p in params])
    assert 1e-6 <= nu.min() and nu.max() <= 1e-4 and 15 <= c.min() and c.max() <= 30
    assert all(p.dt == 1e-4 and p.dim == 2 and p.precision == ffi.FP32 for p in params)


def test_autotune_cache_round_trip(tmp_path, monkeypatch):
    from mps_cuda_amd import autotune
    monkeypatch.setenv("MPS_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    scenes, params = configs.batch_small(3, "fp32")
    key = autotune.workload_key(params, scenes)
    assert key.startswith("d2-fp32-s3-")
    entry = dict(block=128, wave_grid=2048, graph=8, ms_per_step=1.0, trials=9)
    autotune._store(autotune.cache_path(), {key: entry})
    got = autotune.tune(params, scenes)  # served from the cache: no device needed
    assert got["cached"] and got["block"] == 128 and got["graph"] == 8
