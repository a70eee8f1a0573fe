"""Surface_Edge on the GPU (d_Surface_Edge, kernel.cu:553-668): the wall
edge flags a handle computes at creation equal the oracle's and the host
routine's bit for bit.

The GPU kernel takes distances in FP64 and compares normals in the run's
precision, like both CPU routines, and searches the 27 buckets around each
wall; the handle uses it when the edge radius is below the bucket width and
every wall lies inside the bucket grid, and the host routine otherwise.
"""
import numpy as np
import pytest

from mps_cuda_amd import configs, ffi

pytestmark = pytest.mark.gpu


def on_gpu_path(params, scene):
    """The handle's own eligibility test, restated: radius < 0.9 bucket
    widths and no wall in a clamped bucket cell."""
    c = ffi.compute_constants(params, scene)
    if not params.edge_factor * params.l0 < 0.9 * c["db"]:
        return False
    dt = np.float32 if params.precision == ffi.FP32 else np.float64
    w = scene.typ > 0  # WLL, OBJ, OBJ2
    for a in range(params.dim):
        t = (scene.pos[a, w].astype(dt) - dt(c["min_c"][a])) * dt(c["dbinv"])
        if not np.all((t >= 0) & (t < c["nb"][a] - 2)):
            return False
    return True


def check(scene, params, oracle_mod=None, gpu=True):
    print(f"[case] edge n={scene.n} dim={params.dim} prec={params.precision} "
          f"factor={params.edge_factor}", flush=True)
    assert on_gpu_path(params, scene) == gpu
    h = ffi.Handle(params, scene)
    got = h.edge()
    want = ffi.compute_edge(params, scene)
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} flags differ"
    if oracle_mod is not None:
        assert np.array_equal(got, oracle_mod.Oracle(params, scene).edge())
    walls = scene.typ > 0
    assert got[walls].sum() > 0 and got[~walls].sum() == 0
    return got


@pytest.mark.parametrize("make,prec", [(configs.cfg0, "fp64"), (configs.cfg0, "fp32"),
                                       (configs.small3d, "fp64"), (configs.small3d, "fp32"),
                                       (configs.small2d, "fp32")])
def test_gpu_edge_equals_oracle(oracle_mod, make, prec):
    scene, params = make(prec)
    check(scene, params, oracle_mod)


def test_gpu_edge_cfg1():
    scene, params = configs.cfg1("fp32")
    check(scene, params)


def test_gpu_edge_mixed_wall_types(oracle_mod):
    """Only walls of the same type count (kernel.cu:575): recasting the walls
    with a z normal as OBJ drops the edges that came from the side walls
    (1408 edge walls become 936)."""
    scene, params = configs.small3d("fp64")
    typ = scene.typ.copy()
    typ[(typ == ffi.WLL) & (np.abs(scene.normal[2]) > 0.99)] = ffi.OBJ
    mixed = ffi.SceneArrays(typ=typ, pos=scene.pos, normal=scene.normal,
                            range_min=scene.range_min, range_max=scene.range_max)
    e_mixed = check(mixed, params, oracle_mod)
    assert not np.array_equal(e_mixed, ffi.compute_edge(params, scene))


def test_host_path_for_a_wide_radius(oracle_mod):
    """An edge radius wider than the stencil reaches takes the host routine."""
    scene, params = configs.cfg0("fp64")
    params.edge_factor = 3.0  # 0.03 > 0.9 * 0.031
    check(scene, params, oracle_mod, gpu=False)


def test_gpu_edge_batch():
    scenes, params = configs.batch_small(4, "fp32", dim=3)
    h = ffi.Handle(params, scenes)
    got = h.edge()
    for k, (sc, p) in enumerate(zip(scenes, params)):
        assert np.array_equal(got[h.offsets[k]:h.offsets[k + 1]], ffi.compute_edge(p, sc)), f"scene {k}"
