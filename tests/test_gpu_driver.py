"""The command-line driver mps_run on the GPU: its CSV outputs equal the
state of a handle stepped the same way (to the 6 decimals of "%f"), in the
built-in step-count mode and in the reference's file mode with its
output_time / FIN_TIM clock (kernel.cu:1872-1945)."""
import os
import subprocess

import numpy as np
import pytest

from mps_cuda_amd import configs, ffi
from tests.test_driver import DRIVER, read_reference

pytestmark = pytest.mark.gpu


def drive(*args):
    print(f"[case] mps_run {' '.join(args)}", flush=True)
    res = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


def csv(path):
    return np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)


def check_output(path, st, pav_div):
    tab = csv(path)
    fl = st["typ"] == ffi.FLD
    assert tab.shape[0] == fl.sum()
    assert np.allclose(tab[:, 0:3], st["pos"][:, fl].T, atol=6e-7, rtol=0)
    assert np.allclose(tab[:, 3:6], st["vel"][:, fl].T, atol=6e-7, rtol=0)
    assert np.allclose(tab[:, 7], st["prs"][fl], atol=6e-7, rtol=1e-6)
    assert np.allclose(tab[:, 8], st["pav"][fl] / pav_div, atol=6e-7, rtol=1e-6)


def test_driver_step_mode_outputs(tmp_path):
    drive("--dambreak", "small2d", "--steps", "40", "--every", "20", "--out", str(tmp_path))
    sc, p = configs.small2d("fp64")
    h = ffi.Handle(p, sc)
    check_output(tmp_path / "output00000.csv", h.state(), 1.0)
    h.step(20)
    check_output(tmp_path / "output00001.csv", h.state(), 20.0)
    h.set_state(pav=np.zeros(h.n))  # WrtDat clears pav after writing
    h.step(20)
    st = h.state()
    check_output(tmp_path / "output00002.csv", st, 20.0)
    # the mirror file keeps the reference's trailing comma on every row
    mrr = np.loadtxt(tmp_path / "outputMRR00002.csv", delimiter=",", skiprows=1, ndmin=2, usecols=range(5))
    ids, mpos, mprs = h.mirrors()
    assert mrr.shape[0] == len(ids)
    assert np.allclose(mrr[:, 0:2], mpos[:2].T, atol=6e-7, rtol=0)
    assert int(open(tmp_path / "seitei.txt").readline()) == int((st["typ"] == ffi.FLD).sum())
    walls = csv(tmp_path / "outputwall00000.csv")
    assert walls.shape[0] == int((sc.typ == ffi.WLL).sum())
    assert np.array_equal(walls[:, 5].astype(np.int8), ffi.compute_edge(p, sc)[sc.typ == ffi.WLL])


def test_driver_tensor_gradient(tmp_path):
    drive("--dambreak", "small2d", "--steps", "20", "--every", "20", "--grad", "tensor", "--out", str(tmp_path))
    sc, p = configs.small2d("fp64", grad_model=ffi.GRAD_TENSOR)
    h = ffi.Handle(p, sc)
    h.step(20)
    check_output(tmp_path / "output00001.csv", h.state(), 20.0)


def test_driver_file_mode_clock(tmp_path):
    ref = tmp_path / "ref"
    out = tmp_path / "out"
    ref.mkdir()
    out.mkdir()
    drive("--dambreak", "small2d", "--write-reference", str(ref), "--dry-run")
    vals, tab = read_reference(ref)
    drive("--physical", str(ref / "MPS_physical.txt"), "--initial", str(ref / "initial.txt"),
          "--precision", "fp64", "--out", str(out))
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    sc0, _ = configs.small2d("fp64")
    scene = ffi.SceneArrays(typ=sc0.typ, pos=f32(sc0.pos), normal=f32(sc0.normal),
                            range_min=tuple(vals[0:3]), range_max=tuple(vals[3:6]))
    p = ffi.default_params(dim=2, precision="fp64", l0=vals[16], range_min=vals[0:3], range_max=vals[3:6],
                           gravity=vals[6:9], mach=vals[11], restitution=vals[13], dst_lmt_rat=vals[14],
                           nu=vals[15], c=0.0, dt=0.0, pmax_mode=ffi.PMAX_REFERENCE)
    h = ffi.Handle(p, scene)
    dt, out_t, fin = h.constants()["dt"], vals[10], vals[9]
    # ClcMPS's clock in the run's precision
    TIM = outtime = 0.0
    steps, outputs = 0, []
    while True:
        if outtime >= out_t:
            outtime -= out_t
            outputs.append(steps)
            if TIM >= fin:
                break
        steps += 1
        outtime += dt
        TIM += dt
    files = sorted(f for f in os.listdir(out) if f.startswith("output0"))
    assert len(files) == len(outputs) + 1
    done, last = 0, 0
    for i, s in enumerate(outputs, start=1):
        h.step(s - done)
        done = s
        st = h.state()
        check_output(out / f"output{i:05d}.csv", st, float(s - last))
        h.set_state(pav=np.zeros(h.n))
        last = s
    assert int(open(out / "seitei.txt").readline()) == int((st["typ"] == ffi.FLD).sum())
