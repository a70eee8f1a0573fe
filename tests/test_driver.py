"""The command-line driver mps_run on the CPU: reference-format input files
(RdDat, kernel.cu:4-143) and the host-side set-up (--dry-run).  GPU runs of
the driver are in test_gpu_driver.py."""
import os
import re
import subprocess

import numpy as np
import pytest

from mps_cuda_amd import configs, ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mps_cuda_amd", "mps_run")


def run(*args):
    res = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


def constants_of(out):
    m = re.search(r"SetPara: n0=(\S+) n0_grad=(\S+) lambda=(\S+) nB=(\d+)x(\d+)x(\d+) dt=(\S+) c=(\S+) "
                  r"umax=(\S+) Pmax=(\S+)", out)
    f = re.search(r"fluid (\d+), walls (\d+)", out)
    return dict(n0=float(m[1]), n0_grad=float(m[2]), nb=[int(m[4]), int(m[5]), int(m[6])], dt=float(m[7]),
                snd=float(m[8]), umax=float(m[9]), pmax=float(m[10]), n_fluid=int(f[1]), n_wall=int(f[2]))


def test_driver_builtin_workload_constants():
    got = constants_of(run("--dambreak", "cfg0", "--dry-run"))
    sc, p = configs.cfg0("fp64")
    k = ffi.compute_constants(p, sc)
    assert got["nb"] == list(k["nb"]) and got["n_fluid"] == 1250 and got["n_wall"] == k["n_wall"]
    for key in ("n0", "n0_grad", "dt", "snd", "umax", "pmax"):
        assert got[key] == pytest.approx(k[key], rel=1e-8), key


def read_reference(d):
    """Independent reader of the two files, as RdDat reads them (float)."""
    vals = np.array(open(os.path.join(d, "MPS_physical.txt")).read().split(), np.float32).astype(np.float64)
    rows = open(os.path.join(d, "initial.txt")).read().split("\n")
    n = int(rows[0])
    tab = np.array([r.split() for r in rows[1:1 + n]], np.float64)
    return vals, tab


def test_reference_files_round_trip(tmp_path):
    run("--dambreak", "small2d", "--write-reference", str(tmp_path), "--dry-run")
    vals, tab = read_reference(tmp_path)
    sc, _ = configs.small2d("fp64")
    assert len(vals) == 17 and tab.shape == (sc.n, 9)
    assert np.array_equal(tab[:, 1].astype(int), sc.typ)
    assert np.allclose(tab[:, 2:5].T, sc.pos, atol=1e-9)
    assert np.allclose(-tab[:, 5:8].T, sc.normal, atol=1e-9)  # files hold outward normals
    out = run("--physical", str(tmp_path / "MPS_physical.txt"), "--initial", str(tmp_path / "initial.txt"),
              "--dry-run", "--precision", "fp64")
    assert "0 outside the domain dropped" in out
    got = constants_of(out)
    # SetPara on the float-rounded inputs, c and dt derived from Ma (kernel.cu:407-415)
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    scene = ffi.SceneArrays(typ=sc.typ, pos=f32(sc.pos), normal=f32(sc.normal), range_min=tuple(vals[0:3]),
                            range_max=tuple(vals[3:6]))
    p = ffi.default_params(dim=2, precision="fp64", l0=vals[16], range_min=vals[0:3], range_max=vals[3:6],
                           gravity=vals[6:9], mach=vals[11], restitution=vals[13], dst_lmt_rat=vals[14],
                           nu=vals[15], c=0.0, dt=0.0, pmax_mode=ffi.PMAX_REFERENCE)
    k = ffi.compute_constants(p, scene)
    for key in ("n0", "n0_grad", "dt", "snd", "umax", "pmax"):
        assert got[key] == pytest.approx(k[key], rel=1e-8), key
    assert got["pmax"] < 0  # the reference's literal Pmax = 1000 G.y H


def test_reader_drops_particles_outside_the_domain(tmp_path):
    run("--dambreak", "small2d", "--write-reference", str(tmp_path), "--dry-run")
    rows = open(tmp_path / "initial.txt").read().split("\n")
    n = int(rows[0])
    rows[0] = str(n + 2)
    rows.insert(1 + n, f"{n} 0 5.0 0.1 0 0 0 0 0")          # beyond MAXc.x
    rows.insert(2 + n, f"{n + 1} 0 0.1 -0.2 0 0 0 0 0")     # below MINc.y = -0.0305 - 3.1 l0
    open(tmp_path / "initial.txt", "w").write("\n".join(rows))
    out = run("--physical", str(tmp_path / "MPS_physical.txt"), "--initial", str(tmp_path / "initial.txt"),
              "--dry-run")
    assert f"RdDat: {n} particles (2 outside the domain dropped)" in out


def test_driver_gradient_model_option():
    assert "pressure gradient: standard" in run("--dambreak", "cfg0", "--dry-run")
    assert "pressure gradient: tensor" in run("--dambreak", "cfg0", "--dry-run", "--grad", "tensor")
    res = subprocess.run([DRIVER, "--dambreak", "cfg0", "--dry-run", "--grad", "nope"], capture_output=True,
                         text=True)
    assert res.returncode == 2


def test_driver_rejects_bad_arguments():
    res = subprocess.run([DRIVER, "--dambreak", "nope", "--dry-run"], capture_output=True, text=True)
    assert res.returncode != 0 and "unknown" in res.stderr
    res = subprocess.run([DRIVER], capture_output=True, text=True)
    assert res.returncode == 2
