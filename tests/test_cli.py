"""scwatterview (the --nogui counterpart of sCWatterView.cpp:679-796, built
from scwatter_amd/cli/scwatterview.cpp): options and error paths on the CPU,
saved planes and volumes against the FP64 oracle on the GPU."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from scwatter_amd import configs
from scwatter_amd.evaluator import axis_cpu, axis_gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "scwatter_amd", "scwatterview")


def _run(*args):
    assert os.path.exists(EXE), "scwatter_amd/scwatterview is built by __graft_entry__.build()"
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    return r


def test_help_lists_reference_options():
    r = _run("--help")
    assert r.returncode == 1  # as the reference viewer
    for opt in ("--input", "--cuda", "--nogui", "--size", "--resolution", "--output", "--axis", "--center", "--slice"):
        assert opt in r.stdout


def test_missing_file_is_reported(tmp_path):
    r = _run(tmp_path / "absent.cw", "--nogui")
    assert r.returncode == 1 and "ERROR: file" in r.stdout and "not found" in r.stdout


@pytest.mark.parametrize("args", [["--bogus"], ["--resolution", "x"], ["--sample"], ["--form", "fast"]])
def test_bad_options_are_rejected(args):
    r = _run(*args)
    assert r.returncode == 2 and "scwatterview:" in r.stderr


@pytest.fixture(scope="module")
def cw_file(tmp_path_factory):
    cw = configs.layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, 9, ("x",))
    path = tmp_path_factory.mktemp("cli") / "c.cw"
    cw.save(path)
    return path, cw.unpack()


def _npy(path, shape):
    a = np.load(path, allow_pickle=False)
    assert a.dtype == np.complex64 and a.shape == shape
    return a


@pytest.mark.gpu
@pytest.mark.parametrize("device", [0, -1])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_plane_matches_oracle(cw_file, tmp_path, axis, device):
    path, t = cw_file
    res, size, center, sl = 5, 12.0, (0.4, -0.3, 0.5), -0.7
    out = tmp_path / "p.npy"
    r = _run(path, "--nogui", "--cuda", device, "--resolution", res, "--size", size, "--center", *center,
             "--axis", axis, "--slice", sl, "--output", out)
    assert r.returncode == 0 and "Evaluation (no gui) takes:" in r.stdout
    N = 1 << res
    E = _npy(out, (N, N, 3))
    f32 = np.float32
    d = f32(f32(size) / f32(N - 1))
    starts = [f32(f32(c) - f32(size) / f32(2)) for c in center]
    ax = axis_gpu if device >= 0 else axis_cpu
    g = [ax(s, d, N) for s in starts]
    g[axis] = np.array([float(f32(sl))])  # the cut sits at the absolute coordinate `slice`
    zl = t.z_layers.astype(np.float32).astype(np.float64)
    ref = oracle.eval_grid_f64(t.W, zl, t.begin, t.end, *g, tol=0.001 if device >= 0 else 0.0)
    # [z][y][x] with the cut axis of length 1: rows y for xy, z for xz and yz;
    # columns x for xy and xz, y for yz (E_xy[iy*N+ix], E_xz[iz*N+ix], E_yz[iz*N+iy])
    ref = ref.reshape(N, N, 3)
    err = oracle.rel_l2(E, ref)
    print(f"axis {axis} device {device}: rel-L2 {err:.2e}", flush=True)
    assert err < 2e-5


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["fp32", "bf16x3"])
def test_volume_matches_oracle(cw_file, tmp_path, dtype):
    path, t = cw_file
    res, size, center = 4, 10.0, (0.0, 0.2, -0.4)
    out = tmp_path / "v.npy"
    r = _run(path, "--nogui", "--volume", "--dtype", dtype, "--resolution", res, "--size", size, "--center", *center,
             "--output", out)
    assert r.returncode == 0
    N = 1 << res
    E = _npy(out, (N, N, N, 3))
    f32 = np.float32
    d = f32(f32(size) / f32(N - 1))
    g = [axis_gpu(f32(f32(c) - f32(size) / f32(2)), d, N) for c in center]
    zl = t.z_layers.astype(np.float32).astype(np.float64)
    ref = oracle.eval_grid_f64(t.W, zl, t.begin, t.end, *g, tol=0.001).reshape(N, N, N, 3)
    err = oracle.rel_l2(E, ref)
    print(f"volume {dtype}: rel-L2 {err:.2e}", flush=True)
    assert err < (2e-5 if dtype == "fp32" else 1e-4)
