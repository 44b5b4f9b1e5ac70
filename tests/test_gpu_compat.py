"""The drop-in C++ API (include/scw_evaluator.hpp) driven by a compiled caller
that follows sCWatterView.cpp's --nogui sequence (tests/compat/viewer_nogui):
.cw load -> cw_allocate -> cw_unpack -> gpu_initialize -> gpu_cw_evaluate and
the cpu_cw_evaluate_* trio, checked against the oracle with each reference
function's own coordinate rounding and region rule."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from scwatter_amd import configs
from scwatter_amd.evaluator import axis_cpu, axis_gpu

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "compat", "viewer_nogui")


def test_viewer_nogui_sequence(tmp_path):
    assert os.path.exists(EXE), "tests/compat/viewer_nogui is built by __graft_entry__.build()"
    cw = configs.layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, 9, ("x",))
    path = tmp_path / "c.cw"
    cw.save(path)
    t = cw.unpack()
    N, extent, center, plane = 32, 12.0, (0.0, 0.0, 0.5), (0.3, -0.2, -0.0005)
    prefix = str(tmp_path / "out")
    print(f"viewer_nogui N={N} extent={extent} plane={plane}", flush=True)
    r = subprocess.run([EXE, str(path), prefix, str(N), str(extent), *map(str, center), *map(str, plane), "0"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    f32 = np.float32
    d = f32(f32(extent) / f32(N - 1))
    starts = [f32(f32(c) - f32(extent) / f32(2)) for c in center]
    zl = t.z_layers.astype(np.float32).astype(np.float64)
    pl = [[float(f32(p))] for p in plane]

    def load(name):
        return np.fromfile(f"{prefix}_{name}.bin", np.complex64).reshape(N, N, 3)

    gx, gy, gz = (axis_gpu(s, d, N) for s in starts)
    for name, grid in (("xy", (gx, gy, pl[2])), ("xz", (gx, pl[1], gz)), ("yz", (pl[0], gy, gz))):
        ref = oracle.eval_grid_f64(t.W, zl, t.begin, t.end, *grid, tol=0.001).reshape(N, N, 3)
        err = oracle.rel_l2(load("gpu_" + name), ref)
        print(f"gpu_cw_evaluate {name}: rel-L2 {err:.2e}", flush=True)
        assert err < 1e-5
    for name, args in (("xy", (starts[0], starts[1], plane[2])), ("xz", (starts[0], starts[2], plane[1])),
                       ("yz", (starts[1], starts[2], plane[0]))):
        ref = oracle.refcpu_plane(name, t.W, t.z_layers, t.begin, t.end, *args, d, N)
        err = oracle.rel_l2(load("cpu_" + name), ref)
        print(f"cpu_cw_evaluate_{name}: rel-L2 vs reference-semantics port {err:.2e}", flush=True)
        assert err < 2e-5
    # the xy plane sits below the interface within 1e-3: GPU rule -> region 1, CPU rule -> region 0
    assert oracle.rel_l2(load("gpu_xy"), load("cpu_xy")) > 0.1
