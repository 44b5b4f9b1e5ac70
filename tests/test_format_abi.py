""".cw format, unpack order and the C ABI surface (CPU only: no device calls)."""
import ctypes
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from scwatter_amd import ScwError, _native, configs
from scwatter_amd.cw import CoupledWaveStructure, HomogeneousLayer, load_tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rand(rng, n):
    return rng.standard_normal((n, 6)) + 1j * rng.standard_normal((n, 6))


def _structure(seed=0, counts=((3, 2), (0, 4), (5, 1))):
    rng = np.random.default_rng(seed)
    layers = [HomogeneousLayer(z, _rand(rng, nr), _rand(rng, nt)) for z, (nr, nt) in zip((-1.5, 0.25, 2.0), counts)]
    return CoupledWaveStructure(_rand(rng, 4), layers)


def test_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "scwatter_amd.h")).read()
    declared = set(re.findall(r"\b(scw_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(_native.EXPORTS)
    L = _native.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.scw_abi_version() == 1


def test_exports_reference_cpp_names():
    so = _native.LIB_PATH
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", so], capture_output=True, text=True).stdout
    for sig in ("cw_allocate(CoupledWaveStructure<double>*)", "cw_unpack(CoupledWaveStructure<double>*)",
                "gpu_initialize()", "gpu_cw_evaluate(std::complex<float>*, std::complex<float>*, "
                "std::complex<float>*, float, float, float, float, float, float, float, int, int)",
                "cpu_cw_evaluate_xy(scw_vec3c32*, float, float, float, float, unsigned long)",
                "cpu_cw_evaluate_yz(scw_vec3c32*, float, float, float, float, unsigned long)",
                "cpu_cw_evaluate_xz(scw_vec3c32*, float, float, float, float, unsigned long)"):
        assert sig in out, sig
    for var in (" layers", " z_layers", " waves_begin", " waves_end", " W", " free_gpu_memory",
                " total_gpu_memory"):
        assert any(line.endswith(var) for line in out.splitlines()), var


def test_roundtrip_and_byte_count(tmp_path):
    cw = _structure()
    p = tmp_path / "s.cw"
    cw.save(p)
    # CoupledWaveStructure.h:72-101: 8 + 8 + 96 nPi + 8 + sum(8 + 8 + 96 nPr + 8 + 96 nPt)
    expect = 24 + 96 * 4 + sum(24 + 96 * (L.Pr.shape[0] + L.Pt.shape[0]) for L in cw.Layers)
    assert p.stat().st_size == expect
    raw = p.read_bytes()
    assert struct.unpack_from("<QQ", raw, 0) == (8, 4)
    assert np.array_equal(np.frombuffer(raw, np.float64, 12, 16), cw.Pi[0].view(np.float64))
    back = CoupledWaveStructure.load(p)
    assert np.array_equal(back.Pi, cw.Pi)
    for a, b in zip(back.Layers, cw.Layers):
        assert a.z == b.z and np.array_equal(a.Pr, b.Pr) and np.array_equal(a.Pt, b.Pt)


def test_native_unpack_matches_region_rule(tmp_path):
    cw = _structure(1)
    p = tmp_path / "s.cw"
    cw.save(p)
    t = load_tables(p)
    u = cw.unpack()
    assert np.array_equal(t.W, u.W) and np.array_equal(t.z_layers, u.z_layers)
    assert np.array_equal(t.begin, u.begin) and np.array_equal(t.end, u.end)
    # region 0 = Pi + Pr_0, region l = Pt_{l-1} + Pr_l, last = Pt_{L-1} (cpuEvaluator.cpp:34-53)
    L = cw.Layers
    assert np.array_equal(t.region(0), np.concatenate([cw.Pi, L[0].Pr]))
    assert np.array_equal(t.region(1), np.concatenate([L[0].Pt, L[1].Pr]))
    assert np.array_equal(t.region(2), np.concatenate([L[1].Pt, L[2].Pr]))
    assert np.array_equal(t.region(3), L[2].Pt)


def test_float_precision_file_is_widened(tmp_path):
    rng = np.random.default_rng(2)
    Pi = rng.standard_normal((2, 12)).astype(np.float32)
    Pr = rng.standard_normal((1, 12)).astype(np.float32)
    raw = struct.pack("<QQ", 4, 2) + Pi.tobytes() + struct.pack("<Q", 1) + struct.pack("<f", 0.5) + \
        struct.pack("<Q", 1) + Pr.tobytes() + struct.pack("<Q", 0)
    p = tmp_path / "f.cw"
    p.write_bytes(raw)
    t = load_tables(p)
    assert t.z_layers[0] == 0.5
    assert np.array_equal(t.W.view(np.float64).reshape(-1, 12), np.concatenate([Pi, Pr]).astype(np.float64))


@pytest.mark.parametrize("mutate,status", [
    (lambda b: b[:20], "SCW_ERR_IO"),                                  # truncated incident block
    (lambda b: struct.pack("<Q", 2) + b[8:], "SCW_ERR_FORMAT"),        # bad precision field
    (lambda b: b[:8] + struct.pack("<Q", 1 << 60) + b[16:], "SCW_ERR_IO"),  # absurd count, no bad_alloc
    (lambda b: b[:-10], "SCW_ERR_IO"),                                 # truncated last layer
])
def test_invalid_files_are_errors(tmp_path, mutate, status):
    p = tmp_path / "s.cw"
    _structure().save(p)
    q = tmp_path / "bad.cw"
    q.write_bytes(mutate(p.read_bytes()))
    with pytest.raises(ScwError, match=status):
        load_tables(q)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(ScwError, match="SCW_ERR_IO"):
        load_tables(tmp_path / "nope.cw")


def test_tile_sample_region_sizes():
    cw = configs.tile_sample_structure(M=3, Z=4, size=(8.0, 8.0, 4.0), seed=3)
    t = cw.unpack()
    MF = 9
    sizes = t.end - t.begin
    assert sizes[0] == 2 * MF and sizes[-1] == MF and all(s == 4 * MF * MF for s in sizes[1:-1])
    # every internal region: kx, ky on the 3 x 3 order lattice, kz shared by all orders of a mode
    R = t.region(2)
    assert len(np.unique(R[:, 3])) == 3 and len(np.unique(R[:, 4])) == 3 and len(np.unique(R[:, 5])) == 4 * MF


def test_layered_structure_continuity():
    import oracle
    rng = np.random.default_rng(4)
    n = [1.0, 1.6 + 0.02j, 1.2 + 0.01j, 1.9 + 0.04j]
    z = np.array([-0.7, 0.4, 1.3])
    t = configs.layered_structure(n, z, 0.6, 0.7, 5, ("x", "y")).unpack()
    pts = np.column_stack([rng.uniform(-3, 3, 10), rng.uniform(-3, 3, 10), np.zeros(10)])
    one = lambda W, P: oracle.eval_points_f64(W, np.zeros(0), [0], [W.shape[0]], P)
    hw = lambda W: np.concatenate([np.cross(W[:, 3:], W[:, :3]), W[:, 3:]], 1)
    for l, zl in enumerate(z):
        P = pts.copy()
        P[:, 2] = zl
        a, b = t.region(l), t.region(l + 1)
        scale = np.abs(one(a, P)).max()
        assert np.abs(one(a, P)[:, :2] - one(b, P)[:, :2]).max() < 1e-12 * max(scale, 1)
        assert np.abs(one(hw(a), P)[:, :2] - one(hw(b), P)[:, :2]).max() < 1e-11 * max(scale, 1) * 20


def test_fresnel_normal_incidence():
    n2 = 1.4 + 0.05j
    t = configs.layered_structure([1.0, n2], [0.0], 1.0, 0.0, 1, ("x",)).unpack()
    assert abs(t.W[1, 0] / t.W[0, 0] - (1 - n2) / (1 + n2)) < 1e-14
    assert abs(t.W[2, 0] / t.W[0, 0] - 2 / (1 + n2)) < 1e-14
