"""Synthetic beams, TetGen I/O and the C ABI surface (no GPU needed)."""
import os
import re
import subprocess

import numpy as np
import pytest

from wrapd_amd import _lib, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_splitmix64_known_values():
    # SplitMix64 reference outputs for seed 1234567 (Steele, Lea & Flood 2014)
    r = meshgen.SplitMix64(1234567)
    got = [int(v) for v in r.next_u64(3)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423]


@pytest.mark.parametrize("cfg,counts", [(1, (1920, 525, 25)), (2, (245760, 46529, 578))])
def test_config_sizes_and_orientation(cfg, counts):
    p = meshgen.config_problem(cfg)
    assert (p.n_elems, p.n_verts, len(p.pins)) == counts
    D = np.stack([p.rest[p.tets[:, k]] - p.rest[p.tets[:, 0]] for k in (1, 2, 3)], axis=2)
    assert np.all(np.linalg.det(D) > 0)
    # pinned faces keep their exact init positions relative to rest (rotation by 0 / 360 deg)
    assert np.allclose(p.init[p.pins], p.rest[p.pins], atol=1e-12)


def test_conforming_kuhn_split():
    p = meshgen.small_beam(3, 2, 2)
    faces = {}
    for t in p.tets:
        for f in ([t[0], t[1], t[2]], [t[0], t[1], t[3]], [t[0], t[2], t[3]], [t[1], t[2], t[3]]):
            k = tuple(sorted(f))
            faces[k] = faces.get(k, 0) + 1
    assert max(faces.values()) == 2   # every interior face is shared by exactly two tets
    # boundary area equals the box surface
    area = 0.0
    for k, c in faces.items():
        if c == 1:
            a, b, d = (p.rest[i] for i in k)
            area += 0.5 * np.linalg.norm(np.cross(b - a, d - a))
    L, W, H = 0.3, 0.2, 0.2
    assert abs(area - 2 * (L * W + L * H + W * H)) < 0.01


def test_elenode_roundtrip_and_one_based(tmp_path):
    p = meshgen.small_beam(3, 2, 2)
    base = str(tmp_path / "m")
    meshgen.save_elenode(base, p.rest, p.tets)
    v, t = meshgen.load_elenode(base)
    assert np.array_equal(v, p.rest) and np.array_equal(t, p.tets)
    # 1-based files (samples/utils/MeshIO.hpp:215-220)
    with open(base + "1.node", "w") as f:
        f.write(f"{len(v)} 3 0 0\n")
        for i, x in enumerate(v):
            f.write(f"{i + 1} {float(x[0])!r} {float(x[1])!r} {float(x[2])!r}\n")
    with open(base + "1.ele", "w") as f:
        f.write(f"{len(t)} 4 0\n")
        for i, x in enumerate(t):
            f.write(f"{i + 1} {x[0] + 1} {x[1] + 1} {x[2] + 1} {x[3] + 1}\n")
    v1, t1 = meshgen.load_elenode(base + "1")
    assert np.array_equal(v1, p.rest) and np.array_equal(t1, p.tets)


def _header_functions():
    names = []
    for h in ("wrapd_gpu.h", "wrapd_solver.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"\b(wrapd_\w+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\s[TW]\s(wrapd_\w+)$", out, flags=re.M))
    declared = _header_functions()
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    # and the Python binding knows every declared function
    assert sorted(set(_lib.exported_symbols())) == declared


def test_library_loads_without_gpu_calls():
    L = _lib.lib()
    import ctypes
    s = _lib.SolverSettings()
    L.wrapd_solver_default_settings(ctypes.byref(s))
    assert s.lbfgs_m == 6 and s.global_max_iters == 100 and abs(s.gamma - 1.5) < 1e-15


def test_quasi_cli_help_and_flag_errors():
    q = _lib.QUASI_PATH
    assert os.path.exists(q)
    r = subprocess.run([q, "-help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "-rw_delay" in r.stdout and "-global_ls" in r.stdout
    r = subprocess.run([q, "-bogus", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Unrecognized command-line argument" in r.stderr
