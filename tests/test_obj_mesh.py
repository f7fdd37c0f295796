"""Unstructured input and physics diagnostics, CPU only.

The OBJ loader and the surface / edge mass rules against the reference's own
Mesh::loadObj and computeMassesFrom* (fixtures tests/golden/ref_obj_*.npz,
written by tests/golden/make_golden.py from oracle/_ref, and the live build
where present), the edge colouring of a loaded mesh against the reference's,
the masses against closed forms, and pdx_diagnostics against the oracle's
strain and energy functions and against closed forms."""
import glob
import os

import numpy as np
import pytest

from pd_ipc_amd import Cloth, Engine, PdxError, SimConfig, scenes
from pd_ipc_amd.pdx import load_library
from oracle_ffi import ROOT, Oracle, ref_mesh_lib

OBJ_FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "ref_obj_*.npz")))


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _write(tmp_path, text, name="mesh.obj"):
    p = tmp_path / name
    p.write_bytes(text if isinstance(text, bytes) else text.encode())
    return str(p)


def test_obj_fixtures_present():
    names = {os.path.basename(p) for p in OBJ_FIXTURES}
    assert {"ref_obj_tri_grid.npz", "ref_obj_forms.npz", "ref_obj_degenerate.npz",
            "ref_obj_bad_index.npz"} <= names


@pytest.mark.parametrize("path", OBJ_FIXTURES, ids=[os.path.basename(p) for p in OBJ_FIXTURES])
def test_obj_loader_matches_reference(path, tmp_path):
    """Vertices, triangles, unique edges in face order, rest lengths, kappa and
    surface masses equal the reference's bit for bit; files the reference
    refuses are refused."""
    d = _load(path)
    obj = _write(tmp_path, d["obj_text"].tobytes())
    if not bool(d["loads"]):
        with pytest.raises(PdxError):
            scenes.obj_cloth(obj, float(d["kappa"]), float(d["density"]))
        return
    c = scenes.obj_cloth(obj, float(d["kappa"]), float(d["density"]))
    assert np.array_equal(c.positions.view(np.uint64), d["positions"].view(np.uint64))
    assert np.array_equal(c.meta["triangles"].reshape(-1), d["triangles"])
    assert np.array_equal(c.edge_a, d["edge_a"]) and np.array_equal(c.edge_b, d["edge_b"])
    assert np.array_equal(c.edge_rest.view(np.uint64), d["rest"].view(np.uint64))
    assert np.array_equal(c.edge_kappa, d["edge_kappa"])
    assert np.array_equal(c.masses.view(np.uint64), d["masses"].view(np.uint64))
    assert not c.pinned.any() and not d["pinned"].any()


@pytest.mark.parametrize("path", [p for p in OBJ_FIXTURES if "tri_grid" in p or "forms" in p],
                         ids=lambda p: os.path.basename(p))
def test_obj_mesh_coloring_matches_reference(path, tmp_path):
    """Jones-Plassmann edge colouring and incidence lists of an unstructured mesh
    equal the reference's (ConstraintColoring.cpp:35-54, :121-259)."""
    d = _load(path)
    c = scenes.obj_cloth(_write(tmp_path, d["obj_text"].tobytes()), float(d["kappa"]), float(d["density"]))
    c.pinned[0] = 1
    e = Engine("fp64")
    e.add(c, SimConfig())
    t = e.topology(0)
    assert t["color_count"] == int(d["color_count"])
    for k in ("color_of", "order", "start", "vertex_start", "vertex_edges"):
        assert np.array_equal(np.asarray(t[k], np.uint32), d[k]), k
    e.close()


@pytest.mark.skipif(ref_mesh_lib() is None, reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("path", OBJ_FIXTURES, ids=[os.path.basename(p) for p in OBJ_FIXTURES])
def test_obj_fixtures_are_current(path):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import extract_obj
    d = _load(path)
    live = extract_obj(ref_mesh_lib(), str(d["name"]), d["obj_text"].tobytes().decode(), float(d["kappa"]),
                       float(d["density"]))
    for k, v in live.items():
        assert np.array_equal(np.asarray(v), d[k]), k


@pytest.mark.skipif(ref_mesh_lib() is None, reason="reference build oracle/_ref absent")
def test_large_obj_matches_live_reference(tmp_path):
    """A 60x45 triangulated sheet with wavy heights, straight from the reference build."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import _tri_grid, extract_obj
    text = _tri_grid(60, 45)
    ref = extract_obj(ref_mesh_lib(), "big", text, 2300.0, 0.9)
    c = scenes.obj_cloth(_write(tmp_path, text), 2300.0, 0.9)
    assert np.array_equal(c.edge_a, ref["edge_a"]) and np.array_equal(c.edge_b, ref["edge_b"])
    assert np.array_equal(c.edge_rest.view(np.uint64), ref["rest"].view(np.uint64))
    assert np.array_equal(c.masses.view(np.uint64), ref["masses"].view(np.uint64))


@pytest.mark.parametrize("text", ["v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",
                                  "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",
                                  "v 0 0 0\nv 1 0 0\nv 0 1 0\nf -1 -2 -3\n"])
def test_obj_out_of_range_index_is_rejected(text, tmp_path):
    """An index outside [1, n] is refused (the reference would read out of bounds)."""
    with pytest.raises(PdxError):
        scenes.obj_cloth(_write(tmp_path, text), 100.0)


def test_obj_missing_file_is_rejected(tmp_path):
    with pytest.raises(PdxError):
        scenes.obj_cloth(str(tmp_path / "absent.obj"), 100.0)


def _tri_areas(pos, tris):
    P = pos.reshape(-1, 3)
    u = P[tris[:, 1]] - P[tris[:, 0]]
    w = P[tris[:, 2]] - P[tris[:, 0]]
    return 0.5 * np.linalg.norm(np.cross(u, w), axis=1)


def test_surface_masses_closed_form(tmp_path):
    """Total mass = density x total area, and each vertex holds a third of the
    area of its triangles (numpy cross products, summed in another order)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import _tri_grid
    c = scenes.obj_cloth(_write(tmp_path, _tri_grid(12, 8)), 1000.0, density=0.8)
    tris = c.meta["triangles"]
    area = _tri_areas(c.positions, tris)
    assert c.masses.sum() == pytest.approx(0.8 * area.sum(), rel=1e-13)
    per = np.zeros(c.n)
    for k in range(3):
        np.add.at(per, tris[:, k], 0.8 * area / 3.0)
    assert np.allclose(c.masses, per, rtol=1e-13, atol=0)


def test_edge_masses_closed_form():
    """Edge-lumped masses on a regular grid of spacing l and density 1: interior
    2 l^2, side 1.5 l^2, corner l^2 (core/mesh/Mesh.cpp:160-166)."""
    l = 0.025
    c = scenes.grid_cloth(7, 5, 1.0e4, sx=l, sz=l, mass="edges")
    M = c.masses.reshape(5, 7)
    assert np.allclose(M[1:-1, 1:-1], 2 * l * l, rtol=1e-14)
    assert np.allclose(M[0, 1:-1], 1.5 * l * l, rtol=1e-14) and np.allclose(M[1:-1, 0], 1.5 * l * l, rtol=1e-14)
    assert np.allclose(M[[0, 0, -1, -1], [0, -1, 0, -1]], l * l, rtol=1e-14)


def test_mass_rules_fill_nonpositive():
    """A vertex without triangles or edges takes the smallest positive mass; with
    no positive mass at all every vertex gets 1."""
    lib = load_library()
    pos = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 5, 5, 5], dtype=np.float64)
    tris = np.array([0, 1, 2], np.int32)
    m = np.empty(4)
    assert lib.pdx_masses_from_surfaces(4, pos.ctypes.data, 1, tris.ctypes.data, 3.0, m.ctypes.data) == 0
    assert np.allclose(m[:3], 3.0 * 0.5 / 3.0) and m[3] == m[0]
    ea, eb, rest = np.array([0], np.int32), np.array([1], np.int32), np.array([0.0])
    assert lib.pdx_masses_from_edges(3, 1, ea.ctypes.data, eb.ctypes.data, rest.ctypes.data, 1.0, m.ctypes.data) == 0
    assert np.array_equal(m[:3], np.ones(3))
    bad = np.array([0, 1, 7], np.int32)
    assert lib.pdx_masses_from_surfaces(4, pos.ctypes.data, 1, bad.ctypes.data, 1.0, m.ctypes.data) != 0


def _perturbed(seed=3):
    c = scenes.grid_cloth(17, 13, 2300.0, shear=400.0, bend=30.0)
    rng = np.random.default_rng(seed)
    c.positions = c.positions + 0.01 * rng.standard_normal(c.positions.shape)
    c.velocities = 0.3 * rng.standard_normal(c.positions.shape)
    c.pin_positions = c.positions.copy()
    return c


def test_diagnostics_match_oracle_bitwise():
    """pdx_diagnostics on a perturbed state equals the oracle's restatement of
    maxRelativeStrain / meanRelativeStrain / surrogateEnergy / totalEnergy bit for bit."""
    c = _perturbed()
    cfg = SimConfig.fixed_iterations(5, 1.0 / 600.0)
    e = Engine("fp64")
    e.add(c, cfg)
    d = e.diagnostics(0)
    f = e.factor_arrays(0)
    o = Oracle(c, cfg, f["perm3"])
    assert d["max_strain"] == o.max_strain()
    assert d["mean_strain"] == o.mean_strain()
    assert d["elastic_energy"] == o.surrogate_energy()
    assert d["total_energy"] == o.total_energy()
    e.close()


def test_diagnostics_closed_forms():
    """At rest: zero strain and elastic energy.  Uniform stretch by 1+a: every
    edge strain is a, elastic energy sum k (a l)^2 / 2, and a linear field keeps
    bend stencils at zero energy.  Kinetic, potential and speed from their
    definitions (numpy, another order)."""
    c = scenes.grid_cloth(9, 7, 1500.0, shear=300.0, bend=20.0)
    g = np.array([0.0, -9.81, 0.0])
    e = Engine("fp64")
    e.add(c, SimConfig())
    d = e.diagnostics(0)
    # the bend second differences of the rest grid are zero up to rounding
    assert d["max_strain"] == 0.0 and d["mean_strain"] == 0.0 and d["elastic_energy"] < 1e-25
    assert d["total_mass"] == pytest.approx(1.0, rel=1e-14)
    a = 0.03
    x = c.positions * (1.0 + a)
    v = np.tile([0.0, 0.5, -0.25], c.n) * np.repeat(np.arange(c.n) % 3, 3)
    e.set_state(0, x, v)
    d = e.diagnostics(0)
    assert d["max_strain"] == pytest.approx(a, rel=1e-12)
    assert d["mean_strain"] == pytest.approx(a, rel=1e-12)
    el = 0.5 * (c.edge_kappa * (a * c.edge_rest) ** 2).sum()
    assert d["elastic_energy"] == pytest.approx(el, rel=1e-10)
    V = v.reshape(-1, 3)
    ke = 0.5 * (c.masses * (V * V).sum(1)).sum()
    pe = -(c.masses * (x.reshape(-1, 3) @ g)).sum()
    assert d["total_energy"] == pytest.approx(ke + el + pe, rel=1e-12)
    assert d["max_speed"] == pytest.approx(np.sqrt((V * V).sum(1)).max(), rel=1e-15)
    e.close()


def test_report_help_and_bad_preset_on_cpu():
    """The report's option parsing runs before any device call."""
    from pd_ipc_amd import report
    text, _ = report.run(["--help"])
    assert "--preset NAME" in text and "--mesh PATH" in text
    assert text.rstrip().endswith("--device D          HIP device (default 0)")  # the whole text
    with pytest.raises(PdxError) as ei:
        report.run(["--preset", "fastest"])
    assert "preview, realtime, accurate" in str(ei.value)


def test_pdx_bench_program_on_cpu():
    """The pdx_bench program (built next to the library) parses pd_bench's command
    line: usage on --help, exit status 2 and a message on an unknown preset."""
    import subprocess
    exe = os.path.join(ROOT, "pd_ipc_amd", "pdx_bench")
    if not os.path.exists(exe):
        pytest.skip("pdx_bench not built")
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--preset NAME" in r.stdout
    r = subprocess.run([exe, "--preset", "fastest"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "preview, realtime, accurate" in r.stderr
