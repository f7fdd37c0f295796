"""The grid builder, edge colouring and incidence lists against the reference's own
Mesh / ConstraintColoring sources (fixtures in tests/golden/ written by
tests/golden/make_golden.py from oracle/_ref), and the C ABI's exported symbols.
CPU only."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from pd_ipc_amd import Engine, SimConfig, scenes
from pd_ipc_amd.pdx import LIB_PATH
from oracle_ffi import ORACLE_LIB, ROOT, Oracle, ref_mesh_lib

FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "ref_mesh_*.npz")))


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _cloth(d):
    nx, ny, s = int(d["nx"]), int(d["ny"]), float(d["spacing"])
    return scenes.grid_cloth(nx, ny, float(d["kappa"]), shear=float(d["shear"]), bend=float(d["bend"]),
                             sx=s, sz=s, mass="edges")


def test_fixtures_present():
    assert len(FIXTURES) >= 7


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_grid_builder_matches_reference(path):
    """Positions, edge order, rest lengths, stiffness, stencils, masses and pin
    mask of the builder equal the reference's makeGrid / addShearDiagonals /
    addBendingStencils / computeMassesFromEdges bit for bit."""
    d = _load(path)
    c = _cloth(d)
    assert np.array_equal(c.positions, d["positions"])
    assert np.array_equal(c.edge_a, d["edge_a"]) and np.array_equal(c.edge_b, d["edge_b"])
    assert np.array_equal(c.edge_rest, d["rest"])
    assert np.array_equal(c.edge_kappa, d["edge_kappa"])
    assert np.array_equal(c.bend_a, d["bend_a"]) and np.array_equal(c.bend_b, d["bend_b"])
    assert np.array_equal(c.bend_c, d["bend_c"]) and np.array_equal(c.bend_k, d["bend_k"])
    assert np.array_equal(c.masses, d["masses"])
    assert np.array_equal(c.pinned, d["pinned"])


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_library_topology_matches_reference(path):
    """The library's colouring (colour of each edge, colour-sorted order, colour
    ranges) and incidence CSRs equal the reference's arrays exactly; they fix the
    summation order of the right-hand side."""
    d = _load(path)
    e = Engine("fp64")
    e.add(_cloth(d), SimConfig.fixed_iterations(1, 1.0 / 600.0))
    t = e.topology(0)
    assert t["color_count"] == int(d["color_count"])
    for k in ("color_of", "order", "start", "vertex_start", "vertex_edges"):
        assert np.array_equal(t[k], d[k]), k
    if len(d["bend_a"]):
        assert np.array_equal(t["bend_vertex_start"], d["bend_vertex_start"])
        assert np.array_equal(t["bend_stencils"], d["bend_stencils"])
        # weights {1, -2, 1} by the vertex's place in its stencil
        sten = d["bend_stencils"].astype(np.int64)
        bs = d["bend_vertex_start"].astype(np.int64)
        owner = np.repeat(np.arange(len(bs) - 1), np.diff(bs))
        w = np.where(d["bend_b"][sten] == owner, -2.0, 1.0)
        assert np.array_equal(w, d["bend_weights"])
    e.close()


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_coloring_matches_reference(path):
    d = _load(path)
    c = _cloth(d)
    e = Engine("fp64")
    e.add(c, SimConfig.fixed_iterations(1, 1.0 / 600.0))
    o = Oracle(c, SimConfig.fixed_iterations(1, 1.0 / 600.0), e.factor_arrays(0)["perm3"])
    color_of, order, start = o.coloring()
    assert np.array_equal(color_of, d["color_of"])
    assert np.array_equal(order, d["order"])
    assert np.array_equal(start, d["start"])
    e.close()


@pytest.mark.skipif(ref_mesh_lib() is None, reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_fixtures_are_current(path):
    """Where the reference build is present, the committed fixtures are what it
    produces now."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import extract
    d = _load(path)
    live = extract(ref_mesh_lib(), int(d["nx"]), int(d["ny"]), float(d["spacing"]), float(d["kappa"]),
                   float(d["shear"]), float(d["bend"]))
    for k, v in live.items():
        assert np.array_equal(np.asarray(v), d[k]), k


def _declared(header, prefix):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def test_c_abi_exports_every_declared_symbol():
    names = _declared(os.path.join(ROOT, "include", "pdx.h"), "pdx_")
    assert len(names) >= 25
    lib = C.CDLL(LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_every_declared_symbol():
    names = _declared(os.path.join(ROOT, "oracle", "pd_oracle.h"), "orc_")
    lib = C.CDLL(ORACLE_LIB)
    assert not [n for n in names if not hasattr(lib, n)]
