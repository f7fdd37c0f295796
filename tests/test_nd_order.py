"""The nested-dissection factor order (pdx_set_ordering 'nd', fast mode) on meshes
beyond the regular grid: a closed 3-D triangle surface, coincident positions,
many pins, two disconnected pieces.  Each factor must be a factor of the same
matrix (L D L^T in rank space, mapped back through rank -> vertex, equals the
oracle's assembled component matrix to rounding) and its order a vertex
permutation with the three components of a vertex kept together.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle_ffi import Oracle
from pd_ipc_amd import Cloth, Engine, SimConfig, scenes


def _check(c, cfg):
    e = Engine("fp64", "fast")
    e.add(c, cfg)
    e.set_ordering("nd")
    f = e.factor_arrays(0)
    n = c.n
    p3 = f["perm3"]
    assert np.array_equal(np.sort(p3), np.arange(3 * n))
    assert np.array_equal(p3[0::3] // 3, f["rank_vertex"])
    assert np.array_equal(p3 % 3, np.tile([0, 1, 2], n))
    L = sp.csc_matrix((f["vals"], f["rows"], f["ptr"]), shape=(n, n)) + sp.identity(n)
    R = (L @ sp.diags(1.0 / f["dinv"]) @ L.T).toarray()
    cp, r, v = Oracle(c, cfg, p3).matrix()
    A = sp.csc_matrix((v, r, cp), shape=(n, n)).toarray()
    rv = f["rank_vertex"]
    assert np.abs(R - A[np.ix_(rv, rv)]).max() <= 1e-12 * np.abs(A).max()
    st = e.schedule_stats(0)
    e.close()
    return st


def _sphere_obj(tmp_path, nlat=14, nlon=20):
    """A closed UV sphere: two poles and nlat-1 rings of nlon vertices."""
    lines = ["v 0 1 0"]
    for i in range(1, nlat):
        th = np.pi * i / nlat
        for j in range(nlon):
            ph = 2 * np.pi * j / nlon
            lines.append(f"v {np.sin(th) * np.cos(ph):.9f} {np.cos(th):.9f} {np.sin(th) * np.sin(ph):.9f}")
    lines.append("v 0 -1 0")
    ring = lambda i, j: 2 + (i - 1) * nlon + (j % nlon)  # 1-based OBJ index
    south = 2 + (nlat - 1) * nlon
    for j in range(nlon):
        lines.append(f"f 1 {ring(1, j)} {ring(1, j + 1)}")
        lines.append(f"f {south} {ring(nlat - 1, j + 1)} {ring(nlat - 1, j)}")
    for i in range(1, nlat - 1):
        for j in range(nlon):
            a, b, c2, d = ring(i, j), ring(i, j + 1), ring(i + 1, j + 1), ring(i + 1, j)
            lines.append(f"f {a} {d} {c2}")
            lines.append(f"f {a} {c2} {b}")
    p = tmp_path / "sphere.obj"
    p.write_text("\n".join(lines) + "\n")
    return p


def test_nd_closed_surface(tmp_path):
    c = scenes.obj_cloth(_sphere_obj(tmp_path), 2300.0, 1.0, pins=(0,))
    _check(c, SimConfig.fixed_iterations(10, 1.0 / 600.0))


def test_nd_coincident_positions():
    """All vertices at one point: every split falls back to index order."""
    c = scenes.grid_cloth(23, 17, 2300.0, shear=500.0, bend=50.0)
    c.positions[:] = 0.25
    _check(c, SimConfig.fixed_iterations(10, 1.0 / 600.0))


def test_nd_many_pins():
    """Pinned rows are identity rows with no coupling: isolated vertices."""
    c0 = scenes.grid_cloth(30, 30, 2300.0, shear=500.0, bend=50.0)
    c = scenes.grid_cloth(30, 30, 2300.0, shear=500.0, bend=50.0, pins=list(range(0, c0.n, 3)))
    _check(c, SimConfig.fixed_iterations(10, 1.0 / 600.0))


def test_nd_disconnected_pieces():
    """Two grids side by side in one instance, no edge between them."""
    a = scenes.grid_cloth(12, 9, 2300.0, shear=500.0, bend=50.0)
    b = scenes.grid_cloth(10, 14, 1500.0)
    off = a.n
    pos_b = b.positions.reshape(-1, 3) + np.array([2.0, 0.0, 0.0])
    c = Cloth(positions=np.concatenate([a.positions, pos_b.reshape(-1)]),
              masses=np.concatenate([a.masses, b.masses]),
              pinned=np.concatenate([a.pinned, b.pinned]),
              edge_a=np.concatenate([a.edge_a, b.edge_a + off]),
              edge_b=np.concatenate([a.edge_b, b.edge_b + off]),
              edge_rest=np.concatenate([a.edge_rest, b.edge_rest]),
              edge_kappa=np.concatenate([a.edge_kappa, b.edge_kappa]),
              bend_a=a.bend_a, bend_b=a.bend_b, bend_c=a.bend_c, bend_k=a.bend_k)
    st = _check(c, SimConfig.fixed_iterations(10, 1.0 / 600.0))
    assert st["fwd"]["steps"] >= 1


@pytest.mark.parametrize("nx,ny", [(2, 2), (3, 5), (9, 1)])
def test_nd_tiny(nx, ny):
    """Meshes smaller than one leaf, and a single strip."""
    if ny == 1:
        c = scenes.grid_cloth(nx, 2, 2300.0)
    else:
        c = scenes.grid_cloth(nx, ny, 2300.0, shear=500.0)
    _check(c, SimConfig.fixed_iterations(10, 1.0 / 600.0))
