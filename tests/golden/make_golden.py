"""Writes the reference-mesh fixtures of tests/golden/ from the reference's own
sources (core/mesh/Mesh.cpp, core/mesh/ConstraintColoring.cpp), built into
oracle/_ref/libref_mesh.so by `make -C oracle ref`.

Each fixture holds, for one grid scene built the way makeScene does
(core/sim/Scene.cpp:39-169: makeGrid, kappa on every edge, edge-lumped masses,
optional shear diagonals and bend stencils, two top corners pinned), the
vertex positions, masses, pin mask, the edge and stencil arrays, the edge
colouring (colorOf, order, start), the vertex -> edge incidence CSR and the
vertex -> stencil CSR with weights.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_ffi import ref_mesh_lib  # noqa: E402

# (nx, ny, spacing, kappa, shear kappa, bend kappa)
SCENES = [
    (2, 2, 1.0, 100.0, 0.0, 0.0),
    (5, 5, 0.25, 2300.0, 0.0, 0.0),
    (20, 20, 1.0 / 19.0, 2300.0, 0.0, 0.0),
    (40, 40, 1.0 / 39.0, 2300.0, 0.0, 0.0),
    (40, 40, 0.025, 1.0e4, 0.0, 0.0),
    (17, 9, 0.05, 2300.0, 500.0, 50.0),
    (24, 31, 1.0 / 23.0, 1000.0, 200.0, 20.0),
]


def name_of(nx, ny, shear, bend):
    return f"ref_mesh_{nx}x{ny}_{'s' if shear > 0 else ''}{'b' if bend > 0 else ''}.npz"


def extract(L, nx, ny, spacing, kappa, shear, bend):
    h = L.ref_grid_scene(nx, ny, spacing, kappa, 1.0, shear, bend)
    cnt = np.zeros(5, np.int64)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    L.ref_counts(h, p(cnt))
    n, E, B, ncol = (int(v) for v in cnt[:4])
    pos, mass, pinned = np.zeros(3 * n), np.zeros(n), np.zeros(n, np.uint8)
    L.ref_vertices(h, p(pos), p(mass), p(pinned))
    ea, eb, rest, kap = np.zeros(E, np.int32), np.zeros(E, np.int32), np.zeros(E), np.zeros(E)
    L.ref_edges(h, p(ea), p(eb), p(rest), p(kap))
    sa, sb, sc, sk = np.zeros(B, np.int32), np.zeros(B, np.int32), np.zeros(B, np.int32), np.zeros(B)
    if B:
        L.ref_bends(h, p(sa), p(sb), p(sc), p(sk))
    color_of, order = np.zeros(E, np.uint32), np.zeros(E, np.uint32)
    start = np.zeros(ncol + 1, np.uint32)
    vstart, vedges = np.zeros(n + 1, np.uint32), np.zeros(2 * E, np.uint32)
    L.ref_coloring(h, p(color_of), p(order), p(start), p(vstart), p(vedges))
    bvstart, bsten, bw = np.zeros(n + 1, np.uint32), np.zeros(3 * B, np.uint32), np.zeros(3 * B)
    if B:
        L.ref_bend_adjacency(h, p(bvstart), p(bsten), p(bw))
    L.ref_free(h)
    return dict(nx=nx, ny=ny, spacing=spacing, kappa=kappa, shear=shear, bend=bend,
                positions=pos, masses=mass, pinned=pinned, edge_a=ea, edge_b=eb, rest=rest,
                edge_kappa=kap, bend_a=sa, bend_b=sb, bend_c=sc, bend_k=sk,
                color_of=color_of, order=order, start=start, vertex_start=vstart,
                vertex_edges=vedges, bend_vertex_start=bvstart, bend_stencils=bsten,
                bend_weights=bw, color_count=ncol)


def main():
    L = ref_mesh_lib()
    if L is None:
        sys.exit("oracle/_ref/libref_mesh.so is missing: run `make -C oracle ref` first")
    for sc in SCENES:
        d = extract(L, *sc)
        path = os.path.join(HERE, name_of(sc[0], sc[1], sc[4], sc[5]))
        if sc[3] == 1.0e4:
            path = path.replace(".npz", "_k1e4.npz")
        np.savez_compressed(path, **d)
        print(path, "colours", d["color_count"], "edges", len(d["edge_a"]), "bends", len(d["bend_a"]))


if __name__ == "__main__":
    main()
