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
    d = read_mesh(L, h)
    L.ref_free(h)
    d.update(nx=nx, ny=ny, spacing=spacing, kappa=kappa, shear=shear, bend=bend)
    return d


# OBJ inputs for Mesh::loadObj (core/mesh/Mesh.cpp:56-125), each with every face
# index inside [1, n] so that the reference reads nothing out of bounds.
def _tri_grid(nx, ny):
    lines = ["# triangulated grid, alternating diagonals, wavy heights"]
    for j in range(ny):
        for i in range(nx):
            y = 0.01 * ((i * 7 + j * 3) % 5) - 0.02
            lines.append(f"v {i * 0.05:.6f} {y:.6f} {j * 0.04:.6f}")
    for j in range(ny - 1):
        for i in range(nx - 1):
            a, b = j * nx + i + 1, j * nx + i + 2
            c, d = a + nx, b + nx
            if (i + j) % 2 == 0:
                lines += [f"f {a} {b} {d}", f"f {a} {d} {c}"]
            else:
                lines += [f"f {a} {b} {c}", f"f {b} {d} {c}"]
    return "\n".join(lines) + "\n"


OBJ_CASES = {
    "tri_grid": (_tri_grid(9, 6), 2300.0, 1.0),
    # slash forms, vt/vn lines, a quad (first three corners), a face with two
    # corners (skipped), a repeated corner, a short vertex line (z = 0), signs and
    # trailing text on indices, CRLF endings, an unreferenced vertex (mass filled)
    "forms": ("# forms\r\n"
              "v 0 0 0\r\n"
              "v 1.0 0.0 0.0\r\n"
              "  v 1 0.5 1\r\n"
              "v 0 0.25 1.5\r\n"
              "v 2.0 1e-1\r\n"
              "v 5 5 5\r\n"
              "vt 0.5 0.5\r\n"
              "vn 0 1 0\r\n"
              "\r\n"
              "f 1/1/1 2/1/1 3/1/1\r\n"
              "f 1//1 3//1 4//1\r\n"
              "f 2 5 3 4\r\n"
              "f 4 1\r\n"
              "f 3 3 5\r\n"
              "f +2 5abc 1/x\r\n"
              "o name\r\n"
              "g group\r\n", 1000.0, 0.7),
    # every triangle has zero area: every mass is filled with 1
    "degenerate": ("v 0 0 0\nv 1 0 0\nv 2 0 0\nv 3 0 0\nf 1 2 3\nf 2 3 4\n", 500.0, 1.0),
    # refused by the reference: an index without digits, no faces, no vertices
    "bad_index": ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 x 3\n", 100.0, 1.0),
    "no_faces": ("v 0 0 0\nv 1 0 0\nv 0 1 0\n", 100.0, 1.0),
    "no_vertices": ("# nothing\n", 100.0, 1.0),
}


def extract_obj(L, name, text, kappa, density):
    import tempfile
    with tempfile.NamedTemporaryFile("wb", suffix=".obj", delete=False) as f:
        f.write(text.encode())
        path = f.name
    try:
        h = L.ref_load_obj(path.encode(), kappa, density)
    finally:
        os.unlink(path)
    out = dict(name=name, obj_text=np.frombuffer(text.encode(), np.uint8), kappa=kappa, density=density,
               loads=h is not None)
    if h is None:
        return out
    T = int(L.ref_triangle_count(h))
    tris = np.zeros(3 * T, np.int32)
    L.ref_triangles(h, tris.ctypes.data_as(C.c_void_p))
    out.update(read_mesh(L, h))
    out["triangles"] = tris
    L.ref_free(h)
    return out


def read_mesh(L, h):
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
    return dict(positions=pos, masses=mass, pinned=pinned, edge_a=ea, edge_b=eb, rest=rest,
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
    for name, (text, kappa, density) in OBJ_CASES.items():
        d = extract_obj(L, name, text, kappa, density)
        path = os.path.join(HERE, f"ref_obj_{name}.npz")
        np.savez_compressed(path, **d)
        print(path, "loads", d["loads"], "edges", len(d.get("edge_a", [])))


if __name__ == "__main__":
    main()
