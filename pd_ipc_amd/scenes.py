"""Benchmark and test scenes: regular-grid cloths of the five workloads, and
triangle-mesh cloths read from OBJ files.

Grid geometry, edge order, shear diagonals and bend stencils come from the
library's own builder (``pdx_grid_fill``, restating core/mesh/Mesh.cpp:12-54,
:183-207, :219-285).  The cloth is 1 m x 1 m in the xz-plane at y = 0 with
spacing 1/(nx-1) along x and 1/(ny-1) along z, total mass 1 kg lumped uniformly
(m_v = 1/n), gravity -9.81 y, damping 0.02 (core/sim/Scene.h:37).  OBJ meshes
go through ``pdx_obj_fill`` and ``pdx_masses_from_surfaces`` (Mesh.cpp:56-153).

Random draws use a seeded splitmix64 stream (uniforms = top 53 bits / 2^53,
normals by Box-Muller), so every instance is reproducible from (seed, index).
"""
import ctypes as C
import math
import os

import numpy as np

from .pdx import Cloth, PdxError, SimConfig, load_library

H600 = 1.0 / 600.0
_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) with Box-Muller normals."""

    def __init__(self, seed):
        self.s = seed & _MASK

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def uniforms(self, count):
        """`count` consecutive uniforms of the stream (vectorised splitmix64)."""
        k = np.arange(1, count + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.s) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.s = (self.s + count * 0x9E3779B97F4A7C15) & _MASK
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normals(self, count):
        """Box-Muller over pairs of consecutive uniforms."""
        m = (count + 1) // 2
        u = self.uniforms(2 * m)
        u1 = np.maximum(u[0::2], 1e-300)
        r = np.sqrt(-2.0 * np.log(u1))
        t = 2.0 * np.pi * u[1::2]
        out = np.empty(2 * m)
        out[0::2] = r * np.cos(t)
        out[1::2] = r * np.sin(t)
        return out[:count]


def grid_cloth(nx, ny, kappa, shear=0.0, bend=0.0, sx=None, sz=None, pins="top", mass="uniform"):
    """Regular grid cloth.  pins: 'top' (two top corners, core/sim/Scene.cpp:161-168),
    'corners' (four corners), or an explicit index list.  mass: 'uniform' (1 kg
    total) or 'edges' (Mesh::computeMassesFromEdges with density 1, Mesh.cpp:155-181)."""
    lib = load_library()
    sx = 1.0 / (nx - 1) if sx is None else sx
    sz = 1.0 / (ny - 1) if sz is None else sz
    n, E, B = C.c_int(), C.c_int(), C.c_int()
    lib.pdx_grid_counts(nx, ny, int(shear > 0), int(bend > 0), C.byref(n), C.byref(E), C.byref(B))
    n, E, B = n.value, E.value, B.value
    pos = np.empty(3 * n)
    ea, eb = np.empty(E, np.int32), np.empty(E, np.int32)
    rest, kap = np.empty(E), np.empty(E)
    sa, sb, sc = np.empty(B, np.int32), np.empty(B, np.int32), np.empty(B, np.int32)
    bk = np.empty(B)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = lib.pdx_grid_fill(nx, ny, sx, sz, kappa, shear, bend, p(pos), p(ea), p(eb), p(rest), p(kap),
                           p(sa) if B else None, p(sb) if B else None, p(sc) if B else None,
                           p(bk) if B else None)
    assert rc == 0
    if mass == "uniform":
        masses = np.full(n, 1.0 / n)
    else:
        # edge-lumped masses over the structural edges, before shear is added
        # (core/sim/Scene.cpp:62-68 then :80-92)
        masses = np.empty(n)
        ns = 2 * nx * ny - nx - ny
        rc = lib.pdx_masses_from_edges(n, ns, p(ea), p(eb), p(rest), 1.0, p(masses))
        assert rc == 0
    pinned = np.zeros(n, np.uint8)
    if isinstance(pins, str):
        top = (ny - 1) * nx
        idx = [top, top + nx - 1] if pins == "top" else [0, nx - 1, top, top + nx - 1]
    else:
        idx = list(pins)
    pinned[idx] = 1
    return Cloth(positions=pos, masses=masses, pinned=pinned, edge_a=ea, edge_b=eb, edge_rest=rest,
                 edge_kappa=kap, bend_a=sa, bend_b=sb, bend_c=sc, bend_k=bk,
                 meta=dict(nx=nx, ny=ny, kappa=kappa, shear=shear, bend=bend))


def obj_cloth(path, kappa, density=1.0, pins=()):
    """Triangle-mesh cloth from a Wavefront OBJ file, as makeScene builds it for a
    mesh path (core/sim/Scene.cpp:44-68): Mesh::loadObj (unique triangle edges,
    kappa on every edge), masses from the triangle areas with the given density,
    no shear diagonals or bend stencils, and no pins unless given.  The triangle
    corners (0-based) are kept in meta['triangles']."""
    lib = load_library()
    path_b = os.fsencode(path)
    n, E, T = C.c_int(), C.c_int(), C.c_int()
    rc = lib.pdx_obj_counts(path_b, C.byref(n), C.byref(E), C.byref(T))
    if rc != 0:
        raise PdxError(rc, f"{path}: not a loadable OBJ triangle mesh")
    n, E, T = n.value, E.value, T.value
    pos, tris = np.empty(3 * n), np.empty(3 * T, np.int32)
    ea, eb, rest, kap = np.empty(E, np.int32), np.empty(E, np.int32), np.empty(E), np.empty(E)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = lib.pdx_obj_fill(path_b, kappa, p(pos), p(tris), p(ea), p(eb), p(rest), p(kap))
    if rc != 0:
        raise PdxError(rc, f"{path}: not a loadable OBJ triangle mesh")
    masses = np.empty(n)
    rc = lib.pdx_masses_from_surfaces(n, p(pos), T, p(tris), density, p(masses))
    if rc != 0:
        raise PdxError(rc, "masses_from_surfaces")
    pinned = np.zeros(n, np.uint8)
    pinned[list(pins)] = 1
    return Cloth(positions=pos, masses=masses, pinned=pinned, edge_a=ea, edge_b=eb, edge_rest=rest,
                 edge_kappa=kap, meta=dict(kappa=kappa, density=density, triangles=tris.reshape(-1, 3),
                                           path=str(path)))


def bench_scene_40():
    """The reference's pd_bench scene (app/bench/main.cpp:61, core/sim/Scene.h):
    40x40, spacing 0.025, kappa 1e4, edge-lumped masses, dt 1/120, damping 0.02,
    two top corners pinned; 'realtime' preset = 40 iterations, residual tol 0.3."""
    c = grid_cloth(40, 40, 1.0e4, sx=0.025, sz=0.025, mass="edges")
    cfg = SimConfig(dt=1.0 / 120.0, max_iterations=40, residual_tol=0.3)
    return c, cfg


def config1():
    """40x40, distance only kappa 2300, h 1/600, 10 iterations (600 substeps)."""
    return grid_cloth(40, 40, 2300.0), SimConfig.fixed_iterations(10, H600), 600


def config2(seed=2):
    """100x100, kappa 2300 / shear 500 / bend 50, sigma 1e-3 m position noise,
    h 1/600, 40 iterations (300 substeps)."""
    c = grid_cloth(100, 100, 2300.0, shear=500.0, bend=50.0)
    rng = SplitMix64(seed)
    c.positions = c.positions + 1e-3 * rng.normals(3 * c.n)
    c.pin_positions = c.positions.copy()
    c.meta["seed"] = seed
    return c, SimConfig.fixed_iterations(40, H600), 300


def config3():
    """512x512, kappa 2300 / shear 500 / bend 50, four corners, 40 iterations (100 substeps)."""
    return grid_cloth(512, 512, 2300.0, shear=500.0, bend=50.0, pins="corners"), \
        SimConfig.fixed_iterations(40, H600), 100


def boundary_vertices(nx, ny):
    b = [j * nx + i for j in range(ny) for i in range(nx) if i in (0, nx - 1) or j in (0, ny - 1)]
    return b


def config4_instance(seed, index):
    """Instance `index` of the 512-instance mix: sides ~ U{16..96}, kappa
    log-uniform [1e2, 1e4], shear (0.2 kappa) and bend (0.02 kappa) each with p=0.5,
    1-4 distinct boundary pins, v0 ~ N(0, 0.1^2) per component."""
    rng = SplitMix64((seed * 0x100000001B3) ^ (index * 0x9E3779B97F4A7C15))
    nx = rng.randint(16, 96)
    ny = rng.randint(16, 96)
    kappa = 10.0 ** (2.0 + 2.0 * rng.uniform())
    shear = 0.2 * kappa if rng.uniform() < 0.5 else 0.0
    bend = 0.02 * kappa if rng.uniform() < 0.5 else 0.0
    npin = rng.randint(1, 4)
    bnd = boundary_vertices(nx, ny)
    pins = []
    while len(pins) < npin:
        v = bnd[rng.randint(0, len(bnd) - 1)]
        if v not in pins:
            pins.append(v)
    c = grid_cloth(nx, ny, kappa, shear=shear, bend=bend, pins=pins)
    v0 = 0.1 * rng.normals(3 * c.n)
    v0.reshape(-1, 3)[c.pinned != 0] = 0.0
    c.velocities = v0
    c.meta.update(seed=seed, index=index)
    return c


def config4(seed=4, count=512):
    cloths = [config4_instance(seed, i) for i in range(count)]
    return cloths, SimConfig.fixed_iterations(40, H600), 200


def pin_trajectory(cloth, nsub, h, amplitude=0.05, freq=2.0, phase=0.0, axis=(0.0, 1.0, 0.0)):
    """Pin targets of substeps 1..nsub: q(t) = q0 + amplitude sin(2 pi freq t + phase) axis,
    t = (s + 1) h, pins in ascending vertex order.  Shape (nsub, pins, 3)."""
    pins = np.flatnonzero(cloth.pinned)
    q0 = cloth.pin_positions.reshape(-1, 3)[pins]
    t = (np.arange(nsub) + 1) * h
    off = amplitude * np.sin(2.0 * np.pi * freq * t + phase)
    return q0[None, :, :] + off[:, None, None] * np.asarray(axis, np.float64)[None, None, :]


def config5(seed=5, count=64, side=128, nsub=120, iters=20):
    """Differentiable batch: `count` side x side cloths, distance + shear + bend
    with one stiffness scale kappa per instance (edges kappa, shear 0.2 kappa,
    bends 0.02 kappa, the batch's fractions), two top corners pinned.

    Returns (cloths at the starting guess, config, nsub, hidden) where the
    starting guess is kappa = 2300, v0 = 0 and static pins, and `hidden` holds
    per instance the seeded kappa ~ U[1e3, 5e3] and the seeded pin trajectory
    (amplitude 0.05 m, 2 Hz, vertical, seeded phase) that produce the targets."""
    cfg = SimConfig.fixed_iterations(iters, H600)
    cloths, hidden = [], []
    for i in range(count):
        rng = SplitMix64((seed * 0x100000001B3) ^ (i * 0x9E3779B97F4A7C15))
        k_hidden = 1.0e3 + 4.0e3 * rng.uniform()
        phase = 2.0 * np.pi * rng.uniform()
        c = grid_cloth(side, side, 2300.0, shear=0.2 * 2300.0, bend=0.02 * 2300.0)
        c.meta.update(seed=seed, index=i, kappa=2300.0)
        cloths.append(c)
        hidden.append(dict(kappa=k_hidden, traj=pin_trajectory(c, nsub, H600, phase=phase)))
    return cloths, cfg, nsub, hidden
