"""Synthetic meshes and workload definitions for the BEM path.

* ``icosphere``      -- sound-hard sphere test body (outward normals, n = ab x ac).
* ``relief_box``     -- closed box with a heightmap top (the relief), y axis up, the top grid
                        welded to the walls so height changes keep the surface closed.
* ``smooth_heights`` -- seeded white noise, Gaussian-smoothed, min-max scaled.
* ``read_obj`` / ``write_obj`` -- minimal ``v``/``f`` OBJ I/O (read_obj.cpp:8-64 semantics:
                        1-based faces, triangles only).
* incidence directions and the 5 m receiver semicircle used by the relief workloads
  (misc/pressure_calcs.py:40-49 protocol).

Elements use the reference's conventions: vertices (V,3) float64 row-major, faces (N,3) int32,
normals ab x ac pointing out of the body (constants.h:30-31; helpers.cpp:16-32).
"""
import numpy as np

THIRD_OCTAVE_BANDS = [250, 315, 400, 500, 630, 800, 1000, 1250, 1600, 2000, 2500, 3150, 4000]


def icosphere(subdiv=3, radius=1.0):
    """Icosphere with 20 * 4**subdiv triangles and 10 * 4**subdiv + 2 vertices."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                mid[key] = len(V) - 1
            return mid[key]

        nF = []
        for a, b, c in F:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nF += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nF
    P = np.array(V) * radius
    E = np.array(F, dtype=np.int32)
    return P, _orient_outward_convex(P, E)


def _orient_outward_convex(P, E):
    a, b, c = P[E[:, 0]], P[E[:, 1]], P[E[:, 2]]
    n = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", n, (a + b + c) / 3.0) < 0
    E = E.copy()
    E[flip] = E[flip][:, [0, 2, 1]]
    return E


def smooth_heights(n, seed=0, sigma=2.0, hmax=0.1):
    """Seeded N(0,1) n x n noise, Gaussian-smoothed (sigma cells, symmetric padding, radius
    int(4 sigma + 0.5)), min-max scaled to [0, hmax] metres."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((n, n))
    rad = int(4.0 * sigma + 0.5)
    xs = np.arange(-rad, rad + 1, dtype=np.float64)
    w = np.exp(-0.5 * (xs / sigma) ** 2)
    w /= w.sum()
    for axis in (0, 1):
        zp = np.pad(z, [(rad, rad) if a == axis else (0, 0) for a in range(2)], mode="symmetric")
        acc = np.zeros_like(z)
        for i, wi in enumerate(w):
            sl = [slice(None), slice(None)]
            sl[axis] = slice(i, i + n)
            acc += wi * zp[tuple(sl)]
        z = acc
    z = (z - z.min()) / (z.max() - z.min())
    return z * hmax


def relief_box(width=0.6, depth=0.6, base=0.05, grid=64, heights=None, wall_rows=2, bottom_grid=None):
    """Closed box whose top is a grid x grid vertex heightmap at y = base + heights[iz, ix].

    Top: (grid-1)^2 * 2 triangles. Walls: wall_rows rows of (grid-1) quads per side, their top
    row welded to the top grid's edge vertices. Bottom: bottom_grid^2 * 2 triangles at y = 0
    (T-junctions with the wall bottoms keep the surface geometrically closed).
    Returns P (V,3), E (N,3) int32 and top (grid*grid,) vertex ids, row-major over (iz, ix).
    """
    if heights is None:
        heights = np.zeros((grid, grid))
    heights = np.asarray(heights, dtype=np.float64)
    assert heights.shape == (grid, grid)
    if bottom_grid is None:
        bottom_grid = max(2, (grid - 1) // 3)
    xs = np.linspace(-width / 2, width / 2, grid)
    zs = np.linspace(-depth / 2, depth / 2, grid)
    P = []

    def add(p):
        P.append(p)
        return len(P) - 1

    top = np.zeros((grid, grid), dtype=np.int64)
    for iz in range(grid):
        for ix in range(grid):
            top[iz, ix] = add((xs[ix], base + heights[iz, ix], zs[iz]))

    quads, outs = [], []

    def quad(a, b, c, d, outward):
        # quad a-b-c-d (in order around the boundary), split into (a, b, c) and (a, c, d); the winding
        # of each triangle is chosen at the end so that its normal matches outward
        quads.append((a, b, c, d))
        outs.append(outward)

    for iz in range(grid - 1):
        for ix in range(grid - 1):
            quad(top[iz, ix], top[iz, ix + 1], top[iz + 1, ix + 1], top[iz + 1, ix], np.array([0.0, 1.0, 0.0]))

    # walls: (edge vertex ids along the edge, outward direction)
    edges = [(top[0, :], np.array([0.0, 0.0, -1.0])), (top[-1, :], np.array([0.0, 0.0, 1.0])),
             (top[:, 0], np.array([-1.0, 0.0, 0.0])), (top[:, -1], np.array([1.0, 0.0, 0.0]))]
    for ids, out in edges:
        cols = []
        for v in ids:
            x, ytop, z = P[v]
            col = [add((x, ytop * r / wall_rows, z)) for r in range(wall_rows)] + [int(v)]
            cols.append(col)
        for c in range(len(ids) - 1):
            for r in range(wall_rows):
                quad(cols[c][r], cols[c + 1][r], cols[c + 1][r + 1], cols[c][r + 1], out)

    bx = np.linspace(-width / 2, width / 2, bottom_grid + 1)
    bz = np.linspace(-depth / 2, depth / 2, bottom_grid + 1)
    bot = np.zeros((bottom_grid + 1, bottom_grid + 1), dtype=np.int64)
    for iz in range(bottom_grid + 1):
        for ix in range(bottom_grid + 1):
            bot[iz, ix] = add((bx[ix], 0.0, bz[iz]))
    for iz in range(bottom_grid):
        for ix in range(bottom_grid):
            quad(bot[iz, ix], bot[iz, ix + 1], bot[iz + 1, ix + 1], bot[iz + 1, ix], np.array([0.0, -1.0, 0.0]))
    P = np.array(P, dtype=np.float64)
    q = np.array(quads, dtype=np.int64)
    out = np.repeat(np.array(outs), 2, axis=0)
    tris = np.empty((2 * len(q), 3), dtype=np.int64)
    tris[0::2] = q[:, [0, 1, 2]]
    tris[1::2] = q[:, [0, 2, 3]]
    n = np.cross(P[tris[:, 1]] - P[tris[:, 0]], P[tris[:, 2]] - P[tris[:, 0]])
    flip = np.einsum("ij,ij->i", n, out) <= 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return P, tris.astype(np.int32), top.reshape(-1).astype(np.int32)


def relief_config(size=0.6, grid=64, hmax=0.1, seed=0, base=0.05):
    """The relief workloads: size x size box, base 0.05 m, grid x grid top from smoothed noise."""
    h = smooth_heights(grid, seed=seed, sigma=2.0, hmax=hmax)
    return relief_box(width=size, depth=size, base=base, grid=grid, heights=h)


def incidence_directions(n=37):
    """Plane-wave propagation directions: sources on the upper x-y half circle at 0..180 deg
    (5 deg steps for n=37), travelling towards the origin (d = -(cos t, sin t, 0))."""
    t = np.linspace(0.0, np.pi, n)
    return -np.column_stack([np.cos(t), np.sin(t), np.zeros_like(t)])


def semicircle_receivers(radius=5.0, n=181):
    """Receivers (r cos t, r sin t, 0), t = 0..180 deg in 1 deg steps (misc/pressure_calcs.py:46-49)."""
    t = np.linspace(0.0, np.pi, n)
    return np.column_stack([np.cos(t) * radius, np.sin(t) * radius, np.zeros_like(t)])


def read_obj(path):
    """Minimal OBJ reader (read_obj.cpp:8-45): 'v x y z' and 'f i j k' (1-based) lines only."""
    P, E = [], []
    with open(path) as f:
        for line in f:
            if not line:
                continue
            if line[0] == "v" and (len(line) > 1 and line[1] in " \t"):
                parts = line.split()
                if len(parts) < 4:
                    raise ValueError("Invalid vertex specification")
                P.append((float(parts[1]), float(parts[2]), float(parts[3])))
            elif line[0] == "f":
                parts = line.split()
                if len(parts) < 4:
                    raise ValueError("Invalid face specification (only triangles supported)")
                E.append(tuple(int(p.split("/")[0]) - 1 for p in parts[1:4]))
    return np.array(P, dtype=np.float64), np.array(E, dtype=np.int32)


def write_obj(path, P, E):
    """Minimal OBJ writer (read_obj.cpp:47-60), full precision."""
    with open(path, "w") as f:
        for p in P:
            f.write("v %.17g %.17g %.17g\n" % tuple(p))
        for e in E:
            f.write("f %d %d %d\n" % tuple(int(i) + 1 for i in e))
