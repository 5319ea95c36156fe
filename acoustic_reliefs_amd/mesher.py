"""Relief meshes as the reference's optimiser builds them (pyoptim/mesher.py:87-231, acoustics_opt.py:175-210).

The vertex and face numbering is kept identical to the reference's, because the differentiable vertices,
their uvs and DiffBEM's reordering all follow from it:

* generate_base_mesh -- a w x h x b box whose six faces are separate vertex grids (edge vertices are not
  shared), each quad split into two triangles (pyoptim/mesher.py:87-155);
* face_subdivision   -- midpoint subdivision of selected faces; neighbouring faces that gain a midpoint on an
  edge are split into two or three triangles so the surface stays conforming (pyoptim/mesher.py:159-231);
* relief_setup       -- the optimiser's set-up: box, top vertices, optional top subdivision, uvs.
"""
import numpy as np


def _face_grid(start, u_vec, v_vec, nu, nv, outward, base):
    """One face: (nu + 1) x (nv + 1) vertices row by row along v, two triangles per cell."""
    i = np.arange(nu + 1)
    j = np.arange(nv + 1)
    jj, ii = np.meshgrid(j, i, indexing="ij")
    P = (np.asarray(start, float)[None, :] + ii.reshape(-1, 1) * np.asarray(u_vec, float)[None, :]
         + jj.reshape(-1, 1) * np.asarray(v_vec, float)[None, :])
    cj, ci = np.meshgrid(np.arange(nv), np.arange(nu), indexing="ij")
    a = (base + cj * (nu + 1) + ci).reshape(-1)
    b, c = a + 1, a + (nu + 1)
    d = c + 1
    if outward > 0:
        t1, t2 = np.column_stack([a, c, d]), np.column_stack([a, d, b])
    else:
        t1, t2 = np.column_stack([a, d, c]), np.column_stack([a, b, d])
    F = np.empty((2 * len(a), 3), dtype=np.int64)
    F[0::2], F[1::2] = t1, t2
    return P, F


def generate_base_mesh(w, h, b, Nx, Ny, Nz):
    """Box of width w (x), height h (y) and breadth b (z), faces split into Nx, Ny, Nz segments.
    Returns vertices (V, 3) and triangles (F, 3) in the reference's order (pyoptim/mesher.py:87-155)."""
    dx, dy, dz = w / Nx, h / Ny, b / Nz
    faces = [((-w / 2, 0, -b / 2), (dx, 0, 0), (0, dy, 0), Nx, Ny, 1),
             ((-w / 2, 0, b / 2), (dx, 0, 0), (0, dy, 0), Nx, Ny, -1),
             ((-w / 2, h, -b / 2), (dx, 0, 0), (0, 0, dz), Nx, Nz, 1),
             ((-w / 2, 0, -b / 2), (dx, 0, 0), (0, 0, dz), Nx, Nz, -1),
             ((-w / 2, 0, -b / 2), (0, dy, 0), (0, 0, dz), Ny, Nz, 1),
             ((w / 2, 0, -b / 2), (0, dy, 0), (0, 0, dz), Ny, Nz, -1)]
    Ps, Fs, base = [], [], 0
    for start, u, v, nu, nv, sgn in faces:
        P, F = _face_grid(start, u, v, nu, nv, sgn, base)
        Ps.append(P)
        Fs.append(F)
        base += len(P)
    return np.vstack(Ps), np.vstack(Fs)


def face_subdivision(Ps_, Es_, faces_, n_subdivisions=1, ret_faces=False):
    """Midpoint-subdivide the triangles `faces_` n times (pyoptim/mesher.py:159-231).

    Midpoints are appended in order of first appearance over the selected faces' edges (edge order
    (i1,i2), (i2,i3), (i1,i3) per face). Every face is then rewritten in order: a selected face as
    four triangles, a face with one or two split edges as two or three, others unchanged; the
    triangle patterns are the reference's, so the numbering matches it."""
    Ps = np.asarray(Ps_, dtype=np.float64).copy()
    Es = np.asarray(Es_).copy()
    faces = list(faces_)
    for _ in range(n_subdivisions):
        mid = {}
        for f in faces:
            i1, i2, i3 = (int(x) for x in Es[f])
            for e in ((min(i1, i2), max(i1, i2)), (min(i2, i3), max(i2, i3)), (min(i1, i3), max(i1, i3))):
                if e not in mid:
                    mid[e] = len(Ps) + len(mid)
        if mid:
            ends = np.array(list(mid.keys()), dtype=np.int64)
            Ps = np.vstack([Ps, (Ps[ends[:, 0]] + Ps[ends[:, 1]]) / 2])
        selected = np.zeros(len(Es), dtype=bool)
        selected[np.asarray(faces, dtype=np.int64)] = True
        out, nfaces = [], []
        for f, (i1, i2, i3) in enumerate(Es.tolist()):
            m12 = mid.get((min(i1, i2), max(i1, i2)), -1)
            m23 = mid.get((min(i2, i3), max(i2, i3)), -1)
            m13 = mid.get((min(i1, i3), max(i1, i3)), -1)
            if selected[f]:
                nfaces.extend(range(len(out), len(out) + 4))
            split = (m12 >= 0, m23 >= 0, m13 >= 0)
            if split == (True, False, False):
                out += [[m12, i2, i3], [i1, m12, i3]]
            elif split == (False, True, False):
                out += [[i1, i2, m23], [i1, m23, i3]]
            elif split == (False, False, True):
                out += [[i1, i2, m13], [i2, i3, m13]]
            elif split == (False, True, True):
                out += [[i1, i2, m13], [i2, m23, m13], [m23, i3, m13]]
            elif split == (True, False, True):
                out += [[i1, m12, m13], [i2, m13, m12], [i2, i3, m13]]
            elif split == (True, True, False):
                out += [[i1, m12, i3], [i3, m12, m23], [m12, i2, m23]]
            elif split == (True, True, True):
                out += [[i1, m12, m13], [m12, i2, m23], [m13, m23, i3], [m12, m23, m13]]
            else:
                out.append([i1, i2, i3])
        Es = np.array(out, dtype=np.int64)
        faces = nfaces
    return (Ps, Es, faces) if ret_faces else (Ps, Es)


def rect_points_to_uv(points, x_min=-0.3, x_max=0.3, z_min=-0.3, z_max=0.3, uvmin=0.0001, uvmax=0.9999):
    """Top-face points to heightfield uvs (pyoptim/diffmesh.py:248-249)."""
    points = np.asarray(points, dtype=np.float64)
    return np.column_stack([(points[:, 0] - x_min) / (x_max - x_min) * (uvmax - uvmin) + uvmin,
                            (points[:, 2] - z_min) / (z_max - z_min) * (uvmax - uvmin) + uvmin])


def relief_setup(w=0.6, b=0.6, h=0.15, subdiv_top=0, Nx=30, Ny=7, Nz=30):
    """The optimiser's mesh set-up (acoustics_opt.py:181-194, 206-208): the box, its top vertices (the
    differentiable points), optional subdivision of the top faces, and the top vertices' uvs.
    Returns Ps, Es (input order: pass them to DiffBEM.precompute), diff_pts_idx, diff_uvs."""
    Ps, Es = generate_base_mesh(w, h, b, Nx, Ny, Nz)
    diff = np.where(np.isclose(Ps[:, 1], np.max(Ps[:, 1])))[0]
    if subdiv_top > 0:
        top = np.where(np.all(np.isin(Es, diff), axis=1))[0]
        Ps, Es = face_subdivision(Ps, Es, top, subdiv_top)
        diff = np.where(np.isclose(Ps[:, 1], np.max(Ps[:, 1])))[0]
    lo, hi = Ps.min(axis=0), Ps.max(axis=0)
    uvs = rect_points_to_uv(Ps[diff], lo[0], hi[0], lo[2], hi[2])
    return Ps, Es.astype(np.int32), diff.astype(np.int32), uvs
