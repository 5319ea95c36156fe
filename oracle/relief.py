"""CPU restatement of the relief parameterisation around DiffBEM.gradient (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product never imports it.

* eval_uv          -- bilinear heightfield sample at the differentiable vertices' uvs
                      (acoustics_opt.py:24-34), including its np.modf handling of u * w - 0.5 < 0;
* eval_uv_adjoint  -- the scatter of acoustic_gradient (acoustics_opt.py:45-56): per vertex, in vertex
                      order, the four weighted contributions added to their texels;
* rect_points_to_uv (pyoptim/diffmesh.py:248-249).

Pinned in tests/test_relief.py by: exact reproduction of bilinear functions away from the border, the
adjoint identity <eval(T), g> = <T, adjoint(g)>, hand-computed taps for the negative-fraction and
clamped-border cases, and a per-vertex brute-force loop on small inputs.
"""
import numpy as np


def taps(uvs, h, w):
    """(rows, cols, weights) of the four taps per vertex, as acoustics_opt.py:26-34 forms them."""
    uvs = np.asarray(uvs, dtype=np.float64)
    uf, ui = np.modf(uvs[:, 0] * w - 0.5)
    vf, vi = np.modf(uvs[:, 1] * h - 0.5)
    uf = np.where(uf < 0, 1 + uf, uf)
    vf = np.where(vf < 0, 1 + vf, vf)
    ui = np.maximum(ui.astype(int), 0)
    vi = np.maximum(vi.astype(int), 0)
    u1 = np.minimum(ui + 1, w - 1)
    v1 = np.minimum(vi + 1, h - 1)
    rows = np.column_stack([vi, vi, v1, v1])
    cols = np.column_stack([ui, u1, ui, u1])
    wts = np.column_stack([(1 - vf) * (1 - uf), (1 - vf) * uf, vf * (1 - uf), vf * uf])
    return rows, cols, wts


def eval_uv(tex, uvs):
    """Heights at uvs (acoustics_opt.py:24-34): the four terms summed left to right."""
    tex = np.asarray(tex, dtype=np.float64)
    h, w = tex.shape
    rows, cols, wts = taps(uvs, h, w)
    vals = tex[rows, cols]
    return wts[:, 0] * vals[:, 0] + wts[:, 1] * vals[:, 1] + wts[:, 2] * vals[:, 2] + wts[:, 3] * vals[:, 3]


def eval_uv_adjoint(grad, uvs, h, w):
    """Texture gradient from per-vertex gradients (acoustics_opt.py:45-56), same accumulation order."""
    rows, cols, wts = taps(uvs, h, w)
    out = np.zeros((h, w))
    for v in range(len(grad)):
        for k in range(4):
            out[rows[v, k], cols[v, k]] += wts[v, k] * grad[v]
    return out


def rect_points_to_uv(points, x_min=-0.3, x_max=0.3, z_min=-0.3, z_max=0.3, uvmin=0.0001, uvmax=0.9999):
    """pyoptim/diffmesh.py:248-249."""
    points = np.asarray(points, dtype=np.float64)
    return np.column_stack([(points[:, 0] - x_min) / (x_max - x_min) * (uvmax - uvmin) + uvmin,
                            (points[:, 2] - z_min) / (z_max - z_min) * (uvmax - uvmin) + uvmin])
