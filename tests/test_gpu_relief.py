"""Relief parameterisation on the GPU (arh_relief_*) against the oracle's restatement of eval_uv and its
adjoint (acoustics_opt.py:24-56): bit for bit, including the negative-fraction and clamped-border taps;
the autograd Function; one acoustic gradient through DiffBEM and two optimiser iterations."""
import numpy as np
import pytest
import torch

from oracle import relief as R

pytestmark = pytest.mark.gpu


def _uvs(rng, n):
    uv = rng.uniform(0.0, 1.0, (n, 2))
    uv[:8] = [[0.0, 0.0], [1.0, 1.0], [0.0, 1.0], [1.0, 0.0], [0.0001, 0.9999], [0.9999, 0.0001], [0.5, 0.5],
              [0.01, 0.02]]
    return uv


@pytest.mark.parametrize("h,w,n", [(16, 16, 300), (130, 97, 5000), (1, 5, 40)])
def test_eval_and_adjoint_bit_exact(ctx, capi, h, w, n):
    from acoustic_reliefs_amd import relief
    rng = np.random.default_rng(h * 1000 + w)
    uv = _uvs(rng, n)
    print("relief taps h=%d w=%d n=%d" % (h, w, n), flush=True)
    s = relief.HeightfieldSampler(uv, (h, w), ctx=ctx)
    tex = rng.standard_normal((h, w))
    g = rng.standard_normal(n)
    hv = s.heights(torch.from_numpy(tex)).cpu().numpy()
    assert np.array_equal(hv, R.eval_uv(tex, uv))
    tg = s.tex_gradient(torch.from_numpy(g)).cpu().numpy()
    assert np.array_equal(tg, R.eval_uv_adjoint(g, uv, h, w))


def test_autograd_function(ctx, capi):
    from acoustic_reliefs_amd import relief
    rng = np.random.default_rng(3)
    uv = _uvs(rng, 200)
    s = relief.HeightfieldSampler(uv, (20, 24), ctx=ctx)
    tex = torch.from_numpy(rng.standard_normal((20, 24))).cuda().requires_grad_(True)
    g = rng.standard_normal(200)
    hts = relief.sample_heights(tex, s)
    (hts * torch.from_numpy(g).cuda()).sum().backward()
    assert np.array_equal(tex.grad.cpu().numpy(), R.eval_uv_adjoint(g, uv, 20, 24))


def test_acoustic_gradient_and_optimiser(ctx, capi):
    """acoustic_gradient = (1 - value, -adjoint(DiffBEM gradient)) on a small box relief, and two
    optimiser iterations run with finite terms."""
    from acoustic_reliefs_amd import acoustics3d_impl as m
    from acoustic_reliefs_amd import mesher, relief
    P, E, diff, uvs = mesher.relief_setup(0.3, 0.3, 0.08, subdiv_top=0, Nx=10, Ny=2, Nz=10)
    d = m.DiffBEM(32, 1.5, [1000.0], 1, 1e-6, 1e-6, 1e-8, np.array([0.0, 100.0, 0.0]), 50, 15, False,
                  speed_of_sound=343.0)
    d.silent = True
    d.precompute(P, E, diff)
    res, border = 12, 2
    s = relief.HeightfieldSampler(uvs, (res, res), ctx=ctx)
    rng = np.random.default_rng(4)
    tex = np.zeros((res, res))
    tex[border:-border, border:-border] = 0.01 * rng.uniform(size=(res - 2 * border, res - 2 * border))
    v, tg = relief.acoustic_gradient(torch.from_numpy(tex), d, s)
    val, grad = d.gradient(R.eval_uv(tex, uvs))
    assert v == 1 - val
    assert np.array_equal(tg.cpu().numpy(), -R.eval_uv_adjoint(grad, uvs, res, res))
    final, hist = relief.optimise(d, s, iters=2, lr=1e-3, edge_border=border, vmax=0.05)
    assert final.shape == (res, res) and np.all(final[:border] == 0) and np.all(np.isfinite(final))
    assert all(np.all(np.isfinite(hist[k])) for k in hist) and set(hist) == {"ac", "sm", "ng", "ba"}
