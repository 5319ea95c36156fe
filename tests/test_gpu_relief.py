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


def _small_optimiser_setup(ctx):
    from acoustic_reliefs_amd import acoustics3d_impl as m
    from acoustic_reliefs_amd import mesher, relief
    P, E, diff, uvs = mesher.relief_setup(0.3, 0.3, 0.08, subdiv_top=0, Nx=10, Ny=2, Nz=10)
    d = m.DiffBEM(32, 1.5, [1000.0], 1, 1e-6, 1e-6, 1e-8, np.array([0.0, 100.0, 0.0]), 50, 15, False,
                  speed_of_sound=343.0)
    d.silent = True
    d.precompute(P, E, diff)
    return d, relief.HeightfieldSampler(uvs, (12, 12), ctx=ctx), uvs


def test_gradient_device_same_bits_as_host(ctx, capi):
    """DiffBEM.gradient_device (heights and gradient on the GPU) gives gradient(x)'s value and gradient bit for
    bit, for one band and for a multi-band, three-frequency DiffBEM."""
    from acoustic_reliefs_amd import acoustics3d_impl as m
    d, s, uvs = _small_optimiser_setup(ctx)
    rng = np.random.default_rng(5)
    x = 0.005 * rng.standard_normal(len(uvs))
    for dd in (d, None):
        if dd is None:
            from acoustic_reliefs_amd import mesher
            P, E, diff, _ = mesher.relief_setup(0.3, 0.3, 0.08, subdiv_top=0, Nx=10, Ny=2, Nz=10)
            dd = m.DiffBEM(32, 1.5, [800.0, 1600.0], 3, 1e-6, 1e-6, 1e-8, speed_of_sound=343.0)
            dd.silent = True
            dd.precompute(P, E, diff)
        v, g = dd.gradient(x)
        xt = torch.from_numpy(x).cuda()
        gt = torch.empty_like(xt)
        torch.cuda.synchronize()
        vd = dd.gradient_device(xt.data_ptr(), gt.data_ptr())
        assert vd == v and np.array_equal(gt.cpu().numpy(), g)


def test_optimiser_device_route_same_history(ctx, capi):
    """optimise() with the acoustic term on the GPU (no host copy of heights and gradients) gives the numpy
    route's history and final heightfield bit for bit."""
    from acoustic_reliefs_amd import relief
    d, s, _ = _small_optimiser_setup(ctx)
    out = {}
    for device in (True, False):
        final, hist = relief.optimise(d, s, iters=3, lr=1e-3, init_val=0.01, edge_border=2, vmax=0.05, device=device)
        out[device] = (final, hist)
    assert np.array_equal(out[True][0], out[False][0])
    assert all(np.array_equal(out[True][1][k], out[False][1][k]) for k in out[True][1])


def _laplacian(t):
    p = np.pad(t, 1)
    return p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:] - 4 * t


def test_optimiser_composite_gradient_definition(ctx, capi):
    """One optimiser iteration against its definition (acoustics_opt.py:296-345), every term computed
    independently: the acoustic term from DiffBEM.gradient and the restated eval_uv adjoint, the smoothness
    gradient (2/n) L(L(t)) of the mean squared zero-padded Laplacian (L symmetric), the negativity term (zero for
    positive heights), the barrier gradient 1 / (n (v_l - t)) for 0 < t < v_l; each normalised over the interior,
    weighted (ac 5, sm 10, ng 0.5, ba 0.5) and summed; then Adam's first step var - lr g / (|g| + 1e-8)."""
    from acoustic_reliefs_amd import relief
    d, s, uvs = _small_optimiser_setup(ctx)
    res, border, lr, h0, vmax = 12, 2, 1e-3, 0.01, 0.05
    crop = (slice(border, res - border), slice(border, res - border))
    full = np.zeros((res, res))
    full[crop] = h0
    n = res * res
    val, grad = d.gradient(R.eval_uv(full, uvs))
    terms = {"ac": -R.eval_uv_adjoint(grad, uvs, res, res)[crop],
             "sm": (2.0 / n) * _laplacian(_laplacian(full))[crop],
             "ng": np.zeros((res - 2 * border, res - 2 * border)),
             "ba": np.full((res - 2 * border, res - 2 * border), 1.0 / (n * (vmax + 1e-3 - h0)))}
    w = {"ac": 5.0, "sm": 10.0, "ng": 0.5, "ba": 0.5}
    total = np.zeros_like(terms["ac"])
    for k, g in terms.items():
        nrm = np.linalg.norm(g)
        total += w[k] * (g / nrm if nrm else g)
    expect = h0 - lr * total / (np.abs(total) + 1e-8)
    final, hist = relief.optimise(d, s, iters=1, lr=lr, init_val=h0, edge_border=border, vmax=vmax)
    assert hist["ac"][0] == 1 - val
    np.testing.assert_allclose(final[crop], expect, rtol=0, atol=1e-12)
    assert np.all(final[:border] == 0) and np.all(final[:, -border:] == 0)
