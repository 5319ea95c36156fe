"""Relief parameterisation and the acoustics-only optimiser around DiffBEM (acoustics_opt.py).

* HeightfieldSampler -- eval_uv (acoustics_opt.py:24-34) and its adjoint, the texture-gradient scatter of
  acoustic_gradient (acoustics_opt.py:45-56), on the device through the C-ABI (arh_relief_*), bit for bit;
* sample_heights     -- the same as a torch.autograd.Function (texture -> vertex heights);
* acoustic_gradient  -- acoustics_opt.py:37-60: heights from the texture, DiffBEM.gradient, the texture
  gradient; returns (1 - value, -texture gradient);
* smoothness, neg_relu, barrier_loss, preprocess_tex, normalize_gradients -- the regularisers and
  helpers of the optimiser (acoustics_opt.py:67-104);
* optimise           -- the acoustics-only loop of optim_proc (acoustics_opt.py:238-347): weighted,
  normalised gradients of the acoustic term and the regularisers, one Adam step per iteration.

Textures are float64 torch tensors on the GPU. Rendering and image/text guidance (Mitsuba, CLIP, LPIPS)
are not part of this package.
"""
import numpy as np
import torch

from . import capi


class HeightfieldSampler:
    """Heights of the differentiable vertices from a tex_h x tex_w heightfield, for fixed uvs."""

    def __init__(self, uvs, tex_shape, ctx=None, device=0):
        self.ctx = ctx or capi.Context(device)
        self.uvs = np.ascontiguousarray(uvs, dtype=np.float64).reshape(-1, 2)
        self.shape = (int(tex_shape[0]), int(tex_shape[1]))
        self.device = torch.device("cuda", getattr(self.ctx, "device", device))
        self.r = capi.Relief(self.ctx, self.uvs, *self.shape)

    def _in(self, t, shape):
        t = t.detach().to(self.device, torch.float64).contiguous()
        if tuple(t.shape) != shape:
            raise ValueError("expected shape %s, got %s" % (shape, tuple(t.shape)))
        # the library runs on its own stream: whatever produced t on torch's stream must be done
        torch.cuda.current_stream(self.device).synchronize()
        return t

    def heights(self, tex):
        """(n,) heights at the uvs (eval_uv)."""
        tex = self._in(tex, self.shape)
        out = torch.empty(len(self.uvs), dtype=torch.float64, device=self.device)
        self.r.eval(tex, out)
        return out

    def tex_gradient(self, grad):
        """(tex_h, tex_w) texture gradient from the (n,) per-vertex gradient."""
        grad = self._in(grad, (len(self.uvs),))
        out = torch.empty(self.shape, dtype=torch.float64, device=self.device)
        self.r.adjoint(grad, out)
        return out


class SampleHeights(torch.autograd.Function):
    """texture (tex_h, tex_w) -> heights (n,) of the differentiable vertices."""

    @staticmethod
    def forward(ctx, tex, sampler):
        ctx.sampler = sampler
        return sampler.heights(tex)

    @staticmethod
    def backward(ctx, grad):
        return ctx.sampler.tex_gradient(grad), None


def sample_heights(tex, sampler):
    return SampleHeights.apply(tex, sampler)


def acoustic_gradient(tex, diffbem, sampler, device=True):
    """acoustics_opt.py:37-60 with the device sampler: returns (1 - value, -dvalue/dtex).

    device=True (and a DiffBEM with gradient_device, as this package's module has): the heights and the vertex
    gradient stay on the GPU (DiffBEM.gradient_device); otherwise they go through host numpy arrays as in the
    reference (acoustics_opt.py:264). Both routes give the same bits."""
    heights = sampler.heights(tex)
    if device and hasattr(diffbem, "gradient_device"):
        g = torch.empty_like(heights)
        torch.cuda.current_stream(sampler.device).synchronize()  # g's memory is free for the library's stream
        val = diffbem.gradient_device(heights.data_ptr(), g.data_ptr())
    else:
        val, grad = diffbem.gradient(heights.cpu().numpy())
        g = torch.from_numpy(np.ascontiguousarray(grad, dtype=np.float64)).to(sampler.device)
    return 1 - val, -sampler.tex_gradient(g)


def smoothness(tex):
    """Mean squared 5-point Laplacian with zero padding (acoustics_opt.py:67-72)."""
    p = torch.nn.functional.pad(tex, (1, 1, 1, 1), value=0)
    lap = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:] - 4 * p[1:-1, 1:-1]
    return torch.mean(lap * lap)


def neg_relu(tex):
    """Mean of relu(-h) (acoustics_opt.py:78-80)."""
    return torch.mean(torch.nn.functional.relu(-tex))


def barrier_loss(tex, vmax, buffer=1e-3):
    """Log barrier keeping heights inside (-vmax, vmax) (acoustics_opt.py:82-86)."""
    vl = vmax + buffer
    return -torch.mean(torch.minimum(torch.log(tex + vl), torch.log(vl - tex)))


def preprocess_tex(tex, edge_border):
    """Zero border of edge_border texels (acoustics_opt.py:90-92)."""
    return torch.nn.functional.pad(tex, (edge_border, edge_border, edge_border, edge_border), value=0)


def normalize_gradients(grad, edge=0):
    """grad / ||grad|| (its interior when edge > 0), zero stays zero (acoustics_opt.py:94-104)."""
    g = grad[edge:-edge, edge:-edge] if edge != 0 else grad[:, :]
    n = torch.linalg.norm(g)
    return g if n == 0 else g / n


def optimise(diffbem, sampler, iters, lr=1e-3, init_val=0.0, edge_border=2, weights=None, vmax=None,
             normalize_grads=True, band_schedule=None, callback=None, device=True):
    """The acoustics-only optimiser (acoustics_opt.py:238-347): per iteration, the padded heightfield's
    acoustic value and gradient, the smoothness and negativity (and, with vmax, barrier) terms, each
    gradient normalised and weighted, summed into the variable's gradient, one Adam step.

    sampler's texture shape is the padded heightfield (hfield_res x hfield_res); the variable is its
    interior. band_schedule (optional): a band per iteration passed to diffbem.set_band (sample_freq).
    device: the acoustic term's heights and gradient stay on the GPU (acoustic_gradient).
    Returns the final padded heightfield (numpy) and the per-iteration values {term: array}."""
    w = {"ac_wt": 5.0, "sm_wt": 10.0, "ng_wt": 0.5, "ba_wt": 0.5}
    w.update(weights or {})
    res = sampler.shape[0]
    inner = res - 2 * edge_border
    var = torch.full((inner, inner), float(init_val), dtype=torch.float64, device=sampler.device, requires_grad=True)
    opt = torch.optim.Adam([var], lr=lr)
    terms = ["ac", "sm", "ng"] + (["ba"] if vmax else [])
    hist = {t: np.zeros(iters) for t in terms}
    crop = (slice(edge_border, res - edge_border), slice(edge_border, res - edge_border))
    for it in range(iters):
        opt.zero_grad()
        full = preprocess_tex(var, edge_border)
        if band_schedule is not None:
            diffbem.set_band(float(band_schedule[it]))
        grads = {}
        ac_v, ac_g = acoustic_gradient(full.detach(), diffbem, sampler, device=device)
        hist["ac"][it], grads["ac"] = ac_v, ac_g[crop]
        for name, fn in (("sm", smoothness), ("ng", neg_relu)) + ((("ba", lambda t: barrier_loss(t, vmax)),)
                                                                  if vmax else ()):
            var.grad = None
            v = fn(full)
            v.backward(retain_graph=True)
            hist[name][it], grads[name] = v.item(), var.grad.detach().clone()
        total = torch.zeros_like(var)
        for name in terms:
            g = normalize_gradients(grads[name]) if normalize_grads else grads[name]
            total += w[name + "_wt"] * g
        var.grad = total
        opt.step()
        if callback:
            callback(it, {t: hist[t][it] for t in terms})
    return preprocess_tex(var.detach(), edge_border).cpu().numpy(), hist
