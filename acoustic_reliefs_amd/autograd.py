"""torch.autograd.Function over the relief workload: heights -> per-band diffusion coefficients.

Forward, per band (one frequency per band, n_freqs = 1):

* assemble the H-matrix (FP64, or the complex-FP32 copy with FP64 refinement);
* solve all incidences as one batch;
* evaluate the receivers and the diffusion coefficient per incidence;
* take the band value as the mean over the incidences.

Backward (adjoint.cpp:135-239, batched over the incidences):

* vf = (dL/dd_b / n_inc) u conj(y) / |y|;
* adjoint right-hand sides Q^T vf;
* one transposed batched solve (M - I/2)^T lambda = Q^T vf;
* the shape-gradient kernels summed over the incidences (A - B + C).

recompute=True frees each band's H-matrix after the forward and re-assembles it in the backward (the
reference's recompute_matrices, adjoint.cpp:104-109, 173-193), trading assembly time for HBM.

Every forward call returns its own state (heights, per-band solutions, adjoint weights and kept
operators), and the backward of an output uses the state of the forward that produced it: further
forwards in between (a line search, finite-difference probes, gradcheck) do not change it.

Heights are the y offsets of the differentiable vertices (the relief's top grid), as in
DiffBEM.gradient (acoustics3d.cpp:210-213). Everything runs through the C-ABI (capi.py).
"""
import time

import numpy as np
import torch

from . import capi


class ReliefBEM:
    """Device state of one relief: geometry, receivers, incidences, bands and solver settings."""

    def __init__(self, P, E, diff, bands, directions, receivers, c0=343.0, leaf=32, factor=1.5, aca_tol=1e-4,
                 tol=1e-6, storage=capi.ARH_FP64, recompute=False, device=0, ctx=None, all_dense=False):
        """all_dense=True: every leaf pair dense (the exact operator). ACA pivots and ranks change
        discontinuously with the vertex positions, so finite differences are only meaningful
        against a forward without ACA."""
        self.ctx = ctx or capi.Context(device)
        self.g = capi.Geometry.from_mesh(self.ctx, P, E, leaf, factor, all_dense=all_dense)
        self.P0 = np.array(P, dtype=np.float64)
        self.diff = np.ascontiguousarray(diff, dtype=np.int32)
        self.bands = [float(b) for b in bands]
        self.dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 3)
        self.L = np.ascontiguousarray(receivers, dtype=np.float64).reshape(-1, 3)
        self.c0, self.aca_tol, self.tol, self.storage, self.recompute = c0, aca_tol, tol, storage, recompute
        self.last_state = None
        self.last_stats = []
        self.timing = []  # per band: assemble / solve / receivers (forward), adjoint / grad (backward), ms
        self.log = None   # optional callable(str): a line per band as it completes

    def k(self, f):
        return 2 * np.pi * f / self.c0

    def set_heights(self, h):
        P = self.P0.copy()
        P[self.diff, 1] += np.asarray(h, dtype=np.float64)
        self.g.set_vertices(P)

    def forward(self, h):
        """Per-band mean diffusion coefficient over the incidences (and the per-incidence values).
        The call's state is kept as last_state for backward(grad_d)."""
        d_b, d_all, _ = self.forward_state(h)
        return d_b, d_all

    def forward_state(self, h):
        """forward(h) that also returns the call's state, for backward(grad_d, state)."""
        h = np.array(h, dtype=np.float64, copy=True)
        self.set_heights(h)
        R, LL = len(self.dirs), len(self.L)
        saved, d_b, d_all = [], [], []
        self.last_stats = []
        self.timing = []
        for f in self.bands:
            k = self.k(f)
            H = self.g.assemble(k, self.aca_tol, 100, self.storage)
            info = H.info()
            B = self.g.rhs(k, capi.ARH_RHS_PLANE, self.dirs)
            Bn = B.numpy()
            B.upload(-Bn)
            X = self.ctx.empty((R, self.g.N))
            st = H.solve(B, X, R, tol=self.tol, true_residual=True)
            t0 = time.perf_counter()
            Y = self.g.receivers(k, self.L, X, nrhs=R)
            d, _, vf = self.ctx.diffusion(Y, LL, R, want_vf=True)
            t_rec = 1000 * (time.perf_counter() - t0)
            d_all.append(d)
            d_b.append(float(np.mean(d)))
            if not self.recompute:
                H.compact()  # keep the assembled operator only; the backward rebuilds its product copies
            saved.append({"k": k, "X": X, "VF": vf, "H": None if self.recompute else H,
                          "bytes": info["bytes_total"]})
            self.last_stats.append(st)
            self.timing.append({"band": f, "assemble_ms": info["assemble_ms"], "solve_ms": st["solve_ms"],
                                "receivers_ms": t_rec, "iters": st["iters"], "hmat_bytes": info["bytes_total"]})
            if self.log:
                self.log("forward band %g Hz: %s" % (f, self.timing[-1]))
        state = {"h": h, "bands": saved}
        self.last_state = state
        return np.array(d_b), np.array(d_all), state

    def backward(self, grad_d, state=None):
        """dL/dh from dL/dd_b (one value per band) at the heights of state's forward (default: the
        latest forward)."""
        state = state if state is not None else self.last_state
        if state is None:
            raise RuntimeError("forward must run first")
        # the shape gradient (and a recomputed operator) need the geometry at this forward's heights
        self.set_heights(state["h"])
        R, LL, HH = len(self.dirs), len(self.L), len(self.diff)
        g = np.zeros(HH)
        for b, s in enumerate(state["bands"]):
            w = float(grad_d[b]) / R
            if w == 0.0:
                continue
            vf = s["VF"].numpy() * w
            VF = self.ctx.to_device(vf)
            ADJ = self.g.receivers(s["k"], self.L, VF, nrhs=R, transpose=True)
            H = s["H"]
            t_asm = 0.0
            if H is None:
                H = self.g.assemble(s["k"], self.aca_tol, 100, self.storage)
                t_asm = H.info()["assemble_ms"]
            LAM = self.ctx.empty((R, self.g.N))
            st = H.solve(ADJ, LAM, R, transpose=True, tol=self.tol, true_residual=True)
            t0 = time.perf_counter()
            A, Bv, C = self.g.shape_grad(s["k"], self.L, self.diff, R, s["X"], LAM, VF, capi.ARH_RHS_PLANE, self.dirs)
            t_grad = 1000 * (time.perf_counter() - t0)
            g += A - Bv + C
            s["H"] = None  # a band's operator is released once its backward is done
            if b < len(self.timing):
                self.timing[b].update({"bw_assemble_ms": t_asm, "adjoint_ms": st["solve_ms"],
                                       "adjoint_iters": st["iters"], "shape_grad_ms": t_grad})
            if self.log:
                self.log("backward band %d: assemble %.1f ms, adjoint %.1f ms (%d iters), shape gradient %.1f ms" % (
                    b, t_asm, st["solve_ms"], st["iters"], t_grad))
        return g


class DiffusionCoefficients(torch.autograd.Function):
    """heights (HH,) -> d (n_bands,): the mean diffusion coefficient of each band."""

    @staticmethod
    def forward(ctx, heights, engine):
        d_b, _, state = engine.forward_state(heights.detach().cpu().double().numpy())
        ctx.engine = engine
        ctx.state = state
        ctx.hdev = heights.device
        return torch.from_numpy(d_b).to(heights.device)

    @staticmethod
    def backward(ctx, grad_d):
        g = ctx.engine.backward(grad_d.detach().cpu().double().numpy(), ctx.state)
        return torch.from_numpy(g).to(ctx.hdev), None


def diffusion_coefficients(heights, engine):
    return DiffusionCoefficients.apply(heights, engine)
