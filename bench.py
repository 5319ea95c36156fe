"""Benchmark: BEM solves/sec on the 0.6 m relief, all 13 bands x 37 plane waves (481 solves).

A step is the whole workload: for each one-third-octave band 250-4000 Hz, assemble the H-matrix
(dense near-field leaves + ACA, leaf 32, eta 2 / radius factor 1.5, ACA tol 1e-4), solve the 37
plane-wave right-hand sides to FP64 true relative residual <= 1e-6, evaluate 181 receivers on the
5 m semicircle and the diffusion coefficient of every (band, incidence).

* value        -- the mixed-precision batched path: complex-FP32 operator, 37 rhs per solve,
                  FP64 true-residual refinement (solves/s, whole job, one GPU)
* fp64_plain   -- the plain path: complex-FP64 operator, one rhs per solve
* cpu_baseline -- the CPU oracle (FP64, OpenMP) on a bounded sample of the same workload
* roofline     -- live HIP-event timing of the dominant H-matvec kernel over the timed region

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time


def set_hw_queues(argv):
    """Hardware queues of this process (HIP's GPU_MAX_HW_QUEUES; 4 when unset). The served run's four worker
    streams share the process with the main thread's context stream and the runtime's own stream: on 4 queues
    two of the workers end up on one queue and take turns (two bands finish early, two late). 8 gives every
    stream a queue of its own. The HIP runtime reads the variable when it starts, so this runs before anything
    touches the GPU; --hw-queues chooses another value."""
    q = "8"
    for i, a in enumerate(argv):
        if a == "--hw-queues" and i + 1 < len(argv):
            q = argv[i + 1]
        elif a.startswith("--hw-queues="):
            q = a.split("=", 1)[1]
    os.environ["GPU_MAX_HW_QUEUES"] = q

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BEM solves/sec, 0.6m relief all bands (mixed vs FP64); p50/p99 solve latency"
C0 = 343.0
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(grid=64, size=0.6, obj=""):
    """The synthetic 0.6 m relief, or (obj) a relief mesh read from an OBJ file (.xz allowed)."""
    from acoustic_reliefs_amd import meshes
    if obj:
        P, E = read_obj_any(obj)
    else:
        P, E, top = meshes.relief_config(size=size, grid=grid, hmax=0.1, seed=0)
    return P, E, meshes.incidence_directions(37), meshes.semicircle_receivers(5.0, 181), meshes.THIRD_OCTAVE_BANDS


def read_obj_any(path):
    import lzma
    import tempfile
    from acoustic_reliefs_amd import meshes
    if not path.endswith(".xz"):
        return meshes.read_obj(path)
    with tempfile.TemporaryDirectory() as d:
        tmp = os.path.join(d, "mesh.obj")
        with lzma.open(path, "rb") as f, open(tmp, "wb") as g:
            g.write(f.read())
        return meshes.read_obj(tmp)


def run_mixed(g, bands, dirs, L, capi, lat):
    ds = []
    for f in bands:
        k = 2 * np.pi * f / C0
        t0 = time.perf_counter()
        d, _, _, st, info = g.band_solve(k, capi.ARH_RHS_PLANE, dirs, L, aca_tol=1e-4, storage=capi.ARH_MIXED,
                                         tol=1e-6, true_residual=True)
        dt = time.perf_counter() - t0
        lat.extend([dt] * len(dirs))
        log("  mixed band %g Hz: %.3f s (assembly %.1f ms device, solve %.1f ms), iters %d, refinements %d, "
            "true residual %.2e, max rank %d" % (f, dt, info["assemble_ms"], st["solve_ms"], st["iters"],
                                                 st["refinements"], st["max_true_res"], info["max_rank"]))
        ds.append((d, st, info))
    return ds


def run_served(srv, mid, bands, dirs, L, capi, lat):
    """All bands submitted at once to the server, highest frequency (longest solve) first, so the
    FIFO server starts the long jobs early; latency = completion - submission per request."""
    rids = {}
    for f in sorted(bands, reverse=True):
        k = 2 * np.pi * f / C0
        rids[f] = srv.submit(mid, k, dirs, L, kind=capi.ARH_RHS_PLANE, storage=capi.ARH_MIXED, aca_tol=1e-4,
                             tol=1e-6)
    out = []
    for f in bands:
        rid = rids[f]
        d, info = srv.wait(rid)
        lat.extend([info["t_done"] - info["t_submit"]] * len(dirs))
        log("  served band %g Hz: queued %.3f s, service %.3f s (assembly %.1f ms device, solve %.1f ms, receivers "
            "%.1f ms), iters %d, true residual %.2e" % (
                f, info["t_start"] - info["t_submit"], info["t_done"] - info["t_start"], info["assemble_ms"],
                info["solve_ms"], info["post_ms"], info["iters"], info["max_true_res"]))
        out.append((d, info, None))
    return out


def run_fp64_plain(g, bands, dirs, L, capi, ctx, lat):
    """One rhs per solve, complex FP64 operator."""
    out = []
    for f in bands:
        k = 2 * np.pi * f / C0
        t0 = time.perf_counter()
        h = g.assemble(k, 1e-4, 100, capi.ARH_FP64)
        ta = time.perf_counter() - t0
        B = g.rhs(k, capi.ARH_RHS_PLANE, dirs)
        Bn = -B.numpy()
        dband = np.zeros(len(dirs))
        iters = []
        for j in range(len(dirs)):
            t1 = time.perf_counter()
            b = ctx.to_device(Bn[j])
            x = ctx.empty(g.N)
            st = h.solve(b, x, 1, tol=1e-6, true_residual=True)
            y = g.receivers(k, L, x)
            d, _, _ = ctx.diffusion(y, len(L), 1)
            dband[j] = d[0]
            iters.append(st["iters"])
            lat.append(time.perf_counter() - t1 + ta / len(dirs))
        out.append((dband, iters))
        log("  fp64 band %g Hz: assembly %.3f s, %d solves %.3f s, mean iters %.1f" % (
            f, ta, len(dirs), time.perf_counter() - t0 - ta, np.mean(iters)))
        del h
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(P, E, dirs, L, bands, threads=None):
    """CPU oracle (FP64, OpenMP) on a sample of the workload: every band in `bands` is assembled (H-matrix +
    exact listener matrix) and one of its plane waves solved to the workload's tolerance (direction b mod 37
    for the b-th band, so the sample spreads over the fan); a band's assembly is amortised over its 37 solves
    as in the full workload. solves/s = bands / sum over bands of (assembly / 37 + one solve).
    threads=None: all the threads OpenMP is given."""
    from oracle import oracle as O
    if threads:
        O.set_threads(threads)
    threads = O.num_threads()
    Er, dirb, appb, _ = O.setup(P, E, 32, 1.5)
    oc = O.Ctx(P, Er, dirb, appb)
    per, t_asm_all, t_sol_all, iters = [], 0.0, 0.0, []
    for b, f in enumerate(bands):
        k = 2 * np.pi * f / C0
        t0 = time.perf_counter()
        oc.assemble(k, 1e-4, 100)
        Q = oc.listener_matrix(k, L)
        t_asm = time.perf_counter() - t0
        t1 = time.perf_counter()
        x, it, err, _ = oc.bicgstab(-oc.rhs_plane(k, dirs[b % len(dirs)]), 1e-6)
        O.diffusion(np.abs(Q @ x))
        t_sol = time.perf_counter() - t1
        per.append(t_asm / len(dirs) + t_sol)
        t_asm_all += t_asm
        t_sol_all += t_sol
        iters.append(int(it))
    return {"value": len(bands) / sum(per), "unit": "solves/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": "oracle (CPU restatement, FP64, OpenMP %d threads), same relief: bands %s Hz, each assembled "
                      "(%.1f s in all, amortised over 37 solves) and one plane wave solved to 1e-6 (%.1f s in all, "
                      "BiCGSTAB iterations %s)" % (threads, "/".join("%g" % f for f in bands), t_asm_all, t_sol_all,
                                                   iters)}


def measured_traffic(kernel="k_rows_mfma<float, 5, 2>"):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 summary (FETCH_SIZE and
    WRITE_SIZE passes, each scaled by its reading for a plain device copy); None when absent."""
    path = os.path.join(ROOT, "profiles", "k_rows_mfma_relief0.6m_4000Hz_37rhs.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0] == "`%s`" % kernel:
            try:
                return float(cells[3]) * 1e6
            except ValueError:
                return None
    return None


def isolated_roofline(g, capi, ctx, f=4000.0, reps=20):
    """The dominant kernel on its own stream, nothing else in flight: the 37-rhs FP32 MFMA product of
    the f Hz band's operator (and the FP64 one), HIP-event timed."""
    k = 2 * np.pi * f / C0
    h = g.assemble(k, 1e-4, 100, capi.ARH_MIXED)
    R = 37
    rng = np.random.default_rng(0)
    X = ctx.to_device(rng.standard_normal((R, g.N)) + 1j * rng.standard_normal((R, g.N)))
    Y = ctx.empty((R, g.N))
    out = {}
    N, T = g.N, h.info()["sum_rank"]
    for name, prec, kern, peak in (("fp32", capi.ARH_MIXED, "rows_fp32", FP32_PEAK_TFLOPS),
                                   ("fp64", capi.ARH_FP64, "rows_fp64", FP64_PEAK_TFLOPS)):
        h.matvec(X, Y, nrhs=R, precision=prec)
        ctx.synchronize()
        ctx.profile(True)
        for _ in range(reps):
            h.matvec(X, Y, nrhs=R, precision=prec)
        p = ctx.get_profile()
        ctx.profile(False)
        rw = p[kern]
        lr = p[kern.replace("rows", "lowrank")]
        ach = rw["flops"] / (rw["ms"] * 1e-3) / 1e12
        out[name] = {"kernel": "k_rows_mfma<%s, 5, 2> (%g Hz band, %d rhs, N=%d, %d intermediates rows)" % (
                         "float" if name == "fp32" else "double", f, R, N, T),
                     "bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                     "traffic": measured_traffic("k_rows_mfma<%s, 5, 2>" % ("float" if name == "fp32" else "double")),
                     "traffic_source": "profiles/k_rows_mfma_relief0.6m_4000Hz_37rhs.md (rocprofv3 FETCH_SIZE x 2.000 "
                                       "+ WRITE_SIZE x 1.000 per dispatch, same kernel and workload)",
                     "avg_launch_us": 1000 * rw["ms"] / rw["launches"], "launches": rw["launches"],
                     "flops_per_launch": rw["flops"] / rw["launches"], "bytes_per_launch": rw["bytes"] / rw["launches"],
                     "lowrank_phase_us": 1000 * lr["ms"] / max(1, lr["launches"]),
                     "hbm_GBs": rw["bytes"] / (rw["ms"] * 1e-3) / 1e9,
                     "timed": "HIP events around each launch on its stream, %d launches, nothing else in flight"
                              % rw["launches"]}
    del h
    return out


# Declared FP64 flops per kernel evaluation, used for the FP64-VALU rooflines of the assembly and the shape
# gradient (transcendental functions, square roots and divisions counted as one flop each, an FMA as two):
# * one quadrature point of an entry (the 6-point rule of integrate.cpp:5-14): rule point 12, distance 11,
#   s = -(d.n) 5, k r 1, sincos 1, w s / (r^2 r) 3, the two accumulated parts 8: 41 (+ 3 for the sqrt, the
#   division and the rule-point select) = 44;
# * one quadrature point of a shape derivative (dM_ie/dp_i or dM_ei/dq_(i,c), adjoint.cpp:244-352): rule point
#   12, distance 12, s 5, F and F' 31, the collocation term 12 or the three corner terms 48: 90 on average.
FLOPS_PER_ENTRY_POINT = 44
FLOPS_PER_DERIV_POINT = 90


def kernel_counts(g, capi, ctx, top, f=4000.0, R=37):
    """Algorithmic work of the FP64-VALU kernels on the f Hz operator: the ACA kernel (entry evaluations of the
    crosses it computes: rank + 1 crosses per block, the last one discarded, or the rank when the limit
    min(100, m - 1, n - 1) stops it) and the dense-leaf kernel, then one shape-gradient call with R right-hand
    sides (12 derivative points per (touched element, element) pair and pass, plus 8 flops per complex
    multiply-add of the R right-hand sides). Run under rocprofv3 --kernel-trace for the kernel durations."""
    k = 2 * np.pi * f / C0
    h = g.assemble(k, 1e-4, 100, capi.ARH_FP64)
    ranks = h.ranks()
    m = g.lowrank[:, 1] - g.lowrank[:, 0]
    n = g.lowrank[:, 3] - g.lowrank[:, 2]
    lim = np.minimum(100, np.minimum(m - 1, n - 1))
    crosses = np.where(ranks < lim, ranks + 1, ranks)
    aca_entries = int(np.sum(crosses * (m + n)))
    dense_entries = int(np.sum((g.dense[:, 1] - g.dense[:, 0]) * (g.dense[:, 3] - g.dense[:, 2])))
    del h
    P, E = g.P, g.E
    top = np.asarray(top, np.int32)
    touched = np.any(np.isin(E, top), axis=1).sum()
    rng = np.random.default_rng(0)
    X = ctx.to_device(rng.standard_normal((R, g.N)) + 1j * rng.standard_normal((R, g.N)))
    LAM = ctx.to_device(rng.standard_normal((R, g.N)) + 1j * rng.standard_normal((R, g.N)))
    L = np.column_stack([5 * np.cos(np.linspace(0, np.pi, 181)), 5 * np.sin(np.linspace(0, np.pi, 181)), np.zeros(181)])
    VF = ctx.to_device(rng.standard_normal((R, 181)) + 1j * rng.standard_normal((R, 181)))
    dirs = -np.column_stack([np.cos(np.linspace(0, np.pi, R)), np.sin(np.linspace(0, np.pi, R)), np.zeros(R)])
    g.shape_grad(k, L, top, R, X, LAM, VF, capi.ARH_RHS_PLANE, dirs)
    ctx.synchronize()
    passes = -(-R // 20)
    pairs = int(touched) * g.N
    return {"frequency": f, "N": g.N, "aca_entries": aca_entries, "aca_points": 6 * aca_entries,
            "aca_flops": 6 * aca_entries * FLOPS_PER_ENTRY_POINT, "dense_entries": dense_entries,
            "dense_flops": 6 * dense_entries * FLOPS_PER_ENTRY_POINT, "grad_touched": int(touched),
            "grad_pairs_per_pass": pairs, "grad_passes": passes,
            "grad_B_flops": passes * pairs * 12 * FLOPS_PER_DERIV_POINT + pairs * 4 * 8 * R,
            "flops_per_entry_point": FLOPS_PER_ENTRY_POINT, "flops_per_deriv_point": FLOPS_PER_DERIV_POINT}


def copy_calibration(nbytes=1 << 30):
    """A plain device copy of nbytes (torch), for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes}), flush=True)


def run_gradient(args):
    """Adjoint gradient of loss = mean_b (1 - d_b) (d_b: band mean over the 37 incidences) through the
    torch.autograd.Function: forward + backward over all bands, mixed precision, with stored H-matrices
    and with recompute-on-backward; optionally central differences (h = 1e-5 m) on 16 seeded top
    vertices against the exact (all-dense) operator in FP64 and mixed precision."""
    import torch
    from acoustic_reliefs_amd import autograd, capi, meshes
    P, E, top = meshes.relief_config(size=args.size, grid=args.grid, hmax=args.hmax, seed=0)
    bands = [float(f) for f in args.freqs.split(",")] if args.freqs else meshes.THIRD_OCTAVE_BANDS[:args.bands]
    dirs, L = meshes.incidence_directions(37), meshes.semicircle_receivers(5.0, 181)
    ctx = capi.Context(0)
    out = {"workload": "%.1f m relief, %dx%d top, N=%d, %d diff vertices, %d bands x 37 plane waves, 181 receivers"
                       % (args.size, args.grid, args.grid, len(E), len(top), len(bands))}
    for recompute in (False, True):
        t_init = time.perf_counter()
        eng = autograd.ReliefBEM(P, E, top, bands, dirs, L, c0=C0, storage=capi.ARH_MIXED, recompute=recompute,
                                 ctx=ctx)
        eng.log = log
        ctx.synchronize()
        log("engine (partition, geometry) %.2f s" % (time.perf_counter() - t_init))
        capi.device_trim(0)  # blocks kept for reuse are not held by the engine
        free0 = torch.cuda.mem_get_info()[0]
        ctx.profile(True)  # HIP-event totals of the H-matrix product kernels over forward + backward
        h = torch.zeros(len(top), dtype=torch.float64, requires_grad=True)
        t0 = time.perf_counter()
        d = autograd.diffusion_coefficients(h, eng)
        ctx.synchronize()
        t1 = time.perf_counter()
        capi.device_trim(0)
        free1 = torch.cuda.mem_get_info()[0]
        loss = (1.0 - d).mean()
        loss.backward()
        ctx.synchronize()
        t2 = time.perf_counter()
        prof = ctx.get_profile()
        ctx.profile(False)
        kern = {}
        for name, peak in (("rows_fp32", FP32_PEAK_TFLOPS), ("lowrank_fp32", FP32_PEAK_TFLOPS),
                           ("rows_fp64", FP64_PEAK_TFLOPS), ("lowrank_fp64", FP64_PEAK_TFLOPS)):
            kp = prof[name]
            if kp["launches"] and kp["ms"] > 0:
                tf = kp["flops"] / (kp["ms"] * 1e-3) / 1e12
                kern[name] = {"launches": kp["launches"], "ms": kp["ms"], "TFLOPs": tf, "mfma_frac": tf / peak,
                              "GBs": kp["bytes"] / (kp["ms"] * 1e-3) / 1e9,
                              "hbm_frac": kp["bytes"] / (kp["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
        key = "recompute" if recompute else "stored"
        out[key] = {"forward_s": t1 - t0, "backward_s": t2 - t1, "loss": float(loss.detach()),
                    "product_kernels": kern,
                    "device_bytes_held_after_forward": free0 - free1,
                    "hmat_bytes_per_band": [b["hmat_bytes"] for b in eng.timing],
                    "per_band": eng.timing}
        log("gradient %s: forward %.2f s, backward %.2f s, %.2f GB held after the forward" % (
            key, t1 - t0, t2 - t1, (free0 - free1) / 1e9))
        del eng, d, loss, h
        ctx.synchronize()
    if args.fd:
        rng = np.random.default_rng(1)
        pick = rng.choice(len(top), args.fd, replace=False)
        step = 1e-5
        for name, storage, tol in (("fp64", capi.ARH_FP64, 1e-10), ("mixed", capi.ARH_MIXED, 1e-10)):
            eng = autograd.ReliefBEM(P, E, top, bands, dirs, L, c0=C0, storage=storage, tol=tol, ctx=ctx,
                                     all_dense=True)
            h = torch.zeros(len(top), dtype=torch.float64, requires_grad=True)
            (1.0 - autograd.diffusion_coefficients(h, eng)).mean().backward()
            g = h.grad.numpy()
            fds = []
            for t in pick:
                hp, hm = np.zeros(len(top)), np.zeros(len(top))
                hp[t], hm[t] = step, -step
                fds.append((np.mean(1.0 - eng.forward(hp)[0]) - np.mean(1.0 - eng.forward(hm)[0])) / (2 * step))
            fds, ga = np.array(fds), g[pick]
            errs = np.abs(fds - ga) / np.maximum(np.abs(fds), 1e-12)
            scale = np.abs(fds).max()
            out["fd_" + name] = {"vertices": [int(t) for t in pick], "fd": fds.tolist(), "adjoint": ga.tolist(),
                                 "max_rel_err": float(errs.max()), "median_rel_err": float(np.median(errs)),
                                 "max_err_over_max_fd": float(np.abs(fds - ga).max() / scale), "step_m": step,
                                 "operator": "exact (every leaf pair dense), BiCGSTAB tol %g" % tol}
            log("fd %s: max rel err %.2e, max |fd - adjoint| / max |fd| %.2e" % (
                name, errs.max(), np.abs(fds - ga).max() / scale))
            del eng
    print(json.dumps(out), flush=True)


def run_diffbem(args):
    """The drop-in module at a reference optimiser's settings (configs/base.toml: cluster_size 128, all tolerances
    1e-5, bands 1000-3150 Hz, 1297 hemisphere listeners at 50 m with the listener ACA, point source at (0, 100, 0)):
    DiffBEM.value() and DiffBEM.gradient(x) over every top vertex, one band alone, and the CPU oracle on that band."""
    from acoustic_reliefs_amd import acoustics3d_impl as m
    from acoustic_reliefs_amd import meshes
    from oracle import oracle as O
    P, E, top = meshes.relief_config(size=args.size, grid=args.grid, hmax=args.hmax, seed=0)
    bands = [1000.0, 1250.0, 1600.0, 1600.0, 2000.0, 2500.0, 3150.0]
    d = m.DiffBEM(args.leaf, 1.5, bands, 1, 1e-5, 1e-5, 1e-5, np.array([0.0, 100.0, 0.0]), 50, 5, False)
    d.silent = True
    Ps, Es = d.precompute(P, E, top)
    x = np.zeros(len(top))
    out = {"workload": "%.1f m relief, %dx%d top, N=%d, %d diff vertices, cluster_size %d, bands %s, 1297 listeners"
                       % (args.size, args.grid, args.grid, len(E), len(top), args.leaf, bands)}
    t0 = time.perf_counter()
    d.value(x)
    out["warmup_value_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    v = d.value(x)
    out["value_s"] = time.perf_counter() - t0
    log("value %.6f: %.2f s" % (v, out["value_s"]))
    t0 = time.perf_counter()
    v2, g = d.gradient(x)
    out["gradient_s"] = time.perf_counter() - t0
    log("gradient: %.2f s, |g| %.3e" % (out["gradient_s"], np.linalg.norm(g)))
    f = 3150.0
    t0 = time.perf_counter()
    d.surface_vals(f)
    t1 = time.perf_counter()
    d.pvals(f)
    t2 = time.perf_counter()
    out["surface_vals_3150_s"], out["pvals_3150_s"] = t1 - t0, t2 - t1
    log("3150 Hz: surface_vals %.3f s, pvals (+ listener ACA) %.3f s" % (t1 - t0, t2 - t1))
    d.set_band(1000.0)
    t0 = time.perf_counter()
    vb, gb = d.gradient(x)
    out["band_gradient_1000_s"] = time.perf_counter() - t0
    log("1000 Hz band gradient: %.3f s" % out["band_gradient_1000_s"])
    if not args.no_cpu:
        threads = O.num_threads()
        Er, dirb, appb, _ = O.setup(P, E, args.leaf, 1.5)
        oc = O.Ctx(P, Er, dirb, appb)
        k = 2 * np.pi * 1000.0 / 344.0
        t0 = time.perf_counter()
        vo, go = O.band_value_grad(oc, [k], d.get_listeners(), top, aca_tol=1e-5, solver_tol=1e-5, q_aca_tol=1e-5,
                                   want_grad=True)
        tc = time.perf_counter() - t0
        out["cpu_oracle_band_gradient_1000"] = {"s": tc, "threads": threads, "value": vo,
                                                "abs_value_diff": abs(vo - vb),
                                                "grad_rel_diff": float(np.linalg.norm(go - gb) / np.linalg.norm(go))}
        log("CPU oracle 1000 Hz band gradient: %.2f s on %d threads; |d - d_gpu| %.2e, grad rel diff %.2e" % (
            tc, threads, abs(vo - vb), out["cpu_oracle_band_gradient_1000"]["grad_rel_diff"]))
    print(json.dumps(out), flush=True)


def make_stream(n, pool, seed=0, rate=5.0):
    """A seeded serving stream: `pool` distinct reliefs (size 0.3-0.9 m, top grid 16-160, heights scaled to
    size / 6 as 0.1 m at 0.6 m and 0.15 m at 0.9 m) and n requests, each a relief drawn from the pool, a random
    one-third-octave band and 1-37 incidence directions from the fan, arriving as a Poisson process."""
    from acoustic_reliefs_amd import meshes
    rng = np.random.default_rng(seed)
    reliefs = []
    for p in range(pool):
        size = float(rng.uniform(0.3, 0.9))
        grid = int(rng.integers(16, 161))
        P, E, _ = meshes.relief_config(size=size, grid=grid, hmax=size / 6.0, seed=1000 + p)
        reliefs.append({"size": size, "grid": grid, "P": P, "E": E})
    fan = meshes.incidence_directions(37)
    t, reqs = 0.0, []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        nr = int(rng.integers(1, 38))
        reqs.append({"t": t, "relief": int(rng.integers(pool)), "band": float(rng.choice(meshes.THIRD_OCTAVE_BANDS)),
                     "dirs": fan[np.sort(rng.choice(37, nr, replace=False))]})
    return reliefs, reqs


def fifo_latencies(arrivals, service):
    """Completion - arrival of each request for one server taking requests in arrival order."""
    t, out = 0.0, []
    for a, sv in zip(arrivals, service):
        t = max(t, a) + sv
        out.append(t - a)
    return np.array(out)


def run_stream(args):
    """Serving stream (--phase served): the requests are submitted to the continuous-batching server at
    their Poisson arrival times (mixed precision, true residual 1e-6); latency = completion - arrival.
    --phase fp64: the same requests one at a time on the complex-FP64 path (each request's directions in
    one batched solve, a geometry set up whenever the relief changes, as the server does); its latency is
    that of a one-at-a-time server fed the same arrivals. --phase both runs the two. Both paths stop a
    request's BiCGSTAB at --maxit iterations (near the closed box's interior resonances a few requests
    do not reach 1e-6 within 2N); such requests are counted, not retried."""
    import threading
    from acoustic_reliefs_amd import capi, meshes
    reliefs, reqs = make_stream(args.requests, args.pool, seed=0, rate=args.rate)
    L = meshes.semicircle_receivers(5.0, 181)
    nmin, nmax = min(len(r["E"]) for r in reliefs), max(len(r["E"]) for r in reliefs)
    solves = sum(len(q["dirs"]) for q in reqs)
    out = {"workload": "serving stream: %d Poisson requests (rate %.2f/s) over %d seeded reliefs 0.3-0.9 m, "
                       "top grids 16-160 (N %d-%d), random band 250-4000 Hz, 1-37 plane waves (%d solves), "
                       "181 receivers" % (len(reqs), args.rate, len(reliefs), nmin, nmax, solves)}
    recs = [{"t": q["t"], "n": len(reliefs[q["relief"]]["E"]), "nrhs": len(q["dirs"]), "band": q["band"]}
            for q in reqs]
    d_served = None
    if args.phase in ("served", "both"):
        srv = capi.Server(workers=args.workers, maxit=args.maxit)
        t0 = time.perf_counter()
        mids = [srv.add_mesh(r["P"], r["E"], leaf=32, factor=1.5) for r in reliefs]
        log("stream: %d requests over %d reliefs (N %d-%d), rate %.2f/s; partitions %.1f s" % (
            len(reqs), len(reliefs), nmin, nmax, args.rate, time.perf_counter() - t0))
        rids = [None] * len(reqs)
        stop = threading.Event()

        def client():
            start = time.perf_counter()
            for i, q in enumerate(reqs):
                dt = q["t"] - (time.perf_counter() - start)
                if dt > 0 and stop.wait(dt):
                    return
                rids[i] = srv.submit(mids[q["relief"]], 2 * np.pi * q["band"] / C0, q["dirs"], L,
                                     kind=capi.ARH_RHS_PLANE, storage=capi.ARH_MIXED, aca_tol=1e-4, tol=1e-6)

        th = threading.Thread(target=client)
        th.start()
        res = []
        try:
            last = time.perf_counter()
            for i in range(len(reqs)):
                while rids[i] is None:
                    time.sleep(0.001)
                # wait in order, printing a line at least every 30 s while a long request runs
                while srv.poll(rids[i])["status"] in (0, 1):
                    time.sleep(0.01)
                    if time.perf_counter() - last > 30.0:
                        log("  waiting for request %d (N %d, %g Hz, %d rhs)" % (
                            i, recs[i]["n"], recs[i]["band"], recs[i]["nrhs"]))
                        last = time.perf_counter()
                try:
                    d, info = srv.wait(rids[i])
                except RuntimeError as e:
                    # a failed request is reported and counted; the stream goes on
                    log("  request %d (N %d, %g Hz, %d rhs) failed: %s" % (
                        i, recs[i]["n"], recs[i]["band"], recs[i]["nrhs"], e))
                    info = srv.poll(rids[i])
                    info["max_true_res"] = float("inf")
                    d = np.full(recs[i]["nrhs"], np.nan)
                    recs[i]["error"] = str(e)
                res.append((d, info))
                if (i + 1) % 50 == 0 or i + 1 == len(reqs):
                    log("  %d / %d done (latency %.3f s)" % (i + 1, len(reqs), info["t_done"] - info["t_submit"]))
                    last = time.perf_counter()
        finally:
            # the server's workers stop (after their current request) before the interpreter goes
            stop.set()
            th.join()
            srv.close()
        t_first = min(i["t_submit"] for _, i in res)
        t_last = max(i["t_done"] for _, i in res)
        lat = np.array([i["t_done"] - i["t_submit"] for _, i in res])
        out["served"] = {"workers": args.workers, "maxit": args.maxit,
                         "failed_requests": int(sum("error" in r for r in recs)),
                         "unconverged_requests": int(sum(i["max_true_res"] > 1e-6 for _, i in res)),
                         "span_s": t_last - t_first,
                         "solves_per_s": solves / (t_last - t_first), "requests_per_s": len(reqs) / (t_last - t_first),
                         "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                         "max_true_residual": max(i["max_true_res"] for _, i in res)}
        for r, (_, i) in zip(recs, res):
            r.update({"latency_s": i["t_done"] - i["t_submit"], "service_s": i["t_done"] - i["t_start"],
                      "geom_ms": i["geom_ms"], "assemble_ms": i["assemble_ms"], "solve_ms": i["solve_ms"],
                      "post_ms": i["post_ms"], "iters": i["iters"]})
        d_served = [d for d, _ in res]
        log("served: %d solves in %.2f s = %.1f solves/s; latency p50 %.3f s, p99 %.3f s" % (
            solves, t_last - t_first, out["served"]["solves_per_s"], out["served"]["p50_latency_s"],
            out["served"]["p99_latency_s"]))
    if args.phase in ("fp64", "both"):
        ctx = capi.Context(0)
        parts = {}
        cur, g = None, None
        svc, dmax, nconv = [], 0.0, 0
        for i, q in enumerate(reqs):
            r = q["relief"]
            if r not in parts:
                P, E = reliefs[r]["P"], reliefs[r]["E"]
                perm, dense, lowrank = capi.partition(capi.centroids(P, E), 32, 1.5)
                parts[r] = (P, np.asarray(E, np.int32)[perm], dense, lowrank)
            ctx.synchronize()
            t1 = time.perf_counter()
            if cur != r:
                g = None
                g = capi.Geometry(ctx, *parts[r])
                cur = r
            d64, _, _, st, _ = g.band_solve(2 * np.pi * q["band"] / C0, capi.ARH_RHS_PLANE, q["dirs"], L,
                                            aca_tol=1e-4, storage=capi.ARH_FP64, tol=1e-6, true_residual=True,
                                            maxit=args.maxit)
            svc.append(time.perf_counter() - t1)
            nconv += st["max_true_res"] > 1e-6
            recs[i]["fp64_s"] = svc[-1]
            if d_served is not None:
                dmax = max(dmax, float(np.abs(d64 - d_served[i]).max()))
            if (i + 1) % 50 == 0:
                log("  fp64 %d / %d" % (i + 1, len(reqs)))
        g = None
        lat = fifo_latencies([q["t"] for q in reqs], svc)
        busy = float(np.sum(svc))
        out["fp64_one_at_a_time"] = {"busy_s": busy, "solves_per_s_capacity": solves / busy,
                                     "unconverged_requests": int(nconv),
                                     "p50_latency_s": float(np.percentile(lat, 50)),
                                     "p99_latency_s": float(np.percentile(lat, 99)),
                                     "latency_note": "one server taking the same arrivals in order"}
        if d_served is not None:
            out["fp64_one_at_a_time"]["max_abs_d_diff_vs_served"] = dmax
        log("fp64 one at a time: busy %.1f s = %.1f solves/s capacity; latency at the same arrivals p50 %.3f s, "
            "p99 %.3f s" % (busy, solves / busy, out["fp64_one_at_a_time"]["p50_latency_s"],
                            out["fp64_one_at_a_time"]["p99_latency_s"]))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(recs, f)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)  # 5 steps of ~0.3 s: less run-to-run noise than one
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bands", type=int, default=13, help="number of bands (from 250 Hz up); 13 = all")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--obj", type=str, default="", help="a relief mesh (OBJ, or .obj.xz) instead of the synthetic one")
    ap.add_argument("--freqs", type=str, default="", help="comma-separated band frequencies (Hz) instead of --bands")
    ap.add_argument("--workers", type=int, default=4, help="server workers (concurrent band solves)")
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="hardware queues of the process (GPU_MAX_HW_QUEUES, set before the GPU is touched)")
    ap.add_argument("--sequential", action="store_true", help="solve the bands one after another (no server)")
    ap.add_argument("--gradient", action="store_true",
                    help="adjoint-gradient workload instead (forward + backward, stored vs recompute)")
    ap.add_argument("--fd", type=int, default=0, help="with --gradient: central differences on this many vertices")
    ap.add_argument("--size", type=float, default=0.6, help="with --gradient: relief width (m)")
    ap.add_argument("--hmax", type=float, default=0.1, help="with --gradient: maximum height (m)")
    ap.add_argument("--stream", action="store_true", help="the serving stream (Poisson requests over random reliefs)")
    ap.add_argument("--requests", type=int, default=2000, help="with --stream: number of requests")
    ap.add_argument("--pool", type=int, default=64, help="with --stream: distinct reliefs")
    ap.add_argument("--rate", type=float, default=5.0, help="with --stream: arrivals per second")
    ap.add_argument("--maxit", type=int, default=10000, help="with --stream: BiCGSTAB iteration limit per request")
    ap.add_argument("--phase", choices=["served", "fp64", "both"], default="both",
                    help="with --stream: the served stream, the one-at-a-time FP64 pass, or both")
    ap.add_argument("--out", type=str, default="", help="with --stream: write per-request records (JSON) here")
    ap.add_argument("--diffbem", action="store_true",
                    help="the drop-in DiffBEM at a reference optimiser's settings (value, gradient, CPU oracle band)")
    ap.add_argument("--leaf", type=int, default=128, help="with --diffbem: cluster_size")
    ap.add_argument("--hleaf", type=int, default=32, help="H-matrix leaf size of the headline workload")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the isolated timing of the dominant kernel (the roofline figures)")
    ap.add_argument("--kernel-counts", action="store_true",
                    help="only the ACA assembly and one shape-gradient call at 4000 Hz, printing their algorithmic work")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (for calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gradient:
        run_gradient(args)
        return
    if args.diffbem:
        run_diffbem(args)
        return
    if args.stream:
        run_stream(args)
        return
    if args.kernel_counts:
        from acoustic_reliefs_amd import capi, meshes
        P, E, top = meshes.relief_config(size=0.6, grid=args.grid, hmax=0.1, seed=0)
        ctx = capi.Context(0)
        g = capi.Geometry.from_mesh(ctx, P, E, leaf=32, factor=1.5)
        print(json.dumps(kernel_counts(g, capi, ctx, top)), flush=True)
        return
    if args.roofline_only:
        from acoustic_reliefs_amd import capi
        P, E, dirs, L, _ = make_workload(args.grid)
        ctx = capi.Context(0)
        g = capi.Geometry.from_mesh(ctx, P, E, leaf=32, factor=1.5)
        print(json.dumps(isolated_roofline(g, capi, ctx)), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from acoustic_reliefs_amd import capi
    P, E, dirs, L, bands = make_workload(args.grid, obj=args.obj)
    bands = [float(f) for f in args.freqs.split(",")] if args.freqs else bands[:args.bands]
    ctx = capi.Context(0)
    g = capi.Geometry.from_mesh(ctx, P, E, leaf=args.hleaf, factor=1.5)
    nsolve = len(bands) * len(dirs)
    log("workload: N=%d, %d dense / %d low-rank blocks, %d bands x %d incidences" % (g.N, len(g.dense),
                                                                                    len(g.lowrank), len(bands),
                                                                                    len(dirs)))
    if args.sequential:
        def step(lat):
            return run_mixed(g, bands, dirs, L, capi, lat)
    else:
        srv = capi.Server(workers=args.workers)
        mid = srv.add_mesh(P, E, leaf=args.hleaf, factor=1.5)

        def step(lat):
            return run_served(srv, mid, bands, dirs, L, capi, lat)
    for w in range(args.warmup):
        t0 = time.perf_counter()
        step([])
        log("warmup step %d: %.2f s" % (w, time.perf_counter() - t0))

    ctx.synchronize()
    if args.sequential:
        ctx.profile(True)
    else:
        srv.profile(True)
    lat = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        res = step(lat)
    ctx.synchronize()
    t_mixed = time.perf_counter() - t0
    if args.sequential:
        prof = ctx.get_profile()
        ctx.profile(False)
    else:
        prof = srv.get_profile()
        srv.profile(False)
        srv.close()
    d_mixed = np.array([r[0] for r in res])
    iters_mixed = [int(r[1]["iters"]) for r in res]
    worst_res = max(float(r[1]["max_true_res"]) for r in res)
    value = args.steps * nsolve / t_mixed
    log("mixed: %.3f s/step, %.1f solves/s, iters per band %s, worst true residual %.2e" % (
        t_mixed / args.steps, value, iters_mixed, worst_res))

    out = {"metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1000.0 * t_mixed / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "c64 operator + c128 refinement",
           "data": "mesh from " + os.path.basename(args.obj) if args.obj else "synthetic",
           "config": {"workload": ("relief %s (N=%d)" % (os.path.basename(args.obj), g.N) if args.obj else
                                   "0.6 m relief, %dx%d top (N=%d)" % (args.grid, args.grid, g.N)) +
                                  ", %d bands 250-4000 Hz x %d plane waves, 181 receivers at 5 m" % (len(bands),
                                                                                                     len(dirs)),
                      "solves_per_step": nsolve, "leaf": args.hleaf, "radius_factor": 1.5, "aca_tol": 1e-4,
                      "true_residual_tol": 1e-6, "c0": C0,
                      "serving": "sequential bands" if args.sequential else "server, %d workers" % args.workers,
                      "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"])},
           "latency_ms": {"p50": 1000 * float(np.percentile(lat, 50)), "p99": 1000 * float(np.percentile(lat, 99)),
                          "note": ("a solve's latency is its band's wall time (37 rhs per batched solve)"
                                   if args.sequential else
                                   "a solve's latency is its band request's completion - submission time, all "
                                   "%d band requests (37 rhs each) submitted together to a %d-worker server"
                                   % (len(bands), args.workers))},
           "max_true_residual": worst_res}

    if not args.sequential:
        tot = {"assembly": 0.0, "solve": 0.0, "receivers": 0.0}
        for r in res:
            tot["assembly"] += r[1]["assemble_ms"]
            tot["solve"] += r[1]["solve_ms"]
            tot["receivers"] += r[1]["post_ms"]
        out["split_ms"] = dict(tot, note="service time of the last step summed over its band requests (they "
                                         "overlap on the device): H-matrix assembly incl. the FP32 copy, batched "
                                         "solve, receivers + diffusion coefficients")

    rw = prof["rows_fp32"]
    if rw["launches"]:
        ach = rw["flops"] / (rw["ms"] * 1e-3) / 1e12
        out["roofline_served"] = {"kernel": "k_rows_mfma<float> (all pass widths, active rhs of each batched solve)",
                                  "bound": "mfma", "achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                  "frac": ach / FP32_PEAK_TFLOPS,
                                  "avg_launch_us": 1000 * rw["ms"] / rw["launches"], "launches": rw["launches"],
                                  "note": "HIP events over the served timed region: the band requests run "
                                          "concurrently, so a launch shares the GPU and its event interval also "
                                          "holds the wait for a free slot"}
    out["kernel_profile"] = prof
    iso = isolated_roofline(g, capi, ctx)
    out["roofline"] = iso["fp32"]
    out["roofline_fp64_37rhs"] = iso["fp64"]

    if not args.no_fp64:
        ctx.profile(True)
        lat64 = []
        t0 = time.perf_counter()
        res64 = run_fp64_plain(g, bands, dirs, L, capi, ctx, lat64)
        t64 = time.perf_counter() - t0
        prof64 = ctx.get_profile()
        ctx.profile(False)
        d64 = np.array([r[0] for r in res64])
        v64 = nsolve / t64
        err_d = float(np.abs(d_mixed - d64).max())
        out["fp64_plain"] = {"solves_per_s": v64, "s_per_step": t64,
                             "p50_ms": 1000 * float(np.percentile(lat64, 50)),
                             "p99_ms": 1000 * float(np.percentile(lat64, 99)),
                             "mean_iters": float(np.mean([np.mean(r[1]) for r in res64]))}
        out["speedup_mixed_vs_fp64"] = value / v64
        out["max_abs_err_diffusion_vs_fp64"] = err_d
        r64 = prof64["rows_fp64"]
        if r64["launches"]:
            gbs = r64["bytes"] / (r64["ms"] * 1e-3) / 1e9
            out["roofline_fp64_matvec"] = {"kernel": "k_rows_mfma<double> (1 rhs)", "bound": "hbm", "achieved": gbs,
                                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                                           "avg_launch_us": 1000 * r64["ms"] / r64["launches"],
                                           "launches": r64["launches"]}
        log("fp64 plain: %.2f s, %.1f solves/s, max |d_mixed - d_fp64| = %.2e" % (t64, v64, err_d))

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(P, E, dirs, L, bands)
        log("cpu baseline:", out["cpu_baseline"])
        out["cpu_baseline_1core"] = cpu_baseline(P, E, dirs, L, [b for b in bands if b in (250.0, 500.0)] or bands[:1],
                                                 threads=1)
        log("cpu baseline, 1 thread:", out["cpu_baseline_1core"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    set_hw_queues(sys.argv)
    main()
