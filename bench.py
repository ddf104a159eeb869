"""Benchmark of the MI355X pair-counting engine (one JSON line on stdout).

Metric: pair evaluations per second of npairs_3d on the ~1M-point clustered
periodic box (200,000 centres uniform in L = 1000 Mpc/h, each with Poisson(4)
NFW satellites, c = 5, r <= 0.5 Mpc/h; 20 log edges 0.1-50 Mpc/h; auto-count),
with inputs resident in HBM.  A step is one full npairs_3d call: GPU cell
build, pair kernel, reduction and the host round trip of the counts.

`value` counts the candidate pair evaluations the *reference* engine performs
for this input (its own RectangularDoubleMesh: 20^3 cells, 3x3x3 window),
divided by our wall time per step, so it is directly comparable with the CPU
baseline leg, which runs the reference loop restated in C (-Ofast, all host
threads the job may use) on the same input.  The roofline of the dominant kernel uses the
pair evaluations that kernel itself performs (8 FP64 flops each: 3 sub, 3 mul,
2 add, no FMA), timed with HIP events around the kernel on its own stream.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector, with FMA (non-FMA issue ceiling is half)
HBM_PEAK_GBPS = 8000.0
ALG_BUILD_BYTES = 76
# HBM bytes per launch of the pair kernel on this workload: rocprofv3
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors being what the same
# counters read for a plain 256 MiB copy (profiles/k_pairs_count3d_clustered_1M.md)
PAIR_KERNEL_TRAFFIC_BYTES = 5.843e8


def _progress(msg):
    """One line on stderr per finished leg (stdout carries only the JSON line)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def reference_candidates(pts, rmax, L):
    """Candidate pair evaluations of the reference engine on an auto-count.

    The reference mesh with its default cell sizes (approx_cell = the search
    length, PC/npairs_3d.py:202-209) has n_d = max(3, min(floor(L/search_d), 50))
    cells per dimension for both samples and a 3x3x3 window of cell-2s around
    each cell-1 (PC/rectangular_mesh.py:25-82, PC/cpairs/npairs_3d_engine.pyx:74-123);
    it evaluates n1(cell) * sum of n2 over the window for every cell.  ``rmax``
    is the search length, or a (x, y, z) triple of them ((rp, rp, pi) for the
    projected counters).
    """
    import numpy as np
    search = np.broadcast_to(np.asarray(rmax, dtype=float), (3,))
    n = [max(3, min(int(np.floor(L / s)), 50)) for s in search]
    idx = np.stack([np.minimum(np.floor(pts[:, d] // (L / n[d])).astype(int), n[d] - 1) for d in range(3)], axis=1)
    h = np.zeros(n, dtype=np.int64)
    np.add.at(h, (idx[:, 0], idx[:, 1], idx[:, 2]), 1)
    win = np.zeros_like(h)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                win += np.roll(h, (dx, dy, dz), axis=(0, 1, 2))
    return int((h * win).sum())


def bench_batched(dev, steps):
    """64 mocks of uneven size: npairs_3d auto + product-marked forward + backward.

    N_m log-uniform in [5e4, 2e6], uniform in a periodic L = 1000 box, log10 M
    uniform in [11, 15], Zheng07 central weights with logMmin_m ~ U[11.5, 12.5]
    and sigma = 0.3; loss_m = sum_k c_k W_mk.  The batched step (one HIP graph:
    weights and their derivative, one forward-mode pair traversal for counts, W
    and dW/dlogMmin, per-mock losses and gradients) is timed against the same
    work done mock by mock through the drop-in API (counts, marked forward with
    Jacobian rows, hand-written backward).  Like-for-like entries separate the
    gain from batching from the gain from the step algorithm: both step kinds
    (forward-mode "tangent", reverse-mode "reverse") batched in one graph and
    issued one mock per request.
    """
    import numpy as np
    import torch

    from halotools_amd import autotune, batch, mocks
    sizes, pts_h, lm_h, logMmin_h, c_h, L = mocks.many_mocks()
    M = len(sizes)
    rbins = mocks.config_rbins(5)
    pts = [torch.tensor(p, device=dev) for p in pts_h]
    lm = [torch.tensor(v, device=dev) for v in lm_h]
    logMmin = torch.tensor(logMmin_h, device=dev)
    c = torch.tensor(c_h, device=dev)
    del pts_h, lm_h

    b = batch.MockBatch(pts, L, rbins, device=dev)
    # autotune: cell size per density bucket and the workgroup shape, timed on
    # the batch's own largest mock of each bucket in the fused counts+Jacobian mode
    t0 = time.perf_counter()
    tuned = b.tune(autotune.TANGENT, reps=1)
    t_tune = time.perf_counter() - t0
    _progress("batched: tuned in %.1f s" % t_tune)

    build_s = _time_batch_build(b, 3)

    lm_all = torch.cat(lm)

    def graph_step(mode):
        st = batch.HODBatchStep(b, lm_all, logMmin, 0.3, c, mode=mode)
        st.run()
        torch.cuda.synchronize()
        assert st.status() == 0
        st.capture()
        st.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st.replay()
        torch.cuda.synchronize()
        _progress("batched %s step: %.2f ms" % (mode, 1e3 * (time.perf_counter() - t0) / steps))
        return st, (time.perf_counter() - t0) / steps

    step, t_batch = graph_step("tangent")
    counts, W, loss, grad = [t.clone() for t in step.outputs()]

    # finite differences for every mock at once (mocks are independent)
    h = 1e-4
    base = step.logMmin.clone()
    step.logMmin.copy_(base + h)
    lp = step.run()[2].clone()
    step.logMmin.copy_(base - h)
    lmn = step.run()[2].clone()
    step.logMmin.copy_(base)
    fd = (lp - lmn) / (2 * h)
    fd_rel = float(((fd - grad).abs() / fd.abs()).max())
    del step

    # the reverse-mode step batched the same way (count pass, Jacobian pass,
    # backward contraction, per-mock sums; one HIP graph)
    rstep, t_batch_rev = graph_step("reverse")
    rev_equal = bool(torch.equal(rstep.outputs()[0], counts))
    rev_grad_rel = float(((rstep.outputs()[3] - grad).abs() / grad.abs()).max())
    del rstep

    # like for like: the same two step kinds, one mock per request, prebuilt
    # outside the timed region and run eagerly mock after mock (no graph)
    def single_steps(mode):
        sts = []
        for m in range(M):
            bm = batch.MockBatch([pts[m]], L, rbins, device=dev)
            sts.append(batch.HODBatchStep(bm, lm[m], logMmin[m:m + 1], 0.3, c, mode=mode))
        for st in sts[:2]:
            st.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for st in sts:
            st.run()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        _progress("single %s steps: %.2f ms" % (mode, 1e3 * t))
        ok = all(bool(torch.equal(sts[m].outputs()[0][0], counts[m])) for m in range(M))
        g = torch.cat([st.outputs()[3] for st in sts])
        return t, ok, float(((g - grad).abs() / grad.abs()).max())

    t_single_tan, tan_ok, tan_rel = single_steps("tangent")
    t_single_rev, rev_ok, rev_rel = single_steps("reverse")

    # the same work one mock at a time through the drop-in API
    batch.one_at_a_time(pts[:2], lm[:2], logMmin[:2], 0.3, rbins, c, L)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref = batch.one_at_a_time(pts, lm, logMmin, 0.3, rbins, c, L)
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    counts_equal = all(bool(torch.equal(counts[m], ref[0][m])) for m in range(M))
    grad_single = torch.stack([g for g in ref[3]])
    grad_rel = float(((grad_single - grad).abs() / grad_single.abs()).max())
    npts = int(sizes.sum())
    return {"mocks": M, "points": npts, "ms_per_mock": 1e3 * t_batch / M,
            "batched_step_ms": 1e3 * t_batch, "one_at_a_time_ms_per_mock": 1e3 * t_single / M,
            "speedup_vs_one_at_a_time": t_single / t_batch, "counts_equal_one_at_a_time": counts_equal,
            "grad_max_rel_diff_vs_one_at_a_time": grad_rel, "grad_max_rel_diff_vs_central_fd": fd_rel,
            "fd_step": h, "graph": "one HIP graph per step (torch.cuda.CUDAGraph around the library calls)",
            # rows: how the 64 mocks are issued; columns: which step algorithm
            "modes_ms_per_step": {
                "batched_tangent": 1e3 * t_batch, "batched_reverse": 1e3 * t_batch_rev,
                "single_tangent": 1e3 * t_single_tan, "single_reverse": 1e3 * t_single_rev,
                "single_drop_in_api": 1e3 * t_single},
            "batching_gain_same_algorithm": {"tangent": t_single_tan / t_batch, "reverse": t_single_rev / t_batch_rev},
            "modes_check": {"batched_reverse_counts_equal": rev_equal, "batched_reverse_grad_rel": rev_grad_rel,
                            "single_tangent_counts_equal": tan_ok, "single_tangent_grad_rel": tan_rel,
                            "single_reverse_counts_equal": rev_ok, "single_reverse_grad_rel": rev_rel},
            "autotune": {"seconds": t_tune, "tile_waves": b.waves_for(autotune.TANGENT),
                         "buckets": {k: {"target_points_per_cell": v[0][0], "waves": v[0][1]}
                                     for k, v in tuned.items()}},
            "cell_build": _build_rate(npts, build_s)}


def _cpu_leg(fn, gpu_seconds, ref_cand, threads, reps=3, note=None):
    """The oracle's -Ofast restatement of the reference loop on the same input:
    one warm-up call, then the median of `reps` calls (whole calls: mesh build,
    engine, reduction)."""
    import numpy as np
    first = fn()
    walls = []
    for _ in range(reps):
        c0 = time.perf_counter()
        fn()
        walls.append(time.perf_counter() - c0)
    wall = float(np.median(walls))
    _progress("cpu leg: %.1f ms per call" % (wall * 1e3))
    out = {"ms_per_call": wall * 1e3, "threads": threads, "kind": "port", "host": host_cpus()[1],
           "reference_candidates": ref_cand, "candidates_per_s": ref_cand / wall,
           "gpu_speedup": wall / gpu_seconds}
    if note:
        out["note"] = note
    return first, out


OTHER_LEGS = ("xy_z", "per_object", "marked_fwd_bwd", "marked_xy_z", "jackknife", "tpcf", "wp")


def bench_other_configs(dev, steps, cpu_threads=0, legs=OTHER_LEGS):
    """GPU-only timings of the other hot-path rows on their full-size inputs.

    2M uniform points in L = 1000: npairs_xy_z with 16 log rp edges 0.1-30 and
    pi edges 0..40 step 1, npairs_projected with pi_max = 40, and wp.  1M
    uniform halos in L = 1000 with log10 M ~ U[11, 15]: product-weighted marked
    sums with Zheng07 central weights (logMmin = 12, sigma = 0.3), forward plus
    backward to (logMmin, sigma), and marked_npairs_xy_z of the same weights.
    The headline's ~1M-point clustered box: npairs_per_object_3d,
    npairs_jackknife_3d with 125 spatial subvolumes, and tpcf with analytic
    randoms.  Inputs are resident in HBM; each entry is the mean wall
    time of whole calls and, where given, the pair kernel's own evaluations per
    second from HIP events.  With ``cpu_threads`` > 0 each entry also carries
    ``cpu``: the oracle's -Ofast C restatement of the reference loop on the same
    input (on a stated subsample where the full one would take minutes), with
    results checked against the GPU's.
    """
    import numpy as np
    import torch

    from halotools_amd import _device as D
    from halotools_amd import _engine as E
    from halotools_amd import mocks
    from halotools_amd.autodiff import marked_npairs_3d_product
    from halotools_amd.mock_observables import npairs_projected, npairs_xy_z

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = fn()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
        _progress("timed %.3f ms per call" % (sec * 1e3))
        return out, sec

    out = {}
    T = cpu_threads
    if T:
        from oracle import mesh as oracle
    if "xy_z" in legs:
        p3, _, L = mocks.config_points(3)
        t3 = torch.from_numpy(np.ascontiguousarray(p3)).to(dev)
        rp, pi = mocks.config_rbins(3), np.arange(0, 41, 1.0)
        c_xyz, s_xyz = timed(lambda: npairs_xy_z(t3, t3, rp, pi, period=L))
        c_prj, s_prj = timed(lambda: npairs_projected(t3, t3, rp, 40.0, period=L))
        cols = tuple(t3[:, i].contiguous() for i in range(3))
        mock = dict(n1=len(p3), n2=len(p3), period=[L] * 3, pbc=True)
        for name, kind, p2, secs, cnt in (("npairs_xy_z", E.HP_NPAIRS_XY_Z, pi * pi, s_xyz, c_xyz),
                                          ("npairs_projected", E.HP_NPAIRS_PROJECTED, np.array([1600.0]), s_prj,
                                           c_prj)):
            ev = _events(torch)
            res = D.run(kind, [mock], cols, cols, rp * rp, p2=p2, same=True, want_evals=True, device=dev,
                        pair_events=(ev[0].cuda_event, ev[1].cuda_event))
            torch.cuda.synchronize()
            assert torch.equal(res.counts[0], cnt)
            ks = ev[0].elapsed_time(ev[1]) * 1e-3
            evals = int(res.evals[0].item())
            out[name] = {"workload": "2M uniform auto, L=1000, 16 log rp edges 0.1-30, "
                                     + ("pi edges 0..40 step 1" if kind == E.HP_NPAIRS_XY_Z else "pi_max 40"),
                         "ms_per_call": secs * 1e3, "pair_kernel_ms": ks * 1e3, "kernel_evals": evals,
                         "kernel_evals_per_s": evals / ks, "fp64_peak_frac": 8.0 * evals / ks / 1e12 / FP64_PEAK_TFLOPS}
            if T:
                cand = reference_candidates(p3, (rp[-1], rp[-1], 40.0), L)
                if kind == E.HP_NPAIRS_XY_Z:
                    fn = lambda: oracle.npairs_xy_z(p3, p3, rp, pi, period=L, num_threads=T, fast=True)
                else:
                    fn = lambda: oracle.npairs_projected(p3, p3, rp, 40.0, period=L, num_threads=T, fast=True)
                ref, out[name]["cpu"] = _cpu_leg(fn, secs, cand, T)
                assert np.array_equal(ref, cnt.cpu().numpy())
        del t3, cols

    if "per_object" in legs:
        # per-object counts on the headline's clustered box: (N, 20) int64 rows
        from halotools_amd.mock_observables import npairs_per_object_3d
        p2 = mocks.nfw_clustered_box(200000, 1000.0, np.random.default_rng([43, 2]))
        t2 = torch.from_numpy(np.ascontiguousarray(p2)).to(dev)
        rb2 = mocks.config_rbins(2)
        rows, s_obj = timed(lambda: npairs_per_object_3d(t2, t2, rb2, period=1000.0))
        cols2 = tuple(t2[:, i].contiguous() for i in range(3))
        ev = _events(torch)
        m2 = dict(n1=len(p2), n2=len(p2), period=[1000.0] * 3, pbc=True)
        res = D.run(E.HP_NPAIRS_PER_OBJECT_3D, [m2], cols2, cols2, rb2 * rb2, same=True, want_evals=True, device=dev,
                    pair_events=(ev[0].cuda_event, ev[1].cuda_event))
        torch.cuda.synchronize()
        assert torch.equal(res.counts, rows)
        ks = ev[0].elapsed_time(ev[1]) * 1e-3
        evals = int(res.evals[0].item())
        out["npairs_per_object_3d"] = {
            "workload": "%d-point NFW-clustered box (as the headline), L=1000, 20 log edges 0.1-50, rows (N, 20)"
                        % len(p2),
            "ms_per_call": s_obj * 1e3, "pair_kernel_ms": ks * 1e3, "kernel_evals": evals,
            "kernel_evals_per_s": evals / ks, "fp64_peak_frac": 8.0 * evals / ks / 1e12 / FP64_PEAK_TFLOPS}
        if T:
            ref, out["npairs_per_object_3d"]["cpu"] = _cpu_leg(
                lambda: oracle.npairs_per_object_3d(p2, p2, rb2, period=1000.0, num_threads=T, fast=True), s_obj,
                reference_candidates(p2, rb2[-1], 1000.0), T)
            assert np.array_equal(ref, rows.cpu().numpy())
        del t2, cols2, rows, res

    if "marked_fwd_bwd" in legs or "marked_xy_z" in legs:
        rng = np.random.default_rng([43, 4])
        p4 = mocks.uniform_box(1000000, 1000.0, rng)
        logm = torch.tensor(rng.uniform(11, 15, len(p4)), device=dev)
        t4 = torch.from_numpy(p4).to(dev)
        rb = mocks.config_rbins(4)
        c = torch.tensor(np.random.default_rng(7).normal(size=len(rb)), device=dev)
        theta = torch.tensor([12.0, 0.3], dtype=torch.float64, device=dev, requires_grad=True)
    if "marked_fwd_bwd" in legs:

        def fwd_bwd():
            theta.grad = None
            w = mocks.zheng07_ncen(logm, theta[0], theta[1])
            W = marked_npairs_3d_product(t4, t4, rb, w, w, period=1000.0)
            (W * c).sum().backward()
            return theta.grad

        _, s_fb = timed(fwd_bwd)
        out["marked_npairs_3d_product_fwd_bwd"] = {
            "workload": "1M uniform halos, L=1000, 20 log edges 0.1-50, Zheng07 weights, d/d(logMmin, sigma)",
            "ms_per_fwd_bwd": s_fb * 1e3}
        if T:
            w4h = mocks.zheng07_ncen(logm, torch.tensor(12.0, dtype=torch.float64, device=dev),
                                     torch.tensor(0.3, dtype=torch.float64, device=dev)).cpu().numpy()
            _, out["marked_npairs_3d_product_fwd_bwd"]["cpu"] = _cpu_leg(
                lambda: oracle.marked_npairs_3d(p4, p4, rb, period=1000.0, weights1=w4h, weights2=w4h,
                                                weight_func_id=1, num_threads=T, fast=True),
                s_fb, reference_candidates(p4, rb[-1], 1000.0), T,
                note="forward sums only: the reference has no backward, so gpu_speedup compares the GPU's "
                     "forward + backward with the CPU's forward")

    if "marked_xy_z" in legs:
        # marked (rp, pi) sums of the same halos, product of the Zheng07 weights
        from halotools_amd.mock_observables import marked_npairs_xy_z
        w4 = mocks.zheng07_ncen(logm, torch.tensor(12.0, dtype=torch.float64, device=dev),
                                torch.tensor(0.3, dtype=torch.float64, device=dev)).contiguous()
        rp4, pi4 = mocks.config_rbins(3), np.arange(0, 41, 1.0)
        wx, s_mx = timed(lambda: marked_npairs_xy_z(t4, t4, rp4, pi4, period=1000.0, weights1=w4, weights2=w4,
                                                    weight_func_id=1))
        cols4 = tuple(t4[:, i].contiguous() for i in range(3))
        ev = _events(torch)
        m4 = dict(n1=len(p4), n2=len(p4), period=[1000.0] * 3, pbc=True)
        res = D.run(E.HP_MARKED_NPAIRS_XY_Z, [m4], cols4, cols4, rp4 * rp4, p2=pi4 * pi4, same=True,
                    w1=w4.view(-1, 1), nw=1, wfunc=1, want_counts=False, want_sums=True, want_evals=True, device=dev,
                    pair_events=(ev[0].cuda_event, ev[1].cuda_event))
        torch.cuda.synchronize()
        assert torch.equal(res.sums[0], wx)
        ks = ev[0].elapsed_time(ev[1]) * 1e-3
        evals = int(res.evals[0].item())
        out["marked_npairs_xy_z"] = {
            "workload": "1M uniform halos, L=1000, 16 log rp edges 0.1-30, pi edges 0..40 step 1, "
                        "weight_func_id 1 with Zheng07 weights",
            "ms_per_call": s_mx * 1e3, "pair_kernel_ms": ks * 1e3, "kernel_evals": evals,
            "kernel_evals_per_s": evals / ks, "fp64_peak_frac": 8.0 * evals / ks / 1e12 / FP64_PEAK_TFLOPS}
        if T:
            w4h = w4.cpu().numpy()
            ref, out["marked_npairs_xy_z"]["cpu"] = _cpu_leg(
                lambda: oracle.marked_npairs_xy_z(p4, p4, rp4, pi4, period=1000.0, weights1=w4h, weights2=w4h,
                                                  weight_func_id=1, num_threads=T, fast=True),
                s_mx, reference_candidates(p4, (rp4[-1], rp4[-1], 40.0), 1000.0), T)
            np.testing.assert_allclose(ref, wx.cpu().numpy(), rtol=1e-11)
        del t4, cols4, w4, res, wx

    if "jackknife" in legs or "tpcf" in legs:
        # jackknife counts and the two-point caller layer on the headline's clustered box
        from halotools_amd.mock_observables import npairs_jackknife_3d, tpcf, wp
        p2c = mocks.config_points(2)[0]
        t2c = torch.from_numpy(np.ascontiguousarray(p2c)).to(dev)
        rb2 = mocks.config_rbins(2)
        sub = np.minimum((p2c / 1000.0 * 5).astype(np.int64), 4)
        tags = torch.from_numpy((1 + (sub[:, 0] * 5 + sub[:, 1]) * 5 + sub[:, 2]).astype(np.int32)).to(dev)
    if "jackknife" in legs:
        jk, s_jk = timed(lambda: npairs_jackknife_3d(t2c, t2c, rb2, tags, tags, 125, period=1000.0))
        out["npairs_jackknife_3d"] = {
            "workload": "%d-point NFW-clustered box (as the headline), L=1000, 20 log edges 0.1-50, "
                        "125 spatial subvolumes, unit weights, table (126, 20)" % len(p2c),
            "ms_per_call": s_jk * 1e3}
        if T:
            # the reference engine walks the bins once per subsample (126 times per
            # candidate pair): timed on every 8th point (1/64 of the candidates),
            # with the GPU on the same subsample beside it
            ps = np.ascontiguousarray(p2c[::8])
            tg = tags.cpu().numpy()[::8]
            ts, tgs = torch.from_numpy(ps).to(dev), torch.from_numpy(np.ascontiguousarray(tg)).to(dev)
            gsub, s_jks = timed(lambda: npairs_jackknife_3d(ts, ts, rb2, tgs, tgs, 125, period=1000.0))
            ref, leg = _cpu_leg(lambda: oracle.npairs_jackknife(ps, ps, rb2, tg, tg, 125, period=1000.0, num_threads=T,
                                                                fast=True),
                                s_jks, reference_candidates(ps, rb2[-1], 1000.0), T, reps=1,
                                note="every 8th point of the box (%d points); GPU on the same subsample: %.3f ms"
                                     % (len(ps), s_jks * 1e3))
            assert np.array_equal(ref, gsub.cpu().numpy())
            out["npairs_jackknife_3d"]["cpu"] = leg
    if "tpcf" in legs:
        xi, s_tp = timed(lambda: tpcf(t2c, rb2, period=1000.0))
        out["tpcf"] = {"workload": "xi(r) of the headline box, Natural estimator, analytic randoms "
                                   "(DD on the device, estimator on the device)", "ms_per_call": s_tp * 1e3}
        if T:
            from oracle import clustering as OC

            def cpu_tpcf():
                dd = np.diff(oracle.npairs_3d(p2c, p2c, rb2, period=1000.0, num_threads=T, fast=True))
                dv = np.diff(4.0 / 3.0 * np.pi * rb2 ** 3)
                n = len(p2c)
                return OC.estimator(dd, None, dv * (n * n / 1e9), n, n, n, n, "Natural")
            ref, out["tpcf"]["cpu"] = _cpu_leg(cpu_tpcf, s_tp, reference_candidates(p2c, rb2[-1], 1000.0), T)
            np.testing.assert_allclose(ref, xi.cpu().numpy(), rtol=1e-12)
    if "wp" in legs:
        from halotools_amd.mock_observables import wp
        p3 = mocks.config_points(3)[0]
        t3 = torch.from_numpy(np.ascontiguousarray(p3)).to(dev)
        rb3 = mocks.config_rbins(3)
        wpv, s_wp = timed(lambda: wp(t3, rb3, 40.0, period=1000.0))
        out["wp"] = {"workload": "wp(rp) of 2M uniform points, L=1000, 16 log rp edges 0.1-30, pi_max 40, "
                                 "analytic randoms", "ms_per_call": s_wp * 1e3}
        if T:
            from oracle import clustering as OC

            def cpu_wp():
                dd = np.diff(np.diff(oracle.npairs_xy_z(p3, p3, rb3, np.array([0.0, 40.0]), period=1000.0,
                                                        num_threads=T, fast=True), axis=0), axis=1)
                dv = np.diff(np.diff(np.pi * np.outer(rb3 ** 2, 2.0 * np.array([0.0, 40.0])), axis=0), axis=1)
                n = len(p3)
                return 2.0 * OC.estimator(dd, None, dv * (n * n / 1e9), n, n, n, n, "Natural")[:, 0] * 40.0
            ref, out["wp"]["cpu"] = _cpu_leg(cpu_wp, s_wp, reference_candidates(p3, (rb3[-1], rb3[-1], 40.0), 1000.0), T)
            np.testing.assert_allclose(ref, wpv.cpu().numpy(), rtol=1e-12)
    return out


def _events(torch):
    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    for e in ev:
        e.record()  # materialise the HIP event handles
    return ev


def _time_batch_build(b, reps):
    """Cell build of a MockBatch (all mocks in one counting sort): HIP events
    around the build of an npairs_3d call; median of reps calls, in seconds."""
    import numpy as np
    import torch
    build_ms = []
    for _ in range(reps):
        ev = _events(torch)
        b.npairs_3d(build_events=(ev[0].cuda_event, ev[1].cuda_event))
        torch.cuda.synchronize()
        build_ms.append(ev[0].elapsed_time(ev[1]))
    return float(np.median(build_ms)) * 1e-3


def bench_build64(dev, steps):
    """The 64-mock batch's cell build alone (the batched leg's mocks and tuned
    cells, without the steps): 37.2M points sorted in both orders."""
    import torch

    from halotools_amd import autotune, batch, mocks
    sizes, pts_h, _, _, _, L = mocks.many_mocks()
    pts = [torch.tensor(p, device=dev) for p in pts_h]
    del pts_h
    b = batch.MockBatch(pts, L, mocks.config_rbins(5), device=dev)
    b.tune(autotune.TANGENT, reps=1)
    build_s = _time_batch_build(b, max(3, steps))
    _progress("build64: %.3f ms" % (1e3 * build_s))
    return _build_rate(int(sizes.sum()), build_s)


def bench_host_phases(dev, t, rbins, L, calls=60):
    """Host time per phase of the headline drop-in call (median over `calls`):
    the argument checks, the column split of the (N, 3) tensor, the request
    build, the library call (planning and kernel enqueues) and the wait for
    the status word -- the same steps as npairs_3d, timed one by one."""
    import numpy as np
    import torch

    from halotools_amd import _device as D
    from halotools_amd import _engine as E
    from halotools_amd.mock_observables.pair_counters import _common as C
    from halotools_amd.mock_observables.pair_counters import _process as P
    names = ("argument_checks", "column_split", "request_build", "library_call", "wait", "total")
    ph = {k: [] for k in names}
    torch.cuda.synchronize()
    for _ in range(calls):
        t0 = time.perf_counter()
        x1, y1, z1 = P.columns(t)
        x2, y2, z2 = P.columns(t)
        rb = np.atleast_1d(rbins).astype("f8")
        rmax = np.max(rb)
        rb = P.process_bins(rb, P.RBINS_MSG)
        P.check_last_edge(rb)
        period = P.process_period(L)
        a1, a2 = P.default_cell_sizes(None, None, (rmax, rmax, rmax))
        P.set_approximate_cell_sizes(a1, a2, period)
        P.check_search_lengths((rmax, rmax, rmax), tuple(period))
        device = D.require_gpu(t.device)
        same = D.same_array(t, t)
        t1 = time.perf_counter()
        s1, s2 = C.upload((x1, y1, z1), (x2, y2, z2), same, device)
        mock = C.mock_dict(len(s1[0]), len(s2[0]), period, True)
        t2 = time.perf_counter()
        res = D.prepare(E.HP_NPAIRS_3D, [mock], s1, s2, P.squared(rb), same=same, device=device)
        t3 = time.perf_counter()
        D.launch(res)
        t4 = time.perf_counter()
        C.finish(res, lambda host: res.counts[0], True, period, True)
        t5 = time.perf_counter()
        for k, v in zip(names, (t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4, t5 - t0)):
            ph[k].append(v * 1e6)
    out = {k + "_us": float(np.median(v)) for k, v in ph.items()}
    _progress("host phases: " + ", ".join("%s %.1f us" % (k, v) for k, v in out.items()))
    return out


def _build_rate(npts, seconds):
    """Cell build of one auto-count sample against the HBM roofline.

    Algorithmic bytes of a counting-sort cell list: read x, y, z twice (48 B),
    write the sorted x, y, z (24 B) and the int32 permutation (4 B): 76 B per
    point; the cell-start table is negligible.
    """
    gbs = ALG_BUILD_BYTES * npts / seconds / 1e9
    return {"points": npts, "ms": seconds * 1e3, "alg_bytes_per_point": ALG_BUILD_BYTES, "achieved_GBps": gbs,
            "peak_GBps": HBM_PEAK_GBPS, "frac": gbs / HBM_PEAK_GBPS}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: every CPU this process may run on (its affinity set, or OMP_NUM_THREADS when set)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-mock batched leg")
    ap.add_argument("--no-others", action="store_true", help="skip the xy_z / projected / marked timings")
    ap.add_argument("--legs", default="all",
                    help="comma list of legs to run besides the headline: batched, " + ", ".join(OTHER_LEGS) +
                         "; build64 (the batch's cell build alone) and hostphases (host time per phase of "
                         "the headline call) run only when named")
    return ap.parse_args()


def host_cpus():
    """CPUs the CPU baseline may use, and a description of the host.

    The threads are the CPUs this process may run on (its affinity set),
    capped by OMP_NUM_THREADS when the environment sets it (a shared host
    that gives this job a share of its cores says so there).
    """
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(allowed, int(omp)) if omp.isdigit() and int(omp) > 0 else allowed
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"os_cpu_count": os.cpu_count(), "affinity_cpus": allowed, "OMP_NUM_THREADS": omp or None,
                     "model_name": model}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the pair counters run on a single MI355X; sharding "
              "across GPUs is not implemented, so --gpus must be 1.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from halotools_amd import _device as D
    from halotools_amd import _engine as E
    from halotools_amd import mocks
    from halotools_amd.mock_observables import npairs_3d

    if not os.path.exists(E.LIB_PATH):
        from halotools_amd import build_ext
        build_ext.build()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng([43, 2])
    pts = mocks.nfw_clustered_box(200000, 1000.0, rng)
    L = 1000.0
    rbins = mocks.config_rbins(2)
    n = len(pts)
    t = torch.from_numpy(np.ascontiguousarray(pts)).to(dev)
    torch.cuda.synchronize()

    # -------- timed region: whole npairs_3d calls on HBM-resident inputs
    for _ in range(args.warmup):
        out = npairs_3d(t, t, rbins, period=L)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = npairs_3d(t, t, rbins, period=L)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    counts = out.cpu().numpy()
    sec_per_step = (t1 - t0) / args.steps
    _progress("headline: %.3f ms per call" % (sec_per_step * 1e3))

    # the same call with NumPy in and NumPy out (host-to-device copy of the
    # points and the copy back of the counts included; not the headline value)
    host_ms = []
    for _ in range(max(3, args.steps // 4)):
        h0 = time.perf_counter()
        hc = npairs_3d(pts, pts, rbins, period=L)
        host_ms.append(1e3 * (time.perf_counter() - h0))
    assert np.array_equal(hc, counts)
    numpy_io_ms = float(np.median(host_ms))

    # -------- dominant kernel: HIP events around the pair kernel on its stream
    cols = (t[:, 0].contiguous(), t[:, 1].contiguous(), t[:, 2].contiguous())
    mock = dict(n1=n, n2=n, period=[L, L, L], pbc=True)
    kern_ms = []
    build_ms = []
    own_evals = None
    for _ in range(max(3, args.steps // 2)):
        ev = _events(torch)
        bev = _events(torch)
        res = D.run(E.HP_NPAIRS_3D, [mock], cols, cols, rbins * rbins, same=True, want_evals=True, device=dev,
                    pair_events=(ev[0].cuda_event, ev[1].cuda_event),
                    build_events=(bev[0].cuda_event, bev[1].cuda_event))
        torch.cuda.synchronize()
        kern_ms.append(ev[0].elapsed_time(ev[1]))
        build_ms.append(bev[0].elapsed_time(bev[1]))
        own_evals = int(res.evals[0].item())
        assert np.array_equal(res.counts[0].cpu().numpy(), counts)
    kern_s = float(np.mean(kern_ms[1:])) * 1e-3 if len(kern_ms) > 1 else kern_ms[0] * 1e-3
    build_s = float(np.mean(build_ms[1:])) * 1e-3 if len(build_ms) > 1 else build_ms[0] * 1e-3
    achieved = 8.0 * own_evals / kern_s / 1e12

    # -------- reference-engine work for this input (its own mesh), and the CPU baseline
    ref_cand = reference_candidates(pts, float(np.max(rbins)), L)
    value = ref_cand / sec_per_step

    cpu = None
    threads, host = host_cpus()
    threads = args.cpu_threads or threads
    if not args.no_cpu_baseline:
        from oracle import mesh as oracle
        # one warm-up, then the median of 5 whole calls (mesh build + engine + reduction)
        ref = oracle.npairs_3d(pts, pts, rbins, period=L, num_threads=threads, fast=True)
        assert np.array_equal(ref, counts), "GPU counts differ from the CPU reference loop"
        walls = []
        for _ in range(5):
            c0 = time.perf_counter()
            oracle.npairs_3d(pts, pts, rbins, period=L, num_threads=threads, fast=True)
            walls.append(time.perf_counter() - c0)
        wall = float(np.median(walls))
        # the two phases of a call timed apart (median of 3): mesh build and
        # argument handling (NumPy argsort), and the engine loop alone
        phases = [oracle.npairs_3d_phases(pts, pts, rbins, period=L, num_threads=threads, fast=True)[1:]
                  for _ in range(3)]
        mesh_s = float(np.median([p[0] for p in phases]))
        engine_s = float(np.median([p[1] for p in phases]))
        # single core: one call on the same input
        c0 = time.perf_counter()
        oracle.npairs_3d(pts, pts, rbins, period=L, num_threads=1, fast=True)
        wall1 = time.perf_counter() - c0
        cpu = {"value": ref_cand / wall, "unit": "pair evals/s", "cores": threads, "kind": "port",
               "sample": "the full workload: reference mesh + engine loop restated in C (oracle/cpu_ref.c, "
                         "-Ofast, no -march), cell-1 range split over %d threads; median of 5 calls after a "
                         "warm-up, %.2f s each; counts checked equal to the GPU's" % (threads, wall),
               "single_core_s": wall1, "single_core_value": ref_cand / wall1,
               "mesh_build_s": mesh_s, "engine_only_s": engine_s, "engine_only_value": ref_cand / engine_s,
               "host": host}

    legs = set(["batched"] + list(OTHER_LEGS)) if args.legs == "all" else set(args.legs.split(","))
    batched = None if (args.no_batched or "batched" not in legs) else bench_batched(dev, max(3, args.steps // 4))
    build64 = bench_build64(dev, max(3, args.steps // 4)) if "build64" in legs else None
    host_phases = bench_host_phases(dev, t, rbins, L) if "hostphases" in legs else None
    other = [g for g in OTHER_LEGS if g in legs]
    others = None if (args.no_others or not other) else bench_other_configs(
        dev, max(3, args.steps // 4), cpu_threads=0 if args.no_cpu_baseline else threads, legs=other)

    line = {
        "metric": "pair evals/s & %FP64 peak (npairs_3d, 1M pts); batched fwd+bwd ms per mock",
        "value": value,
        "unit": "pair evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded NFW-clustered periodic box",
        "config": {"workload": "npairs_3d auto-count, %d points (200k centres + Poisson(4) NFW satellites), "
                               "L=1000 periodic, 20 log rbins 0.1-50" % n,
                   "points": n, "nbins": len(rbins), "period": L},
        "reference_mesh_candidates": ref_cand,
        "kernel_evals": own_evals,
        "kernel_evals_per_s": own_evals / kern_s,
        "inrange_pairs_per_s": float(counts[-1]) / sec_per_step,
        "pair_kernel_ms": kern_s * 1e3,
        "fp64_peak_frac": achieved / FP64_PEAK_TFLOPS,
        "cell_build": _build_rate(n, build_s),
        "numpy_in_out_ms_per_call": numpy_io_ms,
        # compute-bound on the FP64 vector ALU (a distance threshold, not a
        # contraction, so no MFMA); MI355X's FP64 peak is 78.6 TFLOP/s
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": PAIR_KERNEL_TRAFFIC_BYTES,
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this kernel on this "
                                       "workload, scaled by a plain copy's readings "
                                       "(profiles/k_pairs_count3d_clustered_1M.md)"},
        "batched_64_mocks": batched,
        "cell_build_64_mocks": build64,
        "headline_host_phases": host_phases,
        "other_configs": others,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
