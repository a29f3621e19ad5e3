#!/usr/bin/env python3
"""Benchmark: SPD solve GFLOP/s at n = 32768 on one MI355X.

One step = one SPD solve on device-resident inputs: the blocked POTRF of an
n x n f64 matrix followed by the forward and back triangular solves with m
right-hand sides (llh_dev_solvec_d, the device-memory form of solvec).  Inputs
are synthetic and seeded: A = B B^T / n + lambda I with B ~ N(0, 1), built on
the GPU before the timed region; every step factors its own copy of A, so the
timed region contains only the solve.

GFLOP/s counts the algorithmic work n^3/3 + n^2/2 + n/6 + 2 n^2 m.  The
dominant kernel (the trailing SYRK on MFMA, of rank 1024 at n >= 24576) is
timed live with HIP events on the launch stream; its roofline fraction is
against the FP64 peak.
The CPU baseline is the oracle (the reference algorithm restated in C, one
thread) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "linfa-linalg_amd"))

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "f16": 2500.0}  # dense MFMA peaks (f16: 2.5 PFLOP/s)
def outer_width(n):
    """Outer panel width of the library's blocked f64 factor (the rank of its
    trailing SYRK): 1024 from n = 24576 on, else 512."""
    return 1024 if n >= 24576 else 512
PEAK_HBM_GBS = 8000.0


def solve_flops(n, m):
    return n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0 + 2.0 * n * n * m


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(sizes, m, seed, big_n):
    """The oracle (the reference algorithm restated in C, one thread) timed on
    one solvec per size; the largest size is the reported sample, and a c n^3
    fit from it extrapolates the full-size solves (labelled as such)."""
    import numpy as np
    from oracle import oracle as orc
    rng = np.random.default_rng(seed)
    per = {}
    for n in sizes:
        b = rng.standard_normal((n, n))
        a = b @ b.T / n + np.eye(n)
        rhs = rng.standard_normal((n, m))
        times = []
        for _ in range(2 if n <= 2048 else 1):
            work = np.array(a)
            t0 = time.perf_counter()
            orc.solvec(work, rhs)
            times.append(time.perf_counter() - t0)
        per[n] = min(times)
        log(f"bench: CPU oracle n={n}: {per[n]:.3f} s")
    n = max(sizes)
    t = per[n]
    c = t / solve_flops(n, m)
    extrap = {str(k): round(c * solve_flops(k, mm), 1) for k, mm in big_n}
    return {"value": round(solve_flops(n, m) / t / 1e9, 3), "unit": "GFLOP/s", "cores": 1, "kind": "port",
            "sample": f"oracle solvec (reference algorithm in C, 1 thread), A=B*B^T/n+I, {m} RHS; sample n={n}: "
                      f"{t:.2f} s",
            "seconds_by_n": {str(k): round(v, 4) for k, v in per.items()},
            "extrapolated_seconds_by_n": extrap,
            "extrapolation": "c*flops fit to the largest sample (not measured)"}


def oracle_fixture(n, m, lam, seed):
    """The multi-threaded oracle's solution of the seeded workload (n, m, lam,
    seed), as committed under tests/golden/full_size (a column subset), or None."""
    import numpy as np
    fix = os.path.join(ROOT, "tests", "golden", "full_size")
    try:
        with open(os.path.join(fix, "meta.json")) as f:
            meta = json.load(f)
    except OSError:
        return None
    for cfg in meta.values():
        if cfg["n"] == n and cfg["m"] == m and cfg["seed"] == seed and cfg["lam"] == lam:
            return np.load(os.path.join(fix, cfg["x_ref"]))
    return None


def sol_err(x, x_ref):
    """||x - x_ref|| / ||x_ref|| over the fixture's columns."""
    import numpy as np
    k = x_ref.shape[1]
    xs = x[:, :k].double().cpu().numpy()
    return float(np.linalg.norm(xs - x_ref) / np.linalg.norm(x_ref))


def batched_leg(dev, torch, args):
    """20000 SPD systems, n log-uniform in [8, 512], A_i = B_i B_i^T / n + I with
    B_i uniform in [-1, 1], 1-8 right-hand sides ~ N(0, 1): one planned,
    autotuned batch solve (one hipGraph launch), f64; against the same systems
    solved one plain llh_dev_solvec call at a time (a sample, scaled)."""
    import numpy as np
    rng = np.random.default_rng(args.seed + 4)
    cnt = args.batch
    sizes = np.floor(np.exp(rng.uniform(np.log(8), np.log(512), cnt))).astype(np.int32)
    sizes = np.minimum(sizes, 512)
    rhs = rng.integers(1, 9, cnt).astype(np.int32)
    once = [int(v) for v in args.batch_once.split(",")] if args.batch_once else None
    plan = (dev.BatchPlan(sizes, rhs, prec="f64", small_limit=once[0], small_waves=once[1],
                          medium_mode=once[2] if len(once) > 2 else 0) if once
            else dev.BatchPlan(sizes, rhs, prec="f64"))
    a = torch.zeros(plan.a_elems, dtype=torch.float64, device="cuda")
    b = torch.zeros(plan.b_elems, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(args.seed + 5)
    # one scatter per size (not one copy per system): the systems of a size
    # share a leading dimension
    for s in np.unique(sizes):
        s = int(s)
        idx = np.nonzero(sizes == s)[0]
        bm = torch.rand(len(idx), s, s, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        am = torch.bmm(bm, bm.transpose(1, 2).contiguous()) / s
        am += torch.eye(s, dtype=torch.float64, device="cuda")
        ld = int(plan.lda[idx[0]])
        assert np.all(plan.lda[idx] == ld)
        within = (np.arange(s)[:, None] * ld + np.arange(s)[None, :]).reshape(-1)
        flat = (plan.a_off[idx].astype(np.int64)[:, None] + within[None, :]).reshape(-1)
        a[torch.from_numpy(flat).cuda()] = am.reshape(-1)
        bflat = np.concatenate([int(o) + np.arange(s * int(m)) for o, m in zip(plan.b_off[idx], rhs[idx])])
        b[torch.from_numpy(bflat).cuda()] = torch.randn(len(bflat), dtype=torch.float64, device="cuda",
                                                        generator=g)
    status = torch.zeros(cnt, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    n64 = sizes.astype(np.float64)
    flops = float(np.sum(n64 ** 3 / 3 + n64 ** 2 / 2 + n64 / 6 + 2 * n64 ** 2 * rhs))
    byts = float(np.sum(8 * n64 * (n64 + 1) + 16 * n64 * rhs))
    if once:
        # exactly one graph launch with the given configuration and nothing
        # else of the library's, for counter runs; the clone of A is a copy of
        # known size (a_elems x 8 bytes) that scales the HBM counters
        a_copy = a.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.solve(a, b, status)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        res = {"batch_once": True, "count": cnt, "config": plan.config, "ms_first_launch": round(t * 1e3, 3),
               "ok": int((status == -1).sum().item()), "a_copy_bytes": int(a_copy.numel()) * 8,
               "algorithmic_bytes": byts, "algorithmic_flops": flops}
        plan.close()
        return res
    log("bench: batched autotune")
    tuned_ms = plan.autotune(a, b, status)
    a0, b0 = a.clone(), b.clone()
    times = []
    for rep in range(1 + args.steps):
        a.copy_(a0)
        b.copy_(b0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.solve(a, b, status)
        t1 = time.perf_counter()
        if rep:
            times.append(t1 - t0)
    t = float(np.median(times))
    ok = int((status == -1).sum().item())
    roof_t = max(flops / (PEAK_TFLOPS["f64"] * 1e12), byts / (PEAK_HBM_GBS * 1e9))
    # parity sample against the oracle
    from oracle import oracle as orc
    xs = b.cpu().numpy()
    worst = 0.0
    for i in rng.choice(cnt, 40, replace=False):
        s, m = int(sizes[i]), int(rhs[i])
        o, ld = int(plan.a_off[i]), int(plan.lda[i])
        am = a0[o:o + s * ld].view(s, ld)[:, :s].cpu().numpy()
        bm = b0[int(plan.b_off[i]):int(plan.b_off[i]) + s * m].view(s, m).cpu().numpy()
        xr = orc.solvec(np.array(am), bm)
        x = xs[int(plan.b_off[i]):int(plan.b_off[i]) + s * m].reshape(s, m)
        worst = max(worst, float(np.linalg.norm(x - xr) / np.linalg.norm(xr)))
    # the plain per-matrix path on a sample of the same systems
    sample = rng.choice(cnt, min(300, cnt), replace=False)
    mats = []
    for i in sample:
        s, m = int(sizes[i]), int(rhs[i])
        o, ld = int(plan.a_off[i]), int(plan.lda[i])
        mats.append((a0[o:o + s * ld].view(s, ld)[:, :s].contiguous(),
                     b0[int(plan.b_off[i]):int(plan.b_off[i]) + s * m].view(s, m).contiguous()))
    work = [(x.clone(), y.clone()) for x, y in mats]
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for x, y in work:
        t_a = time.perf_counter()
        dev.solvec(x, y)
        lat.append(time.perf_counter() - t_a)
    t1 = time.perf_counter()
    plain_rate = len(work) / (t1 - t0)
    frac = float(np.sum(sizes[sample].astype(np.float64) ** 3)) / float(np.sum(n64 ** 3))
    plain_rate_scaled = len(work) / ((t1 - t0) / frac * len(work) / cnt) if frac > 0 else None
    cfg64 = plan.config
    plan.close()
    del work, mats
    # the same systems with FP32 factors and f64 refinement (f64 buffers, A kept), one graph launch
    plan32 = dev.BatchPlan(sizes, rhs, prec="f32ir", ir_iterations=2)
    assert np.array_equal(plan32.a_off, plan.a_off) and np.array_equal(plan32.b_off, plan.b_off)
    a.copy_(a0)
    b.copy_(b0)
    log("bench: batched FP32+IR autotune")
    tuned32 = plan32.autotune(a, b, status)
    times32 = []
    for rep in range(1 + args.steps):
        b.copy_(b0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan32.solve(a, b, status)
        t1 = time.perf_counter()
        if rep:
            times32.append(t1 - t0)
    t32 = float(np.median(times32))
    ok32 = int((status == -1).sum().item())
    xs32 = b.cpu().numpy()
    worst32 = 0.0
    for i in rng.choice(cnt, 40, replace=False):
        s, m = int(sizes[i]), int(rhs[i])
        o, ld = int(plan32.a_off[i]), int(plan32.lda[i])
        am = a0[o:o + s * ld].view(s, ld)[:, :s].cpu().numpy()
        bm = b0[int(plan32.b_off[i]):int(plan32.b_off[i]) + s * m].view(s, m).cpu().numpy()
        xr = orc.solvec(np.array(am), bm)
        x = xs32[int(plan32.b_off[i]):int(plan32.b_off[i]) + s * m].reshape(s, m)
        worst32 = max(worst32, float(np.linalg.norm(x - xr) / np.linalg.norm(xr)))
    fp32 = {"workload": f"the same {cnt} systems: f32 factors and solves, 2 f64 refinement steps, f64 buffers, "
                        "one hipGraph launch",
            "solves_per_s": round(cnt / t32, 1), "ms": round(t32 * 1e3, 3),
            "gflops": round(flops / t32 / 1e9, 1), "speedup_vs_f64": round(t / t32, 3),
            "autotuned_ms": round(tuned32, 3), "config": plan32.config, "ok": ok32,
            "max_rel_err_vs_oracle_sample": worst32}
    log(f"bench: batched FP32+IR {fp32}")
    plan32.close()
    return {"workload": f"{cnt} SPD systems, n log-uniform [8,512], 1-8 RHS, f64, one hipGraph launch",
            "solves_per_s": round(cnt / t, 1), "ms": round(t * 1e3, 3), "gflops": round(flops / t / 1e9, 1),
            "gbs_algorithmic": round(byts / t / 1e9, 1), "roofline_frac": round(roof_t / t, 4),
            "roofline_bound": "mfma" if flops / (PEAK_TFLOPS["f64"] * 1e12) > byts / (PEAK_HBM_GBS * 1e9)
            else "hbm", "config": cfg64, "autotuned_ms": round(tuned_ms, 3),
            "traffic": load_batched_traffic(cnt, cfg64),
            "ok": ok, "max_rel_err_vs_oracle_sample": worst,
            "plain_per_matrix": {"solves_per_s_sample": round(plain_rate, 1),
                                 "solves_per_s_same_mix_est": round(plain_rate_scaled, 1) if plain_rate_scaled
                                 else None,
                                 "p50_ms": round(float(np.median(lat)) * 1e3, 3),
                                 "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)},
            "fp32_ir": fp32}


def invc_leg(dev, torch, args):
    """InverseC at n = 8192, f64: the dedicated inverse (llh_dev_invc_d: factor,
    W = L^-1 and W^T W on MFMA) against the reference's own route, solvec on the
    identity (llh_dev_solvec_d with n right-hand sides).  GFLOP/s count the
    reference route's algorithmic work, n^3/3 + 2 n^3, for both."""
    import numpy as np
    n = 8192
    a = dev.gen_spd(dev.GEN_NORMAL, args.seed + 5, n, 1.0)
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    out = torch.empty_like(a)
    fl = n ** 3 / 3.0 + 2.0 * n ** 3

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    def dedicated():
        w = a.clone()
        dev.invc(w, out)

    def via_solve():
        w, x = a.clone(), eye.clone()
        dev.solvec(w, x)
        return x

    t_d = timed(dedicated)
    x_s = via_solve()
    t_s = timed(via_solve, reps=2)
    diff = float((out - x_s).norm() / x_s.norm())
    res = float((a @ out - eye).abs().max())
    del a, eye, out, x_s
    torch.cuda.empty_cache()
    row = {"workload": f"invc n={n} f64 (A = B B^T/n + I, seeded)", "ms": round(t_d * 1e3, 3),
           "value": round(fl / t_d / 1e9, 1), "unit": "GFLOP/s (reference route's n^3/3 + 2n^3)",
           "solvec_identity_ms": round(t_s * 1e3, 3), "speedup": round(t_s / t_d, 2),
           "diff_vs_solvec_identity": diff, "max_abs_A_inv_minus_I": res}
    log(f"bench: invc {row}")
    return row


def next_rows_leg(dev, torch, args):
    """LOBPCG's Cholesky consumers and QR's back-substitution (the rows beside the
    solve path): device orthonormalize and apply_constraints on a tall block
    (n = 2^20, k = 64, f64), and the pivot-vector triangular solve through the
    plain host form (n = 2048, 16 RHS).  CPU baselines: the oracle's cholesky and
    solve_triangular (C, 1 thread) with the Gram products in numpy on 1 thread,
    on a 65536-row sample scaled to 2^20 rows."""
    import numpy as np
    import threadpoolctl
    import linfa_linalg_hip as llh
    from oracle import oracle as orc
    out = {}
    n, k, ky = 1 << 20, 64, 64
    g = torch.Generator(device="cuda").manual_seed(args.seed + 11)
    v0 = torch.randn(n, k, dtype=torch.float64, device="cuda", generator=g)
    y = torch.randn(n, ky, dtype=torch.float64, device="cuda", generator=g)
    l_yy = (y.T @ y).contiguous()
    dev.potrf(l_yy, zero_upper=True)
    for name, fn, flops, nbytes in (
            ("orthonormalize", lambda v: dev.orthonormalize(v), 4.0 * n * k * k + k ** 3 / 3.0, 24.0 * n * k),
            ("apply_constraints", lambda v: dev.apply_constraints(v, l_yy, y), 4.0 * n * k * ky,
             8.0 * n * (2 * ky + 3 * k))):
        v = v0.clone()
        fn(v)
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(3, args.steps)):
            v.copy_(v0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn(v)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        out[name] = {"workload": f"n={n} k={k}" + (f" ky={ky}" if name == "apply_constraints" else "") + " f64",
                     "ms": round(t * 1e3, 3), "tflops": round(flops / t / 1e12, 3),
                     "gbs_algorithmic": round(nbytes / t / 1e9, 1),
                     "roofline_frac": round(max(flops / (PEAK_TFLOPS["f64"] * 1e12), nbytes / (PEAK_HBM_GBS * 1e9))
                                            / t, 4)}
        log(f"bench: {name} {out[name]}")
    # CPU baseline on a sample of the rows
    ns = 65536
    rng = np.random.default_rng(args.seed + 12)
    vs = rng.standard_normal((ns, k))
    with threadpoolctl.threadpool_limits(1):
        t0 = time.perf_counter()
        gm = vs.T @ vs
        lm = orc.cholesky(gm)
        orc.solve_triangular(lm, np.array(vs.T), 1)
        t_orth = time.perf_counter() - t0
        ys = rng.standard_normal((ns, ky))
        lys = orc.cholesky(ys.T @ ys)
        t0 = time.perf_counter()
        gy = ys.T @ vs
        u = orc.solve_triangular(lys, gy, 1)
        vs2 = vs - ys @ u
        t_ac = time.perf_counter() - t0
    scale = n / ns
    out["orthonormalize"]["cpu_s_scaled"] = round(t_orth * scale, 3)
    out["apply_constraints"]["cpu_s_scaled"] = round(t_ac * scale, 3)
    out["cpu_sample"] = (f"oracle cholesky + solve_triangular (C, 1 thread), numpy Gram products on 1 thread, "
                         f"{ns} rows, scaled x{scale:.0f}")
    del v0, y, l_yy, vs2
    # QR back-substitution: R x = b with pivots |diag_i|, plain host form
    nq, mq = 2048, 16
    r = np.triu(rng.uniform(-1, 1, (nq, nq)), 1) / np.sqrt(nq)
    d = rng.uniform(0.5, 2.0, nq) * rng.choice([-1.0, 1.0], nq)
    bq = rng.standard_normal((nq, mq))
    llh.solve_triangular_system(r, np.array(bq), llh.UPLO.Upper, d, True)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        llh.solve_triangular_system(r, np.array(bq), llh.UPLO.Upper, d, True)
        ts.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    orc.solve_triangular_system(r, np.array(bq), 0, d, True)
    t_cpu = time.perf_counter() - t0
    out["invc"] = invc_leg(dev, torch, args)
    out["qr_solve"] = {"workload": f"R x = b, n={nq}, {mq} RHS, pivots |diag_i| (llh_solve_triangular_system_d, "
                                   "host arrays through the boundary)",
                       "ms": round(float(np.median(ts)) * 1e3, 3), "cpu_oracle_ms": round(t_cpu * 1e3, 3)}
    log(f"bench: qr_solve {out['qr_solve']}")
    return out


def serving_leg(dev, torch, args):
    """Poisson arrivals of solve requests at several offered rates: n log-uniform
    in [16, 4096], A = B B^T / n + I with B uniform in [-1, 1] (pre-staged in HBM,
    so the request stream is not PCIe-bound), 1-32 right-hand sides, 30% of the
    requests reuse an earlier matrix id, f64 and FP32+IR half and half.  Latency
    is submit -> solution in the caller's host buffer."""
    import numpy as np
    rng = np.random.default_rng(args.seed + 7)
    npool = 64
    pool_n = np.floor(np.exp(rng.uniform(np.log(16), np.log(4096), npool))).astype(int)
    g = torch.Generator(device="cuda").manual_seed(args.seed + 8)
    pool = []
    for n in pool_n:
        bm = torch.rand(int(n), int(n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        a = bm @ bm.T / int(n)
        a.diagonal().add_(1.0)
        pool.append(a)
    torch.cuda.synchronize()
    plan = ((1000, 2.0), (10000, 0.6), (50000, 0.2))
    out = {"workload": "Poisson solve requests, n log-uniform [16,4096] (A pre-staged in HBM), 1-32 RHS, "
                       "30% id reuse, 50/50 f64 and FP32+IR; arrival streams of "
                       + " / ".join(f"{d:g} s at {r // 1000}k/s" for r, d in plan)
                       + " (not the configuration's 60 s: above saturation latency grows with the stream's "
                         "length, so the 10k and 50k figures are for these lengths)", "rates": []}
    for rate, dur in plan:
        log(f"bench: serving at {rate}/s for {dur} s")
        srv = dev.Server()
        cnt = int(rate * dur)
        gaps = rng.exponential(1.0 / rate, cnt)
        reqs = []
        seen = []
        next_id = 1
        for i in range(cnt):
            if seen and rng.random() < 0.3:
                mid, pi = seen[int(rng.integers(0, len(seen)))]
            else:
                mid, pi = next_id, int(rng.integers(0, npool))
                next_id += 1
                seen.append((mid, pi))
            n, m = int(pool_n[pi]), int(rng.integers(1, 33))
            prec = "f32" if rng.random() < 0.5 else "f64"
            reqs.append((mid, pi, n, m, prec, rng.standard_normal((n, m)), np.zeros((n, m))))
        tickets = []
        t0 = time.perf_counter()
        tnext = t0
        for (mid, pi, n, m, prec, b, x), gap in zip(reqs, gaps):
            tnext += gap
            while time.perf_counter() < tnext:
                pass
            tickets.append(srv.submit(mid, pool[pi], b, x, precision=prec))
        t_sub = time.perf_counter() - t0
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
        srv.close()
        lat = np.array([r["latency_s"] for r in res])
        okr = [r for r in res if r["status"] == 0]
        bad = [(reqs[i][2], reqs[i][4], r["status"], r["failed_row"], r["cache_hit"], r["batched"])
               for i, r in enumerate(res) if r["status"] != 0]
        span = max(r["done_s"] for r in res) - min(r["submit_s"] for r in res)
        out["rates"].append({"offered_per_s": rate, "submitted": cnt, "achieved_offered_per_s": round(cnt / t_sub, 1),
                             "completed_ok": len(okr), "solves_per_s": round(len(okr) / span, 1),
                             "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                             "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                             "cache_hits": st["cache_hits"], "batched": st["batched"], "steps": st["steps"],
                             "peak_gb": round(st["bytes_peak"] / 1e9, 2),
                             "failed": [{"n": int(n_), "precision": p_, "status": int(s_), "failed_row": int(f_),
                                         "cache_hit": bool(h_), "batched": bool(b_)}
                                        for n_, p_, s_, f_, h_, b_ in bad[:10]]})
        log(f"bench: serving {out['rates'][-1]}")
        del reqs
    # the plain path: the same kind of requests one at a time (llh_dev_solvec_d on a fresh copy)
    lat = []
    t0 = time.perf_counter()
    for i in range(100):
        pi = int(rng.integers(0, npool))
        n, m = int(pool_n[pi]), int(rng.integers(1, 33))
        b = rng.standard_normal((n, m))
        ta = pool[pi].clone()
        t_a = time.perf_counter()
        tb = torch.from_numpy(b).cuda()
        dev.solvec(ta, tb)
        tb.cpu()
        lat.append(time.perf_counter() - t_a)
    out["plain_sequential"] = {"solves_per_s": round(100 / sum(lat), 1),
                               "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                               "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trailing_syrk_f64_n32768_traffic.json")


def _is_trailing_syrk(name):
    # the square lower-triangle instantiation (TRI, SQ) of the NT MFMA kernel
    return name.startswith("void llh::(anonymous namespace)::gemm_nt_kernel<double,") and ", true, true, " in name


def rocprof_summary(trace_dir, fetch_dir, write_dir, n, out_md, out_json, command):
    """Summarises one `rocprofv3 --kernel-trace --stats` run and two `--pmc`
    runs (FETCH_SIZE, WRITE_SIZE) of this benchmark: per-kernel average
    durations, and HBM bytes per trailing-SYRK launch scaled by what the same
    counters read for the n x n f64 clone of A (a plain copy of known size)."""
    import csv
    import glob

    def one(pattern):
        hits = glob.glob(os.path.join(pattern))
        if not hits:
            raise SystemExit(f"no file matches {pattern}")
        return hits[0]

    stats = list(csv.DictReader(open(one(os.path.join(trace_dir, "*kernel_stats.csv")))))
    trace = [r for r in csv.DictReader(open(one(os.path.join(trace_dir, "*kernel_trace.csv"))))
             if _is_trailing_syrk(r["Kernel_Name"])]
    durs = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in trace]
    syrk_calls, syrk_ns = len(durs), float(sum(durs))
    counters = {}
    for ctr, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = [r for r in csv.DictReader(open(one(os.path.join(d, "*counter_collection.csv"))))
                if r["Counter_Name"] == ctr]
        copy = max(float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"])
        known_kib = n * n * 8 / 1024.0
        factor = known_kib / copy
        vals = [float(r["Counter_Value"]) for r in rows if _is_trailing_syrk(r["Kernel_Name"])]
        counters[ctr] = {"copy_kib": copy, "known_kib": known_kib, "factor": factor, "launches": len(vals),
                         "avg_kib": sum(vals) / len(vals), "bytes_per_launch": sum(vals) / len(vals) * 1024 * factor}
    traffic = counters["FETCH_SIZE"]["bytes_per_launch"] + counters["WRITE_SIZE"]["bytes_per_launch"]
    # algorithmic bytes of the trailing SYRKs of one factorization (outer panel
    # width outer_width(n); with lookahead each SYRK leaves out the next panel's columns)
    alg = []
    k0 = 0
    outer = outer_width(n)
    while k0 < n:
        kb = min(outer, n - k0)
        m3 = n - k0 - kb - min(outer, max(n - k0 - kb, 0))
        if m3 > 0:
            alg.append(8.0 * (m3 * (m3 + 1) + m3 * kb))
        k0 += kb
    alg_avg = sum(alg) / len(alg)
    res = {"n": n, "kernel_class": "trailing SYRK (gemm_nt_kernel<double,...,TRI,SQ>)",
           "avg_launch_us": syrk_ns / syrk_calls / 1e3, "launches": syrk_calls,
           "fetch": counters["FETCH_SIZE"], "write": counters["WRITE_SIZE"],
           "traffic_bytes_per_launch": traffic, "algorithmic_bytes_per_launch": alg_avg,
           "traffic_over_algorithmic": traffic / alg_avg, "command": command}
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    lines = [f"# rocprofv3 summary: f64 SPD solve, n={n}, 16 RHS (bench.py)", "",
             "Commands (run from the repository root on one MI355X):", "", "```", command, "```", "",
             "## Kernel durations (`--kernel-trace --stats`)", "",
             "| kernel | calls | avg us | total ms | % |", "|---|---:|---:|---:|---:|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:14]:
        name = r["Name"]
        short = name.replace("void ", "").replace("(anonymous namespace)::", "").split("(")[0]
        if len(short) > 90:
            short = short[:87] + "..."
        lines.append(f"| `{short}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.1f} | {float(r['Percentage']):.1f} |")
    lines += ["", f"Trailing SYRK (the TRI+SQ instantiations of `gemm_nt_kernel`, launched only by "
                  f"`launch_syrk_lower`: 256x128 tiles while the trailing matrix is at least 4096^2, 128x128 "
                  f"below): {syrk_calls} launches, average {syrk_ns / syrk_calls / 1e3:.1f} us per launch.", "",
              "## HBM traffic per trailing-SYRK launch (`--pmc FETCH_SIZE`, then `--pmc WRITE_SIZE`)", "",
              "| counter | copy of A reads (KiB) | known (KiB) | factor | avg per launch (KiB) | bytes per launch |",
              "|---|---:|---:|---:|---:|---:|"]
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        c = counters[ctr]
        lines.append(f"| {ctr} | {c['copy_kib']:.0f} | {c['known_kib']:.0f} | {c['factor']:.3f} | "
                     f"{c['avg_kib']:.0f} | {c['bytes_per_launch'] / 1e9:.3f} GB |")
    lines += ["", f"Traffic {traffic / 1e9:.3f} GB per launch against {alg_avg / 1e9:.3f} GB algorithmic "
                  f"(read and write the lower triangle of C once, read the panel once): "
                  f"{traffic / alg_avg:.2f}x.", ""]
    with open(out_md, "w") as f:
        f.write("\n".join(lines))
    print(json.dumps(res))


def pmc_summary(pmc_dir, out_md, command, title):
    """Summarises one `rocprofv3 --pmc <SQ/GRBM counters>` run of this benchmark
    per kernel: each counter summed over the kernel's dispatches, the MFMA work
    the hardware counted (SQ_INSTS_VALU_MFMA_MOPS_* x 512 flops), and the MFMA
    busy fraction SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / 8 x 1024 SIMDs).
    GRBM_GUI_ACTIVE is summed over the 8 XCDs (it reads 8x the kernel's cycles
    at 2.4 GHz), so GRBM_GUI_ACTIVE / 8 is the kernel's duration in cycles."""
    import csv
    import glob
    hits = glob.glob(os.path.join(pmc_dir, "**", "*counter_collection.csv"), recursive=True)
    if not hits:
        raise SystemExit(f"no counter_collection.csv under {pmc_dir}")
    per = {}
    disp = {}
    for r in csv.DictReader(open(hits[0])):
        name = r["Kernel_Name"].replace("void ", "").replace("(anonymous namespace)::", "").split("(")[0]
        per.setdefault(name, {}).setdefault(r["Counter_Name"], 0.0)
        per[name][r["Counter_Name"]] += float(r["Counter_Value"])
        disp.setdefault(name, set()).add(r.get("Dispatch_Id", r.get("Correlation_Id", "")))
    ctrs = sorted({c for v in per.values() for c in v})
    rows = sorted(per.items(), key=lambda kv: -kv[1].get("GRBM_GUI_ACTIVE", 0.0))
    lines = [f"# rocprofv3 counters: {title}", "", "Command (run from the repository root on one MI355X):", "",
             "```", command, "```", "",
             "Derived: `mfma_tflop` = (SQ_INSTS_VALU_MFMA_MOPS_F64 + _F32) x 512 flops summed over the "
             "kernel's dispatches; `mfma_busy` = SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / 8 XCDs x 256 CUs x 4 SIMDs).",
             "", "| kernel | dispatches | " + " | ".join(ctrs) + " | mfma_tflop | mfma_busy |",
             "|---|---:|" + "---:|" * (len(ctrs) + 2)]
    for name, v in rows[:16]:
        short = name if len(name) <= 80 else name[:77] + "..."
        mops = v.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0) + v.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0.0)
        busy = v.get("SQ_VALU_MFMA_BUSY_CYCLES")
        gui = v.get("GRBM_GUI_ACTIVE")
        util = f"{busy / (gui / 8 * 256 * 4):.3f}" if busy is not None and gui else "-"
        lines.append(f"| `{short}` | {len(disp[name])} | " + " | ".join(f"{v.get(c, 0.0):.4g}" for c in ctrs)
                     + f" | {mops * 512 / 1e12:.3f} | {util} |")
    with open(out_md, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def trace_summary(trace_dir, out_md, command, title):
    """Summarises one `rocprofv3 --kernel-trace --stats` run of this benchmark:
    per-kernel calls, average and total duration (top 24 by total time)."""
    import csv
    import glob
    hits = glob.glob(os.path.join(trace_dir, "**", "*kernel_stats.csv"), recursive=True)
    if not hits:
        raise SystemExit(f"no kernel_stats.csv under {trace_dir}")
    stats = list(csv.DictReader(open(hits[0])))
    lines = [f"# rocprofv3 summary: {title}", "", "Command (run from the repository root on one MI355X):", "",
             "```", command, "```", "", "| kernel | calls | avg us | total ms | % |", "|---|---:|---:|---:|---:|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:24]:
        short = r["Name"].replace("void ", "").replace("(anonymous namespace)::", "").split("(")[0]
        if len(short) > 95:
            short = short[:92] + "..."
        lines.append(f"| `{short}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.1f} | {float(r['Percentage']):.1f} |")
    with open(out_md, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


BATCHED_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "batched_uneven_20000_traffic.json")


def batched_summary(trace_dir, fetch_dir, write_dir, once_json, command):
    """Summarises three rocprofv3 runs of `bench.py --batch-once` (one graph
    launch of the batched workload): per-kernel durations (`--kernel-trace`),
    and HBM bytes per kernel (`--pmc FETCH_SIZE`, then `--pmc WRITE_SIZE`),
    scaled by what the same counters read for the clone of the packed A (a copy
    of known size).  The launch's traffic is set against its algorithmic bytes
    (read and write each lower triangle once, read and write each B once)."""
    import csv
    import glob

    def one(d, pat):
        hits = glob.glob(os.path.join(d, "**", pat), recursive=True)
        if not hits:
            raise SystemExit(f"no {pat} under {d}")
        return hits[0]

    def short(name):
        s = name.replace("void ", "").replace("(anonymous namespace)::", "").split("(")[0]
        return s if len(s) <= 70 else s[:67] + "..."

    with open(once_json) as f:
        once = json.loads([ln for ln in f.read().splitlines() if ln.startswith("{")][-1])
    known_kib = once["a_copy_bytes"] / 1024.0
    durs = {}
    for r in csv.DictReader(open(one(trace_dir, "*kernel_trace.csv"))):
        if short(r["Kernel_Name"]).startswith("llh::"):
            durs.setdefault(short(r["Kernel_Name"]), []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    per = {}
    factors = {}
    for ctr, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = [r for r in csv.DictReader(open(one(d, "*counter_collection.csv"))) if r["Counter_Name"] == ctr]
        copy = max(float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"])
        factors[ctr] = {"copy_kib": copy, "known_kib": known_kib, "factor": known_kib / copy}
        for r in rows:
            if short(r["Kernel_Name"]).startswith("llh::"):
                per.setdefault(short(r["Kernel_Name"]), {}).setdefault(ctr, 0.0)
                per[short(r["Kernel_Name"])][ctr] += float(r["Counter_Value"]) * 1024 * factors[ctr]["factor"]
    total = sum(v.get("FETCH_SIZE", 0.0) + v.get("WRITE_SIZE", 0.0) for v in per.values())
    dev_ns = sum(sum(v) for v in durs.values())
    res = {"count": once["count"], "config": once["config"], "algorithmic_bytes": once["algorithmic_bytes"],
           "traffic_bytes": total, "traffic_over_algorithmic": total / once["algorithmic_bytes"],
           "kernel_time_ms": dev_ns / 1e6, "factors": factors,
           "kernels": {k: {"dispatches": len(durs.get(k, [])), "ms": sum(durs.get(k, [])) / 1e6,
                           "fetch_bytes": v.get("FETCH_SIZE", 0.0), "write_bytes": v.get("WRITE_SIZE", 0.0)}
                       for k, v in per.items()},
           "command": command}
    with open(BATCHED_TRAFFIC_FILE, "w") as f:
        json.dump(res, f, indent=1)
    lines = [f"# rocprofv3 summary: {once['count']} uneven SPD systems, one hipGraph launch (bench.py --batch-once)",
             "", "Commands (run from the repository root on one MI355X):", "", "```", command, "```", "",
             f"Configuration {once['config']}; the launch: {once['ok']} of {once['count']} systems solved. "
             f"Counter scale factors from the clone of the packed A ({once['a_copy_bytes'] / 2**20:.1f} MiB): "
             f"FETCH_SIZE x{factors['FETCH_SIZE']['factor']:.3f}, WRITE_SIZE x{factors['WRITE_SIZE']['factor']:.3f}.",
             "", "| kernel | dispatches | ms | HBM read (MB) | HBM write (MB) | GB/s |", "|---|---:|---:|---:|---:|---:|"]
    for k, v in sorted(res["kernels"].items(), key=lambda kv: -kv[1]["ms"]):
        gbs = (v["fetch_bytes"] + v["write_bytes"]) / (v["ms"] / 1e3) / 1e9 if v["ms"] > 0 else 0.0
        lines.append(f"| `{k}` | {v['dispatches']} | {v['ms']:.3f} | {v['fetch_bytes'] / 1e6:.1f} | "
                     f"{v['write_bytes'] / 1e6:.1f} | {gbs:.0f} |")
    lines += ["", f"Kernel time {dev_ns / 1e6:.3f} ms in all. HBM traffic {total / 1e9:.3f} GB against "
                  f"{once['algorithmic_bytes'] / 1e9:.3f} GB algorithmic: {total / once['algorithmic_bytes']:.2f}x.", ""]
    with open(os.path.join(ROOT, "profiles", "batched_uneven_20000.md"), "w") as f:
        f.write("\n".join(lines))
    print("\n".join(lines))


def load_batched_traffic(cnt, config):
    """The batched launch's HBM traffic from the committed summary, when it was
    measured on this workload and configuration."""
    try:
        with open(BATCHED_TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    # the small bucket's resident workgroups change its speed, not the bytes it moves
    keys = ("small_limit", "ir_iterations", "medium_mode")
    if t.get("count") != cnt or any(t.get("config", {}).get(k) != config.get(k) for k in keys):
        return None
    return {"bytes": round(t["traffic_bytes"]), "over_algorithmic": round(t["traffic_over_algorithmic"], 3),
            "config_profiled": t["config"],
            "source": "profiles/batched_uneven_20000.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, one launch)"}


def load_traffic(n, m):
    """Per-launch HBM bytes of the trailing SYRK from the committed PMC summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t if t.get("n") == n and m == 16 else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--lam", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=20241202)
    ap.add_argument("--cpu-sizes", default="512,1024,2048,4096")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--batch", type=int, default=20000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-next", action="store_true")
    ap.add_argument("--rocprof-summary", nargs=3, metavar=("TRACE_DIR", "FETCH_DIR", "WRITE_DIR"),
                    help="summarise rocprofv3 runs of this benchmark into profiles/ and exit")
    ap.add_argument("--rocprof-command", default="")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("PMC_DIR", "OUT_MD", "TITLE"),
                    help="summarise a rocprofv3 --pmc run of SQ/GRBM counters into OUT_MD and exit")
    ap.add_argument("--lambdas", default="1,1e-2,1e-4,1e-6",
                    help="lambdas of the mixed-precision sweep (the first is also --lam's leg)")
    ap.add_argument("--no-f16", action="store_true")
    ap.add_argument("--batch-once", default="", metavar="SMALL_LIMIT,SMALL_WAVES[,MEDIUM_MODE]",
                    help="run one f64 graph launch of the batched workload with this configuration, print "
                         "its JSON and exit (for counter runs)")
    ap.add_argument("--trace-summary", nargs=3, metavar=("TRACE_DIR", "OUT_MD", "TITLE"),
                    help="summarise a rocprofv3 --kernel-trace --stats run into OUT_MD and exit")
    ap.add_argument("--batched-summary", nargs=4, metavar=("TRACE_DIR", "FETCH_DIR", "WRITE_DIR", "ONCE_JSON"),
                    help="summarise rocprofv3 runs of --batch-once into profiles/ and exit")
    args = ap.parse_args()

    if args.trace_summary:
        trace_summary(*args.trace_summary[:2], args.rocprof_command, args.trace_summary[2])
        return

    if args.batched_summary:
        batched_summary(*args.batched_summary, args.rocprof_command)
        return

    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.rocprof_command, args.pmc_summary[2])
        return

    if args.rocprof_summary:
        rocprof_summary(*args.rocprof_summary, args.n,
                        os.path.join(ROOT, "profiles", f"trailing_syrk_f64_n{args.n}.md"),
                        os.path.join(ROOT, "profiles", f"trailing_syrk_f64_n{args.n}_traffic.json"),
                        args.rocprof_command)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU layout is a design only; nothing here runs on a second GPU).", file=sys.stderr)
        sys.exit(2)

    import torch
    from linfa_linalg_hip import device as dev

    n, m = args.n, args.nrhs
    torch.cuda.set_device(0)
    if args.batch_once:
        print(json.dumps(batched_leg(dev, torch, args)))
        return
    log(f"bench: generating A = B B^T/n + {args.lam} I, n={n}, f64 (seeded counter generator)")
    a0 = dev.gen_spd(dev.GEN_NORMAL, args.seed, n, args.lam)
    rhs0 = dev.gen_matrix(dev.GEN_NORMAL, args.seed, 1, n, m)
    copies = [(a0.clone(), rhs0.clone()) for _ in range(args.warmup + args.steps)]
    torch.cuda.synchronize()

    for i in range(args.warmup):
        log(f"bench: warmup step {i}")
        dev.solvec(*copies[i])
    torch.cuda.synchronize()

    # The timed steps record live events for the trailing SYRK only (the
    # roofline kernel); an extra profiled step gives the other classes.
    dev.set_profiling(True, classes=["trailing_syrk"])
    torch.cuda.synchronize()
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    for i in range(args.steps):
        dev.solvec(*copies[args.warmup + i])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = dev.kernel_stats()
    dev.set_profiling(False)
    ca, cb = a0.clone(), rhs0.clone()
    dev.set_profiling(True)
    dev.solvec(ca, cb)
    torch.cuda.synchronize()
    stats_all = dev.kernel_stats()
    dev.set_profiling(False)
    del ca, cb
    ms = (t1 - t0) * 1e3 / args.steps
    gflops = solve_flops(n, m) / (ms / 1e3) / 1e9
    log(f"bench: {ms:.2f} ms/step, {gflops:.1f} GFLOP/s")

    kernels = {}
    for name, (cnt, tot_ms, fl, by) in stats_all.items():
        if cnt == 0:
            continue
        kernels[name] = {"launches_per_step": cnt, "ms_per_step": round(tot_ms, 3),
                         "avg_launch_us": round(tot_ms / cnt * 1e3, 2),
                         "tflops": round(fl / (tot_ms / 1e3) / 1e12, 2) if tot_ms > 0 else None,
                         "gbs": round(by / (tot_ms / 1e3) / 1e9, 1) if tot_ms > 0 else None}
    log("bench: kernels " + json.dumps(kernels))
    cnt, tot_ms, fl, by = stats["trailing_syrk"]
    achieved = fl / (tot_ms / 1e3) / 1e12
    tr = load_traffic(n, m)
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_TFLOPS["f64"], 4),
                "traffic": round(tr["traffic_bytes_per_launch"]) if tr else None,
                "traffic_note": ("HBM bytes per launch (rocprofv3 FETCH_SIZE x%.2f + WRITE_SIZE x%.2f, factors from "
                                 "a copy of A), %.2fx the algorithmic bytes; profiles/trailing_syrk_f64_n%d.md"
                                 % (tr["fetch"]["factor"], tr["write"]["factor"], tr["traffic_over_algorithmic"], n))
                if tr else None,
                "kernel": "gemm_nt_kernel<double,256,128,4,2,TRI,SQ> (trailing SYRK, rank %d; 128x128 tiles once the "
                          "trailing matrix is below 4096^2)" % outer_width(n),
                "avg_launch_us": round(tot_ms / cnt * 1e3, 2), "launches_per_step": cnt // args.steps,
                "algorithmic_tflop_per_step": round(fl / args.steps / 1e12, 4)}

    check = {}
    if not args.no_check:
        a_fac, x = copies[-1]
        l = torch.tril(a_fac)
        res = rhs0 - a0 @ x
        check["residual"] = float(res.norm() / (a0.norm() * x.norm()))
        al = torch.tril(a0)
        check["backward_error"] = float(torch.tril(l @ l.T - a0).norm() / al.norm())
        ref = oracle_fixture(n, m, args.lam, args.seed)
        if ref is not None:
            check["solution_error_vs_oracle"] = sol_err(x, ref)
            check["oracle"] = "multi-threaded oracle (reference operation order), tests/golden/full_size"
        log(f"bench: residual {check['residual']:.3e} backward error {check['backward_error']:.3e} "
            f"vs oracle {check.get('solution_error_vs_oracle')}")
        del l, res, al

    # Mixed precision on the same A (A is not modified): FP32-MFMA and
    # FP16-MFMA (f32 accumulation, diagonally scaled) trailing updates with f64
    # iterative refinement, then on A + (lambda - 1) I for smaller lambda
    # (harder conditioning).  The timed runs are unprofiled and compared with an
    # unprofiled f64 solve; a separate profiled run gives the kernel breakdown.
    mixed = None
    mixed16 = None
    sweep = []
    f64_clean_ms = None
    copies = copies[-1:]  # the last step's factor and solution stay for the comparisons below
    torch.cuda.empty_cache()
    if not args.no_mixed:
        # s0 = B B^T / n: s0 + lambda I is bit for bit the generator's A for
        # lambda, the input the oracle's full-size fixtures were computed on
        s0 = dev.gen_spd(dev.GEN_NORMAL, args.seed, n, 0.0)
        x32 = torch.zeros(n, m, dtype=torch.float64, device="cuda")
        x64 = copies[-1][1]
        ca, cb = a0.clone(), rhs0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev.solvec(ca, cb)
        torch.cuda.synchronize()
        f64_clean_ms = (time.perf_counter() - t0) * 1e3
        del ca, cb
        g64_clean = solve_flops(n, m) / (f64_clean_ms / 1e3) / 1e9
        log(f"bench: unprofiled f64 solve {f64_clean_ms:.2f} ms ({g64_clean:.1f} GFLOP/s)")

        def mixed_run(prec, lam):
            info = dev.solve_mixed(a0, rhs0, x32, prec=prec, fallback=True)
            ref = oracle_fixture(n, m, lam, args.seed)
            err_orc = sol_err(x32, ref) if ref is not None else None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                info = dev.solve_mixed(a0, rhs0, x32, prec=prec, fallback=True)
            torch.cuda.synchronize()
            msx = (time.perf_counter() - t0) * 1e3 / args.steps
            gx = solve_flops(n, m) / (msx / 1e3) / 1e9
            res = rhs0 - a0 @ x32
            bres = float(res.norm() / (torch.linalg.matrix_norm(a0, ord=1) * x32.norm()))
            del res
            row = {"precision": prec, "lambda": lam, "cond_estimate": round((4.0 + lam) / lam, 1),
                   "ms_per_step": round(msx, 3), "value": round(gx, 2), "unit": "GFLOP/s",
                   "speedup_vs_f64": round(gx / g64_clean, 3), "refinement": info, "backward_residual": bres,
                   "solution_error_vs_oracle": err_orc}
            log(f"bench: {prec}+IR lambda={lam:g} {msx:.2f} ms/step, {gx:.1f} GFLOP/s, x{gx / g64_clean:.2f} vs "
                f"f64, {info}, backward residual {bres:.2e}")
            return row

        def breakdown(row, prec, syrk_peak):
            dev.set_profiling(True)
            dev.solve_mixed(a0, rhs0, x32, prec=prec, fallback=True)
            torch.cuda.synchronize()
            st = dev.kernel_stats()
            dev.set_profiling(False)
            cnt_s, tot_s, fl_s, _ = st["trailing_syrk"]
            row["solution_diff_vs_f64_gpu"] = float((x32 - x64).norm() / x64.norm())
            row["syrk_tflops"] = round(fl_s / (tot_s / 1e3) / 1e12, 2) if tot_s > 0 else None
            row["syrk_roofline_frac"] = (round(fl_s / (tot_s / 1e3) / 1e12 / PEAK_TFLOPS[syrk_peak], 4)
                                         if tot_s > 0 else None)
            row["kernels_ms_profiled_run"] = {k: round(v[1], 3) for k, v in st.items() if v[0]}

        lams = [float(v) for v in args.lambdas.split(",") if v]
        for lam in [args.lam] + [v for v in lams if v != args.lam]:
            a0.copy_(s0)
            a0.diagonal().add_(lam)  # exactly the generator's A for this lambda
            row = mixed_run("f32", lam)
            if lam == args.lam:
                breakdown(row, "f32", "f32")
                mixed = row
            sweep.append(row)
            if lam >= 1e-2 and not args.no_f16:  # the f16 factor is for cond up to ~1e3
                row16 = mixed_run("f16", lam)
                if lam == args.lam:
                    breakdown(row16, "f16", "f16")
                    mixed16 = row16
                sweep.append(row16)
        a0.copy_(s0)
        a0.diagonal().add_(args.lam)
        del x32, s0

    # C2: n = 16384, 64 right-hand sides, lambda = 0.1, f64
    c2 = None
    if not args.no_c2:
        n2, m2, lam2 = 16384, 64, 0.1
        del copies
        torch.cuda.empty_cache()
        a2 = dev.gen_spd(dev.GEN_NORMAL, args.seed, n2, lam2)
        r2 = dev.gen_matrix(dev.GEN_NORMAL, args.seed, 1, n2, m2)
        cps = [(a2.clone(), r2.clone()) for _ in range(args.warmup + args.steps)]
        torch.cuda.synchronize()
        for i in range(args.warmup):
            dev.solvec(*cps[i])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            dev.solvec(*cps[args.warmup + i])
        torch.cuda.synchronize()
        ms2 = (time.perf_counter() - t0) * 1e3 / args.steps
        xf = cps[-1][1]
        res = r2 - a2 @ xf
        ref2 = oracle_fixture(n2, m2, lam2, args.seed)
        x2 = torch.zeros_like(r2)
        info2 = dev.solve_mixed(a2, r2, x2, prec="f32", fallback=True)
        c2 = {"workload": f"spd_solve n={n2} nrhs={m2} lambda={lam2} f64 (llh_dev_solvec_d)",
              "ms_per_step": round(ms2, 3), "value": round(solve_flops(n2, m2) / (ms2 / 1e3) / 1e9, 2),
              "unit": "GFLOP/s", "roofline_frac_f64": round(solve_flops(n2, m2) / (ms2 / 1e3) / 1e12 /
                                                           PEAK_TFLOPS["f64"], 4),
              "residual": float(res.norm() / (a2.norm() * xf.norm())),
              "solution_error_vs_oracle": sol_err(xf, ref2) if ref2 is not None else None,
              "fp32_ir_solution_error_vs_oracle": sol_err(x2, ref2) if ref2 is not None else None,
              "fp32_ir_refinement": info2}
        del x2
        log(f"bench: C2 {c2}")
        del cps, a2, r2, res
        torch.cuda.empty_cache()

    # C1: n = 512, 1 right-hand side: latency of one solvec, plain (host arrays
    # through the boundary) and device-resident
    c1 = None
    if not args.no_c1:
        import numpy as np
        import linfa_linalg_hip as llh
        rng1 = np.random.default_rng(args.seed + 1)
        b1 = rng1.uniform(-1, 1, (512, 512))
        a1 = b1 @ b1.T / 512 + np.eye(512)
        v1 = rng1.standard_normal((512, 1))
        lat_plain, lat_dev = [], []
        ta = torch.from_numpy(a1).cuda()
        tb = torch.from_numpy(v1).cuda()
        for rep in range(25):
            w1 = np.array(a1)
            t0 = time.perf_counter()
            llh.solvec(w1, v1)
            lat_plain.append(time.perf_counter() - t0)
            da, db = ta.clone(), tb.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dev.solvec(da, db)
            torch.cuda.synchronize()
            lat_dev.append(time.perf_counter() - t0)
        c1 = {"workload": "spd_solve n=512 nrhs=1 f64", "plain_solvec_p50_us": round(
            float(np.percentile(lat_plain[5:], 50)) * 1e6, 1), "dev_solvec_p50_us": round(
            float(np.percentile(lat_dev[5:], 50)) * 1e6, 1)}
        log(f"bench: C1 {c1}")

    batched = None
    if not args.no_batch:
        torch.cuda.empty_cache()
        batched = batched_leg(dev, torch, args)
        log(f"bench: batched {batched['solves_per_s']} solves/s ({batched['ms']} ms), "
            f"plain {batched['plain_per_matrix']}")

    next_rows = None
    if not args.no_next:
        next_rows = next_rows_leg(dev, torch, args)

    serving = None
    if not args.no_serve:
        serving = serving_leg(dev, torch, args)

    cpu = None
    if not args.no_cpu_baseline:
        sizes = [int(v) for v in args.cpu_sizes.split(",")]
        cpu = cpu_baseline(sizes, m, args.seed, [(16384, 64), (n, m)])

    out = {
        "metric": "SPD solve GFLOP/s at n=32768 (f64 POTRF + forward/back solve, 16 RHS)",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based N(0,1) B, A = B B^T/n + lambda I, built on device; "
                "llh_dev_gen_spd_d, reproduced bit for bit by the oracle's generator)",
        "config": {"workload": f"spd_solve n={n} nrhs={m} lambda={args.lam} f64 (llh_dev_solvec_d)",
                   "n": n, "nrhs": m, "lambda": args.lam, "parallelism": "1 GPU"},
        "solve_roofline_frac_f64": round(gflops / 1e3 / PEAK_TFLOPS["f64"], 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "f64_unprofiled_ms": round(f64_clean_ms, 3) if f64_clean_ms else None,
        "fp32_ir": mixed,
        "fp16_ir": mixed16,
        "mixed_lambda_sweep": sweep,
        "c2": c2,
        "c1": c1,
        "batched": batched,
        "serving": serving,
        "lobpcg_qr": next_rows,
        "kernels": kernels,
        "kernels_note": "one extra profiled solve, every kernel class; the timed steps record the trailing SYRK only",
        "check": check,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
