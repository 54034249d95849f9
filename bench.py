#!/usr/bin/env python3
"""H2 matvec benchmark: N = 2,000,000 points uniform in [0,1]^3 (seed 0), Laplace kernel,
equispaced p = 6 interpolation (K = 216), Toeplitz M2L (11^3 spectra), leaf size <= 64,
32 right-hand sides of N(0,1) FP32 entries (seed 1).

One step = one multi-RHS product Y = A X with all 32 columns (near field + far field) in the
batched mixed-precision form (FP32 far field; near values stored in FP32, near products and the
final sum in FP64; hipGraph replay).
The same run times the plain FP64 form one right-hand side at a time, checks the batched
result against it, reports the error against the O(N^2) direct sum on 1,000 sampled targets,
measures the dominant kernel's per-launch time with HIP events for the roofline figure, serves
a seeded Poisson stream over a mix of resident maps, and times the C/OpenMP restatement of the
reference traversal on the host cores (the CPU baseline).

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "h2factory.jl_amd"))

PHASES = ["gather", "near", "p2m", "m2m", "m2l_fft_in", "m2l_apply", "m2l_fft_out", "l2l", "l2p", "scatter"]
PEAK = {"hbm": 8000.0, "fp32": 157.3, "fp64": 78.6, "bf16": 2500.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--nrhs", type=int, default=32)
    ap.add_argument("--p", type=int, default=6)
    ap.add_argument("--m2l", default="toeplitz")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "bf16x2", "bf16x3"],
                    help="batched form: FP32 (mixed), or dense couplings on split BF16 (2 or 3 planes)")
    ap.add_argument("--plain-rhs", type=int, default=2, help="columns timed in the plain FP64 form")
    ap.add_argument("--cpu-reps", type=int, default=2, help="products timed by the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--plan", default="device", choices=["device", "host"],
                    help="build the trees, pair lists and operators on the GPU (h2g_plan_build, "
                         "h2g_method_operators) or with the host builders")
    ap.add_argument("--tile", default=None, help="RHS block / coupling groups / near-field columns [/ persistent [/ dense pairs]], "
                                                  "e.g. 32/256/32 or 32/512/32/1: use this tile instead of autotuning "
                                                  "(0 = the library default)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (RHS block and tile shapes); use the library's defaults")
    ap.add_argument("--serve-requests", type=int, default=4000,
                    help="single-RHS requests of the serving measurement (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrival rate of the requests (1/s)")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0, help="HBM budget of the resident maps (GB)")
    ap.add_argument("--serve-seed", type=int, default=0, help="seed of the serving map mix")
    ap.add_argument("--serve-ws-per-map", type=int, default=1,
                    help="workspaces per resident map (batches of one map in flight at once; one keeps the "
                         "spectra of one batch per map, so more maps fit the budget)")
    ap.add_argument("--serve-helmholtz-m2l", default="toeplitz", choices=["toeplitz", "dense"],
                    help="couplings of the Helmholtz maps of the serving mix: equispaced Toeplitz (spectra, "
                         "as the Laplace maps) or Chebyshev dense (the Helmholtz workload's form, whose "
                         "operator stream costs a product the same from 1 to 16 columns)")
    ap.add_argument("--serve-rate-low", type=float, default=-1.0,
                    help="operating point below capacity (requests/s; negative: two points, 10%% and 25%% of "
                         "the throughput served at --serve-rate; 0: skip)")
    ap.add_argument("--serve-requests-low", type=int, default=1000)
    ap.add_argument("--serve-only", action="store_true", help="only the serving measurement (one JSON line)")
    ap.add_argument("--workload", default="laplace", choices=["laplace", "helmholtz"],
                    help="helmholtz: N = 4M points on the unit sphere, k = 10, Chebyshev p = 6, dense complex "
                         "couplings, 16 complex RHS (ComplexF32 and split-BF16 forms, forward and adjoint)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy five times (HBM counter calibration)")
    return ap.parse_args()


def algorithmic_work(arrays, nrhs, phase, level, mixed):
    """Algorithmic bytes (HBM-bound phases) or flops (compute phases) of one launch."""
    import numpy as np

    tr = arrays.test.tree
    K, p = arrays.K, arrays.p
    L3 = (2 * p - 1) ** 3
    cb = 4 if mixed else 8
    if phase == "near":
        n = tr.points.shape[0]
        return "hbm", arrays.n_near_values * 8.0 + 2 * n * nrhs * 8.0
    if phase == "m2l_apply":
        # level -1: the persistent coupling stage over every level
        lv = tr.level[arrays.pairs[:, 0]]
        npairs = int(np.sum(lv == level)) if level >= 0 else int(np.sum(lv >= 2))
        if arrays.m2l_kind in ("toeplitz", "symtoeplitz"):
            # one complex multiply-add per stored frequency: for real maps the half spectrum
            # without the conjugate half of its plane fx = 0, ((L^2 + 1) / 2 + (p - 1) L^2)
            L = 2 * p - 1
            F = (L * L + 1) // 2 + (p - 1) * L * L if arrays.dtype == "f64" else L3
            return "flop", npairs * F * 8.0 * nrhs
        return "flop", npairs * 2.0 * K * K * nrhs
    if phase in ("p2m", "l2p"):
        leaves = tr.leaves()
        n = int(np.sum((tr.end - tr.begin)[leaves[tr.level[leaves] == level]]))
        return "hbm", n * K * cb + n * nrhs * 8.0 + n * K * 0
    if phase in ("m2l_fft_in", "m2l_fft_out"):
        nb = int(tr.level_offsets[level + 1] - tr.level_offsets[level])
        return "hbm", nb * nrhs * (L3 * 2 * cb + K * cb)
    if phase in ("m2m", "l2l"):
        nb = int(tr.level_offsets[level + 1] - tr.level_offsets[level])
        return "flop", nb * 2.0 * K * K * nrhs
    n = tr.points.shape[0]
    return "hbm", 2.0 * n * nrhs * 8


def traffic_of(kernel, n, p, nrhs, groups):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE measurement of this workload and tile shape (scaled by the copy-kernel
    calibration), else None."""
    path = os.path.join(ROOT, "profiles", "m2l_direct_traffic_n2m_p6_r32.json")
    if (n, p, nrhs) != (2_000_000, 6, 32) or not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    for m in d.get("measurements", []):
        if m.get("kernel") == kernel and m.get("m2l_groups") == (groups or 512):
            return m["traffic_bytes_per_launch"]
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(arrays, x, reps=2):
    """The C/OpenMP restatement of the reference traversal (oracle/h2cpu.c: FP64, one column,
    every OpenMP thread of the host) on the same core data and the same column as the GPU:
    one warm-up product, then the mean of ``reps`` timed products.  Returns (seconds per product,
    threads, the product)."""
    sys.path.insert(0, ROOT)
    from oracle.cpu import CpuH2

    cpu = CpuH2(arrays)
    y = cpu.mul(x)
    t0 = time.perf_counter()
    for _ in range(reps):
        y = cpu.mul(x)
    dt = (time.perf_counter() - t0) / reps
    threads = cpu.threads
    cpu.close()
    return dt, threads, y


def helmholtz_main(args):
    """Helmholtz workload: e^{ikr}/(4 pi r), k = 10, N points uniform on the unit sphere (seed 0),
    octree on [-1, 1]^3 with leaves of at most 64 points, Chebyshev p = 6 (K = 216), dense complex
    couplings, R complex-normal right-hand sides (seed 1).  Times the ComplexF32 form and the split
    BF16 forms (hi + lo and hi + mid + lo), forward and adjoint, each against the plain ComplexF64 form one column
    at a time, and reports every form's error against the O(N^2) FP64 direct sum on 1,000 targets."""
    import numpy as np
    import torch

    import h2g
    from h2g import method as PM

    dev = torch.device("cuda:0")
    N = args.n if args.n != 2_000_000 else 4_000_000
    R = args.nrhs if args.nrhs != 32 else 16
    k = 10.0
    t0 = time.perf_counter()
    pts = np.random.default_rng(0).standard_normal((N, 3))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    kw = dict(p=6, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=k, leaf_size=64,
              origin=(-1.0, -1.0, -1.0), size=2.0)
    arrays = PM.build_map(pts, plan=args.plan, ops=args.plan, **kw)
    t_build = time.perf_counter() - t0
    log(f"map built ({args.plan} plan) in {t_build:.1f}s: {arrays.stats}")
    t0 = time.perf_counter()
    core = h2g.GpuH2Core(arrays, precisions=("plain", "mixed", "bf16x2", "bf16x3"))
    torch.cuda.synchronize()
    t_core = time.perf_counter() - t0
    log(f"core on device in {t_core:.1f}s, {core.device_bytes() / 1e9:.1f} GB")
    g = torch.Generator(device="cpu").manual_seed(1)
    Xh = (torch.randn((R, N), generator=g, dtype=torch.float32)
          + 1j * torch.randn((R, N), generator=g, dtype=torch.float32)) / np.sqrt(2.0)
    X = Xh.to(dev).t()  # column-major N x R complex64
    ncol = max(1, min(args.plain_rhs, R))
    X64 = X[:, :ncol].to(torch.complex128).t().contiguous().t()
    # plain ComplexF64 form, one column at a time
    P = h2g.H2Workspace(core, max_rhs=1, precision="plain")
    plain = {}
    for op, M in (("N", P), ("C", h2g.adjoint(P))):
        Y64 = torch.zeros((ncol, N), dtype=torch.complex128, device=dev).t()
        h2g.mul_(Y64[:, :1], M, X64[:, :1])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for j in range(ncol):
            h2g.mul_(Y64[:, j:j + 1], M, X64[:, j:j + 1])
        torch.cuda.synchronize()
        plain[op] = (Y64, (time.perf_counter() - t0) / ncol)
        log(f"plain C64 {op}: {1e3 * plain[op][1]:.1f} ms per RHS")
    del P
    # O(N^2) direct sum on 1,000 targets: forward G x, adjoint conj(G) x = conj(G conj(x))
    rows = torch.from_numpy(np.random.default_rng(2).choice(N, 1000, replace=False)).to(dev)
    P_d = torch.from_numpy(pts).to(dev)
    yd = {"N": torch.zeros((1000, ncol), dtype=torch.complex128, device=dev)}
    yd["C"] = torch.zeros_like(yd["N"])
    for c0 in range(0, N, 200_000):
        diff = P_d[rows][:, None, :] - P_d[None, c0:c0 + 200_000, :]
        d = torch.sqrt((diff * diff).sum(-1))
        dd = torch.where(d > 0, d, torch.ones_like(d))
        G = torch.where(d > 0, torch.exp(1j * k * dd) / (4 * np.pi * dd), torch.zeros_like(dd, dtype=torch.complex128))
        yd["N"] += G @ X64[c0:c0 + 200_000]
        yd["C"] += torch.conj(G @ torch.conj(X64[c0:c0 + 200_000]))
        del diff, d, dd, G
    rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    out_forms = {}
    for prec in ("mixed", "bf16x2", "bf16x3"):
        A = h2g.H2Workspace(core, max_rhs=R, precision=prec, use_graph=True)
        form = {}
        for op, M in (("N", A), ("C", h2g.adjoint(A))):
            Y = torch.zeros((R, N), dtype=torch.complex64, device=dev).t()
            for _ in range(args.warmup):
                h2g.mul_(Y, M, X)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                h2g.mul_(Y, M, X)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / args.steps
            Yc = Y[:, :ncol].to(torch.complex128)
            form[op] = {"ms_per_product": round(1e3 * dt, 2), "rhs_per_s": round(R / dt, 2),
                        "rel_err_vs_plain_c64": rel(Yc, plain[op][0]),
                        "rel_err_vs_direct_sum_1000": rel(Yc[rows], yd[op]),
                        "speedup_vs_plain": round(plain[op][1] * R / dt, 2)}
            log(f"{prec} {op}: {form[op]}")
        del A
        Tm = h2g.H2Workspace(core, max_rhs=R, precision=prec, timing=True)
        h2g.mul_(Y, Tm, X)
        h2g.mul_(Y, Tm, X)
        ph = {}
        for (phi, lvl, ms) in Tm.timings():
            ph[PHASES[phi]] = round(ph.get(PHASES[phi], 0.0) + ms, 3)
        form["phase_ms"] = ph
        del Tm
        out_forms[prec] = form
    out = {
        "metric": "Helmholtz H2 matvec RHS/s (k=10, unit sphere, Chebyshev p=6, dense complex M2L), forward + adjoint",
        "workload": f"helmholtz k={k} N={N} unit sphere, leaf<=64, p=6 Chebyshev dense, {R} complex RHS",
        "n": N, "nrhs": R, "data": "synthetic", "stats": arrays.stats,
        "plain_c64_ms_per_rhs": {op: round(1e3 * v[1], 2) for op, v in plain.items()},
        "plain_rel_err_vs_direct_sum_1000": {op: rel(plain[op][0][rows], yd[op]) for op in plain},
        "forms": out_forms,
        "build_s": {"map": round(t_build, 2), "plan": args.plan, "device_core": round(t_core, 1)},
        "device_bytes": {"core": core.device_bytes()},
    }
    print(json.dumps(out), flush=True)


def serving_mix(seed=0, n_maps=8, helmholtz_m2l="toeplitz"):
    """The serving workload's resident maps, drawn from a seed: Laplace (equispaced p, Toeplitz
    couplings) or Helmholtz (k in {2, 10}; equispaced p with Toeplitz complex couplings, or with
    helmholtz_m2l="dense" Chebyshev p with dense complex couplings), N log-uniform in [50k, 4M],
    points uniform in the unit cube or on the unit sphere, p in {4, 6}."""
    import numpy as np

    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n_maps):
        kernel = str(rng.choice(["laplace", "helmholtz"]))
        k = float(rng.choice([2.0, 10.0]))
        n = int(round(np.exp(rng.uniform(np.log(5e4), np.log(4e6)))))
        geom = str(rng.choice(["cube", "sphere"]))
        p = int(rng.choice([4, 6]))
        spec = dict(kernel=kernel, k=k if kernel == "helmholtz" else 0.0, n=n, geom=geom, p=p)
        if kernel == "helmholtz":
            spec["m2l"] = helmholtz_m2l
        specs.append(spec)
    return specs


def build_spec(spec, seed):
    import numpy as np

    from h2g import method as PM

    rng = np.random.default_rng(seed)
    if spec["geom"] == "cube":
        pts, box = rng.random((spec["n"], 3)), dict(origin=(0.0, 0.0, 0.0), size=1.0)
    else:
        pts = rng.standard_normal((spec["n"], 3))
        pts /= np.linalg.norm(pts, axis=1, keepdims=True)
        box = dict(origin=(-1.0, -1.0, -1.0), size=2.0)
    if spec["kernel"] == "laplace":
        return PM.build_map(pts, p=spec["p"], nodekind="equi", m2l_kind="toeplitz", leaf_size=64, plan="device",
                            ops="device", **box)
    if spec.get("m2l", "toeplitz") == "toeplitz":
        return PM.build_map(pts, p=spec["p"], nodekind="equi", m2l_kind="toeplitz", kernel="helmholtz",
                            wavenumber=spec["k"], leaf_size=64, plan="device", ops="device", **box)
    return PM.build_map(pts, p=spec["p"], nodekind="cheb", m2l_kind="dense", kernel="helmholtz",
                        wavenumber=spec["k"], leaf_size=64, plan="device", ops="device", **box)


def serving_leg(args):
    """Continuous batching over the resident map mix: a seeded Poisson stream of single-column
    forward / transpose / adjoint requests (FP32 or ComplexF32 N(0,1)), batched per map up to 32
    columns, maps admitted against the HBM budget; against the plain FP64 form serving one request
    at a time (its service time per map, i.e. its throughput capacity with no queueing)."""
    import numpy as np
    import torch

    import h2g
    from h2g.serve import AdmissionError, Request, Server

    specs = serving_mix(args.serve_seed, helmholtz_m2l=args.serve_helmholtz_m2l)
    srv = Server(budget_bytes=int(args.serve_budget_gb * 1e9), max_batch=32, ws_per_map=args.serve_ws_per_map,
                 precision="mixed",
                 n_streams=4, autotune=not args.no_tune)
    admitted, rejected = [], []
    t0 = time.perf_counter()
    for i, s in enumerate(specs):
        arrays = build_spec(s, 100 + i)
        name = f"m{i}"
        try:
            rm = srv.add_map(name, arrays)
            admitted.append(dict(name=name, **s, gb=round(rm.device_bytes() / 1e9, 1), tile=rm.tile,
                                 workspaces=len(rm.slots), max_batch=rm.max_batch,
                                 product_ms={"1": round(1e3 * rm.t1, 2), str(rm.max_batch): round(1e3 * rm.t32, 2)}))
            log(f"serving: admitted {name} {s}: {rm.device_bytes() / 1e9:.1f} GB, total {srv.used_bytes() / 1e9:.1f} GB")
        except AdmissionError as e:
            rejected.append(dict(name=name, **s, reason=str(e)))
            log(f"serving: rejected {name} {s}: {e}")
        del arrays
    t_build = time.perf_counter() - t0
    names = [a["name"] for a in admitted]
    rng = np.random.default_rng(7)

    def stream(n_req):
        reqs = []
        for _ in range(n_req):
            name = names[int(rng.integers(len(names)))]
            rm = srv.maps[name]
            x = rng.standard_normal(rm.n).astype(np.float32)
            if rm.dtype == np.complex64:
                x = (x + 1j * rng.standard_normal(rm.n)).astype(np.complex64)
            op = "NTC"[int(rng.integers(3))]
            n_out, n_in = rm.core.shape
            reqs.append(Request(name, op, x[:n_in if op == "N" else n_out]))
        return reqs

    def per_map_stats(reqs):
        out = {}
        for name in names:
            lat = np.array([r.latency for r in reqs if r.map_name == name]) * 1e3
            bs = [r.batch_size for r in reqs if r.map_name == name]
            out[name] = {"requests": len(lat), "p50_ms": round(float(np.percentile(lat, 50)), 1),
                         "p99_ms": round(float(np.percentile(lat, 99)), 1), "mean_batch": round(float(np.mean(bs)), 1)}
        return out

    # (each map's first workspace ran a 1-column and a full-width product at admission)
    reqs = stream(args.serve_requests)
    srv.done.clear()
    st = srv.run_stream(reqs, rate_hz=args.serve_rate, seed=6)
    per_map = per_map_stats(reqs)
    st["accounting"] = dict(srv.last_accounting)
    log(f"serving: {st}")
    low = []
    # the main rate saturates the server whenever it exceeds what the mix can serve; the points
    # below arrive at fractions of the measured throughput, so their latencies are service times
    # plus batching and stream waits rather than queueing (batches are narrower there, and a
    # narrow product costs nearly a wide one, so the capacity at low load is below the saturated
    # throughput)
    rates = [args.serve_rate_low] if args.serve_rate_low >= 0 else \
        [round(f * st["matvecs_per_s"], 1) for f in (0.1, 0.25)]
    for i, rate_low in enumerate(rates):
        if rate_low <= 0 or args.serve_requests_low <= 0:
            continue
        reqs_low = stream(args.serve_requests_low)
        srv.done.clear()
        sl = srv.run_stream(reqs_low, rate_hz=rate_low, seed=8 + i)
        pt = {"requests": sl["requests"], "arrival_rate_hz": rate_low,
              "matvecs_per_s": round(sl["matvecs_per_s"], 1), "p50_ms": round(sl["p50_ms"], 1),
              "p99_ms": round(sl["p99_ms"], 1), "mean_batch": round(sl["mean_batch"], 2),
              "per_map": per_map_stats(reqs_low), "accounting": dict(srv.last_accounting)}
        low.append(pt)
        log(f"serving at {rate_low}/s: {pt}")
    # plain FP64 one request at a time: service time per map; and each map's served results (two
    # per map) against the plain product
    svc, errs = {}, {}
    for name in names:
        rm = srv.maps[name]
        P = h2g.H2Workspace(rm.core, max_rhs=1, precision="plain")
        mine = [r for r in reqs if r.map_name == name][:2]
        e = 0.0
        for r in mine:
            M = {"N": P, "T": h2g.transpose(P), "C": h2g.adjoint(P)}[r.op]
            y64 = M @ r.x.astype(np.complex128 if rm.dtype == np.complex64 else np.float64)
            e = max(e, float(np.linalg.norm(r.result - y64) / np.linalg.norm(y64)))
        errs[name] = e
        xd = torch.from_numpy(mine[0].x.astype(np.complex128 if rm.dtype == np.complex64 else np.float64)).cuda()
        n_out = rm.core.shape[0] if mine[0].op == "N" else rm.core.shape[1]
        yd = torch.zeros(n_out, dtype=xd.dtype, device="cuda")
        M = {"N": P, "T": h2g.transpose(P), "C": h2g.adjoint(P)}[mine[0].op]
        h2g.mul_(yd, M, xd)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(2):
            h2g.mul_(yd, M, xd)
        torch.cuda.synchronize()
        svc[name] = (time.perf_counter() - ts) / 2
        P.close()
        log(f"serving: {name} plain FP64 {1e3 * svc[name]:.1f} ms per request, served vs plain {e:.2e}")
    srv.close()
    mean_svc = float(np.mean([svc[r.map_name] for r in reqs]))
    return {"maps_admitted": admitted, "maps_rejected": rejected, "budget_gb": args.serve_budget_gb,
            "requests": st["requests"], "arrival_rate_hz": args.serve_rate, "streams": 4, "max_batch": 32,
            "workspaces_per_map": args.serve_ws_per_map,
            "matvecs_per_s": round(st["matvecs_per_s"], 1), "wall_s": round(st["wall_s"], 2),
            "p50_ms": round(st["p50_ms"], 1), "p99_ms": round(st["p99_ms"], 1),
            "mean_batch": round(st["mean_batch"], 2), "per_map": per_map, "accounting": st["accounting"],
            "plain_fp64_one_at_a_time": {"service_ms": {k: round(1e3 * v, 2) for k, v in svc.items()},
                                         "capacity_matvecs_per_s": round(1.0 / mean_svc, 2)},
            "throughput_vs_plain": round(st["matvecs_per_s"] * mean_svc, 2),
            "max_rel_err_vs_plain_fp64": max(errs.values()), "build_s": round(t_build, 1),
            "below_capacity": low}


def main():
    args = parse()
    if args.workload == "helmholtz" and args.gpus == 1:
        return helmholtz_main(args)
    if args.serve_only and args.gpus == 1:
        print(json.dumps({"serving": serving_leg(args)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding "
              "of this path is a design in DESIGN.md, not a measured configuration", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import h2g
    from h2g import method as PM

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        # a plain copy kernel of known size: 2^30 bytes read and 2^30 bytes written per launch
        a = torch.ones(2 ** 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 2 ** 30, "launches": 5}), flush=True)
        return
    N, R = args.n, args.nrhs

    t0 = time.perf_counter()
    pts = np.random.default_rng(0).random((N, 3))
    arrays = PM.build_map(pts, p=args.p, nodekind="equi", m2l_kind=args.m2l, leaf_size=64, plan=args.plan,
                          ops=args.plan)
    t_build = time.perf_counter() - t0
    log(f"map built ({args.plan} plan) in {t_build:.1f}s: {arrays.stats}")
    t0 = time.perf_counter()
    core = h2g.GpuH2Core(arrays, precisions=("plain", args.precision))
    torch.cuda.synchronize()
    t_core = time.perf_counter() - t0
    log(f"core on device in {t_core:.1f}s, {core.device_bytes() / 1e9:.1f} GB")

    g = torch.Generator(device="cpu").manual_seed(1)
    X = torch.randn((R, N), generator=g, dtype=torch.float32).to(dev).t()  # column-major N x R
    Y = torch.zeros((R, N), dtype=torch.float32, device=dev).t()

    # RHS block and tile shapes: measured on this map by the autotuner (outside the timed region)
    tile = {"max_rhs": R, "m2l_groups": 0, "near_cols": 0, "persistent": 0}
    tune_table = []
    if args.tile:
        v = [int(t) for t in args.tile.split("/")]
        tile = {"max_rhs": v[0], "m2l_groups": v[1], "near_cols": v[2], "persistent": v[3] if len(v) > 3 else 0,
                "dense_pairs": v[4] if len(v) > 4 else 0}
    elif not args.no_tune:
        from h2g.autotune import tune

        t0 = time.perf_counter()
        # RHS blocks 1, 4, 8, 16 and the full width (32 columns as 32 passes of 1, 8 of 4, ...)
        tile, tune_table = tune(core, nrhs=R, precision=args.precision,
                                blocks=tuple(sorted({b for b in (1, 4, 8, 16) if b < R} | {R})))
        log(f"autotune in {time.perf_counter() - t0:.1f}s: {tile}; " +
            ", ".join(f"{c['max_rhs']}/{c['m2l_groups']}/{c['near_cols']}/{c.get('persistent', 0)}: {ms * R:.2f} ms"
                      for c, ms in tune_table))
    R_launch = min(R, tile["max_rhs"])
    A = h2g.H2Workspace(core, precision=args.precision, use_graph=not args.no_graph, **tile)
    log(f"batched workspace {A.device_bytes() / 1e9:.1f} GB")
    for i in range(args.warmup):
        h2g.mul_(Y, A, X)
        log(f"warmup {i} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        h2g.mul_(Y, A, X)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    value = R * args.steps / dt
    log(f"batched: {ms_per_step:.2f} ms per {R}-RHS product, {value:.1f} RHS/s")

    # plain FP64 form, one right-hand side at a time
    P = h2g.H2Workspace(core, max_rhs=1, precision="plain")
    ncol = max(1, min(args.plain_rhs, R))
    X64 = X[:, :ncol].double().t().contiguous().t()
    Y64 = torch.zeros((ncol, N), dtype=torch.float64, device=dev).t()
    h2g.mul_(Y64[:, :1], P, X64[:, :1])  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j in range(ncol):
        h2g.mul_(Y64[:, j:j + 1], P, X64[:, j:j + 1])
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / ncol
    plain_rhs_per_s = 1.0 / t_plain
    err_vs_plain = float(torch.linalg.norm(Y[:, :ncol].double() - Y64) / torch.linalg.norm(Y64))
    log(f"plain FP64: {1e3 * t_plain:.2f} ms per RHS ({plain_rhs_per_s:.1f} RHS/s); "
        f"batched vs plain rel err {err_vs_plain:.2e}")

    # the transpose product, same columns and steps, against the plain FP64 transpose
    AT, PT = h2g.transpose(A), h2g.transpose(P)
    YT = torch.zeros_like(Y)
    h2g.mul_(YT, AT, X)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        h2g.mul_(YT, AT, X)
    torch.cuda.synchronize()
    dt_t = (time.perf_counter() - t0) / args.steps
    YT64 = torch.zeros_like(Y64)
    for j in range(ncol):
        h2g.mul_(YT64[:, j:j + 1], PT, X64[:, j:j + 1])
    transpose = {"ms_per_product": round(1e3 * dt_t, 3), "rhs_per_s": round(R / dt_t, 2),
                 "rel_err_vs_plain_fp64": float(torch.linalg.norm(YT[:, :ncol].double() - YT64) /
                                                torch.linalg.norm(YT64))}
    log(f"transpose: {transpose}")

    # O(N^2) direct sum on 1,000 sampled targets (FP64, on the GPU)
    rows = torch.from_numpy(np.random.default_rng(2).choice(N, 1000, replace=False)).to(dev)
    P_d = torch.from_numpy(pts).to(dev)
    yd = torch.zeros((1000, ncol), dtype=torch.float64, device=dev)
    for c0 in range(0, N, 250_000):
        diff = P_d[rows][:, None, :] - P_d[None, c0:c0 + 250_000, :]
        d = torch.sqrt((diff * diff).sum(-1))
        G = torch.where(d > 0, 1.0 / (4 * np.pi * torch.where(d > 0, d, torch.ones_like(d))), torch.zeros_like(d))
        yd += G @ X64[c0:c0 + 250_000]
        del diff, d, G
    err_direct_plain = float(torch.linalg.norm(Y64[rows] - yd) / torch.linalg.norm(yd))
    err_direct_batched = float(torch.linalg.norm(Y[rows, :ncol].double() - yd) / torch.linalg.norm(yd))
    log(f"direct sum on 1000 targets: plain {err_direct_plain:.2e}, batched {err_direct_batched:.2e}")

    # per-phase device times (HIP events on the workspace stream), same product, eager
    Tm = h2g.H2Workspace(core, precision=args.precision, timing=True, **tile)
    h2g.mul_(Y, Tm, X)
    recs = []
    for _ in range(2):
        h2g.mul_(Y, Tm, X)
        recs.append(Tm.timings())
    per = {}
    for rr in recs:
        for (ph, lvl, ms) in rr:
            per.setdefault((PHASES[ph], lvl), []).append(ms)
    totals = {}
    for (ph, lvl), v in per.items():
        totals[ph] = totals.get(ph, 0.0) + float(np.mean(v))
    for ph in sorted({k[0] for k in per}):
        log(f"  {ph} by level: " + ", ".join(f"L{l}={float(np.mean(per[(ph, l)])):.2f}"
                                             for l in sorted(l for (q, l) in per if q == ph)))
    log("per-phase ms (sum over levels): " + ", ".join(f"{k}={v:.2f}" for k, v in sorted(totals.items(), key=lambda t: -t[1])))
    Tp = h2g.H2Workspace(core, max_rhs=1, precision="plain", timing=True)
    h2g.mul_(Y64[:, :1], Tp, X64[:, :1])
    h2g.mul_(Y64[:, :1], Tp, X64[:, :1])
    ptot = {}
    plev = {}
    for (ph, lvl, ms) in Tp.timings():
        ptot[PHASES[ph]] = ptot.get(PHASES[ph], 0.0) + ms
        plev.setdefault(PHASES[ph], []).append(f"L{lvl}={ms:.2f}")
    log("plain m2l_apply by level: " + ", ".join(plev.get("m2l_apply", [])))
    log("plain per-phase ms (1 RHS): " + ", ".join(f"{k}={v:.2f}" for k, v in sorted(ptot.items(), key=lambda t: -t[1])))
    del Tp
    (dph, dlv), dms = max(((k, float(np.mean(v))) for k, v in per.items()), key=lambda t: t[1])
    kind, work = algorithmic_work(arrays, R_launch, dph, dlv, True)
    if kind == "hbm":
        achieved = work / (dms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK["hbm"], "unit": "GB/s",
                "frac": round(achieved / PEAK["hbm"], 4), "traffic": None}
    else:
        achieved = work / (dms * 1e-3) / 1e12
        # the batched coupling stage runs on FP32 MFMA (FP32 operands, FP64 partial-sum accumulation):
        # priced against the dense FP32 MFMA peak; the split-BF16 forms run 3 (hi + lo) or 6
        # (hi + mid + lo) BF16 MFMA products per FP32 product, priced as that many BF16 products
        pk = PEAK["fp32"]
        if args.precision != "mixed" and arrays.m2l_kind in ("dense", "symdense") and dph == "m2l_apply":
            nprod = 3 if args.precision == "bf16x2" else 6
            work *= nprod
            pk = PEAK["bf16"]
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": pk, "unit": "TFLOP/s",
                "frac": round(achieved / pk, 4), "traffic": traffic_of(f"{dph} level {dlv}", N, args.p, R_launch,
                                                                       tile["m2l_groups"])}
    roof["kernel"] = f"{dph} level {dlv}"
    roof["launch_ms"] = round(dms, 3)

    x0_host = X64[:, 0].cpu().numpy().copy()
    y0_plain_host = Y64[:, 0].cpu().numpy().copy()

    # serving: the resident map mix under a Poisson stream of single-column requests
    serve = None
    if args.serve_requests > 0:
        del A, Tm, P, core
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        serve = serving_leg(args)
        log(f"serving: {serve}")

    cpu = None
    if not args.no_cpu:
        import numpy as np

        dt_cpu, threads, y_cpu = cpu_baseline(arrays, x0_host, args.cpu_reps)
        err_cpu = float(np.linalg.norm(y_cpu - y0_plain_host) / np.linalg.norm(y0_plain_host))
        cpu = {"value": round(1.0 / dt_cpu, 4), "unit": "RHS/s", "cores": threads, "kind": "port",
               "nproc": os.cpu_count(), "cpu_model": cpu_model(),
               "sample": f"C/OpenMP restatement of the reference traversal (FP64), the full N={N} forward product "
                         f"of one column on {threads} threads, mean of {args.cpu_reps} after a warm-up "
                         f"({dt_cpu:.2f}s each); same core data and column as the GPU",
               "rel_err_vs_gpu_plain_fp64": err_cpu}
        log(f"cpu restatement: {dt_cpu:.2f}s per product on {threads} threads; GPU plain FP64 vs CPU {err_cpu:.2e}")

    out = {
        "metric": "H2 matvec RHS/s, N=2M p=6 32 RHS, rel.err<=1e-6 vs FP64, vs plain; p99 latency",
        "value": round(value, 2), "unit": "RHS/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "mixed" else args.precision, "data": "synthetic",
        "config": {"workload": f"laplace equispaced p={args.p} {args.m2l} M2L, N={N} uniform cube, leaf<=64, "
                               f"{R} RHS FP32 batched (FP32-stored near values, FP64 near arithmetic + final sum)"
                               + ("" if args.precision == "mixed" else f", couplings on split BF16 ({args.precision})"),
                   "n": N, "nrhs": R, "p": args.p, "m2l": args.m2l, "leaf_size": 64,
                   "tile": tile},
        "autotune": [{"config": c, "ms_per_product": round(ms * R, 3)} for c, ms in tune_table],
        "plain_fp64_rhs_per_s": round(plain_rhs_per_s, 2),
        "transpose": transpose,
        "speedup_vs_plain": round(value / plain_rhs_per_s, 2),
        "rel_err_vs_plain_fp64": err_vs_plain,
        "gate_1e-6_vs_plain": bool(err_vs_plain <= 1e-6),
        "rel_err_vs_direct_sum_1000": {"plain": err_direct_plain, "batched": err_direct_batched},
        "phase_ms": {k: round(v, 3) for k, v in totals.items()},
        "plain_phase_ms": {k: round(v, 3) for k, v in ptot.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "serving": serve,
        "build_s": {"map": round(t_build, 2), "plan": args.plan, "device_core": round(t_core, 1)},
    }
    print(json.dumps(out), flush=True)
    if not err_vs_plain <= 1e-6:
        log(f"INVALID: batched result {err_vs_plain:.2e} from the plain FP64 product (gate 1e-6)")
        sys.exit(3)


if __name__ == "__main__":
    main()
