"""Benchmark of the asqtad hot path on one MI355X.

Workload (32^3 x 64, the single-GPU configuration the metric is quoted on):
seeded SU(3) gauge links U = reunit(exp(0.3 i H)), MILC phases with
antiperiodic time, asqtad fattening with the MILC coefficients on the GPU,
unit-variance Gaussian even-site sources, m = 0.01, FP64 true relative
residual 1e-10 per solve.

A step is one batch of B sources solved by the batched mixed-precision CG
(sloppy FP16 vectors with per-site norms and FP16 links with per-link scales,
FP32 arithmetic, FP64 reductions, reliable updates and true residuals).
value = whole-job solves/s over the K timed steps; p50/p99 are the per-solve
latencies (step start to that solve's completion).  Beside it the plain FP64
GPU CG with the reference algorithm and stopping rule is timed on the same
lattice and sources, and the FP64 true residual of every solve is checked.

roofline: the dominant kernel of the step, the second multi-RHS FP16 hop of
the CG iteration with the CG update fused into it, timed with HIP events on
the library stream; algorithmic bytes per launch = (16*40 + B*112) * Vh
(hop_bytes_per_site).  traffic: HBM bytes per launch measured with rocprofv3
FETCH_SIZE / WRITE_SIZE (profiles/traffic.json).
cpu_baseline: the C oracle's FP64 CG (the reference algorithm) on the host
cores for a bounded number of iterations, extrapolated to the GPU FP64
solve's iteration count.

Other modes print their own JSON line: --multi (multi-mass), --hisq (HISQ
multi-mass at 48^3x96), --serve (request server latency), --force (asqtad
fermion force, see force_bench).
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s + p99 latency, mixed-prec batched vs FP64 CG at 1e-10 true resid"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--prec", type=int, default=16, choices=[16, 32])
    ap.add_argument("--lattice", default="32,32,32,64")
    ap.add_argument("--mass", type=float, default=0.01)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--fp64-solves", type=int, default=2)
    ap.add_argument("--delta", type=float, default=0.1,
                    help="qmx_delta of the batched solves: a reliable update when the iterated "
                         "residual has dropped by this factor")
    ap.add_argument("--ru-join", type=float, default=0.3,
                    help="qmx_ru_join of the batched solves: a slot whose residual dropped by this "
                         "factor joins another slot's reliable update (0: off, bitwise "
                         "batch-invariant results)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--eigcg", action="store_true",
                    help="a sequence of EVENODD solves with eigCG deflation (option cg = 1) "
                         "against plain CG on the same sources; prints one JSON line")
    ap.add_argument("--eigcg-solves", type=int, default=6)
    ap.add_argument("--eigcg-nev", type=int, default=8)
    ap.add_argument("--eigcg-m", type=int, default=40)
    ap.add_argument("--eigcg-l", type=int, default=120)
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the solver iterations one by one instead of replaying "
                         "captured hipGraphs (option qmx_graph = 0)")
    ap.add_argument("--single-rhs-only", action="store_true",
                    help="time the single-RHS hop in FP64/FP32/FP16 (100 launches each) and the "
                         "1 GiB copy calibration, print one JSON line and exit")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only time the dominant hop kernel (for the rocprofv3 cross-check)")
    ap.add_argument("--shape", default=None,
                    help="--roofline-only, --single-rhs-only: force the hop launch shape (8 "
                         "comma-separated ints: "
                         "RHS per thread, block, XCD remap, non-temporal, RHS-innermost, x0 tile, "
                         "LDS staging, plain loads) instead of the autotuned one")
    ap.add_argument("--serve", action="store_true",
                    help="serving mix instead: Poisson arrivals over several lattices")
    ap.add_argument("--serve-lattices", default="16,16,16,16;24,24,24,48;32,32,32,64")
    ap.add_argument("--serve-requests", type=int, default=48, help="requests per rate")
    ap.add_argument("--serve-rates", default="0.5,1,2,4",
                    help="arrival rates as multiples of the FP64 one-at-a-time capacity")
    ap.add_argument("--serve-capacity", type=int, default=16)
    ap.add_argument("--multi", action="store_true",
                    help="batched multi-mass solves (16 sources x 8 masses, 24^3x48) instead")
    ap.add_argument("--hisq", action="store_true",
                    help="HISQ links, one source x 12 masses at 48^3x96 instead")
    ap.add_argument("--force", action="store_true",
                    help="asqtad fermion force (fnmat form) at --lattice instead of the solver bench")
    ap.add_argument("--force-terms", type=int, default=8, help="vectors x_i of the force")
    ap.add_argument("--force-action", default="asqtad", choices=["asqtad", "hisq"],
                    help="--force: asqtad (MILC coefficients) or HISQ (MILC HISQ, one Naik group)")
    ap.add_argument("--fp64-sources", type=int, default=16,
                    help="--multi: sequential FP64 multi-mass solves timed for the baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)
    if args.serve:
        serve_bench(args)
        return
    if args.multi or args.hisq:
        multi_bench(args)
        return
    if args.force:
        force_bench(args)
        return
    if args.eigcg:
        eigcg_bench(args)
        return

    import qopqdp_amd as q

    L = [int(x) for x in args.lattice.split(",")]
    V = int(np.prod(L))
    Vh = V // 2
    B = args.batch
    t0 = time.time()
    U = q.gen_gauge(L, 1000 + 1 * 10 + 0, 0.3)  # gauge seed 1000 + cfg*10 + k
    s = q.Session(L)
    if args.no_graph:
        assert s.set_opts(qmx_graph=0) == 0
    g = s.gauge(U)
    del U
    r0, bc, sign, keep = q.milc_rephase_args(L)
    s.lib.QOP_D3_rephase_G(g, r0, C.byref(bc), C.byref(sign))
    fla, finfo = s.links_from_gauge(g)
    s.lib.QOP_D3_destroy_G(g)
    log(f"setup {time.time() - t0:.1f}s (fattening {finfo.final_sec * 1e3:.1f} ms on GPU)")

    if args.single_rhs_only:
        res = {}
        if args.shape:
            forced = [int(x) for x in args.shape.split(",")]
            assert len(forced) == 8, "--shape takes 8 integers"
            s.lib.QOPX_hop_shape_override((C.c_int * 8)(*forced))
        for prec, sl, sv in ((64, 144, 48), (32, 72, 24), (16, 40, 16)):
            shape = (C.c_int * 8)()
            ms = s.lib.QOPX_D3_asqtad_hop_time(fla, prec, 1, 100, 0, shape)
            byts = (16 * sl + 2 * sv) * Vh
            res[f"fp{prec}"] = {"kernel_us": round(ms * 1e3, 1), "GBs": round(byts / ms / 1e6, 1),
                                "shape": list(shape), "algorithmic_bytes": byts}
        cbytes = 1 << 30
        cms = s.lib.QOPX_copy_time(cbytes, 20)
        res["copy"] = {"bytes_read": cbytes, "us": round(cms * 1e3, 1),
                       "GBs_rw": round(2 * cbytes / cms / 1e6, 1)}
        print(json.dumps(res))
        s.lib.QOP_D3_asqtad_destroy_L(fla)
        s.close()
        return
    if args.roofline_only:
        res = {}
        if args.shape:
            forced = [int(x) for x in args.shape.split(",")]
            assert len(forced) == 8, "--shape takes 8 integers"
            s.lib.QOPX_hop_shape_override((C.c_int * 8)(*forced))
        for red in (1, 2):
            shape = (C.c_int * 8)()
            ms = s.lib.QOPX_D3_asqtad_hop_time(fla, args.prec, B, 50, red, shape)
            byts = hop_bytes_per_site(args.prec, B, red) * Vh
            res[f"red{red}"] = {"kernel_us": round(ms * 1e3, 1), "GBs": round(byts / ms / 1e6, 1),
                                "shape": list(shape), "algorithmic_bytes": byts}
        cbytes = 1 << 30
        cms = s.lib.QOPX_copy_time(cbytes, 20)
        res["copy"] = {"bytes_read": cbytes, "us": round(cms * 1e3, 1),
                       "GBs_rw": round(2 * cbytes / cms / 1e6, 1)}
        print(json.dumps(res))
        s.lib.QOP_D3_asqtad_destroy_L(fla)
        s.close()
        return
    srcs = [q.gen_source(L, 2000 + 100 * 1 + j, q.QOP_EVEN) for j in range(B)]
    vins = [s.vec(x) for x in srcs]
    vouts = [s.zeros() for _ in range(B)]
    P = C.c_void_p
    assert s.set_opts(qmx_ru_join=args.ru_join, qmx_delta=args.delta) == 0

    def batch_step():
        ras = [q.resid_arg_default() for _ in range(B)]
        for r in ras:
            r.rsqmin = args.tol * args.tol
        ia = q.invert_arg_default()
        ia.evenodd = q.QOP_EVEN
        ia.max_iter = 100000
        lat = (C.c_double * B)()
        masses = (C.c_double * B)(*([args.mass] * B))
        info = q.QOPInfo()
        ok = s.lib.QOPX_D3_asqtad_invert_batch(
            C.byref(info), fla, C.byref(ia), (P * B)(*[C.cast(C.pointer(r), P) for r in ras]),
            masses, (P * B)(*vouts), (P * B)(*vins), B, B, args.prec, lat)
        return ok, ras, list(lat), info

    # warmup (also autotunes the hop launch shapes)
    for w in range(args.warmup):
        batch_step()
        log(f"warmup step {w} done")
    lats, iters, steps_ms, flops = [], [], [], 0.0
    t_start = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        ok, ras, lat, info = batch_step()
        steps_ms.append((time.perf_counter() - ts) * 1e3)
        assert ok == B, f"only {ok}/{B} systems converged"
        lats += lat
        iters += [r.final_iter for r in ras]
        flops += info.final_flop
        log(f"step {k}: {steps_ms[-1]:.0f} ms, max final_rsq {max(r.final_rsq for r in ras):.2e}")
    elapsed = time.perf_counter() - t_start
    value = B * args.steps / elapsed

    # FP64 true residual of the last step's solutions, with the FP64 GPU operator
    true_res = []
    info = q.QOPInfo()
    tmp, Ax = s.zeros(), s.zeros()
    for j in range(B):
        x = s.get(vouts[j]) / args.mass
        vx = s.vec(x)
        s.lib.QOP_D3_asqtad_dslash(info, fla, 0.0, tmp, vx, q.QOP_ODD, q.QOP_EVEN)
        s.lib.QOP_D3_asqtad_dslash(info, fla, 0.0, Ax, tmp, q.QOP_EVEN, q.QOP_ODD)
        ax = args.mass ** 2 * x - s.get(Ax)
        r = srcs[j][:Vh] - ax[:Vh]
        true_res.append(float(np.linalg.norm(r) / np.linalg.norm(srcs[j][:Vh])))
        s.free(vx)
    s.free(tmp)
    s.free(Ax)

    # plain FP64 GPU CG, reference algorithm and stopping rule
    s.set_opts(qmx_solver=0)
    fp64_t, fp64_it = [], []
    for j in range(args.fp64_solves):
        ia = q.invert_arg_default()
        ia.evenodd = q.QOP_EVEN
        ia.max_iter = 100000
        ia.max_restarts = 50
        ra = q.resid_arg_default()
        ra.rsqmin = args.tol * args.tol
        vo = s.zeros()
        ts = time.perf_counter()
        s.lib.QOP_D3_asqtad_invert(C.byref(info), fla, C.byref(ia), C.byref(ra), args.mass, vo, vins[j])
        fp64_t.append(time.perf_counter() - ts)
        fp64_it.append(ra.final_iter)
        s.free(vo)
        log(f"FP64 CG solve {j}: {fp64_t[-1]:.3f}s, {ra.final_iter} iterations, rsq {ra.final_rsq:.2e}")
    fp64_rate = len(fp64_t) / sum(fp64_t)

    # dominant kernel roofline: the second hop of the batched CG iteration
    # (fused CG update); the first hop beside it
    shape, shape1 = (C.c_int * 8)(), (C.c_int * 8)()
    nrep = 50
    hop_ms = s.lib.QOPX_D3_asqtad_hop_time(fla, args.prec, B, nrep, 2, shape)
    hop_bytes = hop_bytes_per_site(args.prec, B, 2) * Vh
    achieved = hop_bytes / (hop_ms * 1e-3) / 1e9
    hop1_ms = s.lib.QOPX_D3_asqtad_hop_time(fla, args.prec, B, nrep, 1, shape1)
    hop1_bytes = hop_bytes_per_site(args.prec, B, 1) * Vh
    kname = f"hop_kernel fp{args.prec} x{B} RHS, fused CG update"
    traffic = measured_traffic(kname, L, list(shape))
    # single-RHS bandwidth baseline (FP64 / FP32 / FP16 even->odd Dslash)
    single = {}
    for prec, sl, sv in ((64, 144, 48), (32, 72, 24), (16, 40, 16)):
        ms = s.lib.QOPX_D3_asqtad_hop_time(fla, prec, 1, 100, 0, None)
        single[f"fp{prec}"] = dict(us=round(ms * 1e3, 1),
                                    GBs=round((16 * sl + 2 * sv) * Vh / (ms * 1e-3) / 1e9, 1))
    log("single-RHS hop:", single)

    # CPU baseline: the oracle's FP64 CG, bounded sample
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(s, fla, L, srcs[0], args, float(np.mean(fp64_it)))
    for v in vins + vouts:
        s.free(v)
    s.lib.QOP_D3_asqtad_destroy_L(fla)
    s.close()

    lat_ms = np.array(lats) * 1e3
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(np.mean(steps_ms)), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"fp{args.prec} storage, fp32 arithmetic, fp64 reductions/true residual",
        "data": "synthetic",
        "config": {
            "workload": f"asqtad even-subset solves {'x'.join(map(str, L))}, m={args.mass}, "
                        f"{B} Gaussian sources per step, FP64 true residual {args.tol}",
            "lattice": L, "batch": B, "mass": args.mass, "tol": args.tol,
            "inner_prec": args.prec, "gauge": "reunit(exp(0.3iH)), MILC phases, asqtad smearing",
            "reliable_update_delta": args.delta, "ru_join": args.ru_join,
        },
        "p50_latency_ms": round(float(np.percentile(lat_ms, 50)), 1),
        "p99_latency_ms": round(float(np.percentile(lat_ms, 99)), 1),
        "fp64_cg_solves_per_s": round(fp64_rate, 3),
        "speedup_vs_fp64_cg": round(value / fp64_rate, 2),
        "fp64_cg_iters": int(np.mean(fp64_it)),
        "mixed_iters_mean": round(float(np.mean(iters)), 1),
        "max_fp64_true_resid": max(true_res),
        "gflops_reference_model": round(flops / elapsed / 1e9, 1),
        "single_rhs_dslash": single,
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "launch_us": round(hop_ms * 1e3, 1),
            "algorithmic_bytes": hop_bytes,
            "shape": list(shape),
        },
        "hop1": {
            "kernel": f"hop_kernel fp{args.prec} x{B} RHS, fused |Dp|^2",
            "launch_us": round(hop1_ms * 1e3, 1), "algorithmic_bytes": hop1_bytes,
            "GBs": round(hop1_bytes / (hop1_ms * 1e-3) / 1e9, 1),
            "frac": round(hop1_bytes / (hop1_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "shape": list(shape1),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def hop_bytes_per_site(prec, nrhs, red):
    """Algorithmic bytes per output site of one multi-RHS hop launch: 16
    links (FP16: 9 half2 + a float scale = 40 B; FP32 72 B) read once, and per
    right-hand side the gathered source vector (FP16 12 B + 4 B norm; FP32
    24 B) and the written output.  red 2 (fused CG update) reads p, r and the
    FP32 solution x and writes r and x instead of writing the hop output."""
    s_link, s_vec = (40, 16) if prec == 16 else (72, 24)
    if red == 2:
        return 16 * s_link + nrhs * (s_vec + 3 * s_vec + 2 * 24)
    return 16 * s_link + 2 * nrhs * s_vec


def measured_traffic(kernel, L, shape):
    """HBM bytes per launch of `kernel` with launch `shape` from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/traffic.json; null if
    that lattice and shape were not measured."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)
        e = t[kernel]
        if list(e["lattice"]) != list(L):
            return None
        return e["shapes"][",".join(str(int(v)) for v in shape)]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def multi_bench(args):
    """Multi-mass solves, every (source, mass) to FP64 true residual 1e-10.
    --multi: 24^3x48 asqtad links, 16 Gaussian even-site sources, 8 masses
    log-spaced 0.005..0.2; the batched multi-shift + reliable-update solver
    against sequential FP64 multi-mass solves (reference algorithm) of the
    same sources.  --hisq: 48^3x96 HISQ links (fat7, U(3) projection, level 2,
    epsilon_Naik = 0), FP16 links with per-link scales, one source, 12 masses
    log-spaced 0.002..0.5, the FP64 true residual of every shift reported."""
    import qopqdp_amd as q
    if args.hisq:
        L, nsrc, masses = [48, 48, 48, 96], 1, list(np.geomspace(0.002, 0.5, 12))
    else:
        L, nsrc, masses = [24, 24, 24, 48], 16, list(np.geomspace(0.005, 0.2, 8))
    V = int(np.prod(L))
    Vh = V // 2
    t0 = time.time()
    U = q.gen_gauge(L, 1000 + 10 * 4 + 0, 0.3)
    s = q.Session(L)
    if args.no_graph:
        assert s.set_opts(qmx_graph=0) == 0
    g = s.gauge(U)
    del U
    r0, bc, sign, keep = q.milc_rephase_args(L)
    s.lib.QOP_D3_rephase_G(g, r0, C.byref(bc), C.byref(sign))
    hisq = None
    if args.hisq:
        info = q.QOPInfo()
        hisq = s.lib.QOP_D3_hisq_create_L_from_G(C.byref(info), C.byref(q.hisq_coeffs()), g)
        fla = s.lib.QOP_D3_get_asqtad_links_from_hisq(hisq)[0]
        prep = {"sec": round(info.final_sec, 3), "svd_fallbacks": info.count1,
                "gflops": round(info.final_flop / info.final_sec / 1e9, 1)}
    else:
        fla, finfo = s.links_from_gauge(g)
        prep = {"sec": round(finfo.final_sec, 3)}
    s.lib.QOP_D3_destroy_G(g)
    log(f"setup {time.time() - t0:.1f}s, links {prep}")
    srcs = [q.gen_source(L, 2000 + 100 * 4 + j, q.QOP_EVEN) for j in range(nsrc)]
    vins = [s.vec(x) for x in srcs]
    assert s.set_opts(qmx_ru_join=args.ru_join) == 0  # the refinement's batched CG
    nm = len(masses)
    n = nsrc * nm
    P = C.c_void_p

    def batched():
        vouts = [s.zeros() for _ in range(n)]
        ras = [q.resid_arg_default() for _ in range(n)]
        for r in ras:
            r.rsqmin = args.tol * args.tol
        ia = q.invert_arg_default()
        ia.evenodd = q.QOP_EVEN
        ia.max_iter = 100000
        ms_it = (C.c_int * nsrc)()
        info = q.QOPInfo()
        ts = time.perf_counter()
        ok = s.lib.QOPX_D3_asqtad_invert_multi_batch(
            C.byref(info), fla, C.byref(ia), (P * n)(*[C.cast(C.pointer(r), P) for r in ras]),
            (C.c_double * nm)(*masses), nm, (P * n)(*vouts), (P * nsrc)(*vins), nsrc,
            min(n, 16), args.prec, 0.0, ms_it)
        dt = time.perf_counter() - ts
        return ok, dt, vouts, ras, list(ms_it)

    batched()  # warmup: autotuning and graph capture
    ok, dt, vouts, ras, ms_it = batched()
    log(f"batched: {ok}/{n} converged in {dt:.2f}s, multi-shift iterations {ms_it[:4]}")
    # FP64 true residual of every (source, mass) with the FP64 GPU operator
    tr = []
    tmp, Ax = s.zeros(), s.zeros()
    info = q.QOPInfo()
    for i in range(nsrc):
        for j, m in enumerate(masses):
            x = s.get(vouts[i * nm + j]) / m
            vx = s.vec(x)
            s.lib.QOP_D3_asqtad_dslash(info, fla, 0.0, tmp, vx, q.QOP_ODD, q.QOP_EVEN)
            s.lib.QOP_D3_asqtad_dslash(info, fla, 0.0, Ax, tmp, q.QOP_EVEN, q.QOP_ODD)
            r = srcs[i][:Vh] - (m * m * x - s.get(Ax))[:Vh]
            tr.append(float(np.linalg.norm(r) / np.linalg.norm(srcs[i][:Vh])))
            s.free(vx)
    for v in vouts + [tmp, Ax]:
        s.free(v)
    per_shift = [max(tr[i * nm + j] for i in range(nsrc)) for j in range(nm)]
    # iterations of the reliable-update refinement after the multi-shift stage
    # (0: the multi-shift solution already met the FP64 tolerance)
    refine_it = [max(ras[i * nm + j].final_iter for i in range(nsrc)) for j in range(nm)]
    # FP64 reference-algorithm multi-mass solves, one source at a time
    nfp = min(args.fp64_sources, nsrc)
    fp_t, fp_ok = [], 0
    for i in range(nfp):
        outs = [s.zeros() for _ in range(nm)]
        rl = [q.resid_arg_default() for _ in range(nm)]
        for r in rl:
            r.rsqmin = args.tol * args.tol
        ia = q.invert_arg_default()
        ia.evenodd = q.QOP_EVEN
        ia.max_iter = 100000
        ia.max_restarts = 50
        ra_p = (C.POINTER(q.QOPResidArg) * nm)(*[C.pointer(r) for r in rl])
        out_p = (P * nm)(*outs)
        ts = time.perf_counter()
        s.lib.QOP_D3_asqtad_invert_multi(
            C.byref(info), fla, C.byref(ia), (P * 1)(C.cast(ra_p, P)),
            (P * 1)(C.cast((C.c_double * nm)(*masses), P)), (C.c_int * 1)(nm),
            (P * 1)(C.cast(out_p, P)), (P * 1)(vins[i]), 1)
        fp_t.append(time.perf_counter() - ts)
        fp_ok += sum(r.final_rsq <= args.tol * args.tol * 1.0001 for r in rl)
        log(f"FP64 multi-mass source {i}: {fp_t[-1]:.2f}s, iters {[r.final_iter for r in rl][:3]}, "
            f"max rsq {max(r.final_rsq for r in rl):.2e}")
        for o in outs:
            s.free(o)
    for v in vins:
        s.free(v)
    if hisq:
        s.lib.QOP_D3_hisq_destroy_L(hisq)
    else:
        s.lib.QOP_D3_asqtad_destroy_L(fla)
    s.close()
    fp_rate = nm * len(fp_t) / sum(fp_t)
    rate = n / dt
    print(json.dumps({
        "metric": "multi-mass solves/s at 1e-10 FP64 true residual per (source, mass), "
                  "batched mixed precision vs sequential FP64 multi-shift CG",
        "value": round(rate, 3), "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "data": "synthetic",
        "dtype": f"fp{args.prec} storage, fp32 arithmetic, fp64 reductions/recurrences/true residual",
        "config": {"workload": ("HISQ" if args.hisq else "asqtad") + f" {'x'.join(map(str, L))}, "
                   f"{nsrc} sources x {nm} masses", "lattice": L, "sources": nsrc,
                   "masses": [round(m, 8) for m in masses], "tol": args.tol,
                   "ru_join": args.ru_join},
        "links": prep, "seconds": round(dt, 3), "converged": ok, "solves": n,
        "multi_shift_iters": ms_it, "max_fp64_true_resid_per_mass": per_shift,
        "max_refine_iters_per_mass": refine_it,
        "fp64_sequential": {"solves_per_s": round(fp_rate, 3), "sources_timed": len(fp_t),
                            "seconds_per_source": round(float(np.mean(fp_t)), 3),
                            "converged": int(fp_ok)},
        "speedup_vs_fp64": round(rate / fp_rate, 2),
    }), flush=True)


def eigcg_bench(args):
    """--eigcg-solves EVENODD solves (the reference's cg = 1 eigCG of
    lib/invert_eigcg_p.c inside the EVENODD invert) on --lattice at --mass to
    rsqmin = tol^2, each on a new seeded Gaussian source, then the same
    sources with plain CG (cg = 0); per-solve seconds and iterations of both."""
    import qopqdp_amd as q
    L = [int(x) for x in args.lattice.split(",")]
    U = q.gen_gauge(L, 1000 + 10 * 6, 0.3)
    s = q.Session(L)
    if args.no_graph:
        assert s.set_opts(qmx_graph=0) == 0
    g = s.gauge(U)
    del U
    r0, bc, sign, keep = q.milc_rephase_args(L)
    s.lib.QOP_D3_rephase_G(g, r0, C.byref(bc), C.byref(sign))
    fla, finfo = s.links_from_gauge(g)
    s.lib.QOP_D3_destroy_G(g)
    srcs = [q.gen_source(L, 2600 + k, q.QOP_EVEN) for k in range(args.eigcg_solves)]

    def sequence(cg):
        assert s.set_opts(cg=cg, eigcg_nev=args.eigcg_nev, eigcg_m=args.eigcg_m,
                          eigcg_l=args.eigcg_l) == 0
        fl = s.links_from_raw(*s.extract_links(fla))  # fresh links: an empty eigCG space
        secs, its = [], []
        for b in srcs:
            vin, vout = s.vec(b), s.zeros()
            ia = q.invert_arg_default()
            ia.evenodd = q.QOP_EVENODD
            ra = q.resid_arg_default()
            ra.rsqmin = args.tol * args.tol
            info = q.QOPInfo()
            t = time.perf_counter()
            s.lib.QOP_D3_asqtad_invert(C.byref(info), fl, C.byref(ia), C.byref(ra), args.mass, vout, vin)
            secs.append(round(time.perf_counter() - t, 4))
            its.append(ra.final_iter)
            assert ra.final_rsq <= args.tol * args.tol * 1.0001, ra.final_rsq
            s.free(vin)
            s.free(vout)
        s.lib.QOP_D3_asqtad_destroy_L(fl)
        return secs, its

    sequence(0)  # warmup (autotuning of nothing; first launches)
    e_s, e_it = sequence(1)
    c_s, c_it = sequence(0)
    s.set_opts(cg=0)
    s.lib.QOP_D3_asqtad_destroy_L(fla)
    s.close()
    print(json.dumps({
        "metric": "EVENODD solves with eigCG deflation vs plain CG, seconds per solve",
        "lattice": L, "mass": args.mass, "tol": args.tol,
        "eigcg": {"nev": args.eigcg_nev, "m": args.eigcg_m, "l": args.eigcg_l,
                  "seconds": e_s, "iterations": e_it, "total_s": round(sum(e_s), 3)},
        "cg": {"seconds": c_s, "iterations": c_it, "total_s": round(sum(c_s), 3)},
        "eigcg_over_cg": round(sum(e_s) / sum(c_s), 3)}), flush=True)


def force_bench(args):
    """Fermion force (fnmat form): one step is one force of --force-terms
    full-lattice vectors (Gaussian, seeded) on a seeded SU(3) gauge field at
    --lattice (default 32^3x64), accumulated into a device force field.
    --force-action asqtad: QOP_D3_asqtad_force_multi, MILC coefficients;
    hisq: QOP_D3_hisq_force_multi on HISQ links built from the rephased field
    (MILC HISQ coefficients, one epsilon-Naik group).  cpu_baseline: the C
    oracle's force of the same number of terms on a 16^3x32 sample (1/16 of
    the volume at the default lattice), scaled by the volume ratio (every
    operation of the force is per site)."""
    import qopqdp_amd as q
    L = [int(x) for x in args.lattice.split(",")]
    V = int(np.prod(L))
    n = args.force_terms
    hisq = args.force_action == "hisq"
    s = q.Session(L)
    U = q.gen_gauge(L, 2000, 0.3)
    g = s.gauge(U)
    del U
    flh = None
    if hisq:
        r0, bc, sign, keep = q.milc_rephase_args(L)
        s.lib.QOP_D3_rephase_G(g, r0, C.byref(bc), C.byref(sign))
        hc = q.hisq_coeffs()
        hinfo = q.QOPInfo()
        flh = s.lib.QOP_D3_hisq_create_L_from_G(C.byref(hinfo), C.byref(hc), g)
        assert hinfo.status == 0
    f = s.force([np.zeros((V, 3, 3, 2)) for _ in range(4)])
    xs = [s.vec(q.gen_source(L, 2100 + i, q.QOP_EVENODD)) for i in range(n)]
    eps = [1.0 / (i + 1) for i in range(n)]

    def one():
        if not hisq:
            return s.asqtad_force(g, f, eps, xs)
        info = q.QOPInfo()
        s.lib.QOP_D3_hisq_force_multi(C.byref(info), flh, f, C.byref(hc), (C.c_double * n)(*eps),
                                      (C.c_void_p * n)(*xs), (C.c_int * 1)(n))
        return info

    for _ in range(args.warmup):
        info = one()
        assert info.status == 0
    t = time.perf_counter()
    secs = []
    for _ in range(args.steps):
        info = one()
        assert info.status == 0
        secs.append(info.final_sec)
    dt = time.perf_counter() - t
    flop = info.final_flop
    F = s.get_force(f)
    a = F[0][::101, :, :, 0] + 1j * F[0][::101, :, :, 1]
    antiherm = float(np.abs(a + np.conj(np.transpose(a, (0, 2, 1)))).max() / np.abs(a).max())
    for x in xs:
        s.free(x)
    s.lib.QOP_D3_destroy_F(f)
    if flh:
        s.lib.QOP_D3_hisq_destroy_L(flh)
    s.lib.QOP_D3_destroy_G(g)
    s.close()
    out = {
        "metric": f"{args.force_action} fermion force (fnmat form) per second",
        "value": round(args.steps / dt, 3), "unit": "forces/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.force_action} force {'x'.join(map(str, L))}, {n} terms, MILC "
                               f"{'HISQ ' if hisq else ''}coefficients",
                   "lattice": L, "terms": n},
        "gflops_reference_count": round(flop / float(np.median(secs)) / 1e9, 1),
        "antihermiticity_check": antiherm,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = force_cpu_baseline(L, n, eps, hisq)
    print(json.dumps(out), flush=True)


def force_cpu_baseline(L, n, eps, hisq=False):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle
    import qopqdp_amd as q
    orc = Oracle()
    Ls = [max(2, x // 2) for x in L]
    Vs = int(np.prod(Ls))
    rng = np.random.default_rng(3)
    U = q.gen_gauge(Ls, 2200, 0.3)
    xs = [rng.standard_normal((Vs, 3, 2)) for _ in range(n)]
    coeffs = dict(one_link=5.0 / 8, three_staple=-1.0 / 16, five_staple=1.0 / 64,
                  seven_staple=-1.0 / 384, lepage=-1.0 / 16, naik=-1.0 / 24)
    if hisq:
        orc.rephase_milc(Ls, U)
        fat7 = dict(one_link=1 / 8, three_staple=-1 / 16, five_staple=1 / 64, seven_staple=-1 / 384,
                    lepage=0.0, naik=0.0)
        lev2 = dict(coeffs, one_link=1.0, lepage=-1 / 8)
    t = time.perf_counter()
    if hisq:
        orc.hisq_force(Ls, U, fat7, lev2, 0.0, -1 / 24, [0.0], eps, xs, [n], force_filter=5e-5)
    else:
        orc.asqtad_force(Ls, U, coeffs, eps, xs)
    dt = time.perf_counter() - t
    scale = int(np.prod(L)) / Vs
    log(f"cpu force baseline: {dt:.2f} s at {Ls} on {orc.threads()} threads")
    return {"value": round(1.0 / (dt * scale), 6), "unit": "forces/s", "cores": orc.threads(),
            "kind": "port",
            "sample": f"one force of {n} terms at {'x'.join(map(str, Ls))} ({dt:.2f} s on "
                      f"{orc.threads()} OpenMP threads), scaled by the volume ratio {scale:g}"}


def serve_bench(args):
    """Serving mix: requests over lattices x 2 resident configurations each,
    masses log-uniform on [0.005, 0.1], tolerances in {1e-8, 1e-10}, seeded
    Gaussian even-site sources, Poisson arrivals (seed 31337) at rates swept
    as multiples of the FP64 one-at-a-time capacity.  The batched server
    (mixed precision, continuous batching) and the FP64 one-at-a-time server
    see the same request streams; prints one JSON line."""
    import qopqdp_amd as q
    lattices = [[int(x) for x in t.split(",")] for t in args.serve_lattices.split(";")]
    rng = np.random.default_rng(31337)
    t0 = time.time()
    gauges = []
    for li, L in enumerate(lattices):
        for k in range(2):
            gauges.append((L, 1000 + 10 * li + k))

    def make(prec, cap):
        srv = q.Server(capacity=cap, prec=prec)
        for L, seed in gauges:
            U = q.gen_gauge(L, seed, 0.3)
            srv.add_gauge(L, U)
            del U
        return srv

    n_max = args.serve_requests
    pool = []  # request parameters, drawn once per rate from the same generator
    src_cache = {}

    def source(cfg, seed):
        key = (cfg, seed)
        if key not in src_cache:
            src_cache[key] = q.gen_source(gauges[cfg][0], seed, q.QOP_EVEN)
        return src_cache[key]

    def draw(n):
        reqs = []
        for i in range(n):
            cfg = int(rng.integers(len(gauges)))
            m = float(np.exp(rng.uniform(np.log(0.005), np.log(0.1))))
            tol = [1e-8, 1e-10][int(rng.integers(2))]
            reqs.append((cfg, m, tol, 5000 + int(rng.integers(8))))
        return reqs

    def run(srv, reqs, rate):
        gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
        ids, t_sub = [], []
        start = time.perf_counter()
        t_next = start
        for (cfg, m, tol, sd), gap in zip(reqs, gaps):
            t_next += gap
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            t_sub.append(time.perf_counter())
            ids.append(srv.submit(cfg, source(cfg, sd), m, tol, q.QOP_EVEN, max_iter=100000))
        lat, ok = [], 0
        for rid in ids:
            out, st, res, l = srv.wait(rid)
            ok += st == 0
            lat.append(l)
        done = max(ts + l for ts, l in zip(t_sub, lat))
        return dict(solves_per_s=round(len(reqs) / (done - start), 3),
                    p50_ms=round(float(np.percentile(lat, 50)) * 1e3, 1),
                    p99_ms=round(float(np.percentile(lat, 99)) * 1e3, 1), converged=ok)

    # FP64 one-at-a-time capacity from back-to-back requests
    srv64 = make(64, 1)
    log(f"serve: setup {time.time() - t0:.1f}s")
    warm = draw(len(gauges))
    run(srv64, [(c, 0.1, 1e-8, 5000) for c in range(len(gauges))], 0)
    cap_reqs = draw(max(8, n_max // 4))
    cap = run(srv64, cap_reqs, 0)
    c64 = cap["solves_per_s"]
    log(f"serve: FP64 one-at-a-time capacity {c64} solves/s")
    rates = [float(x) for x in args.serve_rates.split(",")]
    fp64_half = run(srv64, draw(max(8, n_max // 4)), 0.5 * c64)
    log(f"serve: FP64 loop at 0.5x capacity: {fp64_half}")
    srv64.close()
    srv = make(args.prec, args.serve_capacity)
    run(srv, warm, 0)  # builds and tunes every batch
    sweep = []
    for f in rates:
        r = run(srv, draw(n_max), f * c64)
        r["rate_x_fp64"] = f
        r["arrival_per_s"] = round(f * c64, 3)
        sweep.append(r)
        log(f"serve: rate {f}x -> {r}")
    st = srv.stats()
    srv.close()
    best = max(sweep, key=lambda r: r["solves_per_s"])
    print(json.dumps({
        "metric": "serving solves/s + p50/p99 latency, continuous batching vs FP64 one-at-a-time",
        "value": best["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "higher_is_better": True, "data": "synthetic",
        "dtype": f"fp{args.prec} storage, fp32 arithmetic, fp64 reductions/true residual",
        "config": {"workload": "asqtad even-subset solve requests, Poisson arrivals",
                   "lattices": lattices, "configs_per_lattice": 2,
                   "mass": "log-uniform [0.005, 0.1]", "tol": [1e-8, 1e-10],
                   "requests_per_rate": n_max, "capacity": args.serve_capacity},
        "fp64_one_at_a_time": {"capacity_solves_per_s": c64, "at_0.5x": fp64_half},
        "sweep": sweep, "batches": st["batches"], "chunks": st["chunks"],
    }), flush=True)


def cpu_baseline(s, fla, L, src, args, gpu_iters):
    """Oracle FP64 CG on the host, bounded to about args.cpu_seconds: built
    with -march=native on this host when gcc is there (else the shipped
    x86-64-v3 build), OpenMP threads pinned one per core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    import qopqdp_amd as q
    cores = os.cpu_count()
    try:
        cores = len(os.sched_getaffinity(0))  # before OpenMP pins this thread
    except AttributeError:
        pass
    march = "x86-64-v3"
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle_native.so"],
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    if r.returncode == 0 and _oracle.Oracle._lib is None:
        _oracle.ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle_native.so")
        march = "native"
    # read by the OpenMP runtime when the oracle library first loads
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    orc = _oracle.Oracle()
    fat, lng = s.extract_links(fla)
    h = orc.links(L, fat, lng)
    del fat, lng
    # time a growing number of iterations until the sample is long enough
    n, per_it = 2, None
    while True:
        t = time.perf_counter()
        orc.invert(h, args.mass, src, q.QOP_EVEN, 1e-30, max_iter=n, restart=n + 10, max_restarts=0)
        dt = time.perf_counter() - t
        per_it = dt / n
        if dt > args.cpu_seconds / 4 or n >= 64:
            break
        n *= 2
    orc.free_links(h)
    solve_s = per_it * gpu_iters
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    log(f"cpu baseline: {per_it * 1e3:.1f} ms/iteration over {n} iterations on {orc.threads()} threads")
    return {
        "value": round(1.0 / solve_s, 5),
        "unit": "solves/s",
        "cores": orc.threads(),
        "kind": "port",
        "sample": f"{n} FP64 CG iterations of the same solve on {orc.threads()} OpenMP threads "
                  f"pinned one per core ({cores} cores available, {model}, -march={march}); "
                  f"solves/s extrapolated to the GPU FP64 solve's {int(gpu_iters)} iterations",
        "ms_per_iteration": round(per_it * 1e3, 2),
    }


if __name__ == "__main__":
    main()
