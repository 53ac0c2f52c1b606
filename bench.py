"""Benchmark of the HDDM full-wave hot path on one MI355X.

Metric: COCG interface iterations/s (and subdomain LDL^T solves/s) of the
batched GPU form, solving to ||r||inf/||r0||inf < 1e-8, next to the plain
one-subdomain-per-launch GPU form and the CPU oracle.

A step is one complete interface solve of the workload (first step with the
load vector, COCG iterations to 1e-8, last step), with the model already
assembled and factored in HBM.  The headline form (--schur explicit) applies
the interface operator through subdomain Schur complements formed once from
the factors (their formation time is reported under setup_ms); the implicit
form (Dirichlet LDL^T solves in every iteration, the reference's own
structure) is timed beside it on the same model, batched and plain.  Prints
one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "COCG interface iters/s + subdomain LDL solves/s, batched vs plain, at 1e-8 res"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # dense FP64 (vector and matrix cores), MI355X


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pmc_traffic(cfg, waves, split, persist):
    """HBM bytes per k_smv launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE, WRITE_SIZE, scaled by the copy-kernel calibration), used only
    while the kernel source is the one that was profiled and the run uses the
    profiled waves per workgroup, split of large subdomains and grid form
    (one workgroup per item or persistent); else None."""
    import hashlib
    path = os.path.join(ROOT, "profiles", f"k_smv_cfg{cfg}_pmc.json")
    src = os.path.join(ROOT, "adventure_fullwave_amd", "csrc", "fw_schur.hip")
    try:
        d = json.load(open(path))
        if d.get("source_sha256") != hashlib.sha256(open(src, "rb").read()).hexdigest():
            return None
        if d.get("smv_waves") != waves or d.get("smv_split", 0) != split or d.get("smv_persist", 0) != persist:
            return None
        return d["traffic_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def frequency_sweep(args, model, fwhip, expl):
    """n frequencies on the model's mesh: one batch (every launch serves all
    systems) against n separate single-system solves, same options; each
    timed after one warm-up solve.  Setup (assembly, factors, S_s formation)
    is reported apart from the solve."""
    import dataclasses
    import numpy as np
    n = args.sweep
    fac = [0.9 + 0.2 * k / (n - 1) for k in range(n)]
    models = [dataclasses.replace(model, omega=model.omega * f) for f in fac]
    kw = dict(eps=args.eps, max_loop=args.max_loop, graph_iters=args.graph_iters, schur=expl)

    def setup_and_solve(m):
        t0 = time.perf_counter()
        sol = fwhip.Solver(m)
        if expl:
            sol.schur_form()
        t_setup = time.perf_counter() - t0
        sol.solve(**kw)
        t0 = time.perf_counter()
        sol.solve(**kw)
        return sol, t_setup, time.perf_counter() - t0

    b, b_setup, b_dt = setup_and_solve(models)
    b_stats = b.stats_all()
    xb = [b.select(k).x() for k in range(n)]
    b_info = dict(b.info)
    b.close()
    log(f"[bench] sweep batch of {n}: {[st['iters'] for st in b_stats]} iterations, {1000 * b_dt:.1f} ms")
    s_iters, s_dt, s_setup, diff = [], 0.0, 0.0, 0.0
    for k, m in enumerate(models):
        s1, t_set, dt = setup_and_solve(m)
        s_iters.append(s1.stats()["iters"])
        s_dt += dt
        s_setup += t_set
        x1 = s1.x()
        diff = max(diff, float(np.linalg.norm(xb[k] - x1) / np.linalg.norm(x1)))
        s1.close()
    log(f"[bench] sweep separately: {s_iters} iterations, {1000 * s_dt:.1f} ms")
    tot_b = sum(st["iters"] for st in b_stats)
    return {"n_freq": n, "omega_factors": [round(f, 4) for f in fac], "schur": "explicit" if expl else "implicit",
            "batch": {"iters": [st["iters"] for st in b_stats], "solve_ms": round(1000 * b_dt, 2),
                      "iters_per_s": round(tot_b / b_dt, 2), "setup_ms": round(1000 * b_setup, 1),
                      "schur_form_ms": round(b_info["ms_schur_form"], 2), "factor_ms": round(b_info["ms_factor"], 2)},
            "separate": {"iters": s_iters, "solve_ms": round(1000 * s_dt, 2),
                         "iters_per_s": round(sum(s_iters) / s_dt, 2), "setup_ms": round(1000 * s_setup, 1)},
            "solve_speedup": round(s_dt / b_dt, 3),
            "max_rel_diff_batch_vs_separate": diff}


def kernel_extras(args, fwhip, mesh):
    """Kernel rates of other configurations (the model loaded, its S_s formed,
    the kernels timed with HIP events; no solve): the explicit matvec against
    the per-iteration algorithmic bytes (the headline roofline's definition)
    and its own stream, the implicit Schur step and the LDL^T solve."""
    extras = {}
    for c in [int(x) for x in args.extra_cfg.split(",") if x.strip()]:
        if c == args.cfg:
            continue
        t0 = time.time()
        m = mesh.make_config(c, pml_mode=args.pml)
        sol = fwhip.Solver(m)
        sol.schur_form()
        ms_solve, ms_schur, ms_smv = sol.time_kernels(reps=5)
        info = dict(sol.info)
        sol.close()
        b_it = info["bytes_iter"] - 128.0 * info["n_iface"]
        extras[f"cfg{c}"] = {
            "workload": f"{m.meta.get('dims')} cell box, {m.n_edges} edge DOFs, {m.nsub} subdomains",
            "k_smv": {"ms_per_launch": round(ms_smv, 4), "achieved": round(b_it / (ms_smv * 1e-3) / 1e9, 1),
                      "unit": "GB/s", "frac": round(b_it / (ms_smv * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "explicit_stream_GBs": round(info["bytes_smv"] / (ms_smv * 1e-3) / 1e9, 1),
                      "smv_split": info["smv_split"], "smv_waves": info["smv_waves"],
                      "smv_persist": info["smv_persist"],
                      "ms_static_vs_persistent": [round(v, 4) for v in info["ms_smv_persist"]]},
            "k_solve_iter": {"ms_per_launch": round(ms_schur, 4), "bytes_per_launch": b_it,
                             "GBs": round(b_it / (ms_schur * 1e-3) / 1e9, 1)},
            "ldl_solve": {"ms": round(ms_solve, 4), "bytes": info["bytes_solve"],
                          "GBs": round(info["bytes_solve"] / (ms_solve * 1e-3) / 1e9, 1),
                          "solve_depth": info["solve_depth"], "solve_waves": info["solve_waves"],
                          "solve_persist": info["solve_persist"], "solve_large_queue": info["solve_nlarge"],
                          "ms_static_by_waves_1_2_4": [round(v, 4) for v in info["ms_solve_waves"]],
                          "ms_wide_depth_32": round(info["ms_solve_wide"], 4),
                          "ms_persistent_forms_1_2": [round(v, 4) for v in info["ms_solve_persist"]]},
        }
        log(f"[bench] extras cfg{c}: k_smv {ms_smv:.4f} ms, k_solve {ms_schur:.4f} ms, LDL^T {ms_solve:.4f} ms "
            f"({time.time() - t0:.1f}s)")
    return extras


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cfg", type=int, default=4)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--graph-iters", type=int, default=16)
    ap.add_argument("--plain-iters", type=int, default=10)
    ap.add_argument("--cpu-iters", type=int, default=200,
                    help="interface iterations of the CPU baseline sample (cfg 4 on 16 threads: about 10 s)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-loop", type=int, default=200000)
    ap.add_argument("--pml", default="sigma")
    ap.add_argument("--schur", choices=("explicit", "implicit"), default="explicit")
    ap.add_argument("--implicit-steps", type=int, default=1)
    ap.add_argument("--adjoint", action="store_true",
                    help="also time an adjoint permittivity gradient (forward solve, adjoint solve, gradient)")
    ap.add_argument("--extra-cfg", default="3",
                    help="comma-separated configurations whose kernels (explicit matvec, implicit Schur step, "
                         "LDL^T solve) are also timed and reported under extras ('' for none)")
    ap.add_argument("--factor-budget-mb", type=float, default=256.0,
                    help="after the measurements, release the skyline factors (explicit form) and solve again with "
                         "the first/last steps recomputing them in chunks of this many MB (0: skip)")
    ap.add_argument("--sweep", type=int, default=0,
                    help="also solve a sweep of this many frequencies f0*(0.9 + 0.2k/(n-1)) as one batch "
                         "(fwh_load_batch) and as separate solves")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from adventure_fullwave_amd import fwhip, mesh

    t0 = time.time()
    model = mesh.make_config(args.cfg, pml_mode=args.pml)
    log(f"[bench] cfg{args.cfg}: {model.n_edges} edges, {model.n_tets} tets, {model.nsub} subdomains, "
        f"{model.n_interface} interface unknowns ({time.time() - t0:.1f}s)")
    t0 = time.time()
    sol = fwhip.Solver(model)
    info = sol.info
    log(f"[bench] loaded: assemble {info['ms_assemble']:.1f} ms, factor {info['ms_factor']:.1f} ms, "
        f"sum nlen {info['nlen']} ({time.time() - t0:.1f}s)")

    expl = args.schur == "explicit"

    def run(schur, steps, warmup, tag):
        for _ in range(warmup):
            st = sol.solve(eps=args.eps, max_loop=args.max_loop, graph_iters=args.graph_iters, schur=schur)
            log(f"[bench] {tag} warmup: {st['iters']} iterations, converged={st['converged']}")
        iters = solves = 0
        t_start = time.perf_counter()
        for k in range(steps):
            st = sol.solve(eps=args.eps, max_loop=args.max_loop, graph_iters=args.graph_iters, schur=schur)
            iters += st["iters"]
            solves += st["sub_solves"]
            log(f"[bench] {tag} step {k}: {st['iters']} iterations, resid {st['final_resid']:.3e}")
        return iters, solves, time.perf_counter() - t_start

    # warmup + timed batched solves of the headline form
    iters, solves, dt = run(expl, args.steps, args.warmup, args.schur)
    iters_per_s = iters / dt
    info = dict(sol.info)

    # the other batched form, same model
    other = None
    if args.implicit_steps > 0:
        it2, sv2, dt2 = run(not expl, args.implicit_steps, 0, "implicit" if expl else "explicit")
        other = {"iters_per_s": round(it2 / dt2, 2), "ldl_solves_per_s": round(sv2 / dt2, 1),
                 "iters_per_solve": it2 // args.implicit_steps, "ms_per_step": round(1000 * dt2 / args.implicit_steps, 2)}
        info = dict(sol.info)
    imp = other if expl else {"iters_per_s": round(iters_per_s, 2), "ldl_solves_per_s": round(solves / dt, 1)}

    # plain form (one subdomain per launch, implicit, the reference's loop) on
    # a bounded window of iterations
    sol.first(eps=1e-30, max_loop=args.plain_iters + 1, plain=True, graph_iters=0)
    tp = time.perf_counter()
    sol.iterate(args.plain_iters)
    dtp = time.perf_counter() - tp
    plain_iters_per_s = args.plain_iters / dtp
    log(f"[bench] plain: {plain_iters_per_s:.2f} iters/s")
    # the plain form with the explicit Schur complements (one matvec launch
    # per subdomain): the stronger one-subdomain-per-launch comparator
    plain_expl = None
    if expl:
        sol.first(eps=1e-30, max_loop=4 * args.plain_iters + 1, plain=True, graph_iters=0, schur=True)
        tp = time.perf_counter()
        sol.iterate(4 * args.plain_iters)
        plain_expl = 4 * args.plain_iters / (time.perf_counter() - tp)
        log(f"[bench] plain explicit: {plain_expl:.2f} iters/s")
    plain_best = max(plain_iters_per_s, plain_expl or 0.0)

    # dominant kernels (HIP events on the library stream)
    ms_solve, ms_schur, ms_smv = sol.time_kernels(reps=20)
    bytes_iter_kernel = info["bytes_iter"] - 128.0 * info["n_iface"]
    k_solve = {"kernel": "k_solve (implicit Schur step, all subdomains)", "ms_per_launch": round(ms_schur, 4),
               "bytes_per_launch": bytes_iter_kernel,
               "GBs": round(bytes_iter_kernel / (ms_schur * 1e-3) / 1e9, 1),
               "ldl_only": {"ms": round(ms_solve, 4), "bytes": info["bytes_solve"],
                            "GBs": round(info["bytes_solve"] / (ms_solve * 1e-3) / 1e9, 1)}}
    if expl and ms_smv > 0:
        # one k_smv launch is the Schur step of one interface iteration: its
        # algorithmic bytes are the per-iteration unit's (sum over subdomains
        # of 16 nlen + 40 n + 20 nnz(A_IG, A_G.) + 48 ng) without the
        # interface-vector work (128 N_G, done by k_iter_ab); the explicit
        # form's own stream (packed S_s) is reported beside it
        achieved = bytes_iter_kernel / (ms_smv * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(args.cfg, info["smv_waves"], info["smv_split"], info["smv_persist"]),
                    "kernel": "k_smv (explicit Schur matvec, all subdomains: one interface iteration's Schur step)",
                    "ms_per_launch": round(ms_smv, 4), "bytes_per_launch": bytes_iter_kernel,
                    "explicit_stream": {"bytes_per_launch": info["bytes_smv"],
                                        "GBs": round(info["bytes_smv"] / (ms_smv * 1e-3) / 1e9, 1)},
                    "implicit_kernel": k_solve}
    else:
        achieved = bytes_iter_kernel / (ms_schur * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, **k_solve}
        roofline.pop("GBs")

    cpu = None
    if not args.no_cpu:
        from oracle import fw_oracle as fo
        # the host cores of one GPU's share of the box (at most 16)
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        tc, ns = fo.time_iterations(model, args.cpu_iters, threads)
        n1 = max(1, args.cpu_iters // 12)
        t1, _ = fo.time_iterations(model, n1, 1)
        cpu = {"value": round(args.cpu_iters / tc, 4), "unit": "iters/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_iters} interface iterations of the same workload "
                         f"({ns} subdomain solves, {tc:.1f} s on {threads} OpenMP threads, about "
                         f"{tc * threads:.0f} s of CPU work) with the C oracle, subdomain Schur steps "
                         f"in parallel; one-core rate from {n1} iterations ({t1:.1f} s)",
               "ldl_solves_per_s": round(ns / tc, 1), "one_core_iters_per_s": round(n1 / t1, 4)}
    adjoint = None
    if args.adjoint:
        import numpy as np
        # J = sum |E_k|^2 over 16 interior edges; dJ/deps by one adjoint solve
        off = np.concatenate([[0], np.cumsum(np.asarray(model.nmp, np.int64))])
        cand = []
        for s in range(model.nsub):
            bad = np.zeros(model.nmp[s], bool)
            bad[model.ib_nfid[model.ib_off[s]:model.ib_off[s + 1]]] = True
            bad[model.ibc[model.bc_off[s]:model.bc_off[s + 1]]] = True
            cand.append(off[s] + np.nonzero(~bad)[0])
        probes = np.sort(np.random.default_rng(0).choice(np.concatenate(cand), 16, replace=False))
        ta = time.perf_counter()
        st_f = sol.solve(eps=args.eps, max_loop=args.max_loop, graph_iters=args.graph_iters, schur=expl)
        E = sol.x()
        tb = time.perf_counter()
        f = np.zeros_like(E)
        f[probes] = np.conj(E[probes])
        sol.set_load(f)
        st_a = sol.solve(eps=args.eps, max_loop=args.max_loop, graph_iters=args.graph_iters, schur=expl)
        lam = sol.x()
        sol.set_load(None)
        tc_ = time.perf_counter()
        grad = sol.grad_eps(E, lam)
        td = time.perf_counter()
        adjoint = {"objective": "sum |E|^2 over 16 interior probe edges", "forward_ms": round(1000 * (tb - ta), 2),
                   "adjoint_ms": round(1000 * (tc_ - tb), 2), "gradient_ms": round(1000 * (td - tc_), 2),
                   "iters_forward_adjoint": [st_f["iters"], st_a["iters"]], "n_elem": int(grad.size),
                   "note": "host transfers of E and lambda included"}
        log(f"[bench] adjoint: forward {adjoint['forward_ms']} ms, adjoint {adjoint['adjoint_ms']} ms, "
            f"gradient {adjoint['gradient_ms']} ms")
    sweep = None
    if args.sweep > 1:
        sweep = frequency_sweep(args, model, fwhip, expl)
    # factors released after the formation: the explicit iteration runs on
    # the S_s alone, the first and last steps recompute the factors chunk by
    # chunk within the budget; same bits as the resident solve
    released = None
    if expl and args.factor_budget_mb > 0:
        import numpy as np
        kw = dict(eps=args.eps, max_loop=args.max_loop, graph_iters=args.graph_iters, schur=True)
        st0 = sol.solve(**kw)
        x0 = sol.x()
        budget = int(max(args.factor_budget_mb * 1e6, sol.info["factor_need_max"]))
        sol.release_factors(budget)
        t_r = time.perf_counter()
        st1 = sol.solve(**kw)
        dt_r = time.perf_counter() - t_r
        inf_r = sol.info
        released = {"budget_bytes": budget, "bytes_released": inf_r["factor_bytes_released"],
                    "chunks": inf_r["factor_chunks"], "chunk_buffer_bytes": inf_r["factor_chunk_bytes"],
                    "solve_ms": round(1000 * dt_r, 2), "iters": st1["iters"],
                    "first_ms": [round(st0["ms_first"], 2), round(st1["ms_first"], 2)],
                    "last_ms": [round(st0["ms_last"], 2), round(st1["ms_last"], 2)],
                    "same_bits_as_resident": bool(st0["iters"] == st1["iters"] and np.array_equal(x0, sol.x())),
                    "note": "first_ms and last_ms: resident factors, then recomputed in chunks"}
        log(f"[bench] factors released: {released['chunks']} chunks, first step {released['first_ms']} ms, "
            f"last step {released['last_ms']} ms, same bits {released['same_bits_as_resident']}")
        # a streamed load under the same budget: the factors are never all
        # resident (the S_s formed chunk by chunk)
        t_l = time.perf_counter()
        s2 = fwhip.Solver(model, factor_budget=budget)
        s2.schur_form()
        t_f = time.perf_counter()
        st2 = s2.solve(**kw)
        t_s = time.perf_counter()
        x2 = s2.x()
        released["streamed_load"] = {
            "load_and_formation_ms": round(1000 * (t_f - t_l), 1), "solve_ms": round(1000 * (t_s - t_f), 2),
            "iters": st2["iters"], "chunks": s2.info["factor_chunks"],
            "factors_resident": s2.info["factors_resident"],
            "rel_l2_vs_resident": float(np.linalg.norm(x2 - x0) / np.linalg.norm(x0)),
            "note": "the LDL^T kernel of the first/last steps runs at depth 16 with one wave (no resident factors "
                    "to tune it on), so the solution differs from the tuned resident one by rounding"}
        s2.close()
        log(f"[bench] streamed load: {released['streamed_load']}")
    setup = {"assemble": round(info["ms_assemble"], 2), "factor": round(info["ms_factor"], 2)}

    def fp64(flops, ms, kernel):
        if ms <= 0:
            return None
        tf = flops / (ms * 1e-3) / 1e12
        return {"kernel": kernel, "flops": flops, "ms": round(ms, 4), "achieved": round(tf, 3),
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / FP64_PEAK_TFLOPS, 4)}
    fp64_rooflines = {
        "factor": fp64(info["flops_factor"], info["ms_factor_kernel"],
                       "k_factor16 (skyline LDL^T, 16-row blocks on v_mfma_f64_16x16x4f64)"),
        "schur_fwd": fp64(info["flops_schur_fwd"], info["ms_schur_fwd"],
                          "k_schur_fwd (W = L^-1 (-A_IG), blocked forward substitution on the matrix cores)"),
        "schur_gemm": fp64(info["flops_schur_gemm"], info["ms_schur_gemm"],
                           "k_schur_gemm (S_s -= W^T D^-1 W, 16x16 tiles on the matrix cores)"),
        "note": "algorithmic complex-FP64 flops (8 per multiply-add) over the kernel's HIP-event time at load",
    }
    if expl:
        setup["schur_form"] = round(info["ms_schur_form"], 2)
        setup["schur_form_method"] = "W^T D^-1 W on FP64 MFMA" if info["schur_gemm"] else "unit-vector Schur steps"
    setup["factor_method"] = "16-row blocks on FP64 MFMA" if info["factor_blocked"] else "row by row"
    # the whole load (host index work, uploads, assembly, factorisation,
    # load-time tuning), of which the host index work
    setup["load_wall"] = round(info["ms_load"], 1)
    setup["host_index"] = round(info["ms_host_index"], 1)
    out = {
        "metric": METRIC, "value": round(iters_per_s, 2), "unit": "iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1000 * dt / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "complex128",
        "data": "synthetic",
        "config": {"workload": f"cfg{args.cfg}: {model.meta.get('dims')} cell box, {model.n_edges} edge DOFs, "
                               f"{model.nsub} subdomains in {model.npart} parts, PML {model.meta.get('pml_mode')}, "
                               f"COCG+diag to {args.eps:g}, Schur step {args.schur}",
                   "n_iface": info["n_iface"], "iters_per_solve": iters // max(args.steps, 1)},
        "ldl_solves_per_s": imp["ldl_solves_per_s"] if imp else None,
        "batched_ldl_kernel_solves_per_s": round(info["nsub"] / (ms_solve * 1e-3), 1),
        "implicit" if expl else "explicit": other,
        "plain": {"iters_per_s": round(plain_iters_per_s, 3), "window_iters": args.plain_iters,
                  "ldl_solves_per_s": round(plain_iters_per_s * info["nsub"], 1),
                  "explicit_iters_per_s": None if plain_expl is None else round(plain_expl, 3),
                  "form": "one subdomain per launch: iters_per_s implicit (LDL^T steps), explicit_iters_per_s with the S_s"},
        # against the faster plain form (explicit S_s, one launch per subdomain)
        "batched_over_plain": round(iters_per_s / plain_best, 2),
        "batched_over_plain_implicit": round(iters_per_s / plain_iters_per_s, 2),
        "roofline": roofline,
        "fp64_roofline": fp64_rooflines,
        "cpu_baseline": cpu,
        "setup_ms": setup,
        "autotune": {"smv_waves": info["smv_waves"], "ms_smv_by_waves_2_4_8": [round(v, 4) for v in info["ms_smv_tune"]],
                     "solve_buckets": info["solve_buckets"], "solve_bucketed": info["solve_bucketed"],
                     "ms_solve_single_vs_bucketed": [round(info["ms_solve_single"], 4),
                                                     round(info["ms_solve_bucketed"], 4)],
                     "solve_depth": info["solve_depth"], "unit_depth": info["unit_depth"],
                     "solve_waves": info["solve_waves"],
                     "ms_solve_by_waves_1_2_4": [round(v, 4) for v in info["ms_solve_waves"]],
                     "ms_solve_by_depth_8_16": [round(v, 4) for v in info["ms_solve_depth"]],
                     "ms_solve_wide_depth_32": round(info["ms_solve_wide"], 4),
                     "smv_order": info["smv_order"],
                     "ms_smv_by_order": [round(v, 4) for v in info["ms_smv_order"]],
                     "smv_split": info["smv_split"], "smv_items": info["smv_nwork"],
                     "ms_smv_by_split_0_2_4_8": [round(v, 4) for v in info["ms_smv_split"]],
                     "smv_persist": info["smv_persist"], "smv_persistent_grid": info["smv_pgrid"],
                     "ms_smv_static_vs_persistent": [round(v, 4) for v in info["ms_smv_persist"]],
                     "solve_persist": info["solve_persist"], "solve_persistent_grid": info["solve_pgrid"],
                     "solve_large_queue": info["solve_nlarge"],
                     "ms_solve_persistent_forms_1_2": [round(v, 4) for v in info["ms_solve_persist"]]},
        "adjoint": adjoint,
        "sweep": sweep,
        "factors_released": released,
        # from the model in host memory: the whole load (host index work,
        # uploads, assembly, factorisation, load-time tuning), the Schur
        # formation and one solve
        "time_to_solution_ms": round(setup["load_wall"] + setup.get("schur_form", 0.0) + 1000 * dt / args.steps, 1),
    }
    sol.close()
    out["extras"] = kernel_extras(args, fwhip, mesh)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
