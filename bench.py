"""Benchmark: SEPS of the batched + learned-scheduled sampler vs the plain GPU
sampler on the scale-22 mixed-request workload, with bit-exact edge lists.

Workload (one step = one pass over all requests): RMAT scale 22, edge
factor 16, a/b/c/d = .57/.19/.19/.05, symmetrised, FP32 weights in (0,1];
256 requests, seed counts log-uniform in [64, 65536], fanouts {25,10} or
{15,10,5}, bias uniform / degree / FP32 weight (Philox request mix, key 42).
Inputs (graph, seeds, plan buffers) are resident in HBM before timing.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--retrain] [--cpu-sample F]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "c-saw_amd"))
MODELS = os.path.join(ROOT, "c-saw_amd", "models")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_specs(C, g, n, mix_key):
    specs = []
    for i, r in enumerate(C.request_mix(n, 64, 65536, mix_key)):
        r = dict(r)
        r["seeds"] = g.seeds(r["samples"], key=42, request=i)
        specs.append(r)
    return specs


def get_model(C, CM, g, retrain):
    path = os.path.join(MODELS, "cost_model_gfx950.json")
    if os.path.exists(path) and not retrain:
        m, meta = CM.load(C, path)
        return m, meta
    t0 = time.time()
    train_specs = build_specs(C, g, 64, mix_key=7)
    data = CM.collect(C, g, train_specs)
    m, loss = CM.train(C, data)
    meta = {"train_mix_key": 7, "train_requests": 64, "chunks": int(len(data["keys"])), "final_loss": loss,
            "train_seconds": time.time() - t0}
    CM.save(m, path, meta)
    return m, meta


def get_tuning(C, g, specs, mlp, retrain):
    path = os.path.join(MODELS, "tuning_gfx950_c4.json")
    t = C.Tuning()
    if os.path.exists(path) and not retrain:
        d = json.load(open(path))
        for k, v in d.items():
            setattr(t, k, int(v))
        return t
    reqs, keep = C.make_requests(specs)
    opts = C.PlanOpts(mode=C.MODE_BATCHED, sched=C.SCHED_MLP, use_graph=1, mlp=mlp.h, cdf_cache=1)
    ms = C.ctypes.c_double()
    C._check(C.lib().csaw_autotune(g.h, reqs, len(specs), C.ctypes.byref(opts), C.ctypes.byref(t),
                                   C.ctypes.byref(ms)), "autotune")
    os.makedirs(MODELS, exist_ok=True)
    json.dump({k: getattr(t, k) for k, _ in C.Tuning._fields_}, open(path, "w"))
    return t


def uniform_tuning(C):
    """Launch shapes of the plain-uniform instance (a swept grid: 32
    expansion and 16 count workgroups per CU, 1.57 ms against 1.65 ms at the
    defaults of 8 and 8)."""
    t = C.Tuning()
    for k, v in json.load(open(os.path.join(MODELS, "tuning_gfx950_uniform.json"))).items():
        setattr(t, k, int(v))
    return t


def run_timed(C, plan, steps, warmup):
    for _ in range(warmup):
        plan.run(sync=True)
    t0 = time.perf_counter()
    for i in range(steps):
        plan.run(sync=(i == steps - 1))
    return (time.perf_counter() - t0) / steps


def cpu_baseline(C, O, g, specs, gpu_plan, frac, nthreads):
    """Oracle (OpenMP) on the first ceil(frac*S) samples of every request;
    also checks those samples' GPU edges equal the oracle's."""
    off, adj, w = g.csr(weights=True)
    og = O.OracleGraph(off, adj, weights=w)
    edges, secs, mism, checked = 0, 0.0, 0, 0
    for i, s in enumerate(specs):
        m = max(1, int(np.ceil(frac * len(s["seeds"]))))
        seeds = s["seeds"][:m]
        t0 = time.perf_counter()
        r = O.sample(og, seeds, s["fanouts"], bias=s["bias"], request=i, nthreads=nthreads)
        secs += time.perf_counter() - t0
        edges += len(r["src"])
        full = gpu_plan.result(i)
        ho = full["hop_off"]
        for h in range(len(s["fanouts"])):
            a, b = ho[h, 0], ho[h, m]
            ra, rb = r["hop_off"][h, 0], r["hop_off"][h, m]
            checked += 1
            if not (np.array_equal(full["src"][a:b], r["src"][ra:rb])
                    and np.array_equal(full["dst"][a:b], r["dst"][ra:rb])):
                mism += 1
    return edges, secs, mism, checked


def roofline_of(plan, runs=2):
    """Dominant kernel kind of an event-timed plan: algorithmic bytes per launch / avg launch time."""
    plan.run()
    plan.reset_times()
    for _ in range(runs):
        plan.run()
    kt = plan.kernel_times(with_bytes=True)
    dom = max(kt, key=lambda k: kt[k][0] * kt[k][1])
    avg_ms, nl, avg_b = kt[dom]
    ach = avg_b / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None, "kernel": dom, "algo_bytes_per_launch": avg_b, "avg_launch_ms": avg_ms}, kt


def line(metric, value, unit, steps, warmup, ms, config, dtype, roof, cpu, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": config, "roofline": roof, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out))


def oracle_check_sample(O, og, specs, gpu_results, frac, nthreads):
    """Oracle on the first ceil(frac*S) samples of each request; returns
    (edges, seconds, request-hops equal, request-hops checked)."""
    edges, secs, ok, checked = 0, 0.0, 0, 0
    for i, s in enumerate(specs):
        m = max(1, int(np.ceil(frac * len(s["seeds"]))))
        t0 = time.perf_counter()
        r = O.sample(og, s["seeds"][:m], s["fanouts"], bias=s["bias"], request=s.get("request", i),
                     nthreads=nthreads)
        secs += time.perf_counter() - t0
        edges += len(r["src"])
        full = gpu_results(i)
        for h in range(len(s["fanouts"])):
            a, b = full["hop_off"][h, 0], full["hop_off"][h, m]
            ra, rb = r["hop_off"][h, 0], r["hop_off"][h, m]
            checked += 1
            ok += int(np.array_equal(full["src"][a:b], r["src"][ra:rb])
                      and np.array_equal(full["dst"][a:b], r["dst"][ra:rb]))
    return edges, secs, ok, checked


def bench_single(a, C, O, name, scale, ef, weights, seeds_n, fanouts, bias, variant=None):
    """C1 / C2 / plain-uniform instance: one request, plain and batched
    executors (batched: the learned schedule with the shipped cost model, or
    the variant given)."""
    g = C.Graph.rmat(scale, ef, weights=weights)
    specs = [{"seeds": g.seeds(seeds_n), "fanouts": fanouts, "bias": bias, "request": 0}]
    ut = uniform_tuning(C) if (bias == C.BIAS_UNIFORM and seeds_n > 100000) else None
    plain = C.Plan(g, specs, mode=C.MODE_PLAIN, tuning=ut)
    import cost_model as CM
    mlp, _ = CM.load(C, CM.DEFAULT_MODEL)
    sched = C.SCHED_FORCED if variant is not None else C.SCHED_MLP
    bat = C.Plan(g, specs, mode=C.MODE_BATCHED, sched=sched, variant=variant or 0, use_graph=True, cdf_cache=True,
                 mlp=mlp)
    tp = run_timed(C, plain, a.steps, a.warmup)
    tb = run_timed(C, bat, a.steps, a.warmup)
    e = plain.n_edges(0)
    assert plain.digest(0) == bat.digest(0)
    roof, kt = roofline_of(C.Plan(g, specs, mode=C.MODE_PLAIN, timing=True, tuning=ut))
    cpu = None
    if not a.no_cpu:
        off, adj, *w = g.csr(weights=weights)
        og = O.OracleGraph(off, adj, weights=w[0] if w else None)
        ce, cs, ok, chk = oracle_check_sample(O, og, specs, plain.result, a.cpu_sample if seeds_n > 10000 else 1.0,
                                              host_cpus()["threads_used"])
        cpu = {"value": ce / cs, "unit": "edges/s", "cores": host_cpus()["threads_used"], "kind": "port",
               "sample": f"{ce} edges ({cs:.1f} s); oracle == GPU in {ok}/{chk} hops"}
    line(f"SEPS ({name})", e / tp, "edges/s", a.steps, a.warmup, tp * 1e3,
         {"workload": name, "seeds": seeds_n, "fanouts": fanouts, "bias": bias, "edges_per_step": e},
         "u64", roof, cpu, batched_value=e / tb, kernel_times={k: v[0] for k, v in kt.items()})


def bench_walk(a, C, O):
    """C3: RMAT-22 ef16 FP32 weights, 1,048,576 walkers x 80 steps."""
    g = C.Graph.rmat(22, 16, weights=True)
    seeds = g.seeds(1 << 20)
    wp = C.WalkPlan(g, seeds, 80, bias=C.BIAS_WEIGHT)
    for _ in range(a.warmup):
        wp.run()
    ms = []
    for _ in range(a.steps):
        wp.run()
        W, steps, kms = wp.result(walks=False)
        ms.append(kms)
    kms = float(np.median(ms))
    ach = wp.algo_bytes / (kms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None, "kernel": "walk", "algo_bytes_per_launch": wp.algo_bytes, "avg_launch_ms": kms}
    cpu = None
    if not a.no_cpu:
        W, _, _ = wp.result()
        off, adj, w = g.csr(weights=True)
        og = O.OracleGraph(off, adj, weights=w)
        m = int(len(seeds) * a.cpu_sample)
        t0 = time.perf_counter()
        Wo, so = O.walk(og, seeds[:m], 80, bias=O.BIAS_WEIGHT, nthreads=host_cpus()["threads_used"])
        cs = time.perf_counter() - t0
        cpu = {"value": so / cs, "unit": "steps/s", "cores": host_cpus()["threads_used"], "kind": "port",
               "sample": f"first {m} walkers ({so} steps, {cs:.1f} s); equal to the GPU rows: "
                         f"{bool(np.array_equal(Wo, W[:m]))}"}
    line("walk steps/s (C3)", steps / (kms * 1e-3), "steps/s", a.steps, a.warmup, kms,
         {"workload": "RMAT-22 ef16, FP32 weights, 1,048,576 walkers x 80", "steps": int(steps)}, "f32", roof, cpu)


def get_stream_model(C, CM, hg, retrain):
    """Cost model of the streaming sampler's own kernels, trained on its
    per-chunk logs: streamed runs of the C5 graph with a different seed set
    (key 7; the benchmark's seeds use key 42), one per schedule, joined per
    chunk; the held-out regret (seed key 11) of the learned and the fixed
    schedules is kept in its meta."""
    path = CM.STREAM_MODEL
    if os.path.exists(path) and not retrain:
        return CM.load(C, path)
    t0 = time.time()
    fan = [15, 10, 5]
    data = CM.collect_stream(C, hg, C.seeds_v(hg.V, 65536, key=7), fan, C.BIAS_DEGREE)
    m, loss = CM.train(C, data)
    held = CM.collect_stream(C, hg, C.seeds_v(hg.V, 65536, key=11), fan, C.BIAS_DEGREE)
    meta = {"graph": "RMAT-26 ef16 (host-resident, 16 partitions, 4 GB budget)", "train_seed_key": 7,
            "seeds": 65536, "fanouts": fan, "bias": "degree", "chunks": int(len(data["keys"])),
            "final_loss": loss, "held_out_seed_key": 11, "held_out_regret": CM.regret(held, m),
            "train_seconds": time.time() - t0}
    CM.save(m, path, meta)
    return m, meta


def bench_stream(a, C, O, CM):
    """C5: RMAT-26 ef16 in pinned host memory, 4 GB device budget, 16 partitions,
    65,536 seeds, fanouts {15,10,5}, degree bias; each partition's chunks are
    run by the variant the learned cost model picks, the fixed heuristic and
    the plain one-wave-per-entry expansion are timed beside it."""
    t0 = time.time()
    hg = C.HostGraph.rmat(26, 16)
    log(f"host graph V={hg.V} E={hg.E} maxdeg={hg.max_degree} ({time.time() - t0:.1f}s)")
    t0 = time.time()
    packed = hg.pack() if a.packed_slices else 0
    if packed:
        log(f"packed adjacency {packed / 1e9:.2f} GB of {4 * hg.E / 1e9:.2f} GB ({time.time() - t0:.1f}s)")
    mlp, meta = get_stream_model(C, CM, hg, a.retrain)
    seeds = C.seeds_v(hg.V, 65536)
    kw = dict(bias=C.BIAS_DEGREE, budget_bytes=4 << 30, partitions=16)

    # the three schedules take turns in every round (so that drift in the
    # host link's speed does not favour one of them); medians per schedule
    scheds = {"mlp": dict(mode=C.MODE_BATCHED, sched=C.SCHED_MLP, mlp=mlp),
              "heuristic": dict(mode=C.MODE_BATCHED, sched=C.SCHED_HEURISTIC),
              "plain": dict(mode=C.MODE_PLAIN)}
    last, ms, ems = {}, {k: [] for k in scheds}, {k: [] for k in scheds}
    names = list(scheds)
    for i in range(a.warmup + a.steps):
        for j in range(len(names)):
            k = names[(i + j) % len(names)]
            last[k] = C.stream_sample(hg, seeds, [15, 10, 5], **kw, **scheds[k])
            if i >= a.warmup:
                ms[k].append(last[k]["stats"].device_ms)
                ems[k].append(last[k]["stats"].expand_ms)
    log("per-run ms", {k: [round(x, 1) for x in v] for k, v in ms.items()})
    for k in names:
        st_k = last[k]["stats"]
        log(k, "hop ms", [round(x, 1) for x in list(st_k.hop_ms)[:3]], "copy ms", round(st_k.copy_ms, 1),
            "h2d GB", round(st_k.h2d_bytes / 1e9, 2), "direct", st_k.direct_partitions, "reused", st_k.reused_slices)
    res, heur, plain = last["mlp"], last["heuristic"], last["plain"]
    dms, ems_ = float(np.median(ms["mlp"])), float(np.median(ems["mlp"]))
    hms, hems = float(np.median(ms["heuristic"])), float(np.median(ems["heuristic"]))
    pms, pems = float(np.median(ms["plain"])), float(np.median(ems["plain"]))
    ems = ems_
    keys = ("src", "dst", "smp", "hop_off")
    equal = all(np.array_equal(res[k], heur[k]) and np.array_equal(res[k], plain[k]) for k in keys)
    st = res["stats"]
    log(f"stream: mlp {dms:.1f} ms (expand {ems:.1f})  heuristic {hms:.1f} (expand {hems:.1f})  "
        f"plain {pms:.1f} (expand {pems:.1f})  equal {equal}")
    cpu = None
    if not a.no_cpu:
        off, adj = hg.csr_views()
        og = O.OracleGraph(off, adj)
        m = 256
        t1 = time.perf_counter()
        r = O.sample(og, seeds[:m], [15, 10, 5], bias=O.BIAS_DEGREE, nthreads=host_cpus()["threads_used"])
        cs = time.perf_counter() - t1
        okh = all(np.array_equal(res["dst"][res["hop_off"][h, 0]:res["hop_off"][h, m]],
                                 r["dst"][r["hop_off"][h, 0]:r["hop_off"][h, m]]) for h in range(3))
        cpu = {"value": len(r["src"]) / cs, "unit": "edges/s", "cores": host_cpus()["threads_used"],
               "kind": "port", "sample": f"first {m} seeds ({len(r['src'])} edges, {cs:.1f} s); equal to the "
                                         f"streamed GPU edges: {okh}"}
    line("SEPS streaming (C5)", st.edges / (dms * 1e-3), "edges/s", a.steps, a.warmup, dms,
         {"workload": "RMAT-26 ef16 host-resident, 4 GB budget, 16 partitions, 65,536 seeds, {15,10,5}, degree",
          "schedule": "learned cost model per chunk", "edges": int(st.edges), "h2d_bytes": st.h2d_bytes,
          "peak_device_bytes": int(st.peak_device_bytes), "slice_bytes": int(st.slice_bytes),
          "direct_partitions": int(st.direct_partitions), "reused_slices": int(st.reused_slices),
          "packed_adjacency_bytes": int(packed), "h2d_raw_bytes": st.h2d_raw_bytes}, "u64", None, cpu,
         h2d_gbs=st.h2d_bytes / (dms * 1e-3) / 1e9,
         schedules={"mlp": {"ms": dms, "expand_ms": ems, "variant_chunks": list(st.variant_chunks),
                            "chunks": int(st.chunks)},
                    "heuristic": {"ms": hms, "expand_ms": hems,
                                  "variant_chunks": list(heur["stats"].variant_chunks),
                                  "chunks": int(heur["stats"].chunks)},
                    "plain": {"ms": pms, "expand_ms": pems}},
         schedules_bit_equal=bool(equal), cost_model=meta)


def probe_lib():
    import ctypes
    L = ctypes.CDLL(os.path.join(ROOT, "c-saw_amd", "lib", "libcsaw_gather.so"))
    L.csaw_gather_probe.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
    L.csaw_copy_probe.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    return L


def gather_ceiling(array_mib=512, reads=64 << 20, wg_per_cu=8, iters=5):
    """Random-read ceiling: independent random 4-byte reads per second from an
    array of array_mib MiB (512 MiB: the RMAT-22 neighbour array), measured by
    the library's own probe kernel (csrc/gather_probe.hip)."""
    import ctypes
    L = probe_lib()
    ms, n = ctypes.c_double(), ctypes.c_int64()
    rc = L.csaw_gather_probe(array_mib << 20, reads, wg_per_cu, iters, ctypes.byref(ms), ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f"csaw_gather_probe: hip error {rc}")
    return {"array_mib": array_mib, "random_reads": n.value, "ms": ms.value, "wg_per_cu": wg_per_cu,
            "reads_per_s": n.value / (ms.value * 1e-3), "line_bytes_per_s_at_64B": n.value * 64 / (ms.value * 1e-3)}


def probe(a):
    import ctypes
    if a.only == "copy":
        # 256 MiB read + 256 MiB written per launch: scales FETCH_SIZE / WRITE_SIZE
        ms = ctypes.c_double()
        rc = probe_lib().csaw_copy_probe(256 << 20, a.steps, ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(f"csaw_copy_probe: hip error {rc}")
        return {"copy_bytes_read": 256 << 20, "copy_bytes_written": 256 << 20, "launches": a.steps, "ms": ms.value,
                "gbs": 2 * (256 << 20) / (ms.value * 1e-3) / 1e9}
    return gather_ceiling(min(a.gather_mib, 8191), iters=a.steps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--retrain", action="store_true")
    ap.add_argument("--cpu-sample", type=float, default=0.25)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["c4", "c1", "c2", "c3", "uniform", "c5"], default="c4",
                    help="c4 (default) is the headline; the others measure the remaining configs")
    ap.add_argument("--no-graph", action="store_true", help="--only batched: launch eagerly (profilers)")
    ap.add_argument("--gather-mib", type=int, default=512, help="--only gather: array size (MiB, < 8192)")
    ap.add_argument("--packed-slices", action="store_true",
                    help="c5: pack the host adjacency and copy packed slices (fewer bytes, unpacked on the device)")
    ap.add_argument("--serial-expand", action="store_true",
                    help="run each hop's variant kernels one after another on one stream")
    ap.add_argument("--only", choices=["all", "batched", "plain", "copy", "gather"], default="all",
                    help="profiling modes: run one executor only (or a 256 MiB copy for counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus above 1 is not supported "
                                   "(the path would shard by seeds with no exchange between GPUs)"}))
        sys.exit(2)
    if a.only in ("copy", "gather"):
        print(json.dumps(probe(a)))
        return
    import __graft_entry__ as ge
    C = ge.load_csaw()
    import cost_model as CM

    if a.workload != "c4":
        O = ge.load_oracle() if not a.no_cpu else None
        if a.workload == "c1":
            bench_single(a, C, O, "C1: RMAT-16 ef8, 1000 samples, FS 1, NS 2, depth 2, uniform, bitmap",
                         16, 8, False, 1000, [2, 2], C.BIAS_UNIFORM)
        elif a.workload == "c2":
            bench_single(a, C, O, "C2: RMAT-20 ef5, 8192 seeds, {25,10}, degree bias", 20, 5, False, 8192, [25, 10],
                         C.BIAS_DEGREE)
        elif a.workload == "uniform":
            bench_single(a, C, O, "plain uniform: RMAT-22 ef16, 1,048,576 seeds, {25,10}, uniform", 22, 16, False,
                         1 << 20, [25, 10], C.BIAS_UNIFORM)
        elif a.workload == "c3":
            bench_walk(a, C, O)
        else:
            bench_stream(a, C, O, CM)
        return

    t0 = time.time()
    g = C.Graph.rmat(22, 16, weights=True)
    specs = build_specs(C, g, 256, mix_key=42)
    log(f"graph V={g.V} E={g.E} maxdeg={g.max_degree}; requests={len(specs)} "
        f"seeds={sum(len(s['seeds']) for s in specs)} ({time.time() - t0:.1f}s)")
    mlp, meta = get_model(C, CM, g, a.retrain)
    tuning = get_tuning(C, g, specs, mlp, a.retrain)
    log("tuning", {k: getattr(tuning, k) for k, _ in C.Tuning._fields_})
    if a.only in ("batched", "plain"):
        if a.only == "plain":
            p = C.Plan(g, specs, mode=C.MODE_PLAIN)
        else:
            p = C.Plan(g, specs, mode=C.MODE_BATCHED, sched=C.SCHED_MLP, mlp=mlp, use_graph=not a.no_graph,
                       tuning=tuning, cdf_cache=True, serial_expand=a.serial_expand)
        t = run_timed(C, p, a.steps, a.warmup)
        e = sum(p.n_edges(i) for i in range(len(specs)))
        print(json.dumps({"only": a.only, "ms_per_step": t * 1e3, "edges": e, "value": e / t}))
        return

    # one-time preprocessing, timed apart from the steps: the per-bias data
    # both samplers need (FP32 CDF of the weights, hub CDFs), then the CDF
    # cache of the batched sampler (integer CDF cache = the reference's Cp,
    # 16-ary search indexes, vertex records)
    tiny = [{"seeds": specs[0]["seeds"][:1], "fanouts": [1], "bias": b, "request": 0} for b in (1, 2)]
    t0 = time.perf_counter()
    C.Plan(g, tiny, mode=C.MODE_PLAIN).close()
    prep_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    C.Plan(g, tiny, mode=C.MODE_PLAIN, cdf_cache=True).close()
    cache_ms = (time.perf_counter() - t0) * 1e3
    log(f"one-time: bias data {prep_ms:.1f} ms, CDF cache {cache_ms:.1f} ms")

    # the same 256 requests through four executors: what each of the cache,
    # batching and the learned schedule adds
    plain = C.Plan(g, specs, mode=C.MODE_PLAIN)
    t_plain = run_timed(C, plain, max(1, min(a.steps, 3)), 1)
    want = [plain.digest(i) for i in range(len(specs))]
    edges = sum(plain.n_edges(i) for i in range(len(specs)))
    runs = {}

    def measure(name, **kw):
        p = C.Plan(g, specs, **kw)
        t = run_timed(C, p, a.steps, a.warmup)
        eq = sum(p.digest(i) == want[i] for i in range(len(specs)))
        runs[name] = {"ms_per_step": t * 1e3, "value": edges / t, "requests_equal_to_plain": int(eq)}
        log(f"{name}: {t * 1e3:.2f} ms, digests equal {eq}/{len(specs)}")
        return p, t, eq

    p, _, _ = measure("plain_cdf_cache", mode=C.MODE_PLAIN, cdf_cache=True)
    p.close()
    p, _, _ = measure("batched_heuristic", mode=C.MODE_BATCHED, sched=C.SCHED_HEURISTIC, use_graph=True,
                      tuning=tuning, cdf_cache=True, serial_expand=a.serial_expand)
    p.close()
    # the best fixed policy: one variant forced on every chunk (the wave
    # variant with the bitmap taken set is the fastest single variant on C4)
    p, _, _ = measure("batched_forced_wave_bitmap", mode=C.MODE_BATCHED, sched=C.SCHED_FORCED,
                      variant=C.VAR_WAVE_BITMAP, use_graph=True, tuning=tuning, cdf_cache=True,
                      serial_expand=a.serial_expand)
    p.close()
    batched, t_bat, same = measure("batched_learned", mode=C.MODE_BATCHED, sched=C.SCHED_MLP, mlp=mlp,
                                   use_graph=True, tuning=tuning, cdf_cache=True, serial_expand=a.serial_expand)
    log(f"plain {t_plain * 1e3:.1f} ms  batched {t_bat * 1e3:.1f} ms  edges {edges}  digests equal {same}/256")

    # roofline of the dominant kernel kind: event-timed run of the same plan
    # settings, each expansion kernel timed on the stream it runs on (the
    # hop's variant kernels side by side, as in the timed steps)
    prof = C.Plan(g, specs, mode=C.MODE_BATCHED, sched=C.SCHED_MLP, mlp=mlp, timing=True, tuning=tuning,
                  cdf_cache=True, serial_expand=a.serial_expand)
    prof.run()
    prof.reset_times()
    for _ in range(2):
        prof.run()
    kt = prof.kernel_times(with_bytes=True)
    prof.close()
    dom = max(kt, key=lambda k: kt[k][0] * kt[k][1])
    avg_ms, nl, avg_b = kt[dom]
    achieved = avg_b / (avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_c4_batched.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(dom)
    log("kernel kinds:", {k: (round(v[0], 3), v[1], round(v[2] / 1e6, 1)) for k, v in kt.items()})

    # the plain-uniform roofline instance on the same graph: 1,048,576 seeds,
    # fanouts {25,10}, uniform bias, plain sampler
    uspec = [{"seeds": g.seeds(1 << 20), "fanouts": [25, 10], "bias": C.BIAS_UNIFORM, "request": 0}]
    ut = uniform_tuning(C)
    up = C.Plan(g, uspec, mode=C.MODE_PLAIN, tuning=ut)
    t_u = run_timed(C, up, a.steps, a.warmup)
    e_u = up.n_edges(0)
    up.close()
    uroof, ukt = roofline_of(C.Plan(g, uspec, mode=C.MODE_PLAIN, timing=True, tuning=ut))
    ceiling = gather_ceiling()
    log(f"uniform instance {t_u * 1e3:.3f} ms ({e_u} edges), {uroof['kernel']} {uroof['frac']:.3f} of HBM peak; "
        f"random-read ceiling {ceiling['reads_per_s']:.3g}/s")

    cpu = None
    if not a.no_cpu:
        O = ge.load_oracle()
        host = host_cpus()
        nthreads = host["threads_used"]
        ce, cs, mism, checked = cpu_baseline(C, O, g, specs, batched, a.cpu_sample, nthreads)
        cpu = {"value": ce / cs, "unit": "edges/s", "cores": nthreads, "kind": "port",
               "sample": f"first {a.cpu_sample:.0%} of the seeds of each of the 256 requests "
                         f"({ce} edges, {cs:.1f} s); GPU edges of those samples equal the oracle's in "
                         f"{checked - mism}/{checked} request-hops",
               "host": host}
        log("cpu", cpu)

    out = {
        "metric": "SEPS on 1 MI355X: plain vs batched+learned-scheduled, bit-exact edge lists",
        "value": edges / t_bat, "unit": "edges/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t_bat * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": "RMAT-22 ef16 (symmetrised, FP32 weights), 256 requests: seeds log-uniform "
                               "[64,65536], fanouts {25,10}|{15,10,5}, bias uniform|degree|fp32",
                   "requests": len(specs), "seeds": int(sum(len(s["seeds"]) for s in specs)),
                   "edges_per_step": int(edges)},
        "plain_value": edges / t_plain, "plain_ms_per_step": t_plain * 1e3,
        "speedup_vs_plain": t_plain / t_bat,
        # the same requests through each executor; every one bit-equal to plain
        "executors": dict({"plain": {"ms_per_step": t_plain * 1e3, "value": edges / t_plain}}, **runs),
        "speedup_parts": {"cdf_cache (plain -> plain+cache)": t_plain / (runs["plain_cdf_cache"]["ms_per_step"] * 1e-3),
                          "batching (plain+cache -> batched+heuristic)":
                              runs["plain_cdf_cache"]["ms_per_step"] / runs["batched_heuristic"]["ms_per_step"],
                          "learned schedule (batched+heuristic -> batched+learned)":
                              runs["batched_heuristic"]["ms_per_step"] / runs["batched_learned"]["ms_per_step"],
                          "learned schedule vs the best fixed variant (batched, wave bitmap forced)":
                              runs["batched_forced_wave_bitmap"]["ms_per_step"]
                              / runs["batched_learned"]["ms_per_step"]},
        "bit_exact": {"requests_equal": int(same), "requests": len(specs)},
        "one_time_ms": {"bias_data (FP32 CDF, hub CDFs)": prep_ms,
                        "cdf_cache (integer CDF cache, search indexes, vertex records)": cache_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                     "algo_bytes_per_launch": avg_b, "avg_launch_ms": avg_ms,
                     "timing": "hipEvents on each expansion kernel's own stream, concurrent schedule"},
        "uniform_instance": {"workload": "RMAT-22 ef16, 1,048,576 seeds, {25,10}, uniform, plain sampler",
                             "ms_per_step": t_u * 1e3, "value": e_u / t_u, "edges": int(e_u), "roofline": uroof},
        "random_read_ceiling": ceiling,
        "cpu_baseline": cpu,
        "cost_model": meta,
    }
    print(json.dumps(out))


def host_cpus():
    """Host cores: nproc, the CPUs this process may run on, the cgroup's CPU
    quota, and the model name; the baseline uses every CPU the job is
    granted (affinity and quota)."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except OSError:
        info["affinity"] = os.cpu_count()
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    info["cgroup_cpu_quota"] = quota
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                info["model"] = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    # the job's CPU share: affinity, the cgroup quota, and OMP_NUM_THREADS
    # where the machine sets it to the share it grants (16 per GPU on the
    # GPU boxes here)
    n = info["affinity"] or 1
    if quota:
        n = min(n, max(1, int(quota)))
    omp = os.environ.get("OMP_NUM_THREADS")
    info["omp_num_threads"] = omp
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    info["threads_used"] = n
    return info


if __name__ == "__main__":
    main()
