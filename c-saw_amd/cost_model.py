"""Learned kernel-variant scheduler: data collection, GPU training, regret.

1. collect(): runs a batched workload once per forced variant with per-chunk
   logging; a chunk (hop, count block, degree bucket) is the same in every
   run, so joining the logs gives the device time of each valid variant for
   each chunk.  Costs are log2(ticks / workgroups-per-CU of the variant's
   kernel), i.e. the chunk's share of the device.
2. train(): fits the 16 -> 64 -> 64 -> 5 MLP (BF16 weights, FP32 accumulate)
   on the GPU with the library's forward / backward / SGD kernels.
3. regret(): on held-out chunks, the measured cost of the variant a policy
   picks minus the best measured cost, for the fixed C-SAW heuristic and for
   the trained model.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_DIR = os.path.join(HERE, "models")
DEFAULT_MODEL = os.path.join(MODEL_DIR, "cost_model_gfx950.json")
# trained on runs without the integer CDF cache: the kernels the streaming
# sampler runs (it builds integer CDFs per entry, hubs per partition)
NOCACHE_MODEL = os.path.join(MODEL_DIR, "cost_model_gfx950_nocache.json")
# concurrent chunk workers per CU of each variant's kernel (LaunchCfg
# defaults): THREAD and WAVE kernels claim chunks per wave (4 waves per
# workgroup), BLOCK per workgroup
WG_PER_CU = {0: 16, 1: 32, 2: 32, 3: 2, 4: 2}


def _runs(C, g, specs, sched, variant=0, cdf_cache=True, explore_key=0, tuning=None):
    p = C.Plan(g, specs, mode=C.MODE_BATCHED, sched=sched, variant=variant, log_chunks=True, cdf_cache=cdf_cache,
               explore_key=explore_key, tuning=tuning)
    p.run()  # warm-up
    p.run()
    X, var, lt, keys = p.chunk_log()
    p.close()
    return X, var, lt, keys


def collect(C, g, specs, cdf_cache=True, wg_per_cu=None):
    """Returns dict(keys, X[n,16], T[n,5], M[n,5], heur[n]) joined over variants."""
    wg = dict(WG_PER_CU)
    if wg_per_cu:
        wg.update(wg_per_cu)
    X0, hv, hl, hk = _runs(C, g, specs, C.SCHED_HEURISTIC, cdf_cache=cdf_cache)
    order = np.argsort(hk)
    keys = hk[order]
    X = X0[order]
    heur = hv[order]
    T = np.zeros((len(keys), C.NUM_VARIANTS), np.float32)
    M = np.zeros((len(keys), C.NUM_VARIANTS), np.float32)
    for v in range(C.NUM_VARIANTS):
        _, var, lt, k = _runs(C, g, specs, C.SCHED_FORCED, variant=v, cdf_cache=cdf_cache)
        pos = np.searchsorted(keys, k)
        ok = (pos < len(keys)) & (var == v)
        ok[ok] &= keys[pos[ok]] == k[ok]
        T[pos[ok], v] = lt[ok] - np.log2(wg[v])
        M[pos[ok], v] = 1.0
    return {"keys": keys, "X": X, "T": T, "M": M, "heur": heur}


STREAM_MODEL = os.path.join(MODEL_DIR, "cost_model_gfx950_stream.json")


def collect_stream(C, hg, seeds, fanouts, bias, partitions=16, budget_bytes=4 << 30, wg_per_cu=None,
                   log_cap=1 << 20):
    """collect() for the streaming sampler: one streamed run per schedule
    (heuristic, then each forced variant) with per-chunk logging, joined by
    chunk key (hop, partition, chunk)."""
    wg = dict(WG_PER_CU)
    if wg_per_cu:
        wg.update(wg_per_cu)
    kw = dict(bias=bias, partitions=partitions, budget_bytes=budget_bytes, log_cap=log_cap)
    X0, hv, _, hk = C.stream_sample(hg, seeds, fanouts, sched=C.SCHED_HEURISTIC, **kw)["log"]
    order = np.argsort(hk)
    keys = hk[order]
    X = X0[order]
    heur = hv[order]
    T = np.zeros((len(keys), C.NUM_VARIANTS), np.float32)
    M = np.zeros((len(keys), C.NUM_VARIANTS), np.float32)
    for v in range(C.NUM_VARIANTS):
        _, var, lt, k = C.stream_sample(hg, seeds, fanouts, sched=C.SCHED_FORCED, variant=v, **kw)["log"]
        pos = np.searchsorted(keys, k)
        ok = (pos < len(keys)) & (var == v)
        ok[ok] &= keys[pos[ok]] == k[ok]
        T[pos[ok], v] = lt[ok] - np.log2(wg[v])
        M[pos[ok], v] = 1.0
    return {"keys": keys, "X": X, "T": T, "M": M, "heur": heur}


def stream_specs(C, g, n, mix_key, lo=64, hi=65536):
    """A request mix the streaming sampler accepts: the C4 mix with FP32
    weight bias replaced by degree bias."""
    out = []
    for i, r in enumerate(C.request_mix(n, lo, hi, mix_key)):
        r = dict(r)
        if r["bias"] == C.BIAS_WEIGHT:
            r["bias"] = C.BIAS_DEGREE
        r["seeds"] = g.seeds(r["samples"], key=42, request=i)
        out.append(r)
    return out


def train(C, data, epochs=60, batch=256, lr=0.05, init_key=1, shuffle_key=7):
    m = C.Mlp(init_key)
    sel = data["M"].sum(1) > 0
    loss = m.train(data["X"][sel], data["T"][sel], data["M"][sel], epochs=epochs, batch=batch, lr=lr,
                   shuffle_key=shuffle_key)
    return m, loss


def choose(pred, M):
    """argmin of predicted cost over the variants measured valid (M)."""
    p = np.where(M > 0, pred, np.inf)
    return np.argmin(p, axis=1)


def regret(data, mlp):
    """dict(best, heuristic, mlp) total measured cost (2^T summed) over chunks
    whose heuristic choice and every valid variant were measured."""
    T, M = data["T"], data["M"]
    cost = np.where(M > 0, np.exp2(T), np.inf)
    heur = data["heur"].astype(np.int64)
    rows = np.arange(len(T))
    ok = (M.sum(1) > 1) & (M[rows, heur] > 0)
    pred = mlp.forward(data["X"])
    mc = choose(pred, M)
    best = cost[ok].min(1).sum()
    h = cost[ok, heur[ok]].sum()
    m = cost[ok, mc[ok]].sum()
    # one variant forced for every chunk (the heuristic's choice where that
    # variant is not valid, as the forced schedule runs it): the best such
    # fixed policy is the other baseline the learned one must beat
    single = []
    for v in range(M.shape[1]):
        cv = np.where(M[ok, v] > 0, cost[ok, v], cost[ok, heur[ok]])
        single.append(float(cv.sum()))
    bv = int(np.argmin(single))
    return {"chunks": int(ok.sum()), "best": float(best), "heuristic": float(h), "mlp": float(m),
            "heuristic_regret": float(h / best - 1.0), "mlp_regret": float(m / best - 1.0),
            "single_variant": single, "best_single_variant": bv,
            "best_single_regret": float(single[bv] / best - 1.0),
            "mlp_choice_mix": np.bincount(mc[ok], minlength=M.shape[1]).tolist(),
            "heuristic_choice_mix": np.bincount(heur[ok], minlength=M.shape[1]).tolist()}


def save(mlp, path=DEFAULT_MODEL, meta=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    json.dump({"params": [float(x) for x in mlp.params()], "meta": meta or {}}, open(path, "w"))


def load(C, path=DEFAULT_MODEL):
    d = json.load(open(path))
    m = C.Mlp(1)
    m.set_params(np.array(d["params"], np.float32))
    return m, d.get("meta", {})
