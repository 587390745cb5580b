// Sampling plans: the host side of Sampler() (non-stream/main.cu:196-308),
// re-designed around reusable device buffers.
//
// A plan owns every buffer a run needs (allocated once at creation, sized by
// the per-request upper bound S * F_0 * ... * F_h per hop), so csaw_plan_run
// allocates nothing and its launch sequence can be captured in a hipGraph.
// PLAIN mode executes one request after another (one unit per request, a
// wave per frontier entry, synchronising after each request like a call of
// the reference's Sampler()); BATCHED mode puts all requests into one unit:
// one launch sequence per hop over the concatenated frontiers, with degree
// binning and a per-chunk kernel-variant choice.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "graph.hpp"
#include "plan.hpp"

namespace csaw {

void EventTimer::begin(int which, hipStream_t s) {
  if (!enabled) return;
  if (used >= pool.size()) {
    Rec r;
    r.kind = which;
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
    pool.push_back(r);
  }
  pool[used].kind = which;
  (void)hipEventRecord(pool[used].a, s);
}

void EventTimer::end(int which, hipStream_t s) {
  if (!enabled || used >= pool.size()) return;
  (void)which;
  (void)hipEventRecord(pool[used].b, s);
  ++used;
}

void EventTimer::collect() {
  for (size_t i = 0; i < used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pool[i].a, pool[i].b) == hipSuccess) {
      total_ms[pool[i].kind] += ms;
      launches[pool[i].kind] += 1;
    }
  }
  used = 0;
}

EventTimer::~EventTimer() {
  for (auto& r : pool) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
}

static int64_t ub_mul(int64_t a, int64_t b) {
  if (a == 0 || b == 0) return 0;
  if (a > (int64_t)1 << 40 || b > (int64_t)1 << 40 || a > ((int64_t)1 << 62) / b) return (int64_t)1 << 62;
  return a * b;
}

}  // namespace csaw

using namespace csaw;

static int validate_request(const csaw_graph* g, const csaw_request& r) {
  if (r.samples < 0 || r.depth < 0 || r.depth > CSAW_MAX_HOPS) return CSAW_ERR_INVALID_ARG;
  if (r.bias < CSAW_BIAS_UNIFORM || r.bias > CSAW_BIAS_USER) return CSAW_ERR_INVALID_ARG;
  if (r.dedup < 0 || r.dedup > 2) return CSAW_ERR_INVALID_ARG;
  if (r.bias == CSAW_BIAS_WEIGHT && !g->w) return CSAW_ERR_INVALID_ARG;
  if (r.request_id >= (1u << 20)) return CSAW_ERR_INVALID_ARG;
  for (int h = 0; h < r.depth; ++h)
    if (r.fanouts[h] < 0 || r.fanouts[h] > CSAW_MAX_FANOUT) return CSAW_ERR_INVALID_ARG;
  if (r.samples > 0 && !r.seeds) return CSAW_ERR_INVALID_ARG;
  for (int32_t s = 0; s < r.samples; ++s)
    if (r.seeds[s] < 0 || r.seeds[s] >= g->V) return CSAW_ERR_INVALID_ARG;
  return CSAW_OK;
}

static void plan_free(csaw_plan* p) {
  if (!p) return;
  for (auto& u : p->units) {
    (void)hipFree(u.d_req);
    (void)hipFree(u.d_gs2req);
    (void)hipFree(u.d_fv0);
    (void)hipFree(u.d_fs0);
    (void)hipFree(u.d_fo0);
    for (auto* q : u.d_eo) (void)hipFree(q);
    for (auto* q : u.d_src) (void)hipFree(q);
    for (auto* q : u.d_dst) (void)hipFree(q);
    for (auto* q : u.d_smp) (void)hipFree(q);
  }
  (void)hipFree(p->local_excl);
  (void)hipFree(p->order);
  (void)hipFree(p->blk_excl);
  (void)hipFree(p->grp_excl);
  (void)hipFree(p->chunks);
  (void)hipFree(p->counters);
  (void)hipFree(p->queues);
  (void)hipFree(p->totals);
  (void)hipFree(p->chosen);
  (void)hipFree(p->log);
  (void)hipFree(p->d_mlp);
  (void)hipFree(p->bytes_acc);
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
  for (int i = 0; i < csaw_plan::kSide; ++i) {
    if (p->side[i]) (void)hipStreamDestroy(p->side[i]);
    if (p->join[i]) (void)hipEventDestroy(p->join[i]);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  delete p;
}

static int build_unit(csaw_plan* p, Unit& u) {
  // requests sorted by depth descending: the active set at hop h is a prefix
  std::stable_sort(u.reqs.begin(), u.reqs.end(),
                   [&](int a, int b) { return p->reqs[a].depth > p->reqs[b].depth; });
  u.G = 0;
  u.depth_max = 0;
  u.kinds[0] = u.kinds[1] = u.kinds[2] = false;
  u.remap = false;
  std::vector<ReqDev> rd;
  for (int ri : u.reqs) {
    const csaw_request& r = p->reqs[ri];
    ReqDev d;
    memset(&d, 0, sizeof(d));
    for (int h = 0; h < CSAW_MAX_HOPS; ++h) d.fanout[h] = h < r.depth ? r.fanouts[h] : 0;
    d.depth = r.depth;
    d.bias = r.bias;
    d.dedup = r.dedup == 2 ? 2 : (r.dedup ? 1 : 0);
    d.samples = r.samples;
    d.key_lo = (uint32_t)r.seed;
    d.key_hi = (uint32_t)(r.seed >> 32);
    d.request_id = r.request_id;
    d.gbase = u.G;
    u.gbase.push_back(u.G);
    rd.push_back(d);
    u.G += r.samples;
    u.depth_max = std::max(u.depth_max, r.depth);
    if (r.depth > 0 && r.samples > 0) u.kinds[bias_kind(r.bias)] = true;
    if (r.dedup == 2) u.remap = true;
  }
  if (u.G > 0x7fffffffLL) return CSAW_ERR_CAPACITY;
  // per hop: first inactive global sample, edge capacity, max frontier entries
  u.gb.assign(u.depth_max, 0);
  u.cap.assign(u.depth_max, 0);
  u.nmax.assign(u.depth_max, 0);
  std::vector<int64_t> ub(u.reqs.size());
  for (size_t i = 0; i < u.reqs.size(); ++i) ub[i] = p->reqs[u.reqs[i]].samples;
  for (int h = 0; h < u.depth_max; ++h) {
    int64_t gb = 0, cap = 0, nm = 0;
    for (size_t i = 0; i < u.reqs.size(); ++i) {
      const csaw_request& r = p->reqs[u.reqs[i]];
      if (r.depth > h) {
        gb = u.gbase[i] + r.samples;
        nm += ub[i];
        ub[i] = ub_mul(ub[i], r.fanouts[h]);
        cap += ub[i];
      }
    }
    u.gb[h] = (int32_t)gb;
    u.cap[h] = cap;
    u.nmax[h] = nm;
    if (cap > ((int64_t)1 << 34) || nm > ((int64_t)1 << 32) - 1) {
      set_last_error("plan exceeds the per-hop capacity (2^32 frontier entries)");
      return CSAW_ERR_CAPACITY;
    }
  }
  // device buffers
  const int64_t G = u.G;
  CSAW_HIP(hipMalloc(&u.d_req, sizeof(ReqDev) * std::max<size_t>(rd.size(), 1)));
  if (!rd.empty()) CSAW_HIP(hipMemcpy(u.d_req, rd.data(), sizeof(ReqDev) * rd.size(), hipMemcpyHostToDevice));
  std::vector<int32_t> gs2req(G), fv0(G), fs0(G);
  std::vector<int64_t> fo0(G + 1);
  for (size_t i = 0; i < u.reqs.size(); ++i) {
    const csaw_request& r = p->reqs[u.reqs[i]];
    const std::vector<int32_t>& sd = p->seeds[u.reqs[i]];
    for (int32_t s = 0; s < r.samples; ++s) {
      const int64_t g = u.gbase[i] + s;
      gs2req[g] = (int32_t)i;
      fv0[g] = sd[s];
      fs0[g] = (int32_t)g;
    }
  }
  std::iota(fo0.begin(), fo0.end(), (int64_t)0);
  CSAW_HIP(hipMalloc(&u.d_gs2req, sizeof(int32_t) * std::max<int64_t>(G, 1)));
  CSAW_HIP(hipMalloc(&u.d_fv0, sizeof(int32_t) * std::max<int64_t>(G, 1)));
  CSAW_HIP(hipMalloc(&u.d_fs0, sizeof(int32_t) * std::max<int64_t>(G, 1)));
  CSAW_HIP(hipMalloc(&u.d_fo0, sizeof(int64_t) * (G + 1)));
  if (G > 0) {
    CSAW_HIP(hipMemcpy(u.d_gs2req, gs2req.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice));
    CSAW_HIP(hipMemcpy(u.d_fv0, fv0.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice));
    CSAW_HIP(hipMemcpy(u.d_fs0, fs0.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice));
  }
  CSAW_HIP(hipMemcpy(u.d_fo0, fo0.data(), sizeof(int64_t) * (G + 1), hipMemcpyHostToDevice));
  for (int h = 0; h < u.depth_max; ++h) {
    int64_t* eo;
    int32_t *s, *d, *m;
    CSAW_HIP(hipMalloc(&eo, sizeof(int64_t) * (G + 1)));
    u.d_eo.push_back(eo);
    const int64_t c = std::max<int64_t>(u.cap[h], 1);
    CSAW_HIP(hipMalloc(&s, sizeof(int32_t) * c));
    u.d_src.push_back(s);
    CSAW_HIP(hipMalloc(&d, sizeof(int32_t) * c));
    u.d_dst.push_back(d);
    CSAW_HIP(hipMalloc(&m, sizeof(int32_t) * c));
    u.d_smp.push_back(m);
  }
  return CSAW_OK;
}

static HopDev make_hop(csaw_plan* p, Unit& u, int h) {
  HopDev H;
  H.fv = h == 0 ? u.d_fv0 : u.d_dst[h - 1];
  H.fs = h == 0 ? u.d_fs0 : u.d_smp[h - 1];
  H.fo = h == 0 ? u.d_fo0 : u.d_eo[h - 1];
  H.n_ptr = H.fo + u.gb[h];
  H.hop = h;
  H.gb = u.gb[h];
  H.local_excl = p->local_excl;
  H.blk_excl = p->blk_excl;
  H.grp_excl = p->grp_excl;
  H.order = p->order;
  H.chunks = p->chunks;
  H.n_chunks = p->counters + kCtrChunks;
  H.chunk_cap = p->chunk_cap;
  H.src = u.d_src[h];
  H.dst = u.d_dst[h];
  H.smp = u.d_smp[h];
  H.eo = u.d_eo[h];
  H.flags = p->counters + kCtrFlags;
  H.bytes = p->opts.timing ? p->bytes_acc : nullptr;
  H.expand_waves = p->L.num_cu * p->L.expand_per_cu * 4;  // 4 waves per 256-lane workgroup
  H.chunk_min = p->L.chunk_min;
  H.group_min = p->L.group_min;
  H.only_kind = 0;
  return H;
}

static int run_unit(csaw_plan* p, Unit& u, hipStream_t st, TimingHook* th) {
  const GraphDev G = p->g->view(p->opts.cdf_cache != 0);
  const csaw_graph_view hv = p->g->hook_view();
  LaunchCfg L = p->L;
  L.stream = st;
  const bool batched = p->opts.mode == CSAW_MODE_BATCHED;
  for (int h = 0; h < u.depth_max; ++h) {
    HopDev H = make_hop(p, u, h);
    // a small hop of a batched plan (its entry bound known on the host) is
    // not binned: one whole-hop kernel per bias kind, as in the plain sampler
    // (schedules that ask for particular variants, forced or random, and
    // chunk logging keep the binned path)
    const bool small = batched && u.nmax[h] <= kSmallHopEntries && !p->opts.log_chunks &&
                       (p->opts.sched == CSAW_SCHED_HEURISTIC || p->opts.sched == CSAW_SCHED_MLP);
    const bool bin = batched && !small;
    if (bin) {
      if (th) th->begin(kKernFill, st);
      CSAW_TRY(launch_fill(L, p->counters, 0u, kCtrCount - 1));  // keeps the flags word
      if (th) th->end(kKernFill, st);
    }
    if (th) th->begin(kKernCount, st);
    CSAW_TRY(launch_count(L, G, u.d_req, u.d_gs2req, H, bin, u.nmax[h]));
    if (th) th->end(kKernCount, st);
    if (th) th->begin(kKernScan, st);
    CSAW_TRY(launch_scans(L, H, p->totals + h, u.nmax[h]));
    if (th) th->end(kKernScan, st);
    if (bin) {
      ChooseArgs A;
      A.sched = p->opts.sched;
      A.variant = p->opts.variant;
      A.explore_key = p->opts.explore_key;
      A.qlen = p->counters + kCtrQlen;
      A.queues = p->queues;
      A.queue_cap = p->queue_cap;
      A.chosen = p->chosen;
      A.int_cache = p->opts.cdf_cache != 0;
      A.in_place = 0;
      A.mlp.w = p->d_mlp;
      A.mlp_w = p->d_mlp;
      if (th) th->begin(kKernChoose, st);
      CSAW_TRY(launch_choose(L, A, H, (int64_t)p->chunk_cap));
      if (th) th->end(kKernChoose, st);
      ExpandArgs E;
      E.queues = p->queues;
      E.queue_cap = p->queue_cap;
      E.qlen = p->counters + kCtrQlen;
      E.qhead = p->counters + kCtrQhead;
      E.log = p->opts.log_chunks ? p->log : nullptr;
      E.int_cache = p->opts.cdf_cache != 0;
      E.remap = u.remap ? 1 : 0;
      if (p->side[0]) {
        // the hop's expansion kernels (one per variant queue) run side by
        // side on four streams, so one kernel's last chunks overlap the
        // others' work; all join before the next hop (with timing on, each
        // kernel's events are recorded on the stream it runs on, so the
        // times are those of this concurrent schedule)
        hipStream_t ss[1 + csaw_plan::kSide] = {st};
        CSAW_HIP(hipEventRecord(p->fork, st));
        for (int i = 0; i < csaw_plan::kSide; ++i) {
          CSAW_HIP(hipStreamWaitEvent(p->side[i], p->fork, 0));
          ss[1 + i] = p->side[i];
        }
        CSAW_TRY(launch_expand_chunks(L, G, hv, u.d_req, u.d_gs2req, H, E, u.kinds, th, ss, 1 + csaw_plan::kSide));
        for (int i = 0; i < csaw_plan::kSide; ++i) {
          CSAW_HIP(hipEventRecord(p->join[i], p->side[i]));
          CSAW_HIP(hipStreamWaitEvent(st, p->join[i], 0));
        }
      } else {
        CSAW_TRY(launch_expand_chunks(L, G, hv, u.d_req, u.d_gs2req, H, E, u.kinds, th));
      }
      if (p->opts.log_chunks) {
        // keep this hop's chunks and log for the host
        p->hop_logs.emplace_back();
        p->pending_log_hops.push_back(h);
      }
    } else {
      // one launch per bias kind present (one for a plain unit, whose single
      // request has one kind); with several kinds each launch expands only
      // its own kind's entries
      int nk = 0;
      for (int k = 0; k < 3; ++k) nk += u.kinds[k] ? 1 : 0;
      for (int kind = 0; kind < 3; ++kind) {
        if (!u.kinds[kind]) continue;
        // sub-wave group size from the largest fanout of the kind's requests
        int fmax = 0;
        bool remap = false;
        for (int ri : u.reqs) {
          const csaw_request& r = p->reqs[ri];
          if (r.depth > h && bias_kind(r.bias) == kind) {
            fmax = std::max(fmax, r.fanouts[h]);
            remap |= r.dedup == 2;
          }
        }
        // region-remap requests resolve their slots with whole waves (gs = 0)
        const int gs = remap ? 0 : group_size_for(fmax, L.group_min);
        H.only_kind = nk > 1 ? 1 + kind : 0;
        if (th) th->begin(kKernPlain, st);
        CSAW_TRY(launch_expand_plain(L, kind, G, hv, u.d_req, u.d_gs2req, H, u.nmax[h], gs));
        if (th) th->end(kKernPlain, st);
      }
    }
    if (bin && p->opts.log_chunks) {
      // chunk log must be read before the next hop overwrites chunk storage
      CSAW_HIP(hipStreamSynchronize(st));
      uint32_t nch = 0;
      CSAW_HIP(hipMemcpy(&nch, p->counters + kCtrChunks, sizeof(uint32_t), hipMemcpyDeviceToHost));
      nch = std::min(nch, p->chunk_cap);
      HopLog& hl = p->hop_logs.back();
      hl.hop = h;
      hl.chunks.resize(nch);
      hl.log.resize(nch);
      if (nch) {
        CSAW_HIP(hipMemcpy(hl.chunks.data(), p->chunks, sizeof(Chunk) * nch, hipMemcpyDeviceToHost));
        CSAW_HIP(hipMemcpy(hl.log.data(), p->log, sizeof(ChunkLog) * nch, hipMemcpyDeviceToHost));
      }
    }
  }
  return CSAW_OK;
}

extern "C" int csaw_plan_create(csaw_graph* g, const csaw_request* reqs, int32_t n_req, const csaw_plan_opts* opts,
                                csaw_plan** out) {
  if (!g || !out || n_req < 0 || (n_req > 0 && !reqs)) return CSAW_ERR_INVALID_ARG;
  *out = nullptr;
  csaw_plan_opts o;
  memset(&o, 0, sizeof(o));
  if (opts) o = *opts;
  if (o.mode != CSAW_MODE_PLAIN && o.mode != CSAW_MODE_BATCHED) return CSAW_ERR_INVALID_ARG;
  if (o.sched < CSAW_SCHED_HEURISTIC || o.sched > CSAW_SCHED_RANDOM) return CSAW_ERR_INVALID_ARG;
  if (o.sched == CSAW_SCHED_FORCED && (o.variant < 0 || o.variant >= CSAW_NUM_VARIANTS)) return CSAW_ERR_INVALID_ARG;
  CSAW_TRY(validate_tuning(o.tuning));
  for (int i = 0; i < n_req; ++i) {
    int st = validate_request(g, reqs[i]);
    if (st != CSAW_OK) {
      set_last_error("invalid request " + std::to_string(i));
      return st;
    }
  }
  csaw_plan* p = new csaw_plan();
  p->g = g;
  p->opts = o;
  p->L.num_cu = g->num_cu;
  if (o.tuning) apply_tuning(p->L, o.tuning);
  auto body = [&]() -> int {
    for (int i = 0; i < n_req; ++i) {
      CSAW_TRY(graph_prepare_bias(g, reqs[i].bias));
      if (o.cdf_cache) {
        CSAW_TRY(graph_prepare_index(g, reqs[i].bias));
        CSAW_TRY(graph_prepare_vertex_records(g, reqs[i].bias));
      }
      p->reqs.push_back(reqs[i]);
      p->seeds.emplace_back(reqs[i].seeds, reqs[i].seeds + reqs[i].samples);
      p->reqs.back().seeds = nullptr;
    }
    if (o.mode == CSAW_MODE_BATCHED) {
      p->units.emplace_back();
      for (int i = 0; i < n_req; ++i) p->units.back().reqs.push_back(i);
    } else {
      for (int i = 0; i < n_req; ++i) {
        p->units.emplace_back();
        p->units.back().reqs.push_back(i);
      }
    }
    int64_t nmax = 1, depth = 1;
    for (auto& u : p->units) {
      CSAW_TRY(build_unit(p, u));
      for (int64_t v : u.nmax) nmax = std::max(nmax, v);
      depth = std::max<int64_t>(depth, u.depth_max);
    }
    p->req_unit.assign(n_req, {-1, -1});
    for (size_t ui = 0; ui < p->units.size(); ++ui)
      for (size_t k = 0; k < p->units[ui].reqs.size(); ++k) p->req_unit[p->units[ui].reqs[k]] = {(int)ui, (int)k};
    const int64_t nblk = (nmax + kBlockEntries - 1) / kBlockEntries;
    const int64_t ngrp = (nblk + 1023) / 1024;
    CSAW_HIP(hipMalloc(&p->local_excl, sizeof(uint32_t) * nmax));
    CSAW_HIP(hipMalloc(&p->blk_excl, sizeof(uint32_t) * nblk));
    CSAW_HIP(hipMalloc(&p->grp_excl, sizeof(uint64_t) * ngrp));
    CSAW_HIP(hipMalloc(&p->counters, sizeof(uint32_t) * kCtrCount));
    CSAW_HIP(hipMemset(p->counters, 0, sizeof(uint32_t) * kCtrCount));
    CSAW_HIP(hipMalloc(&p->totals, sizeof(int64_t) * depth));
    CSAW_HIP(hipMalloc(&p->bytes_acc, sizeof(unsigned long long) * kNumKernKinds));
    CSAW_HIP(hipMemset(p->bytes_acc, 0, sizeof(unsigned long long) * kNumKernKinds));
    if (o.mode == CSAW_MODE_BATCHED) {
      CSAW_HIP(hipMalloc(&p->order, sizeof(int32_t) * nblk * kBlockEntries));
      const int64_t waves = (int64_t)p->L.num_cu * p->L.expand_per_cu * 4;
      p->chunk_cap = (uint32_t)std::min<int64_t>(nblk * kChunksPerBlock + waves, 0xFFFFFFF0LL);
      CSAW_HIP(hipMalloc(&p->chunks, sizeof(Chunk) * p->chunk_cap));
      p->queue_cap = p->chunk_cap;
      CSAW_HIP(hipMalloc(&p->queues, sizeof(uint32_t) * (int64_t)kNumQueues * p->queue_cap));
      CSAW_HIP(hipMalloc(&p->chosen, p->chunk_cap));
      if (o.log_chunks) CSAW_HIP(hipMalloc(&p->log, sizeof(ChunkLog) * p->chunk_cap));
      if (o.sched == CSAW_SCHED_MLP) {
        if (!o.mlp) {
          set_last_error("CSAW_SCHED_MLP needs a trained csaw_mlp");
          return CSAW_ERR_INVALID_ARG;
        }
        CSAW_HIP(hipMalloc(&p->d_mlp, sizeof(uint16_t) * kMlpParams));
        CSAW_TRY(mlp_upload_bf16(o.mlp, p->d_mlp));
      }
    }
    CSAW_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    p->own_stream = true;
    if (o.mode == CSAW_MODE_BATCHED && !o.serial_expand) {
      CSAW_HIP(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
      for (int i = 0; i < csaw_plan::kSide; ++i) {
        CSAW_HIP(hipStreamCreateWithFlags(&p->side[i], hipStreamNonBlocking));
        CSAW_HIP(hipEventCreateWithFlags(&p->join[i], hipEventDisableTiming));
      }
    }
    p->timer.enabled = o.timing != 0;
    return CSAW_OK;
  };
  int st = body();
  if (st != CSAW_OK) {
    plan_free(p);
    return st;
  }
  *out = p;
  return CSAW_OK;
}

extern "C" int csaw_plan_run(csaw_plan* p, void* stream, int sync) {
  if (!p) return CSAW_ERR_INVALID_ARG;
  hipStream_t st = stream ? (hipStream_t)stream : p->stream;
  p->hop_logs.clear();
  p->pending_log_hops.clear();
  hipEvent_t e0, e1;
  CSAW_HIP(hipEventCreate(&e0));
  CSAW_HIP(hipEventCreate(&e1));
  CSAW_HIP(hipEventRecord(e0, st));
  TimingHook* th = p->timer.enabled ? &p->timer : nullptr;
  if (th) {
    LaunchCfg L = p->L;
    L.stream = st;
    CSAW_TRY(launch_fill(L, (uint32_t*)p->bytes_acc, 0u, 2 * kNumKernKinds));
  }
  const bool use_graph = p->opts.use_graph && p->opts.mode == CSAW_MODE_BATCHED && !th && !p->opts.log_chunks;
  if (use_graph) {
    if (!p->gexec || p->graph_stream != st) {
      if (p->gexec) {
        CSAW_HIP(hipGraphExecDestroy(p->gexec));
        p->gexec = nullptr;
      }
      hipGraph_t graph;
      CSAW_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rs = CSAW_OK;
      for (auto& u : p->units) {
        rs = run_unit(p, u, st, nullptr);
        if (rs != CSAW_OK) break;
      }
      hipError_t ce = hipStreamEndCapture(st, &graph);
      if (rs != CSAW_OK) return rs;
      CSAW_HIP(ce);
      CSAW_HIP(hipGraphInstantiate(&p->gexec, graph, nullptr, nullptr, 0));
      CSAW_HIP(hipGraphDestroy(graph));
      p->graph_stream = st;
    }
    CSAW_HIP(hipGraphLaunch(p->gexec, st));
  } else {
    for (auto& u : p->units) {
      CSAW_TRY(run_unit(p, u, st, th));
      if (p->opts.mode == CSAW_MODE_PLAIN) CSAW_HIP(hipStreamSynchronize(st));
    }
  }
  CSAW_HIP(hipEventRecord(e1, st));
  if (sync || th) {
    CSAW_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CSAW_HIP(hipEventElapsedTime(&ms, e0, e1));
    p->last_ms = ms;
    if (th) {
      p->timer.collect();
      unsigned long long b[kNumKernKinds];
      CSAW_HIP(hipMemcpy(b, p->bytes_acc, sizeof(b), hipMemcpyDeviceToHost));
      for (int k = 0; k < kNumKernKinds; ++k) p->timer.total_bytes[k] += (double)b[k];
    }
    uint32_t flags = 0;
    CSAW_HIP(hipMemcpy(&flags, p->counters + kCtrFlags, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (flags & 1u) {
      set_last_error("chunk descriptor capacity exceeded");
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return CSAW_ERR_CAPACITY;
    }
  }
  CSAW_HIP(hipEventDestroy(e0));
  CSAW_HIP(hipEventDestroy(e1));
  p->runs += 1;
  return CSAW_OK;
}

// Per-request hop boundaries: hop h edges of request r are
// [eo_h[gbase], eo_h[gbase + S]) of its unit's hop-h arrays.
static int request_bounds(const csaw_plan* p, int32_t req, std::vector<int64_t>& lo, std::vector<int64_t>& hi) {
  if (req < 0 || req >= (int32_t)p->reqs.size()) return CSAW_ERR_INVALID_ARG;
  const auto [ui, k] = p->req_unit[req];
  const Unit& u = p->units[ui];
  const csaw_request& r = p->reqs[req];
  lo.assign(r.depth, 0);
  hi.assign(r.depth, 0);
  for (int h = 0; h < r.depth; ++h) {
    CSAW_HIP(hipMemcpy(&lo[h], u.d_eo[h] + u.gbase[k], sizeof(int64_t), hipMemcpyDeviceToHost));
    CSAW_HIP(hipMemcpy(&hi[h], u.d_eo[h] + u.gbase[k] + r.samples, sizeof(int64_t), hipMemcpyDeviceToHost));
  }
  return CSAW_OK;
}

extern "C" int csaw_plan_request_info(const csaw_plan* p, int32_t req, int64_t* n_edges) {
  if (!p || !n_edges) return CSAW_ERR_INVALID_ARG;
  std::vector<int64_t> lo, hi;
  CSAW_TRY(request_bounds(p, req, lo, hi));
  int64_t n = 0;
  for (size_t h = 0; h < lo.size(); ++h) n += hi[h] - lo[h];
  *n_edges = n;
  return CSAW_OK;
}

extern "C" int csaw_plan_copy(const csaw_plan* p, int32_t req, int32_t* src, int32_t* dst, int32_t* sample,
                              int64_t* hop_offsets) {
  if (!p) return CSAW_ERR_INVALID_ARG;
  std::vector<int64_t> lo, hi;
  CSAW_TRY(request_bounds(p, req, lo, hi));
  const auto [ui, k] = p->req_unit[req];
  const Unit& u = p->units[ui];
  const csaw_request& r = p->reqs[req];
  int64_t at = 0;
  for (int h = 0; h < r.depth; ++h) {
    const int64_t n = hi[h] - lo[h];
    if (n > 0) {
      if (src) CSAW_HIP(hipMemcpy(src + at, u.d_src[h] + lo[h], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
      if (dst) CSAW_HIP(hipMemcpy(dst + at, u.d_dst[h] + lo[h], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
      if (sample) {
        CSAW_HIP(hipMemcpy(sample + at, u.d_smp[h] + lo[h], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) sample[at + i] -= (int32_t)u.gbase[k];
      }
    }
    if (hop_offsets) {
      std::vector<int64_t> eo(r.samples + 1);
      CSAW_HIP(hipMemcpy(eo.data(), u.d_eo[h] + u.gbase[k], sizeof(int64_t) * (r.samples + 1), hipMemcpyDeviceToHost));
      for (int32_t s = 0; s <= r.samples; ++s) hop_offsets[(int64_t)h * (r.samples + 1) + s] = at + eo[s] - lo[h];
    }
    at += n;
  }
  return CSAW_OK;
}

namespace csaw {
__global__ void k_sub_i32(int32_t* __restrict__ x, int64_t n, int32_t sub) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] -= sub;
}
__global__ void k_add_i64(const int64_t* __restrict__ in, int64_t n, int64_t add, int64_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i] + add;
}
}  // namespace csaw

// csaw_plan_copy into device buffers, on the device (for callers that keep
// the edge lists on the GPU, e.g. as framework tensors).
extern "C" int csaw_plan_copy_device(const csaw_plan* p, int32_t req, int32_t* src, int32_t* dst, int32_t* sample,
                                     int64_t* hop_offsets, void* stream, int sync) {
  if (!p) return CSAW_ERR_INVALID_ARG;
  std::vector<int64_t> lo, hi;
  CSAW_TRY(request_bounds(p, req, lo, hi));
  const auto [ui, k] = p->req_unit[req];
  const Unit& u = p->units[ui];
  const csaw_request& r = p->reqs[req];
  hipStream_t st = stream ? (hipStream_t)stream : p->stream;
  auto grid = [&](int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); };
  int64_t at = 0;
  for (int h = 0; h < r.depth; ++h) {
    const int64_t n = hi[h] - lo[h];
    if (n > 0) {
      if (src) CSAW_HIP(hipMemcpyAsync(src + at, u.d_src[h] + lo[h], sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
      if (dst) CSAW_HIP(hipMemcpyAsync(dst + at, u.d_dst[h] + lo[h], sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
      if (sample) {
        CSAW_HIP(hipMemcpyAsync(sample + at, u.d_smp[h] + lo[h], sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
        if (u.gbase[k]) {
          k_sub_i32<<<grid(n), 256, 0, st>>>(sample + at, n, (int32_t)u.gbase[k]);
          CSAW_LAUNCH_CHECK();
        }
      }
    }
    if (hop_offsets) {
      k_add_i64<<<grid(r.samples + 1), 256, 0, st>>>(u.d_eo[h] + u.gbase[k], r.samples + 1, at - lo[h],
                                                     hop_offsets + (int64_t)h * (r.samples + 1));
      CSAW_LAUNCH_CHECK();
    }
    at += n;
  }
  if (sync) CSAW_HIP(hipStreamSynchronize(st));
  return CSAW_OK;
}

extern "C" int csaw_plan_digest(const csaw_plan* p, int32_t req, uint64_t* digest) {
  if (!p || !digest) return CSAW_ERR_INVALID_ARG;
  std::vector<int64_t> lo, hi;
  CSAW_TRY(request_bounds(p, req, lo, hi));
  const auto [ui, k] = p->req_unit[req];
  const Unit& u = p->units[ui];
  unsigned long long* d;
  CSAW_HIP(hipMalloc(&d, sizeof(unsigned long long)));
  CSAW_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), p->stream));
  int64_t at = 0;
  int st = CSAW_OK;
  for (size_t h = 0; h < lo.size() && st == CSAW_OK; ++h) {
    st = launch_digest(p->stream, p->g->num_cu, u.d_src[h], u.d_dst[h], u.d_smp[h], lo[h], hi[h], at, u.gbase[k], d);
    at += hi[h] - lo[h];
  }
  unsigned long long v = 0;
  if (st == CSAW_OK) {
    CSAW_HIP(hipMemcpyAsync(&v, d, sizeof(v), hipMemcpyDeviceToHost, p->stream));
    CSAW_HIP(hipStreamSynchronize(p->stream));
  }
  (void)hipFree(d);
  *digest = (uint64_t)v ^ ((uint64_t)at * 0x100000001b3ull);
  return st;
}

extern "C" int csaw_plan_stats(const csaw_plan* p, csaw_stats* s) {
  if (!p || !s) return CSAW_ERR_INVALID_ARG;
  memset(s, 0, sizeof(*s));
  int64_t edges = 0;
  for (int32_t r = 0; r < (int32_t)p->reqs.size(); ++r) {
    int64_t n = 0;
    CSAW_TRY(csaw_plan_request_info(p, r, &n));
    edges += n;
  }
  s->edges = edges;
  s->kernel_ms = p->last_ms;
  if (p->opts.mode == CSAW_MODE_BATCHED && p->chosen) {
    // variant mix of the last hop (chunk storage is per hop)
    uint32_t nch = 0;
    CSAW_HIP(hipMemcpy(&nch, p->counters + kCtrChunks, sizeof(uint32_t), hipMemcpyDeviceToHost));
    (void)nch;
  }
  return CSAW_OK;
}

extern "C" int csaw_plan_kernel_times(const csaw_plan* p, char* names, size_t names_len, double* avg_ms,
                                      int64_t* launches, double* avg_bytes, int32_t max_kinds, int32_t* n_kinds) {
  if (!p) return CSAW_ERR_INVALID_ARG;
  int32_t n = 0;
  size_t pos = 0;
  for (int k = 0; k < kNumKernKinds; ++k) {
    if (p->timer.launches[k] == 0) continue;
    if (n < max_kinds) {
      if (avg_ms) avg_ms[n] = p->timer.total_ms[k] / (double)p->timer.launches[k];
      if (launches) launches[n] = p->timer.launches[k];
      if (avg_bytes) avg_bytes[n] = p->timer.total_bytes[k] / (double)p->timer.launches[k];
      if (names) {
        size_t l = strlen(kKernNames[k]) + 1;
        if (pos + l <= names_len) {
          memcpy(names + pos, kKernNames[k], l);
          pos += l;
        }
      }
    }
    ++n;
  }
  if (n_kinds) *n_kinds = n;
  return CSAW_OK;
}

extern "C" int csaw_plan_reset_times(csaw_plan* p) {
  if (!p) return CSAW_ERR_INVALID_ARG;
  for (int k = 0; k < kNumKernKinds; ++k) {
    p->timer.total_ms[k] = 0;
    p->timer.launches[k] = 0;
    p->timer.total_bytes[k] = 0;
  }
  return CSAW_OK;
}

extern "C" int csaw_plan_destroy(csaw_plan* p) {
  plan_free(p);
  return CSAW_OK;
}

// ------------------------------------------------------------ csaw_sample
struct csaw_edges {
  int32_t depth = 0, samples = 0;
  std::vector<int32_t> src, dst, smp;
  std::vector<int64_t> hop_off;
};

extern "C" int csaw_sample(csaw_graph* g, const csaw_params* prm, const int32_t* seeds, csaw_edges** out,
                           csaw_stats* stats) {
  if (!g || !prm || !out) return CSAW_ERR_INVALID_ARG;
  *out = nullptr;
  if (prm->frontier_size > 1) {
    // FrontierSize > 1 runs the multi-dimensional random walk, as Sampler()
    // dispatches check_layer (main.cu:285-287); the result is one layer.
    if (prm->samples < 0 || prm->depth < 0 || (prm->samples > 0 && !seeds)) return CSAW_ERR_INVALID_ARG;
    csaw_edges* e = new csaw_edges();
    e->depth = 1;
    e->samples = prm->samples;
    const size_t cap = std::max<size_t>((size_t)prm->samples * prm->depth, 1);
    e->src.resize(cap);
    e->dst.resize(cap);
    e->hop_off.resize((size_t)prm->samples + 1);
    csaw_stats st;
    int rc = csaw_mdrw(g, prm->seed, prm->request_id, prm->samples, prm->frontier_size, prm->depth, seeds,
                       e->src.data(), e->dst.data(), e->hop_off.data(), &st);
    if (rc != CSAW_OK) {
      delete e;
      return rc;
    }
    const int64_t n = e->hop_off[prm->samples];
    e->src.resize(n);
    e->dst.resize(n);
    e->smp.resize(n);
    for (int32_t s = 0; s < prm->samples; ++s)
      for (int64_t i = e->hop_off[s]; i < e->hop_off[s + 1]; ++i) e->smp[i] = s;
    if (stats) *stats = st;
    *out = e;
    return CSAW_OK;
  }
  if (prm->frontier_size != 1) return CSAW_ERR_INVALID_ARG;
  if (prm->depth < 0 || prm->depth > CSAW_MAX_HOPS) return CSAW_ERR_INVALID_ARG;
  csaw_request r;
  memset(&r, 0, sizeof(r));
  r.samples = prm->samples;
  r.depth = prm->depth;
  for (int h = 0; h < prm->depth; ++h) r.fanouts[h] = prm->fanouts ? prm->fanouts[h] : prm->neighbor_size;
  r.bias = prm->bias;
  r.dedup = prm->dedup;
  r.seed = prm->seed;
  r.request_id = prm->request_id;
  r.seeds = seeds;
  csaw_plan_opts o;
  memset(&o, 0, sizeof(o));
  const bool binned = prm->sched != CSAW_SCHED_HEURISTIC || prm->variant != 0;
  o.mode = binned ? CSAW_MODE_BATCHED : CSAW_MODE_PLAIN;
  o.sched = prm->sched == CSAW_SCHED_MLP ? CSAW_SCHED_HEURISTIC : prm->sched;
  o.variant = prm->variant;
  csaw_tuning t;
  memset(&t, 0, sizeof(t));
  if (prm->n_blocks > 0) {
    t.grid_override = prm->n_blocks;
    o.tuning = &t;
  }
  csaw_plan* p = nullptr;
  CSAW_TRY(csaw_plan_create(g, &r, 1, &o, &p));
  int st = csaw_plan_run(p, nullptr, 1);
  csaw_edges* e = nullptr;
  if (st == CSAW_OK) {
    e = new csaw_edges();
    e->depth = r.depth;
    e->samples = r.samples;
    int64_t n = 0;
    st = csaw_plan_request_info(p, 0, &n);
    if (st == CSAW_OK) {
      e->src.resize(n);
      e->dst.resize(n);
      e->smp.resize(n);
      e->hop_off.resize((size_t)r.depth * (r.samples + 1));
      st = csaw_plan_copy(p, 0, e->src.data(), e->dst.data(), e->smp.data(), e->hop_off.data());
    }
    if (st == CSAW_OK && stats) {
      memset(stats, 0, sizeof(*stats));
      stats->edges = n;
      stats->kernel_ms = p->last_ms;
    }
  }
  csaw_plan_destroy(p);
  if (st != CSAW_OK) {
    delete e;
    return st;
  }
  *out = e;
  return CSAW_OK;
}

extern "C" int csaw_edges_info(const csaw_edges* e, int64_t* n_edges, int32_t* depth, int32_t* samples) {
  if (!e) return CSAW_ERR_INVALID_ARG;
  if (n_edges) *n_edges = (int64_t)e->src.size();
  if (depth) *depth = e->depth;
  if (samples) *samples = e->samples;
  return CSAW_OK;
}

extern "C" int csaw_edges_copy(const csaw_edges* e, int32_t* src, int32_t* dst, int32_t* sample, int64_t* hop_offsets) {
  if (!e) return CSAW_ERR_INVALID_ARG;
  if (src && !e->src.empty()) memcpy(src, e->src.data(), sizeof(int32_t) * e->src.size());
  if (dst && !e->dst.empty()) memcpy(dst, e->dst.data(), sizeof(int32_t) * e->dst.size());
  if (sample && !e->smp.empty()) memcpy(sample, e->smp.data(), sizeof(int32_t) * e->smp.size());
  if (hop_offsets && !e->hop_off.empty()) memcpy(hop_offsets, e->hop_off.data(), sizeof(int64_t) * e->hop_off.size());
  return CSAW_OK;
}

extern "C" void csaw_edges_free(csaw_edges* e) { delete e; }

extern "C" int csaw_edges_from_host(int32_t depth, int32_t samples, int64_t n, const int32_t* src, const int32_t* dst,
                                    const int32_t* smp, const int64_t* hop_off, csaw_edges** out) {
  if (!out || n < 0 || depth < 0 || samples < 0 || (n > 0 && (!src || !dst || !smp))) return CSAW_ERR_INVALID_ARG;
  csaw_edges* e = new csaw_edges();
  e->depth = depth;
  e->samples = samples;
  e->src.assign(src, src + n);
  e->dst.assign(dst, dst + n);
  e->smp.assign(smp, smp + n);
  if (hop_off) e->hop_off.assign(hop_off, hop_off + (size_t)depth * (samples + 1));
  *out = e;
  return CSAW_OK;
}

// C4 request mix (all draws from Philox with the mix key, domain kDomMix).
extern "C" int64_t csaw_request_mix(uint64_t mix_key, int32_t n, int32_t lo, int32_t hi, csaw_request* out) {
  if (n < 0 || lo < 1 || hi < lo || (n > 0 && !out)) return -1;
  int64_t total = 0;
  const double l0 = std::log2((double)lo), l1 = std::log2((double)hi);
  for (int32_t i = 0; i < n; ++i) {
    const U4 o = philox4x32_10((uint32_t)i, 0u, 0u, kDomMix << 28, (uint32_t)mix_key, (uint32_t)(mix_key >> 32));
    const double u = (double)(o.x >> 8) * (1.0 / 16777216.0);
    int64_t s = (int64_t)std::floor(std::exp2(l0 + u * (l1 - l0)));
    s = std::min<int64_t>(std::max<int64_t>(s, lo), hi);
    csaw_request& r = out[i];
    memset(&r, 0, sizeof(r));
    r.samples = (int32_t)s;
    if (o.y & 1u) {
      r.depth = 3;
      r.fanouts[0] = 15;
      r.fanouts[1] = 10;
      r.fanouts[2] = 5;
    } else {
      r.depth = 2;
      r.fanouts[0] = 25;
      r.fanouts[1] = 10;
    }
    r.bias = (int32_t)(o.z % 3u);  // uniform, degree, weight
    r.dedup = 1;
    r.seed = 42;
    r.request_id = (uint32_t)i;
    r.seeds = nullptr;
    total += s;
  }
  return total;
}
