// Launch-shape autotuner and the per-chunk cost-model log.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "plan.hpp"

namespace csaw {

void apply_tuning(LaunchCfg& L, const csaw_tuning* t) {
  if (!t) return;
  if (t->count_per_cu > 0) L.count_per_cu = t->count_per_cu;
  if (t->expand_per_cu > 0) L.expand_per_cu = t->expand_per_cu;
  if (t->thread_per_cu > 0) L.thread_per_cu = t->thread_per_cu;
  if (t->block_per_cu > 0) L.block_per_cu = t->block_per_cu;
  if (t->grid_override > 0) L.grid_override = t->grid_override;
  if (t->chunk_min > 0) L.chunk_min = t->chunk_min;
  if (t->group_min > 0) L.group_min = t->group_min;
}

int validate_tuning(const csaw_tuning* t) {
  if (!t) return CSAW_OK;
  const int32_t cu[] = {t->count_per_cu, t->expand_per_cu, t->thread_per_cu, t->block_per_cu};
  for (int32_t v : cu)
    if (v < 0 || v > 64) return CSAW_ERR_INVALID_ARG;
  if (t->grid_override < 0) return CSAW_ERR_INVALID_ARG;
  const int32_t c = t->chunk_min;
  if (c != 0 && (c < kChunkMin || c > kBlockEntries || (c & (c - 1)) != 0)) return CSAW_ERR_INVALID_ARG;
  if (t->group_min != 0 && (t->group_min < kGroupMin || t->group_min > 32)) return CSAW_ERR_INVALID_ARG;
  return CSAW_OK;
}

}  // namespace csaw

using namespace csaw;

extern "C" int csaw_plan_chunk_log(const csaw_plan* p, float* features, int32_t* variant, float* log2_ticks,
                                   uint64_t* keys, int64_t max_n, int64_t* n) {
  if (!p || !n) return CSAW_ERR_INVALID_ARG;
  int64_t k = 0;
  for (const HopLog& hl : p->hop_logs) {
    const uint32_t nch = (uint32_t)hl.chunks.size();
    for (uint32_t i = 0; i < nch; ++i) {
      const Chunk& c = hl.chunks[i];
      if (c.bucket == kBucketCopy) continue;
      if (k < max_n) {
        const int kind = bucket_kind(c.bucket);
        if (features) chunk_features(c.len, c.sum_d, c.max_d, c.sum_f, c.max_f, kind, c.hop, c.bucket, nch,
                                     features + k * kMlpIn);
        if (variant) variant[k] = hl.log[i].variant;
        if (log2_ticks) log2_ticks[k] = std::log2((float)std::max<uint32_t>(hl.log[i].ticks, 1u));
        if (keys) keys[k] = ((uint64_t)hl.hop << 32) | c.key;
      }
      ++k;
    }
  }
  *n = k;
  return CSAW_OK;
}

// Tries workgroups-per-CU settings for each kernel family, then the tile
// shapes (smallest chunk, smallest sub-wave group), one knob at a time
// (coordinate descent), timing full runs of the plan with hipEvents.
extern "C" int csaw_autotune(csaw_graph* g, const csaw_request* reqs, int32_t n_req, const csaw_plan_opts* base,
                             csaw_tuning* best, double* best_ms) {
  if (!g || !best) return CSAW_ERR_INVALID_ARG;
  csaw_plan_opts o;
  memset(&o, 0, sizeof(o));
  if (base) o = *base;
  o.timing = 0;
  o.log_chunks = 0;
  csaw_tuning cur;
  memset(&cur, 0, sizeof(cur));
  LaunchCfg def;
  cur.count_per_cu = def.count_per_cu;
  cur.expand_per_cu = def.expand_per_cu;
  cur.thread_per_cu = def.thread_per_cu;
  cur.block_per_cu = def.block_per_cu;
  cur.chunk_min = def.chunk_min;
  cur.group_min = def.group_min;
  auto measure = [&](const csaw_tuning& t, double* ms) -> int {
    o.tuning = &t;
    csaw_plan* p = nullptr;
    CSAW_TRY(csaw_plan_create(g, reqs, n_req, &o, &p));
    int st = csaw_plan_run(p, nullptr, 1);  // warm-up (and graph capture)
    double acc = 0;
    const int reps = 3;
    for (int r = 0; r < reps && st == CSAW_OK; ++r) {
      st = csaw_plan_run(p, nullptr, 1);
      acc += p->last_ms;
    }
    csaw_plan_destroy(p);
    *ms = acc / reps;
    return st;
  };
  double bm = 0;
  CSAW_TRY(measure(cur, &bm));
  const std::vector<int> per_cu = {1, 2, 4, 8, 16};
  const std::vector<int> chunk = {8, 16, 32, 64, 128, 256};
  const std::vector<int> group = {4, 8, 16};
  const std::pair<int32_t*, const std::vector<int>*> knobs[] = {
      {&cur.expand_per_cu, &per_cu}, {&cur.thread_per_cu, &per_cu}, {&cur.block_per_cu, &per_cu},
      {&cur.count_per_cu, &per_cu},  {&cur.chunk_min, &chunk},       {&cur.group_min, &group}};
  for (const auto& kn : knobs) {
    int32_t* f = kn.first;
    const int32_t keep = *f;
    int32_t bv = keep;
    for (int c : *kn.second) {
      if (c == keep) continue;
      *f = c;
      double ms = 0;
      CSAW_TRY(measure(cur, &ms));
      if (ms < bm) {
        bm = ms;
        bv = c;
      }
    }
    *f = bv;
  }
  *best = cur;
  if (best_ms) *best_ms = bm;
  return CSAW_OK;
}
