// On-disk CSR files and the SNAP edge-list converter (host only).
//
// The reference reads a CSR from two files (non-stream/graph.hpp:18-74):
// beg.bin holds V+1 int64 offsets and csr.bin E int64 neighbour ids, both
// narrowed to 32 bits after loading.  csaw_graph_load_bin (graph.hip) is the
// reader; this file adds the writer and a converter from SNAP text edge
// lists (the dataset source named in README.md:14-18).  The converter the
// reference's README points to is an external project that is not part of
// the reference, so the rules below are this library's own, chosen to give
// the canonical graph every sampler here assumes (sorted, unique neighbour
// lists without self loops):
//  - a line whose first non-blank character is '#' or '%' is a comment, a
//    blank line is skipped;
//  - every other line starts with two non-negative decimal ids (source,
//    destination) separated by blanks, tabs or one comma; further columns
//    (weights, timestamps) are ignored;
//  - ids are kept as they are (no relabelling), V = largest id + 1;
//  - with symmetrize every edge (u, v) also adds (v, u);
//  - self loops and repeated edges are dropped, each neighbour list is sorted.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

using namespace csaw;

namespace {

int write_i64(const char* path, const std::vector<int64_t>& v) {
  FILE* f = fopen(path, "wb");
  if (!f) {
    set_last_error(std::string("cannot create ") + path);
    return CSAW_ERR_IO;
  }
  const size_t put = v.empty() ? 0 : fwrite(v.data(), sizeof(int64_t), v.size(), f);
  const int rc = fclose(f);
  if (put != v.size() || rc != 0) {
    set_last_error(std::string("short write of ") + path);
    return CSAW_ERR_IO;
  }
  return CSAW_OK;
}

// Parses one unsigned decimal id at p (after blanks); returns the position
// after it or nullptr when there is none or it exceeds the id range.
const char* parse_id(const char* p, const char* end, int64_t* out) {
  while (p < end && (*p == ' ' || *p == '\t')) ++p;
  if (p >= end || *p < '0' || *p > '9') return nullptr;
  int64_t v = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    v = v * 10 + (*p - '0');
    if (v > 0x7ffffffeLL) return nullptr;
    ++p;
  }
  *out = v;
  return p;
}

}  // namespace

extern "C" int csaw_csr_write_bin(const char* beg_file, const char* csr_file, const int64_t* offsets,
                                  const int32_t* indices, int64_t V, int64_t E) {
  if (!beg_file || !csr_file || !offsets || V < 0 || E < 0 || (E > 0 && !indices)) return CSAW_ERR_INVALID_ARG;
  if (offsets[0] != 0 || offsets[V] != E) return CSAW_ERR_INVALID_GRAPH;
  for (int64_t v = 0; v < V; ++v)
    if (offsets[v + 1] < offsets[v]) return CSAW_ERR_INVALID_GRAPH;
  std::vector<int64_t> beg(offsets, offsets + V + 1), ids(E);
  for (int64_t i = 0; i < E; ++i) {
    if (indices[i] < 0 || indices[i] >= V) return CSAW_ERR_INVALID_GRAPH;
    ids[i] = indices[i];
  }
  CSAW_TRY(write_i64(beg_file, beg));
  return write_i64(csr_file, ids);
}

extern "C" int csaw_snap_to_bin(const char* snap_file, int symmetrize, const char* beg_file, const char* csr_file,
                                int64_t* V_out, int64_t* E_out) {
  if (!snap_file || !beg_file || !csr_file) return CSAW_ERR_INVALID_ARG;
  FILE* f = fopen(snap_file, "rb");
  if (!f) {
    set_last_error(std::string("cannot open ") + snap_file);
    return CSAW_ERR_IO;
  }
  std::vector<uint64_t> keys;  // src << 32 | dst
  int64_t max_id = -1, line_no = 0;
  std::vector<char> buf(1 << 20);
  std::string carry;
  int rc = CSAW_OK;
  auto take_line = [&](const char* p, const char* end) -> int {
    ++line_no;
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    if (p >= end || *p == '#' || *p == '%') return CSAW_OK;
    int64_t u = 0, v = 0;
    const char* q = parse_id(p, end, &u);
    if (q) {
      while (q < end && (*q == ' ' || *q == '\t')) ++q;
      if (q < end && *q == ',') ++q;
      q = parse_id(q, end, &v);
    }
    if (!q || (q < end && *q != ' ' && *q != '\t' && *q != ',' && *q != '\r')) {
      set_last_error("line " + std::to_string(line_no) + " of " + snap_file + " is not an edge 'src dst'");
      return CSAW_ERR_INVALID_GRAPH;
    }
    max_id = std::max(max_id, std::max(u, v));
    if (u != v) {
      keys.push_back((uint64_t)u << 32 | (uint64_t)v);
      if (symmetrize) keys.push_back((uint64_t)v << 32 | (uint64_t)u);
    }
    return CSAW_OK;
  };
  for (;;) {
    const size_t got = fread(buf.data(), 1, buf.size(), f);
    if (got == 0) break;
    const char* p = buf.data();
    const char* end = p + got;
    while (rc == CSAW_OK && p < end) {
      const char* nl = (const char*)memchr(p, '\n', (size_t)(end - p));
      if (!nl) {
        carry.append(p, end);
        break;
      }
      if (!carry.empty()) {
        carry.append(p, nl);
        rc = take_line(carry.data(), carry.data() + carry.size());
        carry.clear();
      } else {
        rc = take_line(p, nl);
      }
      p = nl + 1;
    }
    if (rc != CSAW_OK) break;
  }
  if (rc == CSAW_OK && ferror(f)) {
    set_last_error(std::string("read error on ") + snap_file);
    rc = CSAW_ERR_IO;
  }
  if (rc == CSAW_OK && !carry.empty()) rc = take_line(carry.data(), carry.data() + carry.size());
  fclose(f);
  if (rc != CSAW_OK) return rc;
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const int64_t V = max_id + 1, E = (int64_t)keys.size();
  std::vector<int64_t> beg(V + 1, 0), ids(E);
  for (int64_t i = 0; i < E; ++i) {
    beg[(keys[i] >> 32) + 1] += 1;
    ids[i] = (int64_t)(keys[i] & 0xffffffffu);
  }
  for (int64_t v = 0; v < V; ++v) beg[v + 1] += beg[v];
  CSAW_TRY(write_i64(beg_file, beg));
  CSAW_TRY(write_i64(csr_file, ids));
  if (V_out) *V_out = V;
  if (E_out) *E_out = E;
  return CSAW_OK;
}
