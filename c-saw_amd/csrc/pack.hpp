// Packed neighbour ids for streamed slice copies (host encoder; the device
// decoder is k_unpack in stream.hip).
//
// Block b holds entries [64 b, min(64 b + 64, E)) of the CSR index array as:
// width w (1 byte), the first id (4 bytes, little endian), then the n - 1
// differences to the previous entry, zigzag-coded, w bytes each (little
// endian), w = bytes of the largest (1..4).  Sorted neighbour lists make the
// differences small inside a list; a list start inside a block only widens
// that block.  The device decodes a block with one wave: lane i reads its
// difference and a prefix sum restores the ids.
#pragma once
#include <stdint.h>
#include <string.h>

namespace csaw {

constexpr int kPackBlock = 64;
constexpr int64_t kPackMaxBlockBytes = 1 + 4 + (kPackBlock - 1) * 4;

inline uint32_t zigzag(int32_t d) { return ((uint32_t)d << 1) ^ (uint32_t)(d >> 31); }
inline int bytes_for(uint32_t z) { return z < (1u << 8) ? 1 : (z < (1u << 16) ? 2 : (z < (1u << 24) ? 3 : 4)); }

// Packs x[0 .. n) (1 <= n <= 64) into out (null: size only); returns bytes.
inline int64_t pack_block(const int32_t* x, int n, uint8_t* out) {
  uint32_t mz = 0;
  for (int i = 1; i < n; ++i) {
    const uint32_t z = zigzag((int32_t)((uint32_t)x[i] - (uint32_t)x[i - 1]));
    mz = mz > z ? mz : z;
  }
  const int w = bytes_for(mz);
  if (out) {
    out[0] = (uint8_t)w;
    memcpy(out + 1, &x[0], 4);
    uint8_t* q = out + 5;
    for (int i = 1; i < n; ++i) {
      const uint32_t z = zigzag((int32_t)((uint32_t)x[i] - (uint32_t)x[i - 1]));
      for (int k = 0; k < w; ++k) *q++ = (uint8_t)(z >> (8 * k));
    }
  }
  return 5 + (int64_t)(n - 1) * w;
}

// Host decoder of one block (the format's reference reading).
inline void unpack_block(const uint8_t* blk, int n, int32_t* out) {
  const int w = blk[0];
  uint32_t x;
  memcpy(&x, blk + 1, 4);
  out[0] = (int32_t)x;
  for (int i = 1; i < n; ++i) {
    uint32_t z = 0;
    for (int k = 0; k < w; ++k) z |= (uint32_t)blk[5 + (i - 1) * w + k] << (8 * k);
    x += (z >> 1) ^ (0u - (z & 1u));
    out[i] = (int32_t)x;
  }
}

}  // namespace csaw
