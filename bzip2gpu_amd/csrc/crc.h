// crc.h - device-side CRC-32/BZIP2 register arithmetic shared by the
// compression front end and the decoder.
//
// The register of a byte string B started from zero, reg0(B), is linear:
// reg0(A||B) = reg0(A) * x^(8|B|) + reg0(B) mod P, and the CRC with the
// all-ones start is ~(reg0(D) + 0xFFFFFFFF * x^(8|D|)).  Each translation
// unit that includes this header gets its own copy of the constant tables.
#pragma once

#include "common.h"

namespace bzg {

struct Pow8Table {
  uint32_t p[40];
  constexpr Pow8Table() : p() {
    uint32_t b = 0x100u;  // x^8
    for (int k = 0; k < 40; k++) {
      p[k] = b;
      b = gf2_mulmod(b, b);
    }
  }
};

// x^(8r) for byte shifts: x^(8q), x^(128m), x^(32768u) with q < 16, m, u < 256
struct ShiftTables {
  uint32_t q1[16];
  uint32_t q16[256];
  uint32_t tile[256];
  constexpr ShiftTables() : q1(), q16(), tile() {
    uint32_t a = 1u;
    for (int q = 0; q < 16; q++) { q1[q] = a; a = gf2_mulmod(a, 0x100u); }
    uint32_t b = 1u;
    for (int m = 0; m < 256; m++) { q16[m] = b; b = gf2_mulmod(b, a); }
    uint32_t c = 1u;
    for (int u = 0; u < 256; u++) { tile[u] = c; c = gf2_mulmod(c, b); }
  }
};
static_assert(kTile == 4096, "shift tables assume 4096-byte tiles");

// Slicing-by-4 tables: s[k][b] is the register of byte b followed by k zero
// bytes, so four bytes b0..b3 advance a register r in one step:
// v = r ^ (b0<<24 | b1<<16 | b2<<8 | b3),
// r' = s[3][v>>24] ^ s[2][(v>>16)&255] ^ s[1][(v>>8)&255] ^ s[0][v&255].
struct CrcSlice4 {
  uint32_t s[4][256];
  constexpr CrcSlice4() : s() {
    const CrcTable t;
    for (int b = 0; b < 256; b++) s[0][b] = t.t[b];
    for (int k = 1; k < 4; k++)
      for (int b = 0; b < 256; b++) s[k][b] = (s[k - 1][b] << 8) ^ t.t[s[k - 1][b] >> 24];
  }
};

static __constant__ CrcTable c_crc = CrcTable();
static __constant__ CrcSlice4 c_crc4 = CrcSlice4();
static __constant__ Pow8Table c_pow8 = Pow8Table();
static __constant__ ShiftTables c_shift = ShiftTables();

// x^(8*len) mod P from the constant table
__device__ inline uint32_t xpow8(uint32_t len) {
  uint32_t r = 1u;
  int k = 0;
  while (len) {
    if (len & 1u) r = gf2_mulmod(r, c_pow8.p[k]);
    len >>= 1;
    k++;
  }
  return r;
}

// reg * x^(8r) mod P
__device__ inline uint32_t crc_shift(uint32_t reg, uint32_t r) {
  if (r >= 256u * kTile) return gf2_mulmod(reg, xpow8(r));
  if (r >> 12) reg = gf2_mulmod(reg, c_shift.tile[r >> 12]);
  if ((r >> 4) & 255u) reg = gf2_mulmod(reg, c_shift.q16[(r >> 4) & 255u]);
  if (r & 15u) reg = gf2_mulmod(reg, c_shift.q1[r & 15u]);
  return reg;
}

// CRC-32/BZIP2 of n bytes whose zero-start register is reg0
__device__ inline uint32_t crc_final(uint32_t reg0, uint32_t n) {
  return ~(gf2_mulmod(0xFFFFFFFFu, xpow8(n)) ^ reg0);
}

}  // namespace bzg
