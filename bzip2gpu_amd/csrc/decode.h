// decode.h - host side of the device bzip2 decoder (.bz2 stream -> raw
// bytes).  The reference lists decompression as not built
// (SebastianBoteroLeonik/bzip2gpu TODO.md:28-34, SOURCES.md:5-6); the
// wire format is the one its stream driver writes
// (src/compression/compress.cu:184-375) and stock bzip2 1.0.8 reads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace bzg {

// Largest RLE1 length a block may decode to: the RLE1 output of a 900000-byte
// chunk (every 4 raw bytes can become 5).  Stock bzip2 stops at 100000*level;
// this decoder also takes the longer blocks the reference's fixed
// 900000-byte cut can produce at any level.
constexpr int32_t kDecMaxM = 5 * (900000 / 4);
// Symbols per tile of the inverse MTF passes.
constexpr int32_t kDecSymTile = 4096;
// Rows between list-ranking splitters of the inverse BWT.
constexpr int32_t kDecSplit = 256;
constexpr int32_t kDecMaxNodes = kDecMaxM / kDecSplit + 2;
// Bytes a walker keeps of its stretch (the rest, for the few longer
// stretches, is walked again after ranking).
constexpr int32_t kDecStage = 1024;
// Selectors kept per block (enough for kDecMaxM symbols).
constexpr int32_t kDecSelCap = (kDecMaxM + 49) / 50 + 32;

// One block-magic or end-of-stream-magic occurrence, and what decoding a
// block from it gave.
struct DecCand {
  int64_t bit;      // bit offset of the 48-bit magic
  int32_t kind;     // 0 block header magic, 1 end-of-stream magic
  int32_t sym_cap;  // symbol slots reserved for it
  int64_t sym_off;  // offset of those slots in the symbol arena
};

enum DecStatus : int32_t {
  kDecOk = 0,
  kDecOverflow = 1,    // more symbols than reserved slots (decode again with more)
  kDecBad = 2,         // malformed header or code
  kDecRandomised = 3,  // randomised block (not produced since bzip2 0.9.5)
  kDecTruncated = 4,   // ran past the end of the input
};

struct DecRes {
  int64_t end_bit;     // first bit after the block (after the EOS CRC)
  int32_t s;           // symbols before the end-of-block symbol
  int32_t status;      // DecStatus
  uint32_t crc;        // stored block CRC (block) or combined CRC (EOS)
  int32_t orig_ptr;
  int32_t n_in_use;
  uint32_t next4;      // EOS: the 4 bytes at the next byte boundary
  int64_t rest;        // EOS: bytes left after the byte boundary
  uint8_t seq[256];    // present bytes, ascending (the MTF start list)
};

// One block on the chain, in output order.
struct DecBlk {
  int64_t sym_off;     // symbols (u16) and MTF slots (u8) in their arenas
  int64_t m_off;       // offset in the L / T / FL-map arrays
  int64_t raw_off;     // offset of the raw bytes in the output
  int32_t s, orig_ptr, n_in_use, cand;
  int32_t st0, nst;    // symbol tiles
  int32_t mt0, nmt;    // 4096-element tiles over m
  int32_t node0, nnode;
  int32_t m, n;        // RLE1 and raw lengths (host copies)
  uint32_t crc_want;
  int32_t pad_;
};

// Device-written per-block results.
struct DecBlkRes {
  int32_t m;
  int32_t err;
  int64_t n;
  uint32_t crc;
  int32_t cyc;  // length of the FL-map cycle through origPtr (m unless periodic)
};

struct DecodeTimes {
  float scan = 0, huff = 0, mtf = 0, sort = 0, walk = 0, rle1 = 0;
  int32_t blocks = 0, streams = 0;
};

// Return value of Decoder::run for malformed or corrupt input (the message is
// in Decoder::error()); other nonzero values are hipError_t codes.
constexpr int kDecDataError = -1;

class Decoder {
 public:
  explicit Decoder(int device = 0);
  ~Decoder();
  Decoder(const Decoder&) = delete;
  Decoder& operator=(const Decoder&) = delete;

  // Decodes the concatenated .bz2 streams in d_in[0, len) (device memory) on
  // stream s.  Afterwards out() holds out_len() raw bytes (device memory
  // owned by the decoder, valid until the next run).
  int run(const uint8_t* d_in, size_t len, hipStream_t s, DecodeTimes* t = nullptr);
  uint8_t* out() const { return (uint8_t*)out_.p; }
  size_t out_len() const { return out_len_; }
  const std::string& error() const { return err_; }

 private:
  struct Buf {
    void* p = nullptr;
    size_t cap = 0;
  };
  int need(Buf& b, size_t bytes);
  int fail(const std::string& msg);
  int decode_cands(const std::vector<int32_t>& idx, hipStream_t s);

  int device_;
  std::string err_;
  size_t out_len_ = 0;
  int64_t nbits_ = 0;
  Buf w_, cand_, res_, sym_, slot_, idx_, small_, blk_, bres_, stblk_, f_, p_, tlen_, toff_;
  Buf sel_;  // per candidate: raw selectors, then their MTF-decoded tables (2 x kDecSelCap bytes)
  Buf l_, tt_, mtblk_, hist_, cb_, nodeblk_, nx_, ln_, pos_, t_, tf_, tent_, too_, treg_, out_, stg_, rrow_;
  int32_t* h_small_ = nullptr;  // pinned
  std::vector<DecCand> cands_;
  std::vector<DecRes> res_h_;
  hipEvent_t ev_[7] = {};
  int64_t sym_tot_ = 0;
};

}  // namespace bzg
