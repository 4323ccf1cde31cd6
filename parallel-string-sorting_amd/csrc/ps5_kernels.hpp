// ps5_kernels.hpp -- shared definitions of the gfx950 pS5 kernels.
//
// Device data layout (one sort, n strings, positions 0..n-1):
//   arena      characters as given by the caller (never written)
//   offs[n]    uint32 start offset of input string i
//   key[2][n]  uint64 ping-pong: 8 characters at the current depth,
//              big-endian, zero from the first NUL on
//   idx[2][n]  uint32 ping-pong: input index of the string at a position
//   bktrank[n] uint32 scratch: (bucket << 21) | rank inside its wave chunk
//   bkey[n]    uint64 scratch: depth key of the first/last position of a leaf
//   perm[n], lcp[n]  results (caller memory)
//
// A sort proceeds in rounds; round r compares the 8 characters at depth 8r.
// Inside a round, segments (contiguous position ranges whose strings share
// their first 8r characters) are split by multi-workgroup splitter-tree steps
// until every piece is a leaf: a block (at most kBlkT strings, sorted
// completely in LDS by one workgroup), a tiny leaf (at most 64 strings, one
// wave) or an equal-key bucket.  Leaves finish their strings at this depth
// (order, LCPs, permutation) and turn every run of equal keys without a NUL
// into a segment of round r+1.
#pragma once
#include <cstdint>

namespace ps5 {

typedef uint64_t key64;

constexpr int kWG = 256;                     // threads per workgroup (small kernels)
constexpr int kWaves = kWG / 64;             // wave64
constexpr int kBigWG = 1024;                 // threads of the LDS-heavy kernels
constexpr int kBigWaves = kBigWG / 64;
constexpr uint32_t kXcds = 8;                // MI355X: 8 accelerator complex dies
constexpr int kMaxTreebits = 10;             // reference DefaultTreebits
constexpr int kMaxSpl = (1 << kMaxTreebits) - 1;    // 1023 splitters
constexpr int kMaxBkt = 2 * kMaxSpl + 1;            // 2047 buckets
constexpr int kMaxSamples = 4096;            // sample keys of one splitter tree
constexpr uint32_t kTreeSlots = 2048;        // per segment: level-order tree [0,1024), sorted splitters [1024,2048)
// A large segment is cut into T tiles of kWaves equal wave chunks; a wave
// chunk (the unit of stable ranks and of the scatter's LDS reorder) holds at
// most a per-level limit of keys, and T is the fewest tiles that allows, so
// a level's tiles are as few and as full as possible.  The limit is
// kChunkMax (kChunkMaxIdent in a level of roots, whose scatter stages
// 2-byte positions) unless the level is too small to give kMinTiles tiles that way;
// then it shrinks (not below kChunkMin) so small sorts still spread over
// the GPU.
constexpr int kChunkMax = 9216;              // keys ranked by one wave (multi-workgroup step)
constexpr int kChunkMaxIdent = 12032;        // the same in a level of root segments (indices = positions)
constexpr int kChunkMin = 2048;
constexpr uint32_t kMinTiles = 512;
constexpr uint32_t kTinyT = 64;              // leaves sorted by one wave
constexpr uint32_t kBlkT = 2048;             // leaves sorted on chip by one workgroup
constexpr uint32_t kMaxLeafTarget = 1024;    // largest leaf size the fan-out aims at
constexpr uint32_t kRunChunk = 4096;         // equal-key leaves are processed in chunks
// Blocks are kept in four size classes, each sorted by a kernel whose LDS
// fits its largest block: (64, 256], (256, 512], (512, 1024], (1024, 2048].
constexpr int kBlkClasses = 4;
__host__ __device__ constexpr int block_class(uint32_t m)
{
    return m <= 256 ? 0 : m <= 512 ? 1 : m <= 1024 ? 2 : 3;
}
constexpr int kRankBits = 21;
constexpr uint32_t kRankMask = (1u << kRankBits) - 1;

struct Item {            // a segment or a leaf
    uint32_t begin;      // first position
    uint32_t size;       // number of positions
    uint32_t rbeg;       // first position of the round segment it belongs to
    uint32_t flags;      // F_BUFB | F_SORTED | chunk index << 8
};
constexpr uint32_t F_BUFB = 1u;              // data lives in key[1]/idx[1]
constexpr uint32_t F_SORTED = 2u;            // all keys equal (equal bucket)
constexpr uint32_t F_IDENT = 4u;             // a root: idx[p] = p, not stored
constexpr uint32_t kFlagShift = 8;           // flags >> kFlagShift: chunk index of an equal-key chunk

// device counters
enum : int {
    C_LROUND = 0,  // [2] large segments starting a round (round parity)
    C_LLVL = 2,    // [2] large segments of the next level (level parity)
    C_TINY = 4,    // [2] leaf lists (round parity)
    C_BLOCK = 6,   // [4][2] blocks: C_BLOCK + 2 * size class + round parity
    C_RUN = 14,    // [2]
    C_ERR = 16,    // error flags
    C_BDONE = 17,  // [4] blocks of the current round already sorted, per size class (device-driven levels)
    C_RUNCONT = 21,  // [2] equal-key leaves whose strings continue into the next round (round parity)
    C_NCNT = 23
};
constexpr uint32_t ERR_CAP = 1u, ERR_LCP = 2u, ERR_BOUNDS = 4u;
constexpr uint32_t ERR_LEVELS = 8u;          // a captured round had more levels than it was built with

struct LargeDesc {       // one segment of the multi-workgroup sample-sort step
    Item seg;
    uint32_t T;          // tiles of kWaves * chunk keys (the last one may be short)
    uint32_t chunk;      // keys per wave chunk, <= kChunkMax (kChunkMaxIdent for a root)
    uint32_t pad;
    uint32_t hist_base;  // first histogram entry of this segment
    uint32_t k;          // treebits
    uint32_t oversample; // sample keys per splitter
};

struct TileRef {         // one workgroup of the multi-workgroup step
    uint32_t seg;        // index into the LargeDesc array
    uint32_t t;          // tile inside the segment
};

struct LevelHdr {        // totals of one level's multi-workgroup step
    uint32_t nlarge;     // segments (LargeDesc entries)
    uint32_t ntiles;     // tiles (TileRef entries), kWaves chunks each
    uint32_t hl;         // histogram entries
    uint32_t kmax;       // largest treebits of the level
    uint32_t pad[4];
};

// A level's plan in device memory: the header, then cap_large descriptors,
// then the tile table (fixed offsets, so a plan written on the device needs
// no pointer arithmetic on the host).
struct Plan {
    const LevelHdr* hdr;
    const LargeDesc* desc;
    const TileRef* tiles;
};

struct Sink {            // where new items are appended
    Item* large; Item* block; Item* tiny; Item* run;
    uint32_t* cnt;       // counter base
    int c_large, c_block, c_tiny, c_run;     // counter indices (c_block: of size class 0)
    int c_runcont;       // counter of equal-key leaves that continue (their key holds no NUL)
    uint32_t cap_large, cap_block, cap_tiny, cap_run;   // cap_block: per size class;
                         // class c lives at block + c * cap_block, counter c_block + 2 * c
};

struct Bufs {
    uint32_t n;          // number of strings (bound of every position and index)
    uint32_t* cnt;       // counters (error flags)
    const uint8_t* arena;
    uint64_t alen;
    const uint32_t* offs;      // string start offsets (uint32), or
    const uint64_t* offs64;    // uint64 offsets when non-null (batches over 4 GiB)
    key64* key0; key64* key1;
    uint32_t* idx0; uint32_t* idx1;
    uint32_t* perm;
    uint32_t* lcp;       // may be null
    key64* bkey;
    uint32_t* bktrank;
    const uint32_t* pending;   // non-null and non-zero: large segments remain, the round's leaves wait
    uint32_t pending_err;      // ... and that is an error (ERR_LEVELS: a captured graph cannot continue the round)
    uint32_t leaf_keys;        // keys (8 characters each) a leaf pass fetches per string at a time, 1..8
};

}  // namespace ps5
