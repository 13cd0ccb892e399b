// BLAKE3-256 (unkeyed hash mode), written from the published specification
// (O'Connor, Aumasson, Neves, Wilcox-O'Hearn, "BLAKE3: one function, fast
// everywhere", 2020): the 7-round ChaCha-style compression function, 1 KiB
// chunks of 64-byte blocks, and the binary tree of chaining values merged
// with a stack as chunks complete.  Host code; used for the RQMC direction
// and scramble checksums the result diagnostics carry (the reference hashes
// them with the blake3 crate, crates/pricing/src/engine/sampling/qmc.rs:236-259).
#include <cstring>

#include "pe_plan.hpp"

namespace pe {

namespace {

constexpr uint32_t kIv[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                             0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
constexpr uint8_t kPermutation[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
constexpr uint32_t kChunkStart = 1, kChunkEnd = 2, kParent = 4, kRoot = 8;

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

inline void mix(uint32_t* s, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    s[a] = s[a] + s[b] + x;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + y;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}

// Full 16-word output of the compression function.
void compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter, uint32_t block_len, uint32_t flags,
              uint32_t out[16]) {
    uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      kIv[0], kIv[1], kIv[2], kIv[3],
                      static_cast<uint32_t>(counter), static_cast<uint32_t>(counter >> 32), block_len, flags};
    uint32_t m[16];
    std::memcpy(m, block, sizeof(m));
    for (int r = 0; r < 7; ++r) {
        mix(s, 0, 4, 8, 12, m[0], m[1]);
        mix(s, 1, 5, 9, 13, m[2], m[3]);
        mix(s, 2, 6, 10, 14, m[4], m[5]);
        mix(s, 3, 7, 11, 15, m[6], m[7]);
        mix(s, 0, 5, 10, 15, m[8], m[9]);
        mix(s, 1, 6, 11, 12, m[10], m[11]);
        mix(s, 2, 7, 8, 13, m[12], m[13]);
        mix(s, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) {
            uint32_t p[16];
            for (int i = 0; i < 16; ++i) p[i] = m[kPermutation[i]];
            std::memcpy(m, p, sizeof(m));
        }
    }
    for (int i = 0; i < 8; ++i) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

void load_block(const uint8_t* bytes, size_t len, uint32_t block[16]) {
    uint8_t buf[64] = {0};
    std::memcpy(buf, bytes, len);
    for (int i = 0; i < 16; ++i)
        block[i] = static_cast<uint32_t>(buf[4 * i]) | static_cast<uint32_t>(buf[4 * i + 1]) << 8 |
                   static_cast<uint32_t>(buf[4 * i + 2]) << 16 | static_cast<uint32_t>(buf[4 * i + 3]) << 24;
}

}  // namespace

void Blake3::reset() {
    std::memcpy(cv_, kIv, sizeof(cv_));
    chunk_counter_ = 0;
    block_len_ = 0;
    blocks_compressed_ = 0;
    stack_len_ = 0;
}

Blake3::Blake3() { reset(); }

// Compresses the buffered full block of the current chunk (not its last one).
void Blake3::compress_buffer() {
    uint32_t block[16], out[16];
    load_block(block_, 64, block);
    compress(cv_, block, chunk_counter_, 64, blocks_compressed_ == 0 ? kChunkStart : 0u, out);
    std::memcpy(cv_, out, sizeof(cv_));
    ++blocks_compressed_;
    block_len_ = 0;
}

// The current chunk is complete: its chaining value joins the tree.
void Blake3::finish_chunk() {
    uint32_t block[16], out[16];
    load_block(block_, block_len_, block);
    const uint32_t flags = (blocks_compressed_ == 0 ? kChunkStart : 0u) | kChunkEnd;
    compress(cv_, block, chunk_counter_, block_len_, flags, out);
    uint32_t cv[8];
    std::memcpy(cv, out, sizeof(cv));
    uint64_t total = chunk_counter_ + 1;
    while ((total & 1u) == 0) {   // merge completed subtrees
        uint32_t pm[16], po[16];
        std::memcpy(pm, stack_[--stack_len_], 32);
        std::memcpy(pm + 8, cv, 32);
        compress(kIv, pm, 0, 64, kParent, po);
        std::memcpy(cv, po, sizeof(cv));
        total >>= 1;
    }
    std::memcpy(stack_[stack_len_++], cv, 32);
    ++chunk_counter_;
    std::memcpy(cv_, kIv, sizeof(cv_));
    blocks_compressed_ = 0;
    block_len_ = 0;
}

void Blake3::update(const void* data, size_t n) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    while (n > 0) {
        // a full chunk is closed only once more input arrives (the last chunk is the root's)
        if (blocks_compressed_ == 15 && block_len_ == 64) finish_chunk();
        if (block_len_ == 64) compress_buffer();
        const size_t take = std::min<size_t>(64 - block_len_, n);
        std::memcpy(block_ + block_len_, p, take);
        block_len_ += static_cast<uint32_t>(take);
        p += take;
        n -= take;
    }
}

void Blake3::finalize(uint8_t out[32]) const {
    // root output: the last chunk's final block, or the parents above it
    uint32_t cv[8], block[16];
    std::memcpy(cv, cv_, sizeof(cv));
    load_block(block_, block_len_, block);
    uint64_t counter = chunk_counter_;
    uint32_t block_len = block_len_;
    uint32_t flags = (blocks_compressed_ == 0 ? kChunkStart : 0u) | kChunkEnd;
    for (int i = static_cast<int>(stack_len_) - 1; i >= 0; --i) {
        uint32_t o[16];
        compress(cv, block, counter, block_len, flags, o);
        std::memcpy(block, stack_[i], 32);
        std::memcpy(block + 8, o, 32);
        std::memcpy(cv, kIv, sizeof(cv));
        counter = 0;
        block_len = 64;
        flags = kParent;
    }
    uint32_t o[16];
    compress(cv, block, counter, block_len, flags | kRoot, o);
    for (int i = 0; i < 8; ++i)
        for (int b = 0; b < 4; ++b) out[4 * i + b] = static_cast<uint8_t>(o[i] >> (8 * b));
}

}  // namespace pe
