/*
 * ORACLE (test infrastructure only) -- Philox4x32-10 restated in plain C.
 *
 * J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11.  Round function and Weyl key
 * schedule exactly as published; 10 rounds, key bumped between rounds.
 * Pinned by the Random123 known-answer vectors (tests/test_oracle_philox.py)
 * and cross-checked against rocRAND's header-only philox4x32_10 engine.
 *
 * This file is never linked into the product library; the product carries
 * its own device implementation.
 */
#ifndef CRITMC_ORACLE_PHILOX_H
#define CRITMC_ORACLE_PHILOX_H
#include <stdint.h>

#define ORC_PHILOX_M0 0xD2511F53u
#define ORC_PHILOX_M1 0xCD9E8D57u
#define ORC_PHILOX_W0 0x9E3779B9u
#define ORC_PHILOX_W1 0xBB67AE85u

static inline void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                     uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)ORC_PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)ORC_PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += ORC_PHILOX_W0;
        k1 += ORC_PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/*
 * Counter layout of every random draw made by a chain (the stream contract
 * the GPU kernels follow too):
 *   key = (k0, k1)                       per chain
 *   ctr = (block, sweep_lo, sweep_hi16 | color << 16 | tag << 24, stream_id)
 */
enum {
    ORC_TAG_INIT = 1,
    ORC_TAG_ISING = 2,
    ORC_TAG_POTTS = 3,
    ORC_TAG_EA = 4,
    ORC_TAG_PT = 5,
    ORC_TAG_ON = 6,
    ORC_TAG_TYPEWRITER = 7,
    ORC_TAG_ONHB = 8,
    ORC_TAG_CLUSTER = 9,
    ORC_TAG_FLAT = 10,
    ORC_TAG_SEQ = 11,
};

static inline void orc_draw_block(uint32_t k0, uint32_t k1, uint32_t stream, uint64_t sweep,
                                  int color, int tag, uint32_t block, uint32_t out[4])
{
    uint32_t ctr[4] = {block, (uint32_t)sweep,
                       (uint32_t)((sweep >> 32) & 0xFFFFu) | ((uint32_t)color << 16) |
                           ((uint32_t)tag << 24),
                       stream};
    uint32_t key[2] = {k0, k1};
    orc_philox4x32_10(ctr, key, out);
}

#endif
