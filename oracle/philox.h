/*
 * oracle/philox.h -- CPU Philox4x32-10 (test infrastructure only).
 *
 * This file belongs to the oracle: only tests/, the smoke check and the
 * bench's cpu_baseline leg may use it, and only to check results.
 *
 * Philox4x32-10 as published by Salmon, Moraes, Dror and Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3" (SC'11), Random123 1.09:
 *   multipliers M0 = 0xD2511F53, M1 = 0xCD9E8D57,
 *   Weyl key increments W0 = 0x9E3779B9, W1 = 0xBB67AE85,
 *   round: (c0,c1,c2,c3) -> (hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1, lo(M0*c0)),
 *   key bumped by (W0,W1) between rounds, 10 rounds.
 * Pinned by the Random123 known-answer vectors and by a cross-check against
 * torch's ATen/core/PhiloxRNGEngine.h (tests/test_oracle_philox.py).
 *
 * The reference (MARQOV) uses std::mt19937_64 behind RNGCache
 * (src/libmarqov/rngcache.h:120-265); Philox replaces it on the GPU path
 * because a counter-based generator lets every site draw its own numbers
 * from (key, counter) with no shared state.
 */
#ifndef MQ_ORACLE_PHILOX_H
#define MQ_ORACLE_PHILOX_H
#include <stdint.h>

#define ORC_PHILOX_M0 0xD2511F53u
#define ORC_PHILOX_M1 0xCD9E8D57u
#define ORC_PHILOX_W0 0x9E3779B9u
#define ORC_PHILOX_W1 0xBB67AE85u

static inline void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
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

/* Convenience: counter (a,b,c,d), key = 64-bit seed split lo/hi. */
static inline void orc_philox_u64key(uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                     uint64_t seed, uint32_t out[4])
{
    uint32_t ctr[4] = {a, b, c, d};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

#endif
