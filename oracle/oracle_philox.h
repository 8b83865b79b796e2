/* oracle_philox.h — TEST INFRASTRUCTURE (the oracle), not part of the product.
 *
 * Independent restatement of the Philox counter-based generators, written
 * from the Random123 specification (Salmon et al., SC'11) and kept separate
 * from the library's include/mppi_hip/mppi_philox.h so that the two can be
 * checked against each other bit for bit.  Pinned by the Random123
 * known-answer vectors in tests/test_oracle_philox.py.
 *
 * Reference followed: shaders/philox.h:19-36 (Philox2x32-10, the
 * reference's generator), and its counter-based use in
 * shaders/mppi_rollout.metal:73-87.
 */
#ifndef ORACLE_PHILOX_H
#define ORACLE_PHILOX_H

#include <stdint.h>

static inline void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                        uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 = k0 + 0x9E3779B9u;
            k1 = k1 + 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void oracle_philox2x32_10(const uint32_t ctr_in[2], uint32_t key,
                                        uint32_t out[2]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) key = key + 0x9E3779B9u;
        uint64_t p = (uint64_t)0xD256D193u * (uint64_t)c0;
        uint32_t n0 = (uint32_t)(p >> 32) ^ key ^ c1;
        uint32_t n1 = (uint32_t)p;
        c0 = n0; c1 = n1;
    }
    out[0] = c0; out[1] = c1;
}

#endif /* ORACLE_PHILOX_H */
