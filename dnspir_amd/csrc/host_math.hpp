// Host-side integer arithmetic of the engine: ChaCha12 in rand_chacha word
// order, seed expansion, mod switching, the bit-plane wire codec and the BFV
// client operations (prepare_query / process_reply need them). All exact.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace dnh {

constexpr int N = 2048;
constexpr uint64_t P1 = 67104769ull, P2 = 33550337ull;
constexpr uint64_t Q = P1 * P2;
constexpr uint64_t T = 1025;

typedef unsigned __int128 u128;
typedef __int128 i128;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }
inline uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    a %= m;
    while (e) {
        if (e & 1) r = mulmod(r, a, m);
        a = mulmod(a, a, m);
        e >>= 1;
    }
    return r;
}
inline uint64_t q_add(uint64_t a, uint64_t b) { uint64_t s = a + b; return s >= Q ? s - Q : s; }
inline uint64_t q_sub(uint64_t a, uint64_t b) { return a >= b ? a - b : a + Q - b; }
inline uint64_t q_of(i128 v) { i128 r = v % (i128)Q; return (uint64_t)(r < 0 ? r + (i128)Q : r); }
inline int64_t q_center(uint64_t x) { return x > Q / 2 ? (int64_t)x - (int64_t)Q : (int64_t)x; }

// The six a-part seeds (base_pir.rs:53-71): SHA3-256 of six UTF-8 titles.
static const uint8_t SEEDS[6][32] = {
    {0xda, 0xd2, 0x0c, 0x14, 0x2d, 0xeb, 0x8b, 0xb4, 0xd6, 0x63, 0x20, 0xee, 0x07, 0xb1, 0xc1, 0x90,
     0xec, 0x25, 0xc0, 0xa1, 0x54, 0x1c, 0xca, 0xf5, 0x98, 0xfa, 0x22, 0xff, 0xdf, 0xdd, 0x68, 0xba},
    {0xd3, 0xc5, 0xf6, 0xee, 0xb3, 0x3a, 0x4c, 0xf3, 0x33, 0x51, 0x6c, 0xe2, 0x73, 0xd5, 0x74, 0x76,
     0xb5, 0x79, 0x6c, 0xce, 0x9c, 0xf3, 0x37, 0xa1, 0x9c, 0x84, 0xbc, 0xbe, 0xa4, 0x3d, 0x2c, 0xb5},
    {0x95, 0x27, 0xb8, 0x6c, 0x30, 0xdc, 0x30, 0xcf, 0x1b, 0xb1, 0x74, 0xc3, 0xcd, 0x6c, 0xd7, 0x50,
     0x72, 0x09, 0xda, 0xa5, 0x4b, 0xbd, 0xb9, 0x99, 0x04, 0x14, 0x00, 0x08, 0x3d, 0x68, 0x24, 0x8a},
    {0x4c, 0xa2, 0xf6, 0x47, 0x8a, 0x58, 0x11, 0x91, 0x67, 0x67, 0x41, 0x00, 0x2a, 0x00, 0xed, 0xd1,
     0xac, 0x97, 0xf9, 0x30, 0x98, 0x0b, 0x7f, 0xbe, 0xf0, 0x33, 0xb4, 0x48, 0x23, 0x4b, 0x21, 0x6b},
    {0x98, 0xfc, 0x62, 0xc5, 0x17, 0x52, 0x0e, 0x5d, 0x29, 0xfc, 0x30, 0x7e, 0x73, 0xef, 0xb1, 0x0e,
     0x14, 0x18, 0xaa, 0x07, 0x0e, 0x32, 0x36, 0x4f, 0x89, 0x3f, 0xa8, 0xbd, 0x92, 0x13, 0x03, 0x27},
    {0xb3, 0x61, 0xdc, 0x94, 0x34, 0xdb, 0xba, 0xe9, 0x28, 0xe5, 0xc9, 0x64, 0xdc, 0x18, 0x02, 0x54,
     0x02, 0xfe, 0x81, 0xaa, 0xf7, 0x49, 0x17, 0x20, 0x8a, 0x6e, 0x6f, 0xd9, 0x55, 0x70, 0x2a, 0x21},
};

// ChaCha with 12 rounds; 64-bit block counter in words 12-13, zero stream id,
// four blocks per refill; next_u64 = two consecutive words, low first.
struct ChaCha12 {
    uint32_t st[16];
    uint32_t buf[64];
    int idx;
    explicit ChaCha12(const uint8_t seed[32]) {
        st[0] = 0x61707865u; st[1] = 0x3320646eu; st[2] = 0x79622d32u; st[3] = 0x6b206574u;
        for (int i = 0; i < 8; ++i)
            st[4 + i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) | ((uint32_t)seed[4 * i + 2] << 16) |
                        ((uint32_t)seed[4 * i + 3] << 24);
        st[12] = st[13] = st[14] = st[15] = 0;
        idx = 64;
    }
    static inline uint32_t rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }
    static inline void qr(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
        a += b; d = rotl(d ^ a, 16);
        c += d; b = rotl(b ^ c, 12);
        a += b; d = rotl(d ^ a, 8);
        c += d; b = rotl(b ^ c, 7);
    }
    void block(uint32_t* out) {
        uint32_t x[16];
        memcpy(x, st, sizeof(x));
        for (int r = 0; r < 12; r += 2) {
            qr(x[0], x[4], x[8], x[12]); qr(x[1], x[5], x[9], x[13]);
            qr(x[2], x[6], x[10], x[14]); qr(x[3], x[7], x[11], x[15]);
            qr(x[0], x[5], x[10], x[15]); qr(x[1], x[6], x[11], x[12]);
            qr(x[2], x[7], x[8], x[13]); qr(x[3], x[4], x[9], x[14]);
        }
        for (int i = 0; i < 16; ++i) out[i] = x[i] + st[i];
        if (++st[12] == 0) ++st[13];
    }
    uint32_t next_u32() {
        if (idx >= 64) {
            for (int b = 0; b < 4; ++b) block(buf + 16 * b);
            idx = 0;
        }
        return buf[idx++];
    }
    uint64_t next_u64() {
        uint64_t lo = next_u32();
        uint64_t hi = next_u32();
        return lo | (hi << 32);
    }
    // standard normal (Box-Muller); the client's noise only
    double gauss() {
        double u1 = ((double)(next_u64() >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        double u2 = (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0);
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

// bfv.rs:46-49. feanor-math's random_element over Z_q is taken as rejection
// of next_u64 & (2^51 - 1) >= q, coefficients in order (not verifiable offline).
inline void expand(const uint8_t seed[32], uint64_t* out) {
    ChaCha12 r(seed);
    const uint64_t mask = (1ull << 51) - 1;
    for (int i = 0; i < N; ++i) {
        uint64_t x;
        do { x = r.next_u64() & mask; } while (x >= Q);
        out[i] = x;
    }
}

// bfv.rs:51-58
inline uint64_t mod_switch_encode(uint64_t x, int bits) {
    i128 num = (i128)q_center(x) * ((i128)1 << bits);
    i128 fl = num / (i128)Q;
    i128 rem = num - fl * (i128)Q;
    if (rem < 0) { rem += Q; fl -= 1; }
    if (2 * rem > (i128)Q) fl += 1;
    i128 m = (i128)1 << bits;
    return (uint64_t)(((fl % m) + m) % m);
}
// bfv.rs:60-69
inline uint64_t mod_switch_decode(uint64_t x, int bits) { return (uint64_t)(((u128)x * Q) >> bits) % Q; }

// pir_wrapper.rs:84-115
inline size_t condense(int bits, const uint64_t* d, size_t n, uint8_t* out) {
    size_t k = 0;
    for (int i = 0; i < bits / 8; ++i)
        for (size_t j = 0; j < n; ++j) out[k++] = (uint8_t)(d[j] >> (8 * i));
    if (bits % 8 >= 4) {
        int sh = (bits / 8) * 8;
        for (size_t j = 0; j + 1 < n; j += 2) out[k++] = (uint8_t)((((d[j] >> sh) & 15) << 4) | ((d[j + 1] >> sh) & 15));
    }
    if (bits % 4 >= 2) {
        int sh = (bits / 4) * 4;
        for (size_t j = 0; j + 3 < n; j += 4)
            out[k++] = (uint8_t)((((d[j] >> sh) & 3) << 6) | (((d[j + 1] >> sh) & 3) << 4) | (((d[j + 2] >> sh) & 3) << 2) |
                                 ((d[j + 3] >> sh) & 3));
    }
    if (bits % 2 == 1) {
        int sh = (bits / 2) * 2;
        for (size_t j = 0; j + 7 < n; j += 8) {
            uint8_t v = 0;
            for (int c = 0; c < 8; ++c) v |= (uint8_t)(((d[j + c] >> sh) & 1) << (7 - c));
            out[k++] = v;
        }
    }
    return k;
}
inline size_t condensed_size(int bits, size_t n) { return n / 8 * bits; }
// pir_wrapper.rs:121-158; false = not enough data
inline bool uncondense(int bits, size_t n, const uint8_t* d, size_t len, uint64_t* out) {
    if (len < condensed_size(bits, n)) return false;
    size_t k = 0;
    for (size_t j = 0; j < n; ++j) out[j] = 0;
    for (int i = 0; i < bits / 8; ++i)
        for (size_t j = 0; j < n; ++j) out[j] |= (uint64_t)d[k++] << (8 * i);
    if (bits % 8 >= 4) {
        int sh = (bits / 8) * 8;
        for (size_t j = 0; j + 1 < n; j += 2) {
            uint8_t b = d[k++];
            out[j] |= (uint64_t)(b >> 4) << sh;
            out[j + 1] |= (uint64_t)(b & 15) << sh;
        }
    }
    if (bits % 4 >= 2) {
        int sh = (bits / 4) * 4;
        for (size_t j = 0; j + 3 < n; j += 4) {
            uint8_t b = d[k++];
            for (int c = 0; c < 4; ++c) out[j + c] |= (uint64_t)((b >> (6 - 2 * c)) & 3) << sh;
        }
    }
    if (bits % 2 == 1) {
        int sh = (bits / 2) * 2;
        for (size_t j = 0; j + 7 < n; j += 8) {
            uint8_t b = d[k++];
            for (int c = 0; c < 8; ++c) out[j + c] |= (uint64_t)((b >> (7 - c)) & 1) << sh;
        }
    }
    return true;
}

// ---- BFV client (bfv.rs:36-124), ring Z_q[X]/(X^N + 1) ---------------------
inline void gen_sk(ChaCha12& r, int hwt, int64_t* sk) {
    for (int i = 0; i < N; ++i) sk[i] = 0;
    for (int i = 0; i < hwt; ++i) {
        uint64_t idx = r.next_u64() % N;
        sk[idx] = (int64_t)(r.next_u32() % 3) - 1;
    }
}
inline void mul_sparse(const uint64_t* a, const int64_t* sk, uint64_t* out) {
    for (int i = 0; i < N; ++i) out[i] = 0;
    for (int j = 0; j < N; ++j) {
        if (!sk[j]) continue;
        for (int i = 0; i < N; ++i) {
            int idx = i + j;
            bool neg = sk[j] < 0;
            if (idx >= N) { idx -= N; neg = !neg; }
            out[idx] = neg ? q_sub(out[idx], a[i]) : q_add(out[idx], a[i]);
        }
    }
}
inline uint64_t scale_up(uint64_t m) {
    int64_t mc = (int64_t)(m % T);
    if ((uint64_t)mc > T / 2) mc -= (int64_t)T;
    i128 up = (i128)mc * (i128)Q;
    int64_t d = (int64_t)(up % (i128)T);
    if (d < 0) d += (int64_t)T;
    if ((uint64_t)d > T / 2) d -= (int64_t)T;
    return q_of((up - d) / (i128)T);
}
// b = e - a s + round(q m / t), a = expand(seed)
inline void enc_sym_b(const uint8_t seed[32], ChaCha12& r, const int64_t* sk, const uint16_t* m, uint64_t* b) {
    std::vector<uint64_t> a(N), as(N);
    expand(seed, a.data());
    for (int i = 0; i < N; ++i) b[i] = q_of((int64_t)llround(r.gauss() * 3.2));
    mul_sparse(a.data(), sk, as.data());
    for (int i = 0; i < N; ++i) b[i] = q_add(q_sub(b[i], as[i]), scale_up(m[i]));
}
// Galois keys for X -> X^5: b_i = e - a_i s + 2^(26 i) s(X^5)
inline void gen_gk_b(ChaCha12& r, const int64_t* sk, uint64_t* gk) {
    const int64_t m2 = 2 * N;
    int64_t ginv = 0;
    for (int64_t x = 1; x < m2; x += 2)
        if ((5 * x) % m2 == 1) { ginv = x; break; }
    std::vector<int64_t> gal(N);
    for (int i = 0; i < N; ++i) {
        int64_t ni = ((int64_t)i * ginv) % m2;
        gal[i] = ni > N ? -sk[ni - N] : sk[ni];
    }
    std::vector<uint64_t> a(N), as(N);
    for (int d = 0; d < 2; ++d) {
        expand(SEEDS[d], a.data());
        uint64_t* b = gk + (size_t)d * N;
        for (int i = 0; i < N; ++i) b[i] = q_of((int64_t)llround(r.gauss() * 3.2));
        mul_sparse(a.data(), sk, as.data());
        uint64_t sc = d == 0 ? 1 : (1ull << 26);
        for (int i = 0; i < N; ++i) b[i] = q_add(q_sub(b[i], as[i]), mulmod(q_of(gal[i]), sc, Q));
    }
}
// round(t (a s + b) / q) mod t
inline void dec(const int64_t* sk, const uint64_t* a, const uint64_t* b, uint16_t* m) {
    std::vector<uint64_t> as(N);
    mul_sparse(a, sk, as.data());
    for (int i = 0; i < N; ++i) {
        int64_t c = q_center(q_add(as[i], b[i]));
        i128 up = (i128)c * (i128)T;
        i128 d = up % (i128)Q;
        if (d < 0) d += Q;
        if (d > (i128)(Q / 2)) d -= Q;
        int64_t r = (int64_t)(((up - d) / (i128)Q) % (i128)T);
        if (r < 0) r += (int64_t)T;
        m[i] = (uint16_t)r;
    }
}

}  // namespace dnh
