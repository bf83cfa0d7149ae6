// Shared constants, device tables and exact modular arithmetic for the two
// RNS moduli of the PIR ciphertext ring (crates/pir/src/base_pir.rs:72-78).
//
// Every device value is a residue mod p_r kept in [0, p_r) unless a function
// says otherwise. p_r = 2^K - 2^12 + 1, so 2^K = 2^12 - 1 (mod p_r) and a
// 64-bit value folds with shifts, masks and a multiply by 4095.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dn {

constexpr int LOGN = 11;
constexpr int N = 1 << LOGN;         // ring dimension
constexpr int M2N = 2 * N;           // order of the Galois group's ambient Z/2N
constexpr int HALFN = N / 2;         // order of <5> in (Z/2N)^*
constexpr int LANES = 8;             // SIMD_COUNT: entries packed per ring slot
constexpr uint32_t P1 = 67104769u;   // 2^26 - 2^12 + 1
constexpr uint32_t P2 = 33550337u;   // 2^25 - 2^12 + 1
constexpr uint64_t Q = (uint64_t)P1 * P2;
constexpr uint32_t T_PLAIN = 1025;   // plaintext modulus t
constexpr int LOG2_B = 26;           // gadget base
// one block of the shard slab's limb-block layout: 1024 residues as three
// int8 digit planes (3 x 1024 B), then the top digits: 2 bits each mod p1
// (64 words), 1 bit each mod p2 (64 half-words)
constexpr int MB_BLOCK_BYTES_P1 = 3328;
constexpr int MB_BLOCK_BYTES_P2 = 3200;
// byte offset of block rb (counted inside modulus r) in a slab of 2 * N/8 * S
// * 32 blocks per modulus: the p1 blocks first, then the p2 blocks
__host__ __device__ constexpr size_t mb_offset(int r, size_t rb, int S) {
    return r == 0 ? rb * MB_BLOCK_BYTES_P1 : (size_t)2 * (N / 8) * S * 32 * MB_BLOCK_BYTES_P1 + rb * MB_BLOCK_BYTES_P2;
}

__host__ __device__ constexpr uint32_t mod_p(int r) { return r == 0 ? P1 : P2; }
__host__ __device__ constexpr int mod_k(int r) { return r == 0 ? 26 : 25; }

// Read-only tables built once per device (see DevTables in the host code).
struct DevTables {
    // negacyclic NTT twiddles: psi[r][k] = psi_r^brv(k), inverse psi^-brv(k),
    // each with its Shoup companion floor(w * 2^32 / p)
    uint32_t psi[2][N], psiq[2][N], ipsi[2][N], ipsiq[2][N];
    uint32_t ninv[2], ninvq[2];
    // CRT factors of the chains' digit phase with their Shoup companions:
    // crt_k[0][r] = N^-1 (q/p_r)^-1 mod p_r, crt_k[1][r] = psi_r^-1 crt_k[0][r]
    // (the last inverse stage's twiddle folded in)
    uint32_t crt_k[2][2], crt_kq[2][2];
    // psi_r^k for k < 2N (twists of the subgroup DFTs), with Shoup companions
    uint32_t psipow[2][M2N], psipowq[2][M2N];
    // 5^i and 5^-i mod 2N for i < N/2
    uint16_t pow5[HALFN], inv5pow[HALFN];
    // exponent e of slot s (slot s holds f(psi^e)), and its inverse map
    uint16_t slot_exp[N];
    uint16_t exp_slot[M2N];  // 0xFFFF for even exponents
    // sigma_5 as a gather in slot order: sigma_5(f)[s] = f[sigma5[s]]
    uint16_t sigma5[N];
    // Galois order: position g = sgn * N/2 + a holds exponent (-1)^sgn * 5^a
    uint16_t gpos_slot[N];
    uint16_t slot_gpos[N];
    // NTT forms of the seed-expanded a-parts: gk digits (SEEDS[0..1]) and
    // query a-parts (SEEDS[2..5]), base_pir.rs:193-205, 782-789
    uint32_t gka[2][2][N];   // [digit][modulus][slot]
    uint32_t qa[4][2][N];    // [seed - 2][modulus][slot]
};

// ---------------------------------------------------------------------------
// exact modular arithmetic. A conditional subtraction x >= p ? x - p : x is
// written as the unsigned minimum of x and x - p (x - p wraps above x when
// x < p; the same result for every 32-bit x): one v_min_u32 instead of a
// compare, a select and an add.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t csub(uint32_t x, uint32_t p) {
    return __builtin_elementwise_min(x, x - p);
}
__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) {
    return csub(a + b, p);
}
__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t d = a - b;  // wraps when a < b, and d + p then wraps back below p
    return __builtin_elementwise_min(d, d + p);
}
// a * w mod p for a < 2^32, w < p, wq = floor(w * 2^32 / p) (Shoup)
__device__ __forceinline__ uint32_t mul_shoup(uint32_t a, uint32_t w, uint32_t wq, uint32_t p) {
    uint32_t qt = __umulhi(a, wq);
    uint32_t r = a * w - qt * p;  // in [0, 2p)
    return csub(r, p);
}
// x < 2^52 (a product of two residues) -> x mod p. Two folds of
// 2^K = 4095 (mod p) leave z < 2^24 + 2^K < 2p.
__device__ __forceinline__ uint32_t red52(uint64_t x, uint32_t p, int K) {
    const uint64_t mask = (1ull << K) - 1;
    uint64_t y = (x >> K) * 4095ull + (x & mask);
    uint32_t z = (uint32_t)((y >> K) * 4095ull + (y & mask));
    return z >= p ? z - p : z;
}
// any x < 2^64 -> x mod p (three folds, then at most two subtractions)
__device__ __forceinline__ uint32_t red64(uint64_t x, uint32_t p, int K) {
    const uint64_t mask = (1ull << K) - 1;
    uint64_t y = (x >> K) * 4095ull + (x & mask);   // < 2^51 + 2^K
    y = (y >> K) * 4095ull + (y & mask);            // < 2^38 + 2^K
    uint32_t z = (uint32_t)((y >> K) * 4095ull + (y & mask));  // < 2^26 + 2^K
    // (kept as compares: k1w_kernel's schedule, which this epilogue shares,
    // measured 4% slower with the minimum form)
    z = z >= p ? z - p : z;
    return z >= p ? z - p : z;
}
__device__ __forceinline__ uint32_t mul_mod(uint32_t a, uint32_t b, uint32_t p, int K) {
    return red52((uint64_t)a * b, p, K);
}
// signed integer |v| < 2^62 -> v mod p
__device__ __forceinline__ uint32_t from_i64(int64_t v, uint32_t p, int K) {
    if (v >= 0) return red64((uint64_t)v, p, K);
    uint32_t r = red64((uint64_t)(-v), p, K);
    return r == 0 ? 0 : p - r;
}


// centered lift of a residue mod q: [-(q-1)/2, (q-1)/2]
__device__ __forceinline__ int64_t center_q(uint64_t v) {
    return v > Q / 2 ? (int64_t)v - (int64_t)Q : (int64_t)v;
}
// mod switch to `bits` bits, bfv.rs:51-58: (round(lift_c(x) 2^B / q) + 2^B) mod 2^B
__device__ __forceinline__ uint32_t mod_switch_enc(uint64_t x, int bits) {
    int64_t c = center_q(x);
    if (bits <= 12) {
        // |c| 2^bits < 2^62 and |r| <= 2^(bits-1) + 1, so |r q| < 2^63: the
        // exact remainder fits a 64-bit integer
        const int64_t r0 = (int64_t)floor((double)c * ((double)(1ull << bits) / (double)Q) + 0.5);
        const int64_t rem = c * ((int64_t)1 << bits) - r0 * (int64_t)Q;
        const int64_t r = r0 + (2 * rem > (int64_t)Q ? 1 : 0) - (2 * rem < -(int64_t)Q ? 1 : 0);
        const int64_t m = (int64_t)1 << bits;
        return (uint32_t)((r + m) & (m - 1));
    }
    double est = (double)c * ((double)(1ull << bits) / (double)Q);
    int64_t r = (int64_t)floor(est + 0.5);
    __int128 num = (__int128)c * ((__int128)1 << bits);
    __int128 rem = num - (__int128)r * (__int128)Q;   // exact
    if (2 * rem > (__int128)Q) r += 1;
    else if (2 * rem < -(__int128)Q) r -= 1;
    int64_t m = (int64_t)1 << bits;
    return (uint32_t)((r + m) & (m - 1));
}

}  // namespace dn
