/*
 * CPU oracle for the DNSPIR server response path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py load this
 * library; the product never links or calls it. It restates, with exact
 * integer arithmetic (every value fully reduced mod p after each operation),
 * the reference algorithms in crates/pir/src:
 *   bfv.rs         expand, mod switch, gen_sk / enc_sym_b / gen_gk_b / dec
 *   base_pir.rs    PIRDatabase: create, set_transformed_entries (K5),
 *                  preprocess + partial_gal (K4/K6), perform_pir_split with
 *                  main_pir_loop (K1), get_a / from_simd_element (K7)
 *   double_pir.rs  perform_batched_double_pir, encode_ciphertext_plaintext,
 *                  enc/dec_double_pir
 *   pir_wrapper.rs condense/uncondense, shape rule, process_queries,
 *                  prepare_query, process_reply
 * Each function cites the reference lines it follows.
 *
 * Pins (see tests/test_oracle_*.py): ChaCha core against the RFC 8439 block
 * vectors, SEEDS against SHA3-256 of their titles, the N = 8 known-answer
 * matrices of base_pir.rs:1281-1411, brute-force negacyclic DFTs, closed
 * forms of K1 / K5 / K6 via plain polynomial arithmetic, decryption round
 * trips (base_pir.rs:1413-1537, double_pir.rs:328-451), codec round trips
 * (pir_wrapper.rs:442-452), wire sizes, and root-of-unity independence.
 *
 * Parity unpinned: `expand` (bfv.rs:46-49) depends on feanor-math 3.5.18
 * `random_element` sampling over rand_chacha 0.3.1 ChaCha12; this file uses
 * rejection sampling of 51-bit masked next_u64 values, coefficient order
 * 0..N-1 (an assumption; the crates are not present offline). The client's
 * Gaussian noise uses Box-Muller, not rand_distr's ziggurat (harness only).
 */
#include "dnspir_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;

#define LANES ORC_LANES
#define GK_DIGITS 2      /* base_pir.rs:72 */
#define LOG2_B 26        /* base_pir.rs:74 */
#define GEN_H 5          /* base_pir.rs:180-182 */
#define CTXT_PARTS 3     /* double_pir.rs:36 */
#define BATCH_COUNT 2    /* double_pir.rs:37 */
#define A_BITS 17        /* double_pir.rs:40, pir_wrapper.rs:57 */
#define B_BITS 12        /* double_pir.rs:41, pir_wrapper.rs:59 */
#define GK_BITS 50       /* pir_wrapper.rs:61 */
#define QRY_BITS 34      /* pir_wrapper.rs:64,67 */
#define SK_HWT 128       /* pir_wrapper.rs:72 */
#define SIGMA 3.2        /* pir_wrapper.rs:70 */

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* exact modular arithmetic                                                  */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t p, mu; } md_t;

static md_t md_make(uint64_t p) {
    md_t m;
    m.p = p;
    m.mu = (uint64_t)((((u128)1) << 64) / p);
    return m;
}
/* full reduction of any x < 2^64 (Barrett with mu = floor(2^64/p)) */
static inline uint64_t md_red(const md_t* M, uint64_t x) {
    uint64_t qt = (uint64_t)(((u128)x * M->mu) >> 64);
    x -= qt * M->p;
    return x >= M->p ? x - M->p : x;
}
static inline uint64_t md_mul(const md_t* M, uint64_t a, uint64_t b) { return md_red(M, a * b); }
static inline uint64_t md_add(const md_t* M, uint64_t a, uint64_t b) { uint64_t s = a + b; return s >= M->p ? s - M->p : s; }
static inline uint64_t md_sub(const md_t* M, uint64_t a, uint64_t b) { return a >= b ? a - b : a + M->p - b; }
static uint64_t md_pow(const md_t* M, uint64_t a, uint64_t e) {
    uint64_t r = 1 % M->p;
    a %= M->p;
    while (e) {
        if (e & 1) r = md_mul(M, r, a);
        a = md_mul(M, a, a);
        e >>= 1;
    }
    return r;
}
static uint64_t md_inv(const md_t* M, uint64_t a) { return md_pow(M, a, M->p - 2); } /* p prime */
/* value of a signed integer mod p */
static inline uint64_t md_from_i64(const md_t* M, int64_t v) {
    int64_t r = v % (int64_t)M->p;
    return (uint64_t)(r < 0 ? r + (int64_t)M->p : r);
}
/* centered lift of x in [0,p): (-p/2, p/2] */
static inline int64_t md_center(const md_t* M, uint64_t x) { return x > M->p / 2 ? (int64_t)x - (int64_t)M->p : (int64_t)x; }

/* arithmetic mod q (q < 2^51) */
static inline uint64_t q_add(uint64_t a, uint64_t b) { uint64_t s = a + b; return s >= ORC_Q ? s - ORC_Q : s; }
static inline uint64_t q_sub(uint64_t a, uint64_t b) { return a >= b ? a - b : a + ORC_Q - b; }
static inline uint64_t q_mul(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % ORC_Q); }
static inline uint64_t q_from_i128(i128 v) { i128 r = v % (i128)ORC_Q; return (uint64_t)(r < 0 ? r + (i128)ORC_Q : r); }
/* smallest_lift over Z_q: q odd, so the range is [-(q-1)/2, (q-1)/2] */
static inline int64_t q_center(uint64_t x) { return x > ORC_Q / 2 ? (int64_t)x - (int64_t)ORC_Q : (int64_t)x; }

/* integer arithmetic mod m = 2N (Galois group) */
static inline int64_t zm(int64_t v, int64_t m) { int64_t r = v % m; return r < 0 ? r + m : r; }
static int64_t zm_pow(int64_t a, int64_t e, int64_t m) {
    int64_t r = 1 % m;
    a = zm(a, m);
    while (e > 0) {
        if (e & 1) r = zm(r * a, m);
        a = zm(a * a, m);
        e >>= 1;
    }
    return r;
}
/* inverse of an odd unit mod a power of two */
static int64_t zm_inv(int64_t a, int64_t m) {
    for (int64_t x = 1; x < m; x += 2)
        if (zm(a * x, m) == 1) return x;
    return -1;
}

/* ------------------------------------------------------------------------ */
/* ChaCha, rand_chacha 0.3 word order (block counter in words 12-13, stream   */
/* id in words 14-15, four blocks buffered, next_u64 = two consecutive words */
/* low first).                                                               */
/* ------------------------------------------------------------------------ */
#define ROTL32(a, b) (((a) << (b)) | ((a) >> (32 - (b))))
#define CHACHA_QR(a, b, c, d)                    \
    do {                                         \
        a += b; d ^= a; d = ROTL32(d, 16);       \
        c += d; b ^= c; b = ROTL32(b, 12);       \
        a += b; d ^= a; d = ROTL32(d, 8);        \
        c += d; b ^= c; b = ROTL32(b, 7);        \
    } while (0)

void orc_chacha_block(const uint32_t in[16], uint32_t out[16], int rounds) {
    uint32_t x[16];
    memcpy(x, in, sizeof(x));
    for (int r = 0; r < rounds; r += 2) {
        CHACHA_QR(x[0], x[4], x[8], x[12]);
        CHACHA_QR(x[1], x[5], x[9], x[13]);
        CHACHA_QR(x[2], x[6], x[10], x[14]);
        CHACHA_QR(x[3], x[7], x[11], x[15]);
        CHACHA_QR(x[0], x[5], x[10], x[15]);
        CHACHA_QR(x[1], x[6], x[11], x[12]);
        CHACHA_QR(x[2], x[7], x[8], x[13]);
        CHACHA_QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

struct orc_rng {
    uint32_t st[16];
    uint32_t buf[64];
    int idx;
};

orc_rng* orc_rng_new(const uint8_t seed[32]) {
    orc_rng* r = (orc_rng*)calloc(1, sizeof(orc_rng));
    r->st[0] = 0x61707865u; r->st[1] = 0x3320646eu; r->st[2] = 0x79622d32u; r->st[3] = 0x6b206574u;
    for (int i = 0; i < 8; ++i)
        r->st[4 + i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
                       ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
    r->idx = 64;
    return r;
}
void orc_rng_free(orc_rng* r) { free(r); }
static void rng_refill(orc_rng* r) {
    for (int b = 0; b < 4; ++b) {
        orc_chacha_block(r->st, r->buf + 16 * b, 12); /* StdRng = ChaCha12 */
        if (++r->st[12] == 0) ++r->st[13];
    }
    r->idx = 0;
}
uint32_t orc_rng_next_u32(orc_rng* r) {
    if (r->idx >= 64) rng_refill(r);
    return r->buf[r->idx++];
}
uint64_t orc_rng_next_u64(orc_rng* r) {
    uint64_t lo = orc_rng_next_u32(r);
    uint64_t hi = orc_rng_next_u32(r);
    return lo | (hi << 32);
}
/* standard normal via Box-Muller (client noise only; not rand_distr's ziggurat) */
static double rng_gauss(orc_rng* r) {
    double u1 = ((double)(orc_rng_next_u64(r) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = (double)(orc_rng_next_u64(r) >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* ------------------------------------------------------------------------ */
/* SEEDS, base_pir.rs:53-71: SHA3-256 of six UTF-8 titles (pinned by a test) */
/* ------------------------------------------------------------------------ */
static const char* SEED_HEX[6] = {
    "dad20c142deb8bb4d66320ee07b1c190ec25c0a1541ccaf598fa22ffdfdd68ba",
    "d3c5f6eeb33a4cf333516ce273d57476b5796cce9cf337a19c84bcbea43d2cb5",
    "9527b86c30dc30cf1bb174c3cd6cd7507209daa54bbdb999041400083d68248a",
    "4ca2f6478a581191676741002a00edd1ac97f930980b7fbef033b448234b216b",
    "98fc62c517520e5d29fc307e73efb10e1418aa070e32364f893fa8bd92130327",
    "b361dc9434dbbae928e5c964dc18025402fe81aaf74917208a6e6fd955702a21",
};
void orc_seed(int idx, uint8_t out[32]) {
    const char* h = SEED_HEX[idx];
    int n = 0;
    for (int i = 0; h[i] && n < 32;) {
        if (h[i] == ' ') { ++i; continue; }
        char c0 = h[i], c1 = h[i + 1];
        int v0 = c0 <= '9' ? c0 - '0' : c0 - 'a' + 10;
        int v1 = c1 <= '9' ? c1 - '0' : c1 - 'a' + 10;
        out[n++] = (uint8_t)(v0 * 16 + v1);
        i += 2;
    }
}

/* ------------------------------------------------------------------------ */
/* BFV primitives (bfv.rs)                                                   */
/* ------------------------------------------------------------------------ */
/* bfv.rs:46-49: a-part from a 32-byte seed. Sampling assumption (parity
 * unpinned): rejection of next_u64 & (2^51 - 1) >= q, coefficients 0..N-1. */
void orc_expand(int logN, const uint8_t seed[32], uint64_t* out) {
    int N = 1 << logN;
    orc_rng* r = orc_rng_new(seed);
    const uint64_t mask = (1ULL << 51) - 1;
    for (int i = 0; i < N; ++i) {
        uint64_t x;
        do { x = orc_rng_next_u64(r) & mask; } while (x >= ORC_Q);
        out[i] = x;
    }
    orc_rng_free(r);
}

/* bfv.rs:36-44 */
void orc_gen_sk(int logN, orc_rng* rng, int hwt, int64_t* sk) {
    int N = 1 << logN;
    memset(sk, 0, sizeof(int64_t) * N);
    for (int i = 0; i < hwt; ++i) {
        uint64_t idx = orc_rng_next_u64(rng) % (uint64_t)N;
        sk[idx] = (int64_t)(orc_rng_next_u32(rng) % 3) - 1;
    }
}

/* a * s in Z_q[X]/(X^N + 1) for a ternary s */
void orc_poly_mul_sparse(int logN, const uint64_t* a, const int64_t* sk, uint64_t* out) {
    int N = 1 << logN;
    memset(out, 0, sizeof(uint64_t) * N);
    for (int j = 0; j < N; ++j) {
        if (sk[j] == 0) continue;
        for (int i = 0; i < N; ++i) {
            int idx = i + j;
            int neg = sk[j] < 0;
            if (idx >= N) { idx -= N; neg = !neg; }
            out[idx] = neg ? q_sub(out[idx], a[i]) : q_add(out[idx], a[i]);
        }
    }
}

/* round(lift_c(m) * q / t) mod q, as the rescale closure of bfv.rs:76-81 */
static uint64_t bfv_scale_up(uint64_t m, uint64_t t) {
    int64_t mc = (int64_t)(m % t);
    if ((uint64_t)mc > t / 2) mc -= (int64_t)t;
    i128 up = (i128)mc * (i128)ORC_Q;
    int64_t d = (int64_t)(up % (i128)t);
    if (d < 0) d += (int64_t)t;
    if ((uint64_t)d > t / 2) d -= (int64_t)t;
    return q_from_i128((up - d) / (i128)t);
}

/* bfv.rs:72-85: b = e - a*s + Delta(m) */
void orc_enc_sym_b(int logN, const uint8_t seed_a[32], orc_rng* rng, const int64_t* sk,
                   const uint16_t* m, uint64_t t, double sigma, uint64_t* b) {
    int N = 1 << logN;
    uint64_t* a = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* as = (uint64_t*)malloc(sizeof(uint64_t) * N);
    orc_expand(logN, seed_a, a);
    for (int i = 0; i < N; ++i) {
        int64_t e = (int64_t)round(rng_gauss(rng) * sigma);
        b[i] = q_from_i128(e);
    }
    orc_poly_mul_sparse(logN, a, sk, as);
    for (int i = 0; i < N; ++i) b[i] = q_add(q_sub(b[i], as[i]), bfv_scale_up(m[i], t));
    free(a);
    free(as);
}

/* bfv.rs:103-124: gk_i = e - a_i*s + B^i * s(X^g), a_i = expand(SEEDS[i]) */
void orc_gen_gk_b(int logN, orc_rng* rng, const int64_t* sk, int g, double sigma, uint64_t* gk) {
    int N = 1 << logN;
    int64_t m = 2 * N;
    int64_t ginv = zm_inv(g, m);
    int64_t* gal_s = (int64_t*)malloc(sizeof(int64_t) * N);
    for (int i = 0; i < N; ++i) {
        int64_t ni = zm((int64_t)i * ginv, m);
        gal_s[i] = ni > N ? -sk[ni - N] : sk[ni];
    }
    uint64_t* a = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* as = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (int d = 0; d < GK_DIGITS; ++d) {
        uint8_t seed[32];
        orc_seed(d, seed); /* UsedSeeds::get_gk_seeds, base_pir.rs:195-197 */
        orc_expand(logN, seed, a);
        uint64_t* b = gk + (size_t)d * N;
        for (int i = 0; i < N; ++i) b[i] = q_from_i128((int64_t)round(rng_gauss(rng) * sigma));
        orc_poly_mul_sparse(logN, a, sk, as);
        uint64_t scale = 1;
        for (int k = 0; k < d; ++k) scale = q_mul(scale, 1ULL << LOG2_B);
        for (int i = 0; i < N; ++i) {
            uint64_t pay = q_mul(q_from_i128(gal_s[i]), scale);
            b[i] = q_add(q_sub(b[i], as[i]), pay);
        }
    }
    free(gal_s);
    free(a);
    free(as);
}

/* bfv.rs:88-101: round(t * (a*s + b) / q) mod t */
void orc_dec(int logN, const int64_t* sk, const uint64_t* a, const uint64_t* b, uint64_t t, uint16_t* m) {
    int N = 1 << logN;
    uint64_t* as = (uint64_t*)malloc(sizeof(uint64_t) * N);
    orc_poly_mul_sparse(logN, a, sk, as);
    for (int i = 0; i < N; ++i) {
        int64_t c = q_center(q_add(as[i], b[i]));
        i128 up = (i128)c * (i128)t;
        i128 d = up % (i128)ORC_Q;
        if (d < 0) d += ORC_Q;
        if (d > (i128)(ORC_Q / 2)) d -= ORC_Q;
        i128 down = (up - d) / (i128)ORC_Q;
        int64_t r = (int64_t)(down % (i128)t);
        if (r < 0) r += (int64_t)t;
        m[i] = (uint16_t)r;
    }
    free(as);
}

/* bfv.rs:51-58: ((round(lift_c(c) * 2^B / q) + 2^B) mod 2^B); q odd, no ties */
void orc_mod_switch_encode(int logN, int bits, const uint64_t* x, uint64_t* out) {
    int N = 1 << logN;
    for (int i = 0; i < N; ++i) {
        i128 num = (i128)q_center(x[i]) * ((i128)1 << bits);
        i128 fl = num / (i128)ORC_Q;
        i128 rem = num - fl * (i128)ORC_Q;
        if (rem < 0) { rem += ORC_Q; fl -= 1; }
        if (2 * rem > (i128)ORC_Q) fl += 1;
        i128 r = (fl + ((i128)1 << bits)) % ((i128)1 << bits);
        out[i] = (uint64_t)r;
    }
}
/* bfv.rs:60-69: floor(c * q / 2^B) mod q */
void orc_mod_switch_decode(int logN, int bits, const uint64_t* x, uint64_t* out) {
    int N = 1 << logN;
    for (int i = 0; i < N; ++i) out[i] = (uint64_t)(((u128)x[i] * ORC_Q) >> bits) % ORC_Q;
}

/* ------------------------------------------------------------------------ */
/* NTTs in the reference convention.                                         */
/* weighted_ntt (base_pir.rs:453-467): slot i = f(zeta^-(2i+1)), natural    */
/* order (pinned by the KATs of base_pir.rs:1281-1411). The inner DFT is a   */
/* plain iterative Cooley-Tukey computing out[k] = sum_j v_j root^(jk).      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int logN, N, m;
    md_t M[2];
    uint64_t zeta[2];
    uint64_t* zp[2];   /* zeta^k, k in [0, m) */
    uint64_t* wf[2];   /* (zeta^-2)^k, k < N/2 */
    uint64_t* wi[2];   /* (zeta^2)^k, k < N/2 */
    uint64_t ninv[2];
    int* brvN;
} ring_t;

static uint64_t find_prim_root_2pow(const md_t* M, uint64_t order) {
    for (uint64_t c = 2;; ++c) {
        uint64_t z = md_pow(M, c, (M->p - 1) / order);
        if (md_pow(M, z, order / 2) == M->p - 1) return z;
    }
}

static void ring_init(ring_t* R, int logN, int root_pow) {
    R->logN = logN;
    R->N = 1 << logN;
    R->m = 2 * R->N;
    R->M[0] = md_make(ORC_P1);
    R->M[1] = md_make(ORC_P2);
    for (int r = 0; r < 2; ++r) {
        const md_t* M = &R->M[r];
        uint64_t z0 = find_prim_root_2pow(M, (uint64_t)R->m);
        R->zeta[r] = md_pow(M, z0, (uint64_t)root_pow);
        R->zp[r] = (uint64_t*)malloc(sizeof(uint64_t) * R->m);
        R->zp[r][0] = 1;
        for (int k = 1; k < R->m; ++k) R->zp[r][k] = md_mul(M, R->zp[r][k - 1], R->zeta[r]);
        int h = R->N / 2 > 0 ? R->N / 2 : 1;
        R->wf[r] = (uint64_t*)malloc(sizeof(uint64_t) * h);
        R->wi[r] = (uint64_t*)malloc(sizeof(uint64_t) * h);
        for (int k = 0; k < h; ++k) {
            R->wf[r][k] = R->zp[r][zm(-2LL * k, R->m)];
            R->wi[r][k] = R->zp[r][zm(2LL * k, R->m)];
        }
        R->ninv[r] = md_inv(M, (uint64_t)R->N);
    }
    R->brvN = (int*)malloc(sizeof(int) * R->N);
    for (int i = 0; i < R->N; ++i) {
        int v = 0;
        for (int b = 0; b < logN; ++b)
            if (i & (1 << b)) v |= 1 << (logN - 1 - b);
        R->brvN[i] = v;
    }
}
static void ring_free(ring_t* R) {
    for (int r = 0; r < 2; ++r) {
        free(R->zp[r]);
        free(R->wf[r]);
        free(R->wi[r]);
    }
    free(R->brvN);
}

/* out[k] = sum_j a_j root^(jk), natural order in and out; w[k] = root^k (k < n/2) */
static void dft(const md_t* M, uint64_t* a, int n, const uint64_t* w) {
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { uint64_t t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len)
            for (int j = 0; j < half; ++j) {
                uint64_t u = a[i + j];
                uint64_t v = md_mul(M, a[i + j + half], w[j * step]);
                a[i + j] = md_add(M, u, v);
                a[i + j + half] = md_sub(M, u, v);
            }
    }
}

/* base_pir.rs:453-467 */
static void weighted_ntt(const ring_t* R, int r, uint64_t* x) {
    const md_t* M = &R->M[r];
    for (int j = 0; j < R->N; ++j) x[j] = md_mul(M, x[j], R->zp[r][zm(-(int64_t)j, R->m)]);
    dft(M, x, R->N, R->wf[r]);
}
/* base_pir.rs:472-483 */
static void inv_weighted_ntt(const ring_t* R, int r, uint64_t* x) {
    const md_t* M = &R->M[r];
    dft(M, x, R->N, R->wi[r]);
    for (int j = 0; j < R->N; ++j) x[j] = md_mul(M, md_mul(M, x[j], R->ninv[r]), R->zp[r][j]);
}

/* weighted NTT / inverse of one ring element (checked against a brute-force
 * negacyclic DFT by the tests); returns the root zeta used */
uint64_t orc_ntt(int logN, int modulus, int root_pow, int inverse, uint64_t* x) {
    ring_t R;
    ring_init(&R, logN, root_pow > 0 ? root_pow : 1);
    if (inverse) inv_weighted_ntt(&R, modulus, x);
    else weighted_ntt(&R, modulus, x);
    uint64_t z = R.zeta[modulus];
    ring_free(&R);
    return z;
}

/* ------------------------------------------------------------------------ */
/* shard                                                                     */
/* ------------------------------------------------------------------------ */
enum { ARR_TE = 0, ARR_IP0 = 1, ARR_IP1 = 2 };

struct orc_shard {
    ring_t R;
    int logN, N, m, l, L, conj, k, seeds, root_pow;
    uint64_t* wl;            /* coset DFT twiddles (length l, modulus p1) */
    uint32_t* arr[3][2][2];  /* [array][modulus][stream]: [row][lane][slot] */
    uint64_t* reply_a[2];    /* [modulus]: [lane][slot] */
    int has_db;
};

/* IndexGroup, base_pir.rs:133-183 */
static int ig_loop_len(int N, int l) { return l < N ? l : N / 2; }
static int ig_conj(int N, int l) { return l == N; }
static int ig_qbp(int N, int l) { return N / 2 / ig_loop_len(N, l); }

orc_shard* orc_shard_create(int logN, int group_order, int used_seeds, int root_pow) {
    int N = 1 << logN;
    if (group_order <= 0 || N % group_order != 0) return NULL;
    if (group_order != N && 2 * group_order > N) return NULL; /* base_pir.rs:150-154 */
    orc_shard* s = (orc_shard*)calloc(1, sizeof(orc_shard));
    ring_init(&s->R, logN, root_pow > 0 ? root_pow : 1);
    s->logN = logN;
    s->N = N;
    s->m = 2 * N;
    s->l = group_order;
    s->L = ig_loop_len(N, group_order);
    s->conj = ig_conj(N, group_order);
    s->k = ig_qbp(N, group_order);
    s->seeds = used_seeds;
    s->root_pow = root_pow > 0 ? root_pow : 1;
    size_t rows = (size_t)s->L * LANES * N;
    for (int a = 0; a < 3; ++a)
        for (int r = 0; r < 2; ++r)
            for (int st = 0; st < (s->conj ? 2 : 1); ++st) s->arr[a][r][st] = (uint32_t*)calloc(rows, sizeof(uint32_t));
    for (int r = 0; r < 2; ++r) s->reply_a[r] = (uint64_t*)calloc((size_t)LANES * N, sizeof(uint64_t));
    /* length-l DFT with root zeta^-(m/l): the len_subgroup_order_ntts table of
     * base_pir.rs:343-347 in the convention out[i] = sum_j v_j root^-(ij) */
    int l = group_order, h = l / 2 > 0 ? l / 2 : 1;
    s->wl = (uint64_t*)malloc(sizeof(uint64_t) * h);
    for (int k = 0; k < h; ++k) s->wl[k] = s->R.zp[0][zm(-(int64_t)(s->m / l) * k, s->m)];
    return s;
}

void orc_shard_free(orc_shard* s) {
    if (!s) return;
    for (int a = 0; a < 3; ++a)
        for (int r = 0; r < 2; ++r)
            for (int st = 0; st < 2; ++st) free(s->arr[a][r][st]);
    for (int r = 0; r < 2; ++r) free(s->reply_a[r]);
    free(s->wl);
    ring_free(&s->R);
    free(s);
}
int orc_shard_loop_len(const orc_shard* s) { return s->L; }
int orc_shard_requires_conj(const orc_shard* s) { return s->conj; }

static inline uint32_t* row_ptr(const orc_shard* s, int a, int r, int st, int row, int lane) {
    return s->arr[a][r][st] + ((size_t)row * LANES + lane) * s->N;
}

/* base_pir.rs:397-444 (coset_ntt) on one row of length l, modulus p1 */
static void coset_ntt_row(const orc_shard* s, uint64_t* v, uint64_t* tmp, int64_t zeta_power) {
    const md_t* M = &s->R.M[0];
    int l = s->l, m = s->m, N = s->N;
    int64_t step = m / l;
    int64_t coset = zm(zeta_power, m) % step;
    for (int j = 0; j < l; ++j) v[j] = md_mul(M, v[j], s->R.zp[0][zm(-coset * j, m)]);
    dft(M, v, l, s->wl);
    /* permute: new[i] = v[galois_to_index(...)] */
    int64_t zneg = zm(-zeta_power, m);
    for (int i = 0; i < l; ++i) {
        int64_t ge;
        if (s->conj)
            ge = i < N / 2 ? zm(zneg * zm_pow(GEN_H, i, m), m) : zm(zeta_power * zm_pow(GEN_H, i - N / 2, m), m);
        else
            ge = zm(zneg * zm_pow(GEN_H, i, m), m);
        int64_t lifted = zm(-ge, m) - coset;
        int64_t idx = lifted / step; /* exact: asserted by checked_div in the reference */
        tmp[i] = v[idx];
    }
    memcpy(v, tmp, sizeof(uint64_t) * l);
}

/* base_pir.rs:662-772 */
int orc_shard_set_transformed_entries(orc_shard* s, const uint16_t* entries, size_t n_entries) {
    int N = s->N, l = s->l, m = s->m, L = s->L;
    if (n_entries != (size_t)l * LANES) return -3; /* base_pir.rs:693 */
    const ring_t* R = &s->R;
    const md_t* M1 = &R->M[0];
    uint64_t t = ORC_T;
    /* mat[j][lane][slot] over p1 */
    uint64_t* mat = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)l * LANES * N);
    /* STEP 1+2: columns = coefficient vectors, then weighted NTT (p1) */
#pragma omp parallel for schedule(static)
    for (int jl = 0; jl < l * LANES; ++jl) {
        uint64_t* col = mat + (size_t)jl * N;
        const uint16_t* e = entries + (size_t)jl * N; /* entry 8j+lane -> column j, lane (base_pir.rs:679-683) */
        for (int i = 0; i < N; ++i) col[i] = e[i];
        weighted_ntt(R, 0, col);
    }
    /* STEP 3: row-wise coset NTT, zeta_power = (2s+1) * (-query_base_power) */
#pragma omp parallel
    {
        uint64_t* v = (uint64_t*)malloc(sizeof(uint64_t) * l);
        uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * l);
#pragma omp for schedule(static)
        for (int sl = 0; sl < N * LANES; ++sl) {
            int slot = sl / LANES, lane = sl % LANES;
            for (int j = 0; j < l; ++j) v[j] = mat[((size_t)j * LANES + lane) * N + slot];
            int64_t zp = zm((int64_t)(2 * slot + 1) * -(int64_t)s->k, m);
            coset_ntt_row(s, v, tmp, zp);
            for (int j = 0; j < l; ++j) mat[((size_t)j * LANES + lane) * N + slot] = v[j];
        }
        free(v);
        free(tmp);
    }
    /* STEP 4: per column j, new[i] = old[(g^-j (2i+1) - 1)/2] */
    int64_t ginv = zm_inv(GEN_H, m);
#pragma omp parallel
    {
        uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * N);
#pragma omp for schedule(static)
        for (int jl = 0; jl < l * LANES; ++jl) {
            int j = jl / LANES;
            uint64_t* col = mat + (size_t)jl * N;
            int64_t au = zm_pow(ginv, j, m);
            for (int i = 0; i < N; ++i) tmp[i] = col[(zm(au * (2 * i + 1), m) - 1) / 2];
            memcpy(col, tmp, sizeof(uint64_t) * N);
        }
        free(tmp);
    }
    /* STEP 5 (reverse within <g> and -<g>) folded into the column index of
     * STEP 6: final column c takes column src(c). */
    int64_t scale = 1;
    {
        /* l^-1 mod t */
        for (int64_t x = 1; x < (int64_t)t; ++x)
            if ((x * l) % (int64_t)t == 1) { scale = x; break; }
    }
#pragma omp parallel
    {
        uint64_t* c1 = (uint64_t*)malloc(sizeof(uint64_t) * N);
        uint64_t* c2 = (uint64_t*)malloc(sizeof(uint64_t) * N);
#pragma omp for schedule(static)
        for (int cl = 0; cl < l * LANES; ++cl) {
            int c = cl / LANES, lane = cl % LANES;
            int src;
            if (s->conj) src = c < N / 2 ? N / 2 - 1 - c : N / 2 + (N - 1 - c);
            else src = l - 1 - c;
            memcpy(c1, mat + ((size_t)src * LANES + lane) * N, sizeof(uint64_t) * N);
            /* STEP 6: iNTT, smallest lift, times l^-1 mod t, smallest lift, embed, NTT in p1 and p2 */
            inv_weighted_ntt(R, 0, c1);
            for (int i = 0; i < N; ++i) {
                int64_t x = md_center(M1, c1[i]);
                int64_t xt = zm(x, (int64_t)t);
                int64_t y = zm(xt * scale, (int64_t)t);
                if (y > (int64_t)t / 2) y -= (int64_t)t;
                c1[i] = md_from_i64(&R->M[0], y);
                c2[i] = md_from_i64(&R->M[1], y);
            }
            weighted_ntt(R, 0, c1);
            weighted_ntt(R, 1, c2);
            int st = c < L ? 0 : 1;
            int row = c < L ? c : c - L;
            uint32_t* d1 = row_ptr(s, ARR_TE, 0, st, row, lane);
            uint32_t* d2 = row_ptr(s, ARR_TE, 1, st, row, lane);
            for (int i = 0; i < N; ++i) { d1[i] = (uint32_t)c1[i]; d2[i] = (uint32_t)c2[i]; }
        }
        free(c1);
        free(c2);
    }
    free(mat);
    (void)M1;
    return 0;
}

/* to_simd_element, base_pir.rs:865-877: coefficients mod q -> NTT mod p_r */
static void to_ntt(const ring_t* R, int r, const uint64_t* x, uint64_t* out) {
    for (int i = 0; i < R->N; ++i) out[i] = x[i] % R->M[r].p;
    weighted_ntt(R, r, out);
}

/* Galois lookup table, base_pir.rs:974-978 / partial_gal :509-515:
 * gt[i] = (g^-1 (2i+1) mod 2N - 1) / 2 */
static void galois_table(int N, int* gt) {
    int64_t m = 2 * N, ginv = zm_inv(GEN_H, m);
    for (int i = 0; i < N; ++i) gt[i] = (int)((zm(ginv * (2 * i + 1), m) - 1) / 2);
}

/* base_pir.rs:496-594 for one lane: returns <Decomp(gal(a)), gk_a> and the digits */
static void partial_gal(const ring_t* R, const int* gt, uint64_t* const cur[2], uint64_t* const gka[2][2],
                        uint64_t* out[2], uint64_t* dig[2][2], uint64_t* work[2]) {
    int N = R->N;
    const md_t* M1 = &R->M[0];
    const md_t* M2 = &R->M[1];
    for (int r = 0; r < 2; ++r) {
        for (int i = 0; i < N; ++i) work[r][gt[i]] = cur[r][i];
        inv_weighted_ntt(R, r, work[r]);
    }
    const uint64_t q2inv = md_inv(M1, ORC_P2 % ORC_P1);
    const uint64_t q1inv = md_inv(M2, ORC_P1 % ORC_P2);
    const int64_t Q = (int64_t)ORC_Q;
    for (int i = 0; i < N; ++i) {
        uint64_t e1 = md_mul(M1, work[0][i], q2inv);
        uint64_t e2 = md_mul(M2, work[1][i], q1inv);
        int64_t x = (int64_t)(e1 * ORC_P2 + e2 * ORC_P1);
        if (x > Q) x -= Q;
        if (x > Q / 2) x -= Q;
        int64_t hi = (x + (1LL << (LOG2_B - 1))) >> LOG2_B;
        int64_t lo = x - (hi << LOG2_B);
        int64_t hi2 = (hi + (1LL << (LOG2_B - 1))) >> LOG2_B;
        int64_t lo2 = hi - (hi2 << LOG2_B);
        dig[0][0][i] = md_from_i64(M1, lo);
        dig[0][1][i] = md_from_i64(M2, lo);
        dig[1][0][i] = md_from_i64(M1, lo2);
        dig[1][1][i] = md_from_i64(M2, lo2);
    }
    for (int d = 0; d < 2; ++d)
        for (int r = 0; r < 2; ++r) weighted_ntt(R, r, dig[d][r]);
    for (int r = 0; r < 2; ++r) {
        const md_t* M = &R->M[r];
        for (int i = 0; i < N; ++i)
            out[r][i] = md_add(M, md_mul(M, dig[0][r][i], gka[0][r][i]), md_mul(M, dig[1][r][i], gka[1][r][i]));
    }
}

/* base_pir.rs:782-863 */
int orc_shard_preprocess(orc_shard* s) {
    int N = s->N, L = s->L;
    const ring_t* R = &s->R;
    int nst = s->conj ? 2 : 1;
    uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* gka[2][2];
    uint64_t* qa[2][2];
    for (int d = 0; d < 2; ++d) {
        uint8_t seed[32];
        orc_seed(d, seed);
        orc_expand(s->logN, seed, tmp);
        for (int r = 0; r < 2; ++r) {
            gka[d][r] = (uint64_t*)malloc(sizeof(uint64_t) * N);
            to_ntt(R, r, tmp, gka[d][r]);
        }
    }
    for (int st = 0; st < nst; ++st) {
        uint8_t seed[32];
        orc_seed((s->seeds == 0 ? 2 : 4) + st, seed); /* base_pir.rs:199-204 */
        orc_expand(s->logN, seed, tmp);
        for (int r = 0; r < 2; ++r) {
            qa[st][r] = (uint64_t*)malloc(sizeof(uint64_t) * N);
            to_ntt(R, r, tmp, qa[st][r]);
        }
    }
    int* gt = (int*)malloc(sizeof(int) * N);
    galois_table(N, gt);
    for (int r = 0; r < 2; ++r) memset(s->reply_a[r], 0, sizeof(uint64_t) * LANES * N);
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int st = 0; st < nst; ++st)
        for (int lane = 0; lane < LANES; ++lane) {
            uint64_t* buf = (uint64_t*)calloc((size_t)10 * N, sizeof(uint64_t));
            uint64_t* a[2] = {buf, buf + N};
            uint64_t* ga[2] = {buf + 2 * N, buf + 3 * N};
            uint64_t* dig[2][2] = {{buf + 4 * N, buf + 5 * N}, {buf + 6 * N, buf + 7 * N}};
            uint64_t* work[2] = {buf + 8 * N, buf + 9 * N};
            for (int i = 0; i < L; ++i) {
                partial_gal(R, gt, a, gka, ga, dig, work);
                for (int d = 0; d < 2; ++d)
                    for (int r = 0; r < 2; ++r) {
                        uint32_t* dst = row_ptr(s, ARR_IP0 + d, r, st, i, lane);
                        for (int j = 0; j < N; ++j) dst[j] = (uint32_t)dig[d][r][j];
                    }
                for (int r = 0; r < 2; ++r) {
                    const md_t* M = &R->M[r];
                    const uint32_t* te = row_ptr(s, ARR_TE, r, st, i, lane);
                    for (int j = 0; j < N; ++j) a[r][j] = md_add(M, ga[r][j], md_mul(M, qa[st][r][j], te[j]));
                }
            }
#pragma omp critical
            for (int r = 0; r < 2; ++r) {
                const md_t* M = &R->M[r];
                uint64_t* ra = s->reply_a[r] + (size_t)lane * N;
                for (int j = 0; j < N; ++j) ra[j] = md_add(M, ra[j], a[r][j]);
            }
            free(buf);
        }
    for (int d = 0; d < 2; ++d)
        for (int r = 0; r < 2; ++r) free(gka[d][r]);
    for (int st = 0; st < nst; ++st)
        for (int r = 0; r < 2; ++r) free(qa[st][r]);
    free(gt);
    free(tmp);
    s->has_db = 1;
    return 0;
}

int orc_shard_set_db(orc_shard* s, const uint16_t* entries, size_t n_entries) {
    int rc = orc_shard_set_transformed_entries(s, entries, n_entries);
    if (rc) return rc;
    return orc_shard_preprocess(s);
}

/* base_pir.rs:905-935: iNTT and CRT Z_p1 x Z_p2 -> Z_q, one lane */
static void from_simd(const ring_t* R, uint64_t* x1, uint64_t* x2, uint64_t* out) {
    const md_t* M1 = &R->M[0];
    const md_t* M2 = &R->M[1];
    inv_weighted_ntt(R, 0, x1);
    inv_weighted_ntt(R, 1, x2);
    const uint64_t q2inv = md_inv(M1, ORC_P2 % ORC_P1);
    const uint64_t q1inv = md_inv(M2, ORC_P1 % ORC_P2);
    for (int i = 0; i < R->N; ++i) {
        uint64_t e1 = md_mul(M1, x1[i], q2inv);
        uint64_t e2 = md_mul(M2, x2[i], q1inv);
        out[i] = (e1 * ORC_P2 + e2 * ORC_P1) % ORC_Q;
    }
}

/* base_pir.rs:897-903 */
void orc_shard_get_a(const orc_shard* s, uint64_t* out) {
    int N = s->N;
    uint64_t* x1 = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* x2 = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (int lane = 0; lane < LANES; ++lane) {
        memcpy(x1, s->reply_a[0] + (size_t)lane * N, sizeof(uint64_t) * N);
        memcpy(x2, s->reply_a[1] + (size_t)lane * N, sizeof(uint64_t) * N);
        from_simd(&s->R, x1, x2, out + (size_t)lane * N);
    }
    free(x1);
    free(x2);
}

/* base_pir.rs:958-1078 with main_pir_loop :1085-1247 (all values kept fully
 * reduced; the lazy 341-row window of the reference is an overflow schedule
 * that does not change any value mod p) */
int orc_shard_perform_pir_split(const orc_shard* s, const uint64_t* qb, const uint64_t* qbc,
                                const uint64_t* gkb, uint64_t* out) {
    int N = s->N, L = s->L;
    if (s->conj && !qbc) return -2; /* base_pir.rs:965 */
    const ring_t* R = &s->R;
    int nst = s->conj ? 2 : 1;
    /* query vectors in NTT form: [r][stream][lane][N], gk: [r][d][lane][N] */
    size_t LN = (size_t)LANES * N;
    uint64_t* qv = (uint64_t*)malloc(sizeof(uint64_t) * 2 * 2 * LN);
    uint64_t* gv = (uint64_t*)malloc(sizeof(uint64_t) * 2 * 2 * LN);
    for (int r = 0; r < 2; ++r)
        for (int lane = 0; lane < LANES; ++lane) {
            to_ntt(R, r, qb + (size_t)lane * N, qv + ((size_t)(r * 2 + 0) * LANES + lane) * N);
            if (s->conj) to_ntt(R, r, qbc + (size_t)lane * N, qv + ((size_t)(r * 2 + 1) * LANES + lane) * N);
            for (int d = 0; d < 2; ++d)
                to_ntt(R, r, gkb + ((size_t)lane * 2 + d) * N, gv + ((size_t)(r * 2 + d) * LANES + lane) * N);
        }
    int* gt = (int*)malloc(sizeof(int) * N);
    galois_table(N, gt);
    uint64_t* tot = (uint64_t*)calloc(2 * LN, sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int r = 0; r < 2; ++r)
        for (int lane = 0; lane < LANES; ++lane) {
            const md_t* M = &R->M[r];
            uint64_t* sum = (uint64_t*)calloc(N, sizeof(uint64_t));
            uint64_t* nsum = (uint64_t*)calloc(N, sizeof(uint64_t));
            uint64_t* acc = tot + ((size_t)r * LANES + lane) * N;
            const uint64_t* g0 = gv + ((size_t)(r * 2 + 0) * LANES + lane) * N;
            const uint64_t* g1 = gv + ((size_t)(r * 2 + 1) * LANES + lane) * N;
            for (int st = 0; st < nst; ++st) {
                const uint64_t* qq = qv + ((size_t)(r * 2 + st) * LANES + lane) * N;
                memset(sum, 0, sizeof(uint64_t) * N);
                for (int i = 0; i < L; ++i) {
                    const uint32_t* ip0 = row_ptr(s, ARR_IP0, r, st, i, lane);
                    const uint32_t* ip1 = row_ptr(s, ARR_IP1, r, st, i, lane);
                    const uint32_t* te = row_ptr(s, ARR_TE, r, st, i, lane);
                    for (int j = 0; j < N; ++j) {
                        uint64_t v = sum[j] + md_mul(M, ip0[j], g0[j]) + md_mul(M, ip1[j], g1[j]) + md_mul(M, te[j], qq[j]);
                        nsum[gt[j]] = md_red(M, v);
                    }
                    uint64_t* t = sum; sum = nsum; nsum = t;
                }
                for (int j = 0; j < N; ++j) acc[j] = md_add(M, acc[j], sum[j]);
            }
            /* permute: res[i] = acc[gt[i]] */
            for (int j = 0; j < N; ++j) nsum[j] = acc[gt[j]];
            memcpy(acc, nsum, sizeof(uint64_t) * N);
            free(sum);
            free(nsum);
        }
    for (int lane = 0; lane < LANES; ++lane)
        from_simd(R, tot + (size_t)lane * N, tot + ((size_t)LANES + lane) * N, out + (size_t)lane * N);
    free(qv);
    free(gv);
    free(gt);
    free(tot);
    return 0;
}

/* The operands of main_pir_loop (base_pir.rs:1085-1247) in the reference's
 * memory layout, for the CPU-baseline port (cpu_baseline.c): each slot holds
 * its 8 lanes side by side (CompressedZnx8El, 32 B). arrays: [r][st][IPGK0,
 * IPGK1, TE][row][slot][lane] u32; vecs: [r][st][gk0, gk1, query][slot][lane]
 * u32 (the conjugate query for st = 1) in the NTT domain; gal: the Galois
 * lookup table (base_pir.rs:974-978). Query k goes to lane k, as in
 * perform_pir_split. */
int orc_shard_k1_operands(const orc_shard* s, const uint64_t* qb, const uint64_t* qbc, const uint64_t* gkb,
                          uint32_t* arrays, uint32_t* vecs, int* gal) {
    const int N = s->N, L = s->L, nst = s->conj ? 2 : 1;
    if (s->conj && !qbc) return -2;
    const int amap[3] = {ARR_IP0, ARR_IP1, ARR_TE};
    for (int r = 0; r < 2; ++r)
        for (int st = 0; st < nst; ++st)
            for (int a = 0; a < 3; ++a) {
                uint32_t* dst = arrays + ((size_t)(r * nst + st) * 3 + a) * L * N * LANES;
#pragma omp parallel for
                for (int row = 0; row < L; ++row)
                    for (int lane = 0; lane < LANES; ++lane) {
                        const uint32_t* src = row_ptr(s, amap[a], r, st, row, lane);
                        for (int j = 0; j < N; ++j) dst[((size_t)row * N + j) * LANES + lane] = src[j];
                    }
            }
    uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (int r = 0; r < 2; ++r)
        for (int st = 0; st < nst; ++st)
            for (int v = 0; v < 3; ++v)
                for (int lane = 0; lane < LANES; ++lane) {
                    const uint64_t* x = v < 2 ? gkb + ((size_t)lane * 2 + v) * N
                                              : (st ? qbc : qb) + (size_t)lane * N;
                    to_ntt(&s->R, r, x, t);
                    uint32_t* dst = vecs + ((size_t)(r * nst + st) * 3 + v) * N * LANES;
                    for (int j = 0; j < N; ++j) dst[(size_t)j * LANES + lane] = (uint32_t)t[j];
                }
    free(t);
    galois_table(N, gal);
    return 0;
}

/* the end of perform_pir_split (base_pir.rs:1072-1077) after main_pir_loop:
 * sums [r][slot][lane] (both streams added, reduced mod p_r) are gathered by
 * the Galois table and taken out of the NTT domain: out [8][N] mod q */
int orc_shard_k1_finish(const orc_shard* s, const uint64_t* sums, uint64_t* out) {
    const int N = s->N;
    int* gt = (int*)malloc(sizeof(int) * N);
    galois_table(N, gt);
    uint64_t* x1 = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* x2 = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (int lane = 0; lane < LANES; ++lane) {
        for (int j = 0; j < N; ++j) {
            x1[j] = sums[((size_t)0 * N + gt[j]) * LANES + lane] % s->R.M[0].p;
            x2[j] = sums[((size_t)1 * N + gt[j]) * LANES + lane] % s->R.M[1].p;
        }
        from_simd(&s->R, x1, x2, out + (size_t)lane * N);
    }
    free(gt);
    free(x1);
    free(x2);
    return 0;
}

/* base_pir.rs:945-952: one query broadcast to all lanes */
int orc_shard_perform_pir(const orc_shard* s, const uint64_t* qb, const uint64_t* qbc, const uint64_t* gkb, uint64_t* out) {
    int N = s->N;
    size_t LN = (size_t)LANES * N;
    uint64_t* q8 = (uint64_t*)malloc(sizeof(uint64_t) * LN);
    uint64_t* qc8 = qbc ? (uint64_t*)malloc(sizeof(uint64_t) * LN) : NULL;
    uint64_t* g8 = (uint64_t*)malloc(sizeof(uint64_t) * 2 * LN);
    for (int lane = 0; lane < LANES; ++lane) {
        memcpy(q8 + (size_t)lane * N, qb, sizeof(uint64_t) * N);
        if (qbc) memcpy(qc8 + (size_t)lane * N, qbc, sizeof(uint64_t) * N);
        memcpy(g8 + (size_t)lane * 2 * N, gkb, sizeof(uint64_t) * 2 * N);
    }
    int rc = orc_shard_perform_pir_split(s, q8, qc8, g8, out);
    free(q8);
    free(qc8);
    free(g8);
    return rc;
}

uint64_t orc_shard_export_te_p1(const orc_shard* s, uint32_t* out) {
    size_t rows = (size_t)s->L * LANES * s->N;
    memcpy(out, s->arr[ARR_TE][0][0], rows * sizeof(uint32_t));
    if (s->conj) memcpy(out + rows, s->arr[ARR_TE][0][1], rows * sizeof(uint32_t));
    return s->R.zeta[0];
}

void orc_shard_export_row_coeffs(const orc_shard* s, int array, int stream, int row, int lane, int64_t* out) {
    int N = s->N;
    uint64_t* x = (uint64_t*)malloc(sizeof(uint64_t) * N);
    const uint32_t* src = row_ptr(s, array, 0, stream, row, lane);
    for (int i = 0; i < N; ++i) x[i] = src[i];
    inv_weighted_ntt(&s->R, 0, x);
    for (int i = 0; i < N; ++i) out[i] = md_center(&s->R.M[0], x[i]);
    free(x);
}

/* ------------------------------------------------------------------------ */
/* double PIR, double_pir.rs:126-228                                         */
/* ------------------------------------------------------------------------ */
/* pir_wrapper.rs:166-175 */
int orc_choose_index_group_secondary(int logN, size_t num_shards) {
    (void)logN;
    int log2l = 0;
    while (((size_t)1 << log2l) < num_shards) ++log2l;
    int l = LANES << log2l;
    return l; /* l == N means the full group */
}

/* double_pir.rs:70-81: a*2^12 + b as three base-1025 digits */
static void encode_ct_plain(uint64_t a, uint64_t b, uint16_t out[3]) {
    uint64_t v = (a << B_BITS) + b;
    for (int j = 0; j < CTXT_PARTS; ++j) { out[j] = (uint16_t)(v % 1025); v /= 1025; }
}

static double now_us(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e6 + (double)ts.tv_nsec * 1e-3;
}

static int g_reuse_repeated = 0;
void orc_set_reuse_repeated(int on) { g_reuse_repeated = on; }

int orc_double_pir(orc_shard* const* fleet0, size_t len0, orc_shard* const* fleet1, size_t len1,
                   const uint64_t* q0, const uint64_t* q1, int l_s, uint64_t* out0, uint64_t* out1,
                   double* primary_us, double* secondary_us) {
    orc_shard* const* fleets[2] = {fleet0, fleet1};
    size_t lens[2] = {len0, len1};
    const uint64_t* qs[2] = {q0, q1};
    uint64_t* outs[2] = {out0, out1};
    const orc_shard* s0 = fleet0[0];
    int logN = s0->logN, N = s0->N;
    for (int b = 0; b < 2; ++b)
        if (lens[b] * LANES > (size_t)l_s) return -4; /* double_pir.rs:151-152 */
    double t0 = now_us();
    /* PRIMARY, double_pir.rs:157-177: per shard i, lane k -> 3 plaintexts at index 8i+k */
    uint16_t* parts[2] = {NULL, NULL};
    uint64_t* a = (uint64_t*)malloc(sizeof(uint64_t) * LANES * N);
    uint64_t* b = (uint64_t*)malloc(sizeof(uint64_t) * LANES * N);
    uint64_t* am = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* bm = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (int sl = 0; sl < BATCH_COUNT; ++sl) {
        if (!qs[sl]) continue;
        const uint64_t* q = qs[sl];
        parts[sl] = (uint16_t*)calloc(lens[sl] * LANES * CTXT_PARTS * (size_t)N, sizeof(uint16_t));
        for (size_t i = 0; i < lens[sl]; ++i) {
            const orc_shard* sh = fleets[sl][i];
            if (sh->conj != s0->conj) return -2; /* double_pir.rs:160 */
            if (g_reuse_repeated) {
                /* a shard that repeats in the fleet gives the same plaintexts for the same query */
                size_t prev = i;
                for (size_t k = 0; k < i; ++k)
                    if (fleets[sl][k] == sh) {
                        prev = k;
                        break;
                    }
                if (prev != i) {
                    memcpy(parts[sl] + (i * LANES * CTXT_PARTS) * (size_t)N, parts[sl] + (prev * LANES * CTXT_PARTS) * (size_t)N,
                           sizeof(uint16_t) * LANES * CTXT_PARTS * (size_t)N);
                    continue;
                }
            }
            orc_shard_get_a(sh, a);
            int rc = orc_shard_perform_pir(sh, q + 2 * N, sh->conj ? q + 3 * N : NULL, q, b);
            if (rc) return rc;
            for (int k = 0; k < LANES; ++k) {
                orc_mod_switch_encode(logN, A_BITS, a + (size_t)k * N, am);
                orc_mod_switch_encode(logN, B_BITS, b + (size_t)k * N, bm);
                uint16_t* dst = parts[sl] + ((i * LANES + k) * CTXT_PARTS) * (size_t)N;
                for (int c = 0; c < N; ++c) {
                    uint16_t d[3];
                    encode_ct_plain(am[c], bm[c], d);
                    for (int j = 0; j < CTXT_PARTS; ++j) dst[(size_t)j * N + c] = d[j];
                }
            }
        }
    }
    double t1 = now_us();
    /* SECONDARY, double_pir.rs:195-218 */
    orc_shard* sec = orc_shard_create(logN, l_s, 1, s0->root_pow);
    if (!sec) return -4;
    uint16_t* ent = (uint16_t*)calloc((size_t)l_s * LANES * N, sizeof(uint16_t));
    for (int i = 0; i < l_s; ++i)
        for (int j = 0; j < LANES; ++j) {
            int slot = j / CTXT_PARTS;
            if (j < CTXT_PARTS * BATCH_COUNT && (size_t)i < LANES * lens[slot] && parts[slot]) {
                const uint16_t* src = parts[slot] + ((size_t)i * CTXT_PARTS + j % CTXT_PARTS) * N;
                memcpy(ent + ((size_t)i * LANES + j) * N, src, sizeof(uint16_t) * N);
            }
        }
    orc_shard_set_db(sec, ent, (size_t)l_s * LANES);
    uint64_t* sa = (uint64_t*)malloc(sizeof(uint64_t) * LANES * N);
    uint64_t* sb = (uint64_t*)malloc(sizeof(uint64_t) * LANES * N);
    orc_shard_get_a(sec, sa);
    uint64_t* qb = (uint64_t*)calloc((size_t)LANES * N, sizeof(uint64_t));
    uint64_t* qbc = (uint64_t*)calloc((size_t)LANES * N, sizeof(uint64_t));
    uint64_t* gkb = (uint64_t*)calloc((size_t)LANES * 2 * N, sizeof(uint64_t));
    for (int j = 0; j < LANES; ++j) {
        int slot = j / CTXT_PARTS;
        if (slot < BATCH_COUNT && qs[slot]) {
            const uint64_t* q = qs[slot];
            memcpy(gkb + (size_t)j * 2 * N, q, sizeof(uint64_t) * 2 * N);
            memcpy(qb + (size_t)j * N, q + 4 * N, sizeof(uint64_t) * N);
            if (sec->conj) memcpy(qbc + (size_t)j * N, q + 5 * N, sizeof(uint64_t) * N);
        }
    }
    orc_shard_perform_pir_split(sec, qb, qbc, gkb, sb);
    for (int sl = 0; sl < BATCH_COUNT; ++sl) {
        if (!qs[sl] || !outs[sl]) continue;
        for (int p = 0; p < CTXT_PARTS; ++p) {
            int lane = sl * CTXT_PARTS + p;
            memcpy(outs[sl] + ((size_t)p * 2 + 0) * N, sa + (size_t)lane * N, sizeof(uint64_t) * N);
            memcpy(outs[sl] + ((size_t)p * 2 + 1) * N, sb + (size_t)lane * N, sizeof(uint64_t) * N);
        }
    }
    double t2 = now_us();
    if (primary_us) *primary_us = t1 - t0;
    if (secondary_us) *secondary_us = t2 - t1;
    orc_shard_free(sec);
    free(ent); free(sa); free(sb); free(qb); free(qbc); free(gkb);
    free(a); free(b); free(am); free(bm);
    free(parts[0]); free(parts[1]);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* byte wire format, pir_wrapper.rs                                          */
/* ------------------------------------------------------------------------ */
/* pir_wrapper.rs:84-115 */
size_t orc_condense(int bits, const uint64_t* data, size_t count, uint8_t* out) {
    size_t n = 0;
    for (int i = 0; i < bits / 8; ++i)
        for (size_t j = 0; j < count; ++j) out[n++] = (uint8_t)(data[j] >> (8 * i));
    if (bits % 8 >= 4) {
        int sh = (bits / 8) * 8;
        for (size_t j = 0; j + 1 < count; j += 2)
            out[n++] = (uint8_t)((((data[j] >> sh) & 0xF) << 4) | ((data[j + 1] >> sh) & 0xF));
    }
    if (bits % 4 >= 2) {
        int sh = (bits / 4) * 4;
        for (size_t j = 0; j < count; j += 4)
            out[n++] = (uint8_t)((((data[j] >> sh) & 3) << 6) | (((data[j + 1] >> sh) & 3) << 4) |
                                 (((data[j + 2] >> sh) & 3) << 2) | ((data[j + 3] >> sh) & 3));
    }
    if (bits % 2 >= 1) {
        int sh = (bits / 2) * 2;
        for (size_t j = 0; j < count; j += 8) {
            uint8_t v = 0;
            for (int k = 0; k < 8; ++k) v |= (uint8_t)(((data[j + k] >> sh) & 1) << (7 - k));
            out[n++] = v;
        }
    }
    return n;
}

/* pir_wrapper.rs:121-158; -1 = "not enough data" */
int orc_uncondense(int bits, size_t count, const uint8_t* data, size_t len, size_t* consumed, uint64_t* out) {
    size_t n = 0;
    memset(out, 0, sizeof(uint64_t) * count);
#define NEXT_BYTE(v) do { if (n >= len) return -1; (v) = data[n++]; } while (0)
    for (int i = 0; i < bits / 8; ++i)
        for (size_t j = 0; j < count; ++j) { uint8_t b; NEXT_BYTE(b); out[j] |= (uint64_t)b << (8 * i); }
    if (bits % 8 >= 4) {
        int sh = (bits / 8) * 8;
        for (size_t j = 0; j + 2 <= count; j += 2) {
            uint8_t b; NEXT_BYTE(b);
            out[j] |= (uint64_t)(b >> 4) << sh;
            out[j + 1] |= (uint64_t)(b & 0xF) << sh;
        }
    }
    if (bits % 4 >= 2) {
        int sh = (bits / 4) * 4;
        for (size_t j = 0; j + 4 <= count; j += 4) {
            uint8_t b; NEXT_BYTE(b);
            out[j] |= (uint64_t)(b >> 6) << sh;
            out[j + 1] |= (uint64_t)((b >> 4) & 3) << sh;
            out[j + 2] |= (uint64_t)((b >> 2) & 3) << sh;
            out[j + 3] |= (uint64_t)(b & 3) << sh;
        }
    }
    if (bits % 2 >= 1) {
        int sh = (bits / 2) * 2;
        for (size_t j = 0; j + 8 <= count; j += 8) {
            uint8_t b; NEXT_BYTE(b);
            for (int k = 0; k < 8; ++k) out[j + k] |= (uint64_t)((b >> (7 - k)) & 1) << sh;
        }
    }
#undef NEXT_BYTE
    if (consumed) *consumed = n;
    return 0;
}

/* pir_wrapper.rs:188,219-231 (N = 2048) */
void orc_database_shape(size_t num_entries, int* l_primary, int* l_secondary) {
    const size_t N = 2048;
    int lp = num_entries <= N * N / 16 ? (int)(N / 2) : (int)N;
    size_t shards = (num_entries + (size_t)LANES * lp - 1) / ((size_t)LANES * lp);
    *l_primary = lp;
    *l_secondary = orc_choose_index_group_secondary(11, shards);
}
/* pir_wrapper.rs:238-241 */
size_t orc_num_databases(size_t num_entries) {
    int lp, ls;
    orc_database_shape(num_entries, &lp, &ls);
    return (num_entries + (size_t)LANES * lp - 1) / ((size_t)LANES * lp);
}
size_t orc_query_bytes(int l_primary, int l_secondary) {
    const size_t N = 2048;
    size_t n = 2 * N / 8 * GK_BITS + N / 8 * QRY_BITS + N / 8 * QRY_BITS;
    if (l_primary == (int)N) n += N / 8 * QRY_BITS;
    if (l_secondary == (int)N) n += N / 8 * QRY_BITS;
    return n;
}

/* pir_wrapper.rs:349-409 */
int orc_process_queries(orc_shard* const* fleet0, size_t len0, orc_shard* const* fleet1, size_t len1,
                        const uint8_t* query0, size_t qlen0, const uint8_t* query1, size_t qlen1,
                        uint8_t* reply0, uint8_t* reply1, double* primary_us, double* secondary_us) {
    const int logN = 11, N = 2048;
    if (len0 < len1) return -5; /* pir_wrapper.rs:362 */
    int conj_p = fleet0[0]->conj;
    int l_s = orc_choose_index_group_secondary(logN, len0);
    int conj_s = l_s == N;
    const uint8_t* qb[2] = {query0, query1};
    size_t ql[2] = {qlen0, qlen1};
    uint64_t* qv[2] = {NULL, NULL};
    uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (int s = 0; s < 2; ++s) {
        if (!qb[s]) continue;
        qv[s] = (uint64_t*)calloc((size_t)6 * N, sizeof(uint64_t));
        size_t off = 0, used;
        int bits_of[6] = {GK_BITS, GK_BITS, QRY_BITS, QRY_BITS, QRY_BITS, QRY_BITS};
        int present[6] = {1, 1, 1, conj_p, 1, conj_s};
        for (int e = 0; e < 6; ++e) {
            if (!present[e]) continue;
            if (orc_uncondense(bits_of[e], N, qb[s] + off, ql[s] - off, &used, tmp)) {
                free(tmp); free(qv[0]); free(qv[1]);
                return -1; /* "not enough data", pir_wrapper.rs:124 */
            }
            off += used;
            orc_mod_switch_decode(logN, bits_of[e], tmp, qv[s] + (size_t)e * N);
        }
    }
    uint64_t* out[2] = {NULL, NULL};
    for (int s = 0; s < 2; ++s)
        if (qv[s]) out[s] = (uint64_t*)malloc(sizeof(uint64_t) * 6 * N);
    int rc = orc_double_pir(fleet0, len0, fleet1, len1, qv[0], qv[1], l_s, out[0], out[1], primary_us, secondary_us);
    uint8_t* rep[2] = {reply0, reply1};
    for (int s = 0; s < 2 && rc == 0; ++s) {
        if (!out[s] || !rep[s]) continue;
        size_t n = 0;
        for (int p = 0; p < CTXT_PARTS; ++p) {
            orc_mod_switch_encode(logN, A_BITS, out[s] + ((size_t)p * 2 + 0) * N, tmp);
            n += orc_condense(A_BITS, tmp, N, rep[s] + n);
            orc_mod_switch_encode(logN, B_BITS, out[s] + ((size_t)p * 2 + 1) * N, tmp);
            n += orc_condense(B_BITS, tmp, N, rep[s] + n);
        }
    }
    free(tmp);
    for (int s = 0; s < 2; ++s) { free(qv[s]); free(out[s]); }
    return rc;
}

/* monomial X^e in Z_t[X]/(X^N + 1) */
static void monomial(int N, int64_t e, uint64_t t, uint16_t* out) {
    memset(out, 0, sizeof(uint16_t) * N);
    e = zm(e, 2 * N);
    if (e < N) out[e] = 1;
    else out[e - N] = (uint16_t)(t - 1);
}

/* pir_wrapper.rs:276-317 with enc_double_pir double_pir.rs:245-290 */
int orc_prepare_query(orc_rng* rng, size_t primary_idx, size_t secondary_idx, int l_primary, int l_secondary,
                      uint8_t* query, size_t* qlen, uint8_t sk_seed[32]) {
    const int logN = 11, N = 2048;
    if (primary_idx >= (size_t)l_primary * LANES) return -6;
    if (secondary_idx >= (size_t)l_secondary / LANES) return -6;
    for (int i = 0; i < 32; ++i) sk_seed[i] = (uint8_t)(orc_rng_next_u32(rng) & 0xFFFF);
    int64_t* sk = (int64_t*)malloc(sizeof(int64_t) * N);
    orc_rng* skr = orc_rng_new(sk_seed);
    orc_gen_sk(logN, skr, SK_HWT, sk);
    orc_rng_free(skr);
    int conj_p = l_primary == N, conj_s = l_secondary == N;
    int kp = ig_qbp(N, l_primary), ks = ig_qbp(N, l_secondary);
    int64_t i = (int64_t)(primary_idx / LANES);
    int64_t j = (int64_t)(secondary_idx * LANES + primary_idx % LANES);
    uint64_t* gk = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    uint64_t* ct[4];
    for (int c = 0; c < 4; ++c) ct[c] = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint16_t* mono = (uint16_t*)malloc(sizeof(uint16_t) * N);
    orc_gen_gk_b(logN, rng, sk, GEN_H, SIGMA, gk);
    uint8_t seed[32];
    orc_seed(2, seed);
    monomial(N, i * kp, ORC_T, mono);
    orc_enc_sym_b(logN, seed, rng, sk, mono, ORC_T, SIGMA, ct[0]);
    if (conj_p) {
        orc_seed(3, seed);
        monomial(N, 2 * N - i * kp, ORC_T, mono);
        orc_enc_sym_b(logN, seed, rng, sk, mono, ORC_T, SIGMA, ct[1]);
    }
    orc_seed(4, seed);
    monomial(N, j * ks, ORC_T, mono);
    orc_enc_sym_b(logN, seed, rng, sk, mono, ORC_T, SIGMA, ct[2]);
    if (conj_s) {
        orc_seed(5, seed);
        monomial(N, 2 * N - j * ks, ORC_T, mono);
        orc_enc_sym_b(logN, seed, rng, sk, mono, ORC_T, SIGMA, ct[3]);
    }
    uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * N);
    size_t n = 0;
    for (int d = 0; d < 2; ++d) {
        orc_mod_switch_encode(logN, GK_BITS, gk + (size_t)d * N, tmp);
        n += orc_condense(GK_BITS, tmp, N, query + n);
    }
    int order[4] = {0, 1, 2, 3};
    int pres[4] = {1, conj_p, 1, conj_s};
    for (int c = 0; c < 4; ++c) {
        if (!pres[order[c]]) continue;
        orc_mod_switch_encode(logN, QRY_BITS, ct[order[c]], tmp);
        n += orc_condense(QRY_BITS, tmp, N, query + n);
    }
    *qlen = n;
    free(sk); free(gk); free(mono); free(tmp);
    for (int c = 0; c < 4; ++c) free(ct[c]);
    return 0;
}

/* pir_wrapper.rs:420-440 with dec_double_pir double_pir.rs:292-317 */
int orc_process_reply(const uint8_t* reply, size_t rlen, const uint8_t sk_seed[32], uint16_t* out) {
    const int logN = 11, N = 2048;
    int64_t* sk = (int64_t*)malloc(sizeof(int64_t) * N);
    orc_rng* skr = orc_rng_new(sk_seed);
    orc_gen_sk(logN, skr, SK_HWT, sk);
    orc_rng_free(skr);
    uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* a = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* b = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint16_t* parts = (uint16_t*)malloc(sizeof(uint16_t) * 3 * N);
    size_t off = 0, used;
    int rc = 0;
    for (int p = 0; p < CTXT_PARTS && rc == 0; ++p) {
        if (orc_uncondense(A_BITS, N, reply + off, rlen - off, &used, tmp)) { rc = -1; break; }
        off += used;
        orc_mod_switch_decode(logN, A_BITS, tmp, a);
        if (orc_uncondense(B_BITS, N, reply + off, rlen - off, &used, tmp)) { rc = -1; break; }
        off += used;
        orc_mod_switch_decode(logN, B_BITS, tmp, b);
        orc_dec(logN, sk, a, b, ORC_T, parts + (size_t)p * N);
    }
    if (rc == 0) {
        uint64_t* av = (uint64_t*)malloc(sizeof(uint64_t) * N);
        uint64_t* bv = (uint64_t*)malloc(sizeof(uint64_t) * N);
        for (int c = 0; c < N; ++c) {
            uint64_t v = 0;
            for (int p = CTXT_PARTS - 1; p >= 0; --p) v = v * 1025 + parts[(size_t)p * N + c];
            if ((v >> B_BITS) >= (1ULL << A_BITS)) rc = -7; /* double_pir.rs:88 */
            av[c] = v >> B_BITS;
            bv[c] = v & ((1ULL << B_BITS) - 1);
        }
        orc_mod_switch_decode(logN, A_BITS, av, a);
        orc_mod_switch_decode(logN, B_BITS, bv, b);
        orc_dec(logN, sk, a, b, ORC_T, out);
        free(av);
        free(bv);
    }
    free(sk); free(tmp); free(a); free(b); free(parts);
    return rc;
}

/* ---- bucket <-> coefficient packing (crates/dns-over-pir/src/dns/bucket.rs) ----
 * 2560-byte bucket -> N = 2048 ten-bit coefficients: each group of 5 bytes
 * read as a little-endian 40-bit integer v gives the 4 coefficients
 * v >> 0, >> 10, >> 20, >> 30 (masked to 10 bits), bucket.rs:289-303. */
void orc_pack_bytes_into_coeffs(const uint8_t* bucket, uint16_t* coeffs) {
    for (int g = 0; g < 2048 / 4; ++g) {
        uint64_t v = 0;
        for (int b = 4; b >= 0; --b) v = (v << 8) | bucket[g * 5 + b];
        for (int k = 0; k < 4; ++k) coeffs[g * 4 + k] = (uint16_t)((v >> (10 * k)) & 1023);
    }
}

/* the inverse (bucket.rs:271-283): low 10 bits of each coefficient */
void orc_unpack_coeffs_into_bytes(const uint16_t* coeffs, uint8_t* bucket) {
    for (int g = 0; g < 2048 / 4; ++g) {
        uint64_t v = 0;
        for (int k = 0; k < 4; ++k) v |= (uint64_t)(coeffs[g * 4 + k] & 1023) << (10 * k);
        for (int b = 0; b < 5; ++b) bucket[g * 5 + b] = (uint8_t)(v >> (8 * b));
    }
}
