/*
 * CPU baseline of the primary inner product in the reference's own form --
 * TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and its tests; never
 * the product).
 *
 * main_pir_loop (crates/pir/src/base_pir.rs:1085-1247) restated for timing:
 * operands in the reference's layout (8 lanes of a slot side by side, 32 B
 * per element), for every row i and slot j the three u32 x u32 -> u64 lane
 * products IPGK0*gkb0 + IPGK1*gkb1 + TE*qb added to the running sum and
 * stored at the Galois position gal[j] of the next sum (:1153-1219), with the
 * sums reduced only every `reduce_every` rows (:1001-1005, :1240-1245; 341
 * for these moduli). One call of the loop per (modulus, stream) pair, and
 * those pairs of all shards run in parallel over OpenMP threads as the
 * reference's rayon join / par_iter do (:1007-1064, double_pir.rs:157). The
 * vector width is chosen at run time: AVX-512 (the reference's avx.rs, 8 u64
 * lanes per register), else AVX2 (two registers per slot, as its
 * emulate_avx512 feature), else scalar.
 *
 * Its sums are checked against the oracle's perform_pir (tests/), so the
 * timed loop is an exact one.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define LANES 8

typedef void (*row_fn)(int N, const uint32_t* a0, const uint32_t* a1, const uint32_t* te, const uint32_t* g0,
                       const uint32_t* g1, const uint32_t* q, const int* gal, const uint64_t* sum, uint64_t* sum_new);

__attribute__((target("avx512f"))) static void row_avx512(int N, const uint32_t* a0, const uint32_t* a1,
                                                           const uint32_t* te, const uint32_t* g0, const uint32_t* g1,
                                                           const uint32_t* q, const int* gal, const uint64_t* sum,
                                                           uint64_t* sum_new) {
    for (int j = 0; j < N; ++j) {
        const size_t o = (size_t)j * LANES;
        if ((j & 1) == 0) {  /* the reference's software preloads (:1117-1142) as prefetches, 32 slots ahead */
            _mm_prefetch((const char*)(a0 + o + 32 * LANES), _MM_HINT_T0);
            _mm_prefetch((const char*)(a1 + o + 32 * LANES), _MM_HINT_T0);
            _mm_prefetch((const char*)(te + o + 32 * LANES), _MM_HINT_T0);
        }
        __m512i cur = _mm512_load_si512((const void*)(sum + o));
        __m512i x = _mm512_cvtepu32_epi64(_mm256_loadu_si256((const __m256i*)(a0 + o)));
        __m512i y = _mm512_cvtepu32_epi64(_mm256_loadu_si256((const __m256i*)(g0 + o)));
        cur = _mm512_add_epi64(cur, _mm512_mul_epu32(x, y));
        x = _mm512_cvtepu32_epi64(_mm256_loadu_si256((const __m256i*)(a1 + o)));
        y = _mm512_cvtepu32_epi64(_mm256_loadu_si256((const __m256i*)(g1 + o)));
        cur = _mm512_add_epi64(cur, _mm512_mul_epu32(x, y));
        x = _mm512_cvtepu32_epi64(_mm256_loadu_si256((const __m256i*)(te + o)));
        y = _mm512_cvtepu32_epi64(_mm256_loadu_si256((const __m256i*)(q + o)));
        cur = _mm512_add_epi64(cur, _mm512_mul_epu32(x, y));
        _mm512_store_si512((void*)(sum_new + (size_t)gal[j] * LANES), cur);
    }
}

__attribute__((target("avx2"))) static void row_avx2(int N, const uint32_t* a0, const uint32_t* a1, const uint32_t* te,
                                                      const uint32_t* g0, const uint32_t* g1, const uint32_t* q,
                                                      const int* gal, const uint64_t* sum, uint64_t* sum_new) {
    for (int j = 0; j < N; ++j) {
        const size_t o = (size_t)j * LANES;
        uint64_t* dst = sum_new + (size_t)gal[j] * LANES;
        for (int h = 0; h < 2; ++h) {  /* lanes 4h .. 4h+3 */
            __m256i cur = _mm256_loadu_si256((const __m256i*)(sum + o + 4 * h));
            __m256i x = _mm256_cvtepu32_epi64(_mm_loadu_si128((const __m128i*)(a0 + o + 4 * h)));
            __m256i y = _mm256_cvtepu32_epi64(_mm_loadu_si128((const __m128i*)(g0 + o + 4 * h)));
            cur = _mm256_add_epi64(cur, _mm256_mul_epu32(x, y));
            x = _mm256_cvtepu32_epi64(_mm_loadu_si128((const __m128i*)(a1 + o + 4 * h)));
            y = _mm256_cvtepu32_epi64(_mm_loadu_si128((const __m128i*)(g1 + o + 4 * h)));
            cur = _mm256_add_epi64(cur, _mm256_mul_epu32(x, y));
            x = _mm256_cvtepu32_epi64(_mm_loadu_si128((const __m128i*)(te + o + 4 * h)));
            y = _mm256_cvtepu32_epi64(_mm_loadu_si128((const __m128i*)(q + o + 4 * h)));
            cur = _mm256_add_epi64(cur, _mm256_mul_epu32(x, y));
            _mm256_store_si256((__m256i*)(dst + 4 * h), cur);
        }
    }
}

static void row_scalar(int N, const uint32_t* a0, const uint32_t* a1, const uint32_t* te, const uint32_t* g0,
                       const uint32_t* g1, const uint32_t* q, const int* gal, const uint64_t* sum, uint64_t* sum_new) {
    for (int j = 0; j < N; ++j)
        for (int k = 0; k < LANES; ++k) {
            const size_t o = (size_t)j * LANES + k;
            sum_new[(size_t)gal[j] * LANES + k] =
                sum[o] + (uint64_t)a0[o] * g0[o] + (uint64_t)a1[o] * g1[o] + (uint64_t)te[o] * q[o];
        }
}

static int g_force_isa = -1;  /* -1: by the host CPU */

/* tests: 512, 256 or 0 runs that row loop (if the CPU has it), -1 = auto */
void orc_cpu_force_isa(int isa) { g_force_isa = isa; }

static row_fn pick(int* isa) {
    __builtin_cpu_init();
    if (g_force_isa == 0) {
        *isa = 0;
        return row_scalar;
    }
    if (g_force_isa == 256 && __builtin_cpu_supports("avx2")) {
        *isa = 256;
        return row_avx2;
    }
    if (__builtin_cpu_supports("avx512f")) {
        *isa = 512;
        return row_avx512;
    }
    if (__builtin_cpu_supports("avx2")) {
        *isa = 256;
        return row_avx2;
    }
    *isa = 0;
    return row_scalar;
}

const char* orc_cpu_isa(void) {
    int isa;
    pick(&isa);
    return isa == 512 ? "avx512" : isa == 256 ? "avx2" : "scalar";
}

/* the window of main_pir_loop (base_pir.rs:1001-1005): the most rows whose
 * three products per row, each below (2p)^2 with representatives up to 2p,
 * fit a u64 sum */
int orc_cpu_reduce_every(void) {
    const unsigned __int128 m = ~(uint64_t)0;
    const uint64_t b1 = 2 * 67104769ULL, b2 = 2 * 33550337ULL;
    const uint64_t w1 = (uint64_t)(m / b1 / b1 / 3), w2 = (uint64_t)(m / b2 / b2 / 3);
    return (int)(w1 < w2 ? w1 : w2);
}

/* main_pir_loop for one (modulus, stream): arrays [3][L][N][8], vecs [3][N][8];
 * the final sum (reduced mod p) lands in out [N][8] */
static void main_loop(row_fn fn, int N, int L, int window, uint64_t p, const uint32_t* arrays, const uint32_t* vecs,
                      const int* gal, uint64_t* out, uint64_t* tmp) {
    const size_t plane = (size_t)L * N * LANES;
    uint64_t* sum = out;
    uint64_t* nsum = tmp;
    memset(sum, 0, sizeof(uint64_t) * N * LANES);
    const uint32_t *g0 = vecs, *g1 = vecs + (size_t)N * LANES, *q = vecs + (size_t)2 * N * LANES;
    for (int i0 = 0; i0 < L; i0 += window) {
        const int n = L - i0 < window ? L - i0 : window;
        for (int ii = 0; ii < n; ++ii) {
            const size_t ro = (size_t)(i0 + ii) * N * LANES;
            fn(N, arrays + ro, arrays + plane + ro, arrays + 2 * plane + ro, g0, g1, q, gal, sum, nsum);
            uint64_t* t = sum;
            sum = nsum;
            nsum = t;
        }
        for (size_t k = 0; k < (size_t)N * LANES; ++k) sum[k] %= p;  /* reduce_from_u64 */
    }
    if (sum != out) memcpy(out, sum, sizeof(uint64_t) * N * LANES);
}

/* The primary inner product of n shards (arrays[s] as orc_shard_k1_operands
 * writes them; all shards of a fleet share vecs and gal) on `threads` OpenMP
 * threads: n * 2 * nst main_pir_loop calls, reducing every `window` rows (0:
 * the reference's window; a smaller one only for tests). sums (NULL: not kept) receive per
 * shard [2][N][8] the two streams' sums added and reduced mod p_r, as
 * perform_pir_split forms them before its Galois gather (base_pir.rs:1066-1076).
 * Returns 0, or -1 when memory runs out. */
int orc_cpu_k1(int n, const uint32_t* const* arrays, int N, int L, int nst, const uint32_t* vecs, const int* gal,
               int threads, int window, uint64_t* sums) {
    int isa;
    const row_fn fn = pick(&isa);
    if (window <= 0 || window > orc_cpu_reduce_every()) window = orc_cpu_reduce_every();
    const uint64_t P[2] = {67104769ULL, 33550337ULL};
    const int tasks = n * 2 * nst;
    const size_t words = (size_t)N * LANES;
    uint64_t* part = (uint64_t*)aligned_alloc(64, sizeof(uint64_t) * words * (size_t)tasks);
    if (!part) return -1;
    int fail = 0;
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
    {
        uint64_t* tmp = (uint64_t*)aligned_alloc(64, sizeof(uint64_t) * words);
        if (!tmp) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 1)
        for (int t = 0; t < tasks; ++t) {
            if (!tmp) continue;
            const int sh = t / (2 * nst), rs = t % (2 * nst), r = rs / nst;
            const size_t plane = (size_t)3 * L * N * LANES;
            main_loop(fn, N, L, window, P[r], arrays[sh] + (size_t)rs * plane, vecs + (size_t)rs * 3 * words, gal,
                      part + (size_t)t * words, tmp);
        }
        free(tmp);
    }
    if (!fail && sums)
        for (int sh = 0; sh < n; ++sh)
            for (int r = 0; r < 2; ++r) {
                uint64_t* dst = sums + ((size_t)sh * 2 + r) * words;
                const uint64_t* a = part + ((size_t)sh * 2 * nst + r * nst) * words;
                for (size_t k = 0; k < words; ++k) dst[k] = (a[k] + (nst == 2 ? a[k + words] : 0)) % P[r];
            }
    free(part);
    return fail ? -1 : 0;
}
