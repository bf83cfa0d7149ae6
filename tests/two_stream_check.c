/*
 * CPU check of the two-stream inner product identity used by the device
 * kernels (k1w_*): the IPGK0 array of a shard is never read; instead
 *   out[a] = sum_(st,j) IPGK1_st[i][j] G[a][j] + TE_st[i][j] Q_st[a][j],
 *   i = (a - j + L - 1) mod L  (output-major Galois coordinates, L = N/2),
 *   G[a][j]    = gkb1'[j] + gka1'[j] C[(j - a) mod L][j],
 *   Q_st[a][j] = qb_st[j] + qa_st[j] C[(j - a) mod L][j],
 *   C[M][j]    = sum_(n=1..M) gkb0[j-n] prod_(m=1..n-1) gka0[j-m],
 *   x' = x - 2^26 x0 (gkb1' from gkb1, gkb0; gka1' from gka1, gka0),
 * with Galois position j of half h holding the slot of exponent
 * (-1)^h 5^j mod 2N. Compared with the oracle's main_pir_loop restatement
 * (the three-array sum, then the final gather), modulus by modulus, lane by
 * lane, for a half-group shard (no conjugate) and a full-group shard (main and
 * conjugate streams). With `shift` != 0 the weight index is taken off by
 * `shift` (a planted mistake) and the check must fail.
 *
 * Test infrastructure only: it includes the oracle's source to reach the
 * stored rows. Usage: two_stream_check logN shift
 */
#include <stdio.h>

#include "../oracle/dnspir_oracle.c"

static uint64_t xs(uint64_t* s) {
    *s ^= *s << 13;
    *s ^= *s >> 7;
    *s ^= *s << 17;
    return *s;
}

static int check(int logN, int l, int shift) {
    const int N = 1 << logN, m2 = 2 * N, L = N / 2;
    orc_shard* s = orc_shard_create(logN, l, 0, 1);
    uint64_t seed = 0x9E3779B97F4A7C15ull ^ (uint64_t)(l * 131 + logN);
    uint16_t* ent = malloc(sizeof(uint16_t) * 8 * (size_t)l * N);
    for (size_t i = 0; i < (size_t)8 * l * N; ++i) ent[i] = (uint16_t)(xs(&seed) % 1025);
    if (orc_shard_set_db(s, ent, 8 * (size_t)l)) return -1;
    const int nst = s->conj ? 2 : 1;
    const ring_t* R = &s->R;
    /* Galois position -> oracle slot (slot k holds f(zeta^-(2k+1))) */
    int* gslot = malloc(sizeof(int) * N);
    for (int h = 0; h < 2; ++h) {
        int64_t e = 1;
        for (int a = 0; a < L; ++a) {
            const int64_t ex = h ? zm(-e, m2) : e;
            gslot[h * L + a] = (int)((zm(-ex, m2) - 1) / 2);
            e = e * 5 % m2;
        }
    }
    /* one query broadcast to the lanes: qb, qb_conj, gk_b[2] (coefficients) */
    uint64_t *qb = malloc(8 * N), *qbc = malloc(8 * N), *gkb = malloc(16 * N);
    for (int i = 0; i < N; ++i) {
        qb[i] = xs(&seed) % ORC_Q;
        qbc[i] = xs(&seed) % ORC_Q;
        gkb[i] = xs(&seed) % ORC_Q;
        gkb[N + i] = xs(&seed) % ORC_Q;
    }
    /* seed-expanded a-parts, base_pir.rs:782-789 */
    uint64_t* tmp = malloc(8 * N);
    uint64_t *gka[2][2], *qa[2][2];
    for (int d = 0; d < 2; ++d) {
        uint8_t sd[32];
        orc_seed(d, sd);
        orc_expand(logN, sd, tmp);
        for (int r = 0; r < 2; ++r) {
            gka[d][r] = malloc(8 * N);
            to_ntt(R, r, tmp, gka[d][r]);
        }
    }
    for (int st = 0; st < nst; ++st) {
        uint8_t sd[32];
        orc_seed(2 + st, sd);
        orc_expand(logN, sd, tmp);
        for (int r = 0; r < 2; ++r) {
            qa[st][r] = malloc(8 * N);
            to_ntt(R, r, tmp, qa[st][r]);
        }
    }
    int* gt = malloc(sizeof(int) * N);
    galois_table(N, gt);
    uint64_t* C = malloc(sizeof(uint64_t) * (size_t)L * N); /* C[M][gpos] */
    uint64_t *g0 = malloc(8 * N), *g1 = malloc(8 * N), *qq[2] = {malloc(8 * N), malloc(8 * N)};
    uint64_t *acc = malloc(8 * N), *sum = malloc(8 * N), *nsum = malloc(8 * N);
    int bad = 0;
    for (int r = 0; r < 2; ++r) {
        const md_t* M = &R->M[r];
        const uint64_t t26 = (1ULL << 26) % M->p;
        to_ntt(R, r, gkb, g0);
        to_ntt(R, r, gkb + N, g1);
        to_ntt(R, r, qb, qq[0]);
        if (s->conj) to_ntt(R, r, qbc, qq[1]);
        for (int h = 0; h < 2; ++h)
            for (int j = 0; j < L; ++j) {
                uint64_t c = 0, P = 1;
                C[h * L + j] = 0;
                for (int Mi = 1; Mi < L; ++Mi) {
                    const int y = gslot[h * L + ((j - Mi) % L + L) % L];
                    c = md_add(M, c, md_mul(M, g0[y], P));
                    P = md_mul(M, P, gka[0][r][y]);
                    C[(size_t)Mi * N + h * L + j] = c;
                }
            }
        for (int lane = 0; lane < LANES; ++lane) {
            /* the oracle: three arrays, scatter each row, final gather */
            memset(acc, 0, 8 * N);
            for (int st = 0; st < nst; ++st) {
                memset(sum, 0, 8 * N);
                for (int i = 0; i < L; ++i) {
                    const uint32_t* i0 = row_ptr(s, ARR_IP0, r, st, i, lane);
                    const uint32_t* i1 = row_ptr(s, ARR_IP1, r, st, i, lane);
                    const uint32_t* te = row_ptr(s, ARR_TE, r, st, i, lane);
                    for (int j = 0; j < N; ++j)
                        nsum[gt[j]] = md_red(M, sum[j] + md_mul(M, i0[j], g0[j]) + md_mul(M, i1[j], g1[j]) +
                                                    md_mul(M, te[j], qq[st][j]));
                    memcpy(sum, nsum, 8 * N);
                }
                for (int j = 0; j < N; ++j) acc[j] = md_add(M, acc[j], sum[j]);
            }
            /* the two-array form */
            for (int h = 0; h < 2; ++h)
                for (int a = 0; a < L; ++a) {
                    uint64_t o = 0;
                    for (int st = 0; st < nst; ++st)
                        for (int j = 0; j < L; ++j) {
                            const int i = ((a - j + L - 1) % L + L) % L;
                            const int sl = gslot[h * L + j];
                            const int Mi = ((j - a + shift) % L + L) % L;
                            const uint64_t cv = C[(size_t)Mi * N + h * L + j];
                            const uint64_t gb1 = md_sub(M, g1[sl], md_mul(M, t26, g0[sl]));
                            const uint64_t ga1 = md_sub(M, gka[1][r][sl], md_mul(M, t26, gka[0][r][sl]));
                            const uint64_t G = md_add(M, gb1, md_mul(M, ga1, cv));
                            const uint64_t Q = md_add(M, qq[st][sl], md_mul(M, qa[st][r][sl], cv));
                            o = md_add(M, o,
                                       md_red(M, md_mul(M, row_ptr(s, ARR_IP1, r, st, i, lane)[sl], G) +
                                                     md_mul(M, row_ptr(s, ARR_TE, r, st, i, lane)[sl], Q)));
                        }
                    bad += o != acc[gt[gslot[h * L + a]]];
                }
        }
    }
    printf("logN %d l %d conj %d shift %d: %d mismatches of %d\n", logN, l, s->conj, shift, bad, 2 * LANES * N);
    orc_shard_free(s);
    return bad;
}

int main(int argc, char** argv) {
    const int logN = argc > 1 ? atoi(argv[1]) : 6;
    const int shift = argc > 2 ? atoi(argv[2]) : 0;
    const int N = 1 << logN;
    int bad = check(logN, N / 2, shift);
    bad += check(logN, N, shift);
    return bad ? 1 : 0;
}
