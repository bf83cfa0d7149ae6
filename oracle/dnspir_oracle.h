/*
 * CPU oracle for the DNSPIR server response path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it. It restates the
 * reference algorithms of crates/pir/src/{base_pir,double_pir,pir_wrapper,
 * bfv}.rs with exact integer arithmetic and emits the response values and
 * bytes that the GPU engine must reproduce bit for bit.
 *
 * Ring elements are passed as uint64_t[N] canonical coefficients in [0, q).
 * Database entries are uint16_t[N] coefficients in [0, t), t = 1025.
 */
#ifndef DNSPIR_ORACLE_H
#define DNSPIR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_P1 67104769ULL            /* 2^26 - 2^12 + 1, base_pir.rs:75 */
#define ORC_P2 33550337ULL            /* 2^25 - 2^12 + 1, base_pir.rs:76 */
#define ORC_Q (ORC_P1 * ORC_P2)       /* base_pir.rs:77 */
#define ORC_T 1025ULL                 /* double_pir.rs:38-39 */
#define ORC_LANES 8                   /* SIMD_COUNT, base_pir.rs:73 */

typedef struct orc_rng orc_rng;
typedef struct orc_shard orc_shard;

/* ---- ChaCha (rand_chacha word order) ---------------------------------- */
void orc_chacha_block(const uint32_t in[16], uint32_t out[16], int rounds);
orc_rng* orc_rng_new(const uint8_t seed[32]);
void orc_rng_free(orc_rng*);
uint32_t orc_rng_next_u32(orc_rng*);
uint64_t orc_rng_next_u64(orc_rng*);

/* ---- BFV primitives (bfv.rs) ------------------------------------------ */
void orc_expand(int logN, const uint8_t seed[32], uint64_t* out);
void orc_seed(int idx, uint8_t out[32]);
void orc_gen_sk(int logN, orc_rng* rng, int hwt, int64_t* sk);
void orc_enc_sym_b(int logN, const uint8_t seed_a[32], orc_rng* rng, const int64_t* sk,
                   const uint16_t* m, uint64_t t, double sigma, uint64_t* b);
void orc_gen_gk_b(int logN, orc_rng* rng, const int64_t* sk, int g, double sigma, uint64_t* gk /*[2][N]*/);
void orc_dec(int logN, const int64_t* sk, const uint64_t* a, const uint64_t* b, uint64_t t, uint16_t* m);
void orc_poly_mul_sparse(int logN, const uint64_t* a, const int64_t* sk, uint64_t* out);
void orc_mod_switch_encode(int logN, int bits, const uint64_t* x, uint64_t* out);
void orc_mod_switch_decode(int logN, int bits, const uint64_t* x, uint64_t* out);

/* weighted NTT (inverse = 1 for the inverse) in the reference convention,
 * modulus 0 = p1, 1 = p2; returns zeta */
uint64_t orc_ntt(int logN, int modulus, int root_pow, int inverse, uint64_t* x);

/* ---- shard (base_pir.rs) ------------------------------------------------ */
/* group_order = l; used_seeds 0 = FirstSet, 1 = SecondSet; root_pow selects the
 * primitive 2N-th root zeta0^root_pow (odd), to check root independence. */
orc_shard* orc_shard_create(int logN, int group_order, int used_seeds, int root_pow);
void orc_shard_free(orc_shard*);
int orc_shard_set_transformed_entries(orc_shard*, const uint16_t* entries, size_t n_entries);
int orc_shard_preprocess(orc_shard*);
int orc_shard_set_db(orc_shard*, const uint16_t* entries, size_t n_entries);
void orc_shard_get_a(const orc_shard*, uint64_t* out /*[8][N]*/);
/* qb,qbc: [8][N] (qbc NULL iff no conjugate); gkb: [8][2][N]; out: [8][N] */
int orc_shard_perform_pir_split(const orc_shard*, const uint64_t* qb, const uint64_t* qbc,
                                const uint64_t* gkb, uint64_t* out);
int orc_shard_perform_pir(const orc_shard*, const uint64_t* qb, const uint64_t* qbc,
                          const uint64_t* gkb /*[2][N]*/, uint64_t* out);
/* Transformed entries, modulus p1, slot domain of the reference NTT, layout
 * [row (main rows then conj rows)][lane][N]. Returns the root zeta mod p1. */
uint64_t orc_shard_export_te_p1(const orc_shard*, uint32_t* out);
/* Coefficient-domain export of one stored row (array 0 = TE, 1 = IPGK0, 2 = IPGK1;
 * stream 0 = main, 1 = conj) for lane `lane`: centered values mod p1 -> int64. */
void orc_shard_export_row_coeffs(const orc_shard*, int array, int stream, int row, int lane, int64_t* out);
int orc_shard_loop_len(const orc_shard*);
int orc_shard_requires_conj(const orc_shard*);

/* ---- double PIR (double_pir.rs) ----------------------------------------- */
/* A query block holds 6 ring elements: gk0, gk1, primary, primary_conj,
 * secondary, secondary_conj (conj entries ignored when not required).
 * fleets: shard pointer lists (shards may repeat); q[s] NULL = empty slot.
 * out[s]: [3][2][N] (a then b per part), written only for present slots. */
/* test speed-up: a shard repeated in a fleet reuses the plaintexts computed
 * for its first occurrence (identical by construction); off by default */
void orc_set_reuse_repeated(int on);
int orc_double_pir(orc_shard* const* fleet0, size_t len0, orc_shard* const* fleet1, size_t len1,
                   const uint64_t* q0, const uint64_t* q1, int secondary_group_order,
                   uint64_t* out0, uint64_t* out1, double* primary_us, double* secondary_us);
int orc_choose_index_group_secondary(int logN, size_t num_shards);

/* ---- byte wire format (pir_wrapper.rs, N = 2048 only) ------------------- */
size_t orc_condense(int bits, const uint64_t* data, size_t count, uint8_t* out);
int orc_uncondense(int bits, size_t count, const uint8_t* data, size_t len, size_t* consumed, uint64_t* out);
void orc_database_shape(size_t num_entries, int* l_primary, int* l_secondary);
size_t orc_num_databases(size_t num_entries);
size_t orc_query_bytes(int l_primary, int l_secondary);
int orc_process_queries(orc_shard* const* fleet0, size_t len0, orc_shard* const* fleet1, size_t len1,
                        const uint8_t* query0, size_t qlen0, const uint8_t* query1, size_t qlen1,
                        uint8_t* reply0, uint8_t* reply1, double* primary_us, double* secondary_us);
int orc_prepare_query(orc_rng* rng, size_t primary_idx, size_t secondary_idx, int l_primary, int l_secondary,
                      uint8_t* query, size_t* qlen, uint8_t sk_seed[32]);
int orc_process_reply(const uint8_t* reply, size_t rlen, const uint8_t sk_seed[32], uint16_t* out);

/* ---- CPU baseline in the reference's loop form (cpu_baseline.c) --------- */
/* main_pir_loop's operands in the reference layout: arrays [r][st][IPGK0,
 * IPGK1, TE][L][N][8] u32 (2 * nst * 3 * L * N * 8 words), vecs [r][st][gk0,
 * gk1, query][N][8] u32, gal [N]; query k on lane k (perform_pir_split) */
int orc_shard_k1_operands(const orc_shard* s, const uint64_t* qb /*[8][N]*/, const uint64_t* qbc,
                          const uint64_t* gkb /*[8][2][N]*/, uint32_t* arrays, uint32_t* vecs, int* gal);
/* sums [r][N][8] mod p_r -> Galois gather + iNTT + CRT -> out [8][N] mod q */
int orc_shard_k1_finish(const orc_shard* s, const uint64_t* sums, uint64_t* out);
/* n shards' main_pir_loops (lazy u64 sums reduced every orc_cpu_reduce_every()
 * rows, AVX-512 / AVX2 / scalar) on `threads` threads; sums [n][2][N][8] or NULL */
int orc_cpu_k1(int n, const uint32_t* const* arrays, int N, int L, int nst, const uint32_t* vecs, const int* gal,
               int threads, int window /* 0 = the reference's */, uint64_t* sums);
const char* orc_cpu_isa(void);
void orc_cpu_force_isa(int isa); /* 512 / 256 / 0 (scalar), -1 = by the CPU */
int orc_cpu_reduce_every(void);

/* ---- bucket packing (dns/bucket.rs:271-303), N = 2048 only --------------- */
/* 2560 bucket bytes -> 2048 ten-bit coefficients, and back */
void orc_pack_bytes_into_coeffs(const uint8_t* bucket, uint16_t* coeffs);
void orc_unpack_coeffs_into_bytes(const uint16_t* coeffs, uint8_t* bucket);

/* threads used by the oracle's OpenMP loops (0 = runtime default) */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
