/*
 * dnspir.h -- C ABI of the MI355X PIR response engine.
 *
 * Mirrors the `dnspir-pir` crate API (crates/pir/src/lib.rs:35-41) so that the
 * reference's Rust host (or any C/C++/Python caller) can bind it over FFI:
 * each entry point names the reference item it replaces. All values are
 * exact: ring elements cross the boundary as uint64_t[N] canonical
 * coefficients in [0, q), q = 67104769 * 33550337, N = 2048; database entries
 * as uint16_t[N] coefficients in [0, 1025). Every call returns an int status
 * (0 = OK, negative = DNSPIR_E_*) where the reference panics.
 */
#ifndef DNSPIR_H
#define DNSPIR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DNSPIR_N 2048
#define DNSPIR_LANES 8

#define DNSPIR_OK 0
#define DNSPIR_E_SHORT_QUERY (-1) /* "not enough data", pir_wrapper.rs:124 */
#define DNSPIR_E_CONJ (-2)        /* conj presence mismatch, base_pir.rs:965, double_pir.rs:160 */
#define DNSPIR_E_COUNT (-3)       /* entry count != 8 * l, base_pir.rs:693 */
#define DNSPIR_E_SHAPE (-4)       /* index group / fleet shape, base_pir.rs:150-154, double_pir.rs:148-152 */
#define DNSPIR_E_ARG (-5)         /* null pointer or bad argument */
#define DNSPIR_E_ALIGN (-6)       /* slab misaligned or too small, base_pir.rs:299-300 */
#define DNSPIR_E_OOM (-7)         /* device allocation failed */
#define DNSPIR_E_HIP (-8)         /* a HIP runtime call failed */
#define DNSPIR_E_DECODE (-9)      /* reply digits out of range, double_pir.rs:88 */
#define DNSPIR_E_NOINIT (-10)     /* dnspir_init not called / no device */
#define DNSPIR_E_NOKEY (-11)      /* keyless query, but the client's Galois keys are not cached */
#define DNSPIR_E_CLOSED (-12)     /* the server is shutting down */
#define DNSPIR_E_IO (-13)         /* a shard file could not be opened, written or closed */
#define DNSPIR_E_FORMAT (-14)     /* not a shard file of this build's layout, or corrupt / truncated */
#define DNSPIR_E_SCRATCH (-15)    /* the call needs more than the fixed scratch of dnspir_process_queries_noalloc */
#define DNSPIR_E_FRAME (-16)      /* front end: frame over 64 MiB, empty request or unknown command,
                                     protocol.rs:111-113, server.rs:756-761, 798-803 */

typedef struct dnspir_shard dnspir_shard;
typedef struct dnspir_ctx dnspir_ctx;
typedef struct dnspir_rng dnspir_rng;

/* per-call phase timing; replaces the racy globals PRIMARY_PHASE_MICROS /
 * SECONDARY_PHASE_MICROS / MEM_LOAD_DATA (double_pir.rs:229-243, base_pir.rs:1251) */
typedef struct {
    double primary_us;       /* device time of the primary phase */
    double secondary_us;     /* device time of the secondary phase */
    double total_us;         /* wall time of the call */
    uint64_t bytes_streamed; /* bytes read from the shard slabs by the primary inner product */
    double k1_us;            /* device time of the primary inner-product launches */
    int k1_launches;         /* number of those launches */
    uint64_t ref_bytes;      /* the same pass in the reference's accounting, 6*N*l*32 B per
                              * shard (MEM_LOAD_DATA, base_pir.rs:1242); the slabs here keep
                              * 26 bits per residue mod p1 and 25 mod p2, so
                              * bytes_streamed is 51/64 of it */
} dnspir_timing;

/* Initialise the library on HIP device `device` (builds the NTT / Galois
 * tables). Must precede every other device call. */
int dnspir_init(int device);
const char* dnspir_strerror(int code);
/* text of the last HIP failure behind a DNSPIR_E_HIP (call site and HIP message) */
const char* dnspir_last_error(void);

/* ---- one shard: PIRDatabase (base_pir.rs:225-1247) ------------------------ */

/* Bytes of caller-owned device memory a shard of index-group order
 * `group_order` needs (replaces PIRDatabase::required_memory_general * 32 B,
 * base_pir.rs:369-371): the slab, which keeps 26 bits per residue mod p1 and
 * 25 mod p2 for group orders N/2 and N (51/64 of the reference's size) and
 * 32 bits below,
 * plus 192 KiB for the cached a-part. */
size_t dnspir_required_bytes(int group_order);
/* PIRDatabase::create (base_pir.rs:280-349). used_seeds: 0 = FirstSet,
 * 1 = SecondSet. dev_mem must be 256-byte aligned device memory of at least
 * dnspir_required_bytes(group_order) bytes; the shard borrows it. */
int dnspir_shard_create(int group_order, int used_seeds, void* dev_mem, size_t bytes, dnspir_shard** out);
/* frees the host handle only; the caller owns dev_mem */
int dnspir_shard_destroy(dnspir_shard* shard);
int dnspir_shard_group_order(const dnspir_shard* shard);
/* PIRDatabase::set_db (base_pir.rs:632-637): exactly 8 * group_order entries,
 * host memory [n_entries][N]. */
int dnspir_shard_set_db(dnspir_shard* shard, const uint16_t* entries, size_t n_entries);
/* Same with the entries already in device memory (no host copy). The call
 * runs on the library's own stream: work that produces `entries_dev` on other
 * streams must be complete before the call. */
int dnspir_shard_set_db_device(dnspir_shard* shard, const uint16_t* entries_dev, size_t n_entries);
/* Persisted shards (checkpoint / resume of preprocessing; no counterpart in
 * the reference, whose server preprocesses at start-up, server.rs:664-739).
 * dnspir_shard_save writes a preprocessed shard's device state to `path`: a
 * 64-byte header (magic "DNSPIRSH", version, N, group order, used seeds, a
 * fingerprint of this build's device layout, payload size, checksum), then the
 * dnspir_required_bytes(group_order) bytes of its device memory.
 * dnspir_shard_file_info reads the header (no device needed).
 * dnspir_shard_load checks the header against this build, copies the state
 * into caller-owned device memory as dnspir_shard_create takes it, verifies
 * the checksum and returns a shard ready for queries. DNSPIR_E_IO: the file
 * cannot be opened or written; DNSPIR_E_FORMAT: wrong magic, version, layout
 * or size, a truncated file or a checksum mismatch. */
int dnspir_shard_save(const dnspir_shard* shard, const char* path);
int dnspir_shard_file_info(const char* path, int* group_order, int* used_seeds, size_t* bytes);
int dnspir_shard_load(const char* path, void* dev_mem, size_t bytes, dnspir_shard** out);
/* set_db from the DNS layer's entries: up to 8 * group_order buckets of
 * DNSPIR_BUCKET_BYTES bytes (host memory, back to back), each packed into N
 * ten-bit coefficients on the device as pack_bytes_into_coeffs does
 * (dns/bucket.rs:289-303); the shard's remaining entries are zero, as the
 * last shard of a database is padded (pir_wrapper.rs:235-237,
 * server.rs:664-739). DNSPIR_E_COUNT for more than 8 * group_order buckets. */
#define DNSPIR_BUCKET_BYTES 2560
int dnspir_shard_set_db_buckets(dnspir_shard* shard, const uint8_t* buckets, size_t n_buckets);
/* pack_bytes_into_coeffs (dns/bucket.rs:289-303) on the host: n buckets ->
 * coeffs [n][N] */
int dnspir_pack_bytes_into_coeffs(const uint8_t* buckets, size_t n, uint16_t* coeffs);
/* PIRDatabase::get_a (base_pir.rs:897-903): out [8][N] host */
int dnspir_shard_get_a(const dnspir_shard* shard, uint64_t* out);
/* PIRDatabase::perform_pir (base_pir.rs:945-952): qb [N], qb_conj [N] (NULL
 * iff the index group needs no conjugate), gkb [2][N]; out_b [8][N] host. */
int dnspir_shard_perform_pir(const dnspir_shard* shard, const uint64_t* qb, const uint64_t* qb_conj,
                             const uint64_t* gkb, uint64_t* out_b);
/* PIRDatabase::perform_pir_split (base_pir.rs:958-1078): query k on lane k;
 * qb [8][N], qb_conj [8][N] or NULL, gkb [8][2][N]; out_b [8][N] host. */
int dnspir_shard_perform_pir_split(const dnspir_shard* shard, const uint64_t* qb, const uint64_t* qb_conj,
                                   const uint64_t* gkb, uint64_t* out_b);

/* ---- query contexts: one per stream / serving slot ------------------------ */
/* Device scratch of the per-query secondary database (the thread-local
 * SECONDARY_DB_MEMORY of double_pir.rs:43-56), grow-only. `stream` is a
 * hipStream_t (NULL = a new stream owned by the context). */
int dnspir_ctx_create(void* stream, dnspir_ctx** out);
int dnspir_ctx_destroy(dnspir_ctx* ctx);
void* dnspir_ctx_stream(dnspir_ctx* ctx);
/* Batched calls (dnspir_process_batch*) run their device work as a captured
 * HIP graph, replayed while shapes and buffers stay the same (default on).
 * Returns 1 if capture failed once on this context (it then launches
 * eagerly), else 0. */
int dnspir_ctx_set_graphs(dnspir_ctx* ctx, int on);
/* Few-query calls run each secondary chain on four CUs (up to 16 chains:
 * one per modulus and half of the slots) or two CUs (up to 128 chains: one
 * per modulus) that exchange data every step, while the blocks of all such
 * launches in flight on the device fit its CUs, one block per CU; otherwise
 * on fewer CUs per chain. Each wait for a partner CU ends after `max_polls`
 * polls (default 2^21, about a second); the call then runs its secondary
 * phase again on one CU per chain, which never waits, and still answers
 * exactly. 0 makes every wait give up at once (to exercise that path). No
 * counterpart in the reference. */
int dnspir_ctx_set_exchange_polls(dnspir_ctx* ctx, int max_polls);
/* The most CUs per secondary chain on this context: 0 (default) by the
 * chain count as above, 1, 2 or 4. Every setting gives the same bytes. No
 * counterpart in the reference. */
int dnspir_ctx_set_chain_cus(dnspir_ctx* ctx, int cus);
/* counters of a context's secondary phases (NULL ctx = the default context) */
typedef struct {
    uint64_t pair_runs;      /* phases run on two or four CUs per chain */
    uint64_t pair_refused;   /* phases run on one CU per chain because the device's
                                multi-CU budget was taken by other calls in flight */
    uint64_t pair_retries;   /* multi-CU phases whose partner wait timed out and that
                                were run again on one CU per chain */
    uint64_t quad_runs;      /* the phases of pair_runs run on four CUs per chain */
} dnspir_ctx_stats;
int dnspir_ctx_get_stats(dnspir_ctx* ctx, dnspir_ctx_stats* out);
/* Inner product of single queries on shards with loop length N/2. mode 3:
 * stream all three shard arrays (TE, IPGK0, IPGK1), as main_pir_loop does
 * (base_pir.rs:1153-1208). mode 2: stream TE and IPGK1 only and fold IPGK0
 * into per-query weights (IPGK0 = sigma(a_i) - 2^26 IPGK1, with the chain
 * a_i linear in the stored rows), a third fewer bytes. mode 0 (default):
 * mode 2 for fleets of 8 or more shards, else mode 3. Every mode gives the
 * same bytes. Two queries over one fleet (both slots of a call on the same
 * shards, or a two-query batch) share one pass; batches of 3 or more queries
 * always stream three arrays. No counterpart in the reference. */
int dnspir_ctx_set_k1_mode(dnspir_ctx* ctx, int mode);

/* ---- two-level composition: perform_batched_double_pir (double_pir.rs:126-228)
 * fleets: random-access shard lists (shards may repeat). q0/q1: query blocks
 * [6][N] = gk0, gk1, primary, primary_conj, secondary, secondary_conj (conj
 * entries ignored when not required), NULL = empty batch slot.
 * out0/out1: [3][2][N] = (a, b) of the three reply ciphertexts. ctx NULL =
 * the library's default context. */
int dnspir_double_pir(dnspir_ctx* ctx, const dnspir_shard* const* fleet0, size_t len0,
                      const dnspir_shard* const* fleet1, size_t len1, const uint64_t* q0, const uint64_t* q1,
                      uint64_t* out0, uint64_t* out1, dnspir_timing* timing);

/* ---- byte API: pir_wrapper.rs ---------------------------------------------- */
/* process_queries (pir_wrapper.rs:349-409): query bytes in, reply bytes out
 * (dnspir_reply_bytes() each), host buffers; query1 NULL = empty slot. */
int dnspir_process_queries(dnspir_ctx* ctx, const dnspir_shard* const* fleet0, size_t len0,
                           const dnspir_shard* const* fleet1, size_t len1, const uint8_t* query0, size_t qlen0,
                           const uint8_t* query1, size_t qlen1, uint8_t* reply0, uint8_t* reply1,
                           dnspir_timing* timing);
/* Same with query and reply bytes in device memory on the context's stream
 * (no host copies; the call returns after the stream has finished). Work on
 * other streams that writes query0_dev/query1_dev must be complete first. */
int dnspir_process_queries_device(dnspir_ctx* ctx, const dnspir_shard* const* fleet0, size_t len0,
                                  const dnspir_shard* const* fleet1, size_t len1, const uint8_t* query0_dev,
                                  size_t qlen0, const uint8_t* query1_dev, size_t qlen1, uint8_t* reply0_dev,
                                  uint8_t* reply1_dev, dnspir_timing* timing);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dnspir_process_queries_device with every scratch buffer
 * carved from one fixed device array of DNSPIR_NOALLOC_SCRATCH_BYTES per GPU,
 * set up by dnspir_init and held until the process ends. The call allocates
 * and frees nothing (no hipMalloc / hipFree) and runs on a stream the library
 * owns for these calls. It answers one query (query1 / reply1 NULL) against
 * any fleet the reference allows (up to N^2 entries), or both slots when their
 * scratch fits; larger calls fail with DNSPIR_E_SCRATCH before any device work.
 * Every such call on a GPU uses that GPU's one array: calls are serialised by a
 * lock, and overlapping them from several threads only waits; device work of
 * other streams must not read the array. Replies equal
 * dnspir_process_queries_device's byte for byte. */
#define DNSPIR_NOALLOC_SCRATCH_BYTES ((size_t)224 << 20)
int dnspir_process_queries_noalloc(const dnspir_shard* const* fleet0, size_t len0, const dnspir_shard* const* fleet1,
                                   size_t len1, const uint8_t* query0_dev, size_t qlen0, const uint8_t* query1_dev,
                                   size_t qlen1, uint8_t* reply0_dev, uint8_t* reply1_dev, dnspir_timing* timing);
/* Batched serving: nq independent queries (each answered exactly as
 * process_query, pir_wrapper.rs:328-337) against one fleet of half- or
 * full-size shards, sharing one pass over the shards (an int8-MFMA contraction
 * per 64 queries). Queries are packed back to back, qlen bytes each (all the
 * same size for one fleet); replies are packed dnspir_reply_bytes() each.
 * No counterpart in the reference, which batches at most 2 queries
 * (BATCH_COUNT, double_pir.rs:37). Host buffers: */
int dnspir_process_batch(dnspir_ctx* ctx, const dnspir_shard* const* fleet, size_t len, const uint8_t* queries,
                         size_t qlen, int nq, uint8_t* replies, dnspir_timing* timing);
/* Same with queries and replies in device memory. */
int dnspir_process_batch_device(dnspir_ctx* ctx, const dnspir_shard* const* fleet, size_t len, const uint8_t* queries_dev,
                                size_t qlen, int nq, uint8_t* replies_dev, dnspir_timing* timing);
/* Autotuner of the batched inner product: times every candidate kernel
 * (query-tile staging, prefetch depth, unrolling) for batches of nq queries
 * (3..64) on this fleet with random operands, checks that all of them give
 * identical results, and makes the fastest the one used for that batch class
 * (column tiles = nq > 32 ? 2 : 1, conjugated streams or not) from then on.
 * The limb split is not tuned: 4 x 4 signed base-256 digits is the only exact
 * int8 split of 26-bit residues. entries: up to max_entries results. */
typedef struct {
    char name[48];
    int column_tiles;
    double ms;               /* mean device time of one launch */
    int exact;               /* 1: output identical to the first candidate's */
    int chosen;
} dnspir_tune_entry;
int dnspir_autotune_batch(dnspir_ctx* ctx, const dnspir_shard* const* fleet, size_t len, int nq, dnspir_tune_entry* entries,
                          int max_entries, int* n_entries);
/* Diagnostic for the batched inner product's roofline (no counterpart in the
 * reference): runs its int8 MFMA instruction mix (14 accumulator tiles per
 * wave, 32 v_mfma_i32_32x32x32_i8 per K-step, 8192 blocks of 4 waves, 96
 * K-steps) with operands held in registers, so no memory traffic, on the
 * context's stream, and returns the rate in int8 TOPS (the mean of `reps`
 * launches after one warm-up). zero_operands = 0 uses hashed digits that change
 * every step, as real shards and queries do; 1 keeps every digit zero. The
 * clock the GPU holds under int8 MFMA load depends on how much the operand
 * bits switch, so the first is the ceiling the batched kernels face on real
 * data and the second the switching-free rate. Allocates 8 MiB of device
 * memory for the call. */
int dnspir_probe_mfma_rate(dnspir_ctx* ctx, int zero_operands, int reps, double* tops);
/* Diagnostic for the single-query inner product's roofline (no counterpart in
 * the reference): a plain streaming read, with the 16-byte loads the K1
 * kernels issue, of every listed shard's slab (dnspir_required_bytes less the
 * cached a-part), `reps` times after one warm-up, on the context's stream;
 * returns GB/s. Shards must hold a database; a repeated shard is read again.
 * Allocates a few MiB of device memory for the call. */
int dnspir_probe_read_rate(dnspir_ctx* ctx, const dnspir_shard* const* fleet, size_t len, int reps, double* gbs);
/* get_database_shape (pir_wrapper.rs:219-231): index-group orders.
 * DNSPIR_E_SHAPE (outputs untouched) where the reference asserts: num_entries
 * of 0 or above N^2 (pir_wrapper.rs:221-222). */
int dnspir_database_shape(size_t num_entries, int* l_primary, int* l_secondary);
/* num_databases (pir_wrapper.rs:238-241); 0 where get_database_shape refuses */
size_t dnspir_num_databases(size_t num_entries);
/* choose_index_group_secondary (pir_wrapper.rs:166-175): the group order, or
 * DNSPIR_E_SHAPE for more than N/8 shards (pir_wrapper.rs:168) */
int dnspir_choose_index_group_secondary(size_t num_shards);
size_t dnspir_query_bytes(int l_primary, int l_secondary);
size_t dnspir_reply_bytes(void);

/* ---- serving: continuous batching over resident databases -----------------
 * No counterpart in the reference, whose server answers each query on its
 * connection thread (dns/server.rs:747-807 -> process_query). Requests queue
 * per database; `slots` workers (one query context and HIP stream each) take
 * up to max_batch waiting requests of the database with the oldest request
 * whenever they are free, if the batch's device scratch fits in the free part
 * of a `scratch_budget`-byte arena allocated once (admission by HBM budget;
 * no hipMalloc while serving). Batches of different slots overlap on the
 * device. Each reply is byte-identical to the query's process_queries reply.
 * Per-client Galois keys stay in HBM (key_cache_entries clients, LRU): after
 * one full query a client may send keyless ones (its query without the
 * dnspir_key_bytes() key prefix, flag DNSPIR_QUERY_KEYLESS); once its entry is
 * evicted they fail with DNSPIR_E_NOKEY and the client sends the full query. */
typedef struct dnspir_server dnspir_server;

#define DNSPIR_QUERY_KEYLESS 1

typedef struct {
    int slots;               /* batches on the device at once, one stream each (1..8) */
    int max_batch;           /* queries per batch (1..128) */
    size_t scratch_budget;   /* device bytes all in-flight batches may hold together */
    int key_cache_entries;   /* clients whose Galois keys stay on the device (0 = none) */
    double batch_window;     /* while other batches run, a request may wait up to this many times
                                its database's measured primary-pass time for more requests to
                                join its batch (0 = start whenever a slot is free) */
} dnspir_server_config;

typedef struct {
    int status;              /* 0 or DNSPIR_E_* */
    int db;
    int batch_size;          /* queries in the batch that answered the request */
    int key_cached;          /* 1: the Galois keys came from the device cache */
    int resent;              /* dnspir_server_replay: 1 = keys sent after DNSPIR_E_NOKEY */
    double submit_us;        /* monotonic clock, microseconds */
    double start_us;         /* batch admitted */
    double done_us;          /* reply in host memory */
} dnspir_request_info;

typedef struct {
    uint64_t completed, failed, batches;
    uint64_t key_hits;        /* keyless queries served from the cache */
    uint64_t key_inserts;     /* keys stored from full queries */
    uint64_t key_evictions;
    uint64_t key_misses;      /* keyless queries answered DNSPIR_E_NOKEY */
    uint64_t admission_waits; /* times waiting requests had to wait for scratch */
    size_t scratch_in_use_peak;
    int inflight_peak;        /* most batches on the device at once */
    uint64_t pair_refused;    /* batches whose secondary chains took one CU each because
                                 the device's two-CU budget was in use (dnspir_ctx_stats) */
    uint64_t pair_retries;    /* batches whose two-CU chains timed out and ran again on one CU */
} dnspir_server_stats;

/* one request of an open-loop load (dnspir_server_replay) */
typedef struct {
    int db;
    uint64_t client;
    double arrival_us;        /* submit time relative to the start of the replay */
    const uint8_t* query;     /* keyless query bytes */
    size_t qlen;
    const uint8_t* keys;      /* the client's key bytes, sent after DNSPIR_E_NOKEY */
} dnspir_arrival;

/* length of the Galois-key prefix of a wire query (2 x 12,800 B) */
size_t dnspir_key_bytes(void);
/* device scratch of one batch of nq queries against num_shards shards of
 * order l_primary (for sizing scratch_budget); 0 = unsupported shape */
size_t dnspir_batch_scratch_bytes(int l_primary, size_t num_shards, int nq);
int dnspir_server_create(const dnspir_server_config* cfg, dnspir_server** out);
/* registers a fleet (half- or full-size shards) before the first submit */
int dnspir_server_add_database(dnspir_server* s, const dnspir_shard* const* fleet, size_t len, int* db);
/* the largest batch of database `db` whose scratch fits the budget */
int dnspir_server_max_batch(dnspir_server* s, int db);
/* queues one query (bytes copied); returns at once */
int dnspir_server_submit(dnspir_server* s, int db, uint64_t client, const uint8_t* query, size_t qlen, int flags,
                         uint64_t* ticket);
/* blocks until the request is answered; reply: dnspir_reply_bytes() or NULL */
int dnspir_server_wait(dnspir_server* s, uint64_t ticket, uint8_t* reply, dnspir_request_info* info);
/* submits arrivals[i] at arrival_us (keyless first, full after DNSPIR_E_NOKEY),
 * then waits for all; replies: n * dnspir_reply_bytes() or NULL */
int dnspir_server_replay(dnspir_server* s, const dnspir_arrival* arrivals, size_t n, uint8_t* replies,
                         dnspir_request_info* infos);
int dnspir_server_get_stats(dnspir_server* s, dnspir_server_stats* out);
/* stops the workers after their current batches; queued requests end with
 * DNSPIR_E_CLOSED. Call after every dnspir_server_wait has returned. */
int dnspir_server_destroy(dnspir_server* s);

/* ---- DNS-over-PIR front end: the reference server's TCP protocol ----------
 * (crates/dns-over-pir/src/dns/protocol.rs, server.rs:152-173, 747-807).
 * Frames are a 4-byte big-endian length and at most 64 MiB of payload; a
 * request is [cmd][body]. CMD_INFO answers encode_info's bytes (flags,
 * num_entries u32, zone name length u32 + bytes, dump size u32, salt u64, all
 * big endian); CMD_QUERY answers the reply bytes of the query in its body
 * (the serving loop's dnspir_server_submit / _wait in place of process_query,
 * so the queries of all connections are batched together); CMD_DUMP answers
 * the dump bytes; a capability the front end lacks answers the single byte
 * DNSPIR_RESP_UNAVAILABLE. A connection ends cleanly when the peer closes
 * (also inside a frame, as read_exact's UnexpectedEof does there); an
 * oversized frame, an empty request or an unknown command end it with
 * DNSPIR_E_FRAME, a query the serving loop refuses with that loop's error.
 * The zone parser, bucket builder and resolver of the reference's DNS layer
 * are not part of this library. */
#define DNSPIR_CMD_INFO 1
#define DNSPIR_CMD_QUERY 2
#define DNSPIR_CMD_DUMP 3
#define DNSPIR_FLAG_PIR_AVAILABLE 1
#define DNSPIR_FLAG_DUMP_AVAILABLE 2
#define DNSPIR_RESP_UNAVAILABLE 0xFF
#define DNSPIR_FRAME_MAX ((size_t)64 << 20)

typedef struct dnspir_frontend dnspir_frontend;
typedef struct {
    dnspir_server* server;   /* answers CMD_QUERY; NULL = no PIR (RESP_UNAVAILABLE); borrowed:
                                destroy the front end before this server */
    int db;                  /* its database (dnspir_server_add_database) */
    size_t num_entries;      /* entries advertised by INFO; get_database_shape of it must
                                give that database's group order and shard count */
    const char* zone_name;   /* INFO zone name ("" = the root zone) */
    uint64_t hash_salt;      /* INFO hash salt */
    const uint8_t* dump;     /* bytes answered to CMD_DUMP (the gzipped zone), NULL = none */
    size_t dump_len;
} dnspir_frontend_config;

typedef struct {
    uint64_t connections, frames_in, frames_out, bytes_in, bytes_out;  /* bytes include length prefixes */
    uint64_t queries, infos, dumps;
    uint64_t unavailable;    /* RESP_UNAVAILABLE answers */
    uint64_t errors;         /* listener connections that ended with an error */
} dnspir_frontend_stats;

int dnspir_frontend_create(const dnspir_frontend_config* cfg, dnspir_frontend** out);
/* handle (server.rs:747-807) on one connected stream socket (or pipe pair
 * read and written through one fd) in the calling thread: 0 when the peer
 * closed, else the error that ended it. The fd is not closed. */
int dnspir_frontend_handle_fd(dnspir_frontend* fe, int fd);
/* run (server.rs:152-173): listens on host:port (IPv4, "0.0.0.0" = any; port 0
 * = a free one, written to *bound_port) and handles each connection on a
 * thread of its own until dnspir_frontend_destroy */
int dnspir_frontend_listen(dnspir_frontend* fe, const char* host, int port, int* bound_port);
int dnspir_frontend_get_stats(dnspir_frontend* fe, dnspir_frontend_stats* out);
/* stops the listener, ends open connections and waits for their threads */
int dnspir_frontend_destroy(dnspir_frontend* fe);

/* ---- client side (host only): prepare_query / process_reply --------------- */
/* StdRng stand-in: ChaCha12 in rand_chacha 0.3 word order */
dnspir_rng* dnspir_rng_new(const uint8_t seed[32]);
void dnspir_rng_free(dnspir_rng* rng);
/* prepare_query_general (pir_wrapper.rs:276-317); query buffer of
 * dnspir_query_bytes(l_primary, l_secondary) bytes */
int dnspir_prepare_query(dnspir_rng* rng, size_t primary_idx, size_t secondary_idx, int l_primary, int l_secondary,
                         uint8_t* query, size_t* qlen, uint8_t sk_seed[32]);
/* The two halves of prepare_query for a client with long-lived keys: draws
 * the secret-key seed and writes the Galois-key bytes (dnspir_key_bytes());
 * then each dnspir_client_query writes one keyless query for that key.
 * keys || client_query bytes equal prepare_query's bytes for the same rng. */
int dnspir_client_keys(dnspir_rng* rng, uint8_t sk_seed[32], uint8_t* keys);
int dnspir_client_query(dnspir_rng* rng, const uint8_t sk_seed[32], size_t primary_idx, size_t secondary_idx,
                        int l_primary, int l_secondary, uint8_t* query, size_t* qlen);
/* process_reply (pir_wrapper.rs:420-440): out [N] coefficients mod 1025 */
int dnspir_process_reply(const uint8_t* reply, size_t rlen, const uint8_t sk_seed[32], uint16_t* out);

#ifdef __cplusplus
}
#endif
#endif /* DNSPIR_H */
