// Internal types shared by the C ABI (api.cpp) and the serving loop
// (serve.cpp): shard and context objects, device scratch buffers, and the
// batched query pipeline.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <mutex>
#include <vector>

#include "../../include/dnspir.h"
#include "dn_common.hpp"

namespace dne {

using dn::LANES;
using dn::N;

// wire widths (pir_wrapper.rs:57-72)
constexpr int A_BITS = 17, B_BITS = 12, GK_BITS = 50, QRY_BITS = 34;
constexpr size_t REPLY_BYTES = 3 * (N * A_BITS / 8 + N * B_BITS / 8);
// bytes of one database entry as the DNS layer stores it: N ten-bit
// coefficients, 4 per 5 bytes (dns/bucket.rs:54-73)
constexpr size_t BUCKET_BYTES = N / 4 * 5;
// the two Galois keys at the head of every query (pir_wrapper.rs:370-372)
constexpr size_t KEY_BYTES = 2 * (N / 8) * GK_BITS;
// a device-cached key: NTT rows of gk0, gk1 in both moduli, in slot order
// (secondary phase) and in Galois-position order (primary phase)
constexpr size_t KEY_ENTRY_WORDS = 2 * 2 * 2 * N;

int hip_fail(hipError_t e, int line);
#define HIP_OK(x)                                            \
    do {                                                     \
        hipError_t e_ = (x);                                 \
        if (e_ != hipSuccess) return dne::hip_fail(e_, __LINE__); \
    } while (0)

inline int loop_len(int l) { return l < N ? l : N / 2; }
inline int nstreams(int l) { return l == N ? 2 : 1; }
// the reference's slab size, (GK_DIGITS+1)*2*N*l elements of 32 B: the bytes
// one query's inner product reads there (MEM_LOAD_DATA, base_pir.rs:1242)
inline size_t ref_slab_bytes(int l) { return (size_t)6 * N * l * 32; }
// bytes of this engine's slab: 26 bits per residue mod p1 and 25 mod p2 in
// the limb-block layout (loop length N/2: 2 halves x N/8 a-groups x 3*nst
// arrays x 32 j-chunks of blocks per modulus), 32-bit residues in the
// rotated layout
inline size_t slab_bytes(int l) {
    if (loop_len(l) != N / 2) return ref_slab_bytes(l);
    return (size_t)2 * (N / 8) * 3 * nstreams(l) * 32 * (dn::MB_BLOCK_BYTES_P1 + dn::MB_BLOCK_BYTES_P2);
}

// A bump region of device memory that a context's scratch buffers are carved
// from (instead of hipMalloc) while it is bound. base == nullptr with a huge
// cap only counts bytes (sizing pass).
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
};

// grow-only device buffer, from hipMalloc or from the bound arena
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    Arena* arena = nullptr;
    int ensure(size_t bytes) {
        if (bytes <= n) return 0;
        if (arena) {
            const size_t off = (arena->used + 255) & ~(size_t)255;
            if (off + bytes > arena->cap) return DNSPIR_E_OOM;
            p = arena->base + off;
            n = bytes;
            arena->used = off + bytes;
            return 0;
        }
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) return DNSPIR_E_OOM;
        n = bytes;
        return 0;
    }
    // forget the arena storage (the arena owns it)
    void detach() {
        if (arena) {
            p = nullptr;
            n = 0;
        }
    }
    template <class X>
    X* as() const {
        return reinterpret_cast<X*>(p);
    }
    ~DevBuf() {
        if (p && !arena) (void)hipFree(p);
    }
};

// Per-query Galois-key sources of one batch: for query q, key_src[q] (a
// cache entry to read its keys from instead of its bytes) and key_dst[q] (a
// cache entry to store the keys decoded from its bytes), each may be null.
struct BatchKeys {
    const uint32_t* const* src = nullptr;
    uint32_t* const* dst = nullptr;
};

int ensure_init();
const dn::DevTables* tables();
double now_us();

// Offsets of the six query vectors in the wire format (pir_wrapper.rs:370-384):
// gk0 gk1 (50 bits) || primary (34) [|| conj] || secondary (34) [|| conj];
// offs[v] = -1 for an absent conjugate. Returns the total length.
size_t query_layout(int l_p, int l_s, int offs[6], int bits[6]);

}  // namespace dne

struct dnspir_shard {
    int l, L, nst, seeds;
    uint32_t* slab;
    uint64_t* a_coef;  // [8][N] mod q, the get_a result
    uint32_t* a17;     // [8][N] a mod-switched to 17 bits
    bool has_db;
};

struct dnspir_ctx {
    std::mutex mu;  // one call at a time per context
    hipStream_t stream = nullptr;
    bool own = false;
    // size-dependent scratch (arena-capable)
    dne::DevBuf se, M, te, aout, cout, qcoef, qs, qp, acc, outb, ql, acoef, bcoef, a17o, b12o, qbytes, rbytes;
    // small persistent buffers (pointer lists, lane lists, the two-CU chain's
    // exchange words and timeout flag)
    dne::DevBuf ptrs, idx, keyp, xch;
    // two-stream inner product of single queries: the C table and the weight tiles
    dne::DevBuf wscan, wtile;
    int k1_mode = 0;  // 0: two streams from K1W_MIN_SHARDS shards on, 2: always, 3: never
    int chain_cus = 0;  // most CUs per secondary chain (0: by chain count; 1, 2 or 4)
    unsigned* h_err = nullptr;  // pinned copy of the chain timeout flag
    // pinned staging of dnspir_process_queries (host buffers): two queries of
    // up to HOST_QCAP bytes, two replies
    static constexpr size_t HOST_QCAP = 80000;
    uint8_t* h_q = nullptr;
    uint8_t* h_r = nullptr;
    int exchange_polls = 1 << 21;  // bound of each exchange wait of the two-CU chain
    bool pair_pending = false;  // h_err is to be checked after the next sync
    int pair_held = 0;          // multi-CU chain blocks this context has reserved on the device
    // the last secondary phase's arguments and the reply encodings that
    // followed it (repeated with one CU per chain when a partner wait timed out)
    struct SecondaryCall {
        const uint16_t* entries = nullptr;
        long long col_stride = 0;
        const long long* d_off = nullptr;
        const int *d_ncols = nullptr, *d_slot = nullptr;
        int nlanes = 0, l_s = 0;
        int n_enc = 0, enc_lane0[2] = {0, 0}, enc_nq[2] = {0, 0};
        uint8_t* enc_reply[2] = {nullptr, nullptr};
    } last_sec;
    dnspir_ctx_stats stats{};
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // start, primary end, end, k1 start, k1 end
    int k1_launches = 0;
    // the slab / a17 pointer lists last uploaded to ptrs (per fleet slot) and
    // the device address they were uploaded to
    std::vector<const uint32_t*> cached[2];
    const void* cached_base[2] = {nullptr, nullptr};
    std::vector<uint64_t> stage;
    std::vector<long long> h_idx;  // host staging of the lane lists (outlives the async copy)
    const void* idx_uploaded = nullptr;  // idx buffer that holds h_idx
    std::vector<const void*> h_keyp;
    bool k1_timed = true;  // ev[3]/ev[4] bracket the first inner-product launch
    // batched pipeline captured as two HIP graphs (primary, secondary phase),
    // replayed while its key matches
    bool graphs = true;
    bool graphs_broken = false;  // capture failed once: stay eager
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    std::vector<uintptr_t> graph_key;
    void drop_graphs() {
        for (auto& g : graph_exec)
            if (g) (void)hipGraphExecDestroy(g), g = nullptr;
        graph_key.clear();
    }
    ~dnspir_ctx() {
        drop_graphs();
        if (h_err) (void)hipHostFree(h_err);
        if (h_q) (void)hipHostFree(h_q);
        if (h_r) (void)hipHostFree(h_r);
    }

    dne::DevBuf* scratch_bufs[19] = {&se,    &M,     &te,    &aout, &cout, &qcoef,  &qs,     &qp,    &acc,  &outb,
                                     &ql,    &acoef, &bcoef, &a17o, &b12o, &qbytes, &rbytes, &wscan, &wtile};
    // carve all scratch from `a` from now on (nullptr = back to hipMalloc);
    // storage the buffers held before is released or forgotten
    void bind_arena(dne::Arena* a) {
        for (auto* b : scratch_bufs) {
            if (!b->arena && b->p) (void)hipFree(b->p);
            b->p = nullptr;
            b->n = 0;
            b->arena = a;
        }
    }
    // the same for the small persistent buffers too (a context whose every
    // buffer lives in a caller's region; never called on one that used hipMalloc)
    void bind_all(dne::Arena* a) {
        bind_arena(a);
        for (auto* b : {&ptrs, &idx, &keyp, &xch}) {
            b->p = nullptr;
            b->n = 0;
            b->arena = a;
        }
        cached[0].clear();
        cached[1].clear();
        idx_uploaded = nullptr;
    }
};

namespace dne {

int ctx_init(dnspir_ctx* c, hipStream_t s);
// chains at most this many run on two CUs each (chain_pair_kernel): up to one
// block per CU of the 256; at most QUAD_MAX_CHAINS on four (chain_quad_kernel)
constexpr int PAIR_MAX_CHAINS = 128;
constexpr int QUAD_MAX_CHAINS = 16;
// fleets of at least this many limb-block shards take the two-stream inner
// product for single queries (below, preparing the weights costs about what
// the third stream would)
constexpr int K1W_MIN_SHARDS = 8;
// whether a primary pass of nq <= 2 queries over `len` shards of group order
// l_p on this context takes the two-stream inner product
inline bool use_k1w(const dnspir_ctx* c, int l_p, size_t len, int nq) {
    return (nq == 1 || nq == 2) && 2 * l_p >= N &&
           (c->k1_mode == 2 || (c->k1_mode == 0 && len >= (size_t)K1W_MIN_SHARDS));
}
// after the stream sync that follows a call: releases the call's two-CU
// chain reservation; if that chain timed out waiting for a partner block,
// runs the secondary phase again with one CU per chain (*retried = true)
int check_pair(dnspir_ctx* c, bool* retried = nullptr);
// sizes the persistent pointer / lane-list buffers for batches of up to max_nq
int ctx_presize(dnspir_ctx* c, int max_nq);

// nq queries (packed at stride qlen in device memory, full wire format)
// against one fleet; replies packed REPLY_BYTES each at rdev. Caller holds
// the context's lock. sync = false: returns once everything is enqueued on
// the context's stream; c->ev[1] then marks the end of the primary phase
// (the last read of the shards) and tm is not filled. graph = true (and the
// context's graphs on): the device work runs as a captured HIP graph.
int batch_impl(dnspir_ctx* c, const dnspir_shard* const* fleet, size_t len, const uint8_t* qdev, size_t qlen, int nq,
               uint8_t* rdev, dnspir_timing* tm, double t0, const BatchKeys* keys, bool sync = true, bool graph = false);

// primary group order and fleet length of a serving loop's database
int server_db_shape(dnspir_server* s, int db, int* l_p, size_t* len);

// device scratch bytes batch_impl takes from an arena for nq queries
// against a fleet of `len` shards of order l_p, plus the query and reply
// staging (qbytes, rbytes) of a serving slot
size_t batch_scratch_bytes(int l_p, size_t len, int nq, size_t qlen);
// carves the staging and every scratch buffer batch_impl needs
int batch_reserve(dnspir_ctx* c, int l_p, size_t len, int nq, size_t qlen);

}  // namespace dne
