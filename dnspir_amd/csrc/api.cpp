// C ABI of the MI355X PIR response engine (include/dnspir.h): shard
// lifecycle, query contexts, the two-level composition and the byte API,
// plus the host-side client operations.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "dn_common.hpp"
#include "engine.hpp"
#include "host_math.hpp"
#include "kernels.hpp"

using namespace dn;
using namespace dne;

namespace {
thread_local char g_last_error[256] = "";
}  // namespace

// Fixed device scratch of dnspir_process_queries_noalloc, one copy per GPU,
// held until the process ends. In host code `scratch` only names the symbol:
// its device address comes from hipGetSymbolAddress at dnspir_init
// (Global::NoAlloc::base), and only that address is ever passed on.
static __device__ __align__(256) unsigned char scratch[DNSPIR_NOALLOC_SCRATCH_BYTES];

int dne::hip_fail(hipError_t e, int line) {
    snprintf(g_last_error, sizeof(g_last_error), "api.cpp:%d: %s", line, hipGetErrorString(e));
    return DNSPIR_E_HIP;
}

namespace {

constexpr size_t EXTRA_BYTES = (size_t)LANES * N * 8 + (size_t)LANES * N * 4;  // a_coef + a17
constexpr size_t PTR_SLOTS = 2 * 2 * 256 + 1;  // two fleets x (slabs, a17s) + one scratch slot

bool valid_group(int l) {
    if (l <= 0 || l > N || (N % l) != 0) return false;
    return l == N || 2 * l <= N;  // base_pir.rs:150-154
}

struct Global {
    std::mutex mu;
    int device = -1;
    DevTables* tables = nullptr;
    uint64_t layout = 0;  // fingerprint of the device shard layout (persisted shard files)
    hipStream_t pre_stream = nullptr;
    DevBuf pre_entries, pre_M, pre_te, pre_ip, pre_aout, pre_idx, pre_buckets;
    dnspir_ctx* default_ctx = nullptr;
    // blocks of two-CU chain launches in flight (all contexts) and the CU count
    std::mutex pair_mu;
    int pair_used = 0, pair_cap = 0;
    // per GPU: the device address of `scratch` and the context whose buffers
    // are carved from it (dnspir_process_queries_noalloc)
    struct NoAlloc {
        char* base = nullptr;
        dnspir_ctx* ctx = nullptr;
        Arena arena;
    };
    static constexpr int MAX_DEVICES = 64;
    NoAlloc na[MAX_DEVICES];
};
Global g;

// ---- device tables ----------------------------------------------------------
uint64_t find_root_2n(uint64_t p) {
    for (uint64_t c = 3;; c += 2) {
        uint64_t z = dnh::powmod(c, (p - 1) / M2N, p);
        if (dnh::powmod(z, N, p) == p - 1) return z;
    }
}
uint32_t shoup(uint64_t w, uint64_t p) { return (uint32_t)(((unsigned __int128)w << 32) / p); }
int brv11(int v) {
    int r = 0;
    for (int b = 0; b < LOGN; ++b)
        if (v & (1 << b)) r |= 1 << (LOGN - 1 - b);
    return r;
}
// host replica of the device forward NTT (same butterflies)
void host_ntt_fwd(std::vector<uint64_t>& a, uint64_t p, const uint32_t* psi) {
    for (int s = 0; s < LOGN; ++s) {
        int m = 1 << s, t = N >> (s + 1);
        for (int i = 0; i < m; ++i)
            for (int j = 2 * i * t; j < 2 * i * t + t; ++j) {
                uint64_t U = a[j], V = a[j + t] * psi[m + i] % p;
                a[j] = (U + V) % p;
                a[j + t] = (U + p - V) % p;
            }
    }
}

int build_tables(DevTables* T) {
    memset(T, 0, sizeof(*T));
    for (int r = 0; r < 2; ++r) {
        const uint64_t p = mod_p(r);
        const uint64_t psi = find_root_2n(p);
        const uint64_t ipsi = dnh::powmod(psi, p - 2, p);
        for (int k = 0; k < N; ++k) {
            uint64_t w = dnh::powmod(psi, brv11(k), p), wi = dnh::powmod(ipsi, brv11(k), p);
            T->psi[r][k] = (uint32_t)w;
            T->psiq[r][k] = shoup(w, p);
            T->ipsi[r][k] = (uint32_t)wi;
            T->ipsiq[r][k] = shoup(wi, p);
        }
        uint64_t ni = dnh::powmod(N, p - 2, p);
        T->ninv[r] = (uint32_t)ni;
        T->ninvq[r] = shoup(ni, p);
        // (q/p_r)^-1 mod p_r = (the other modulus)^-1 mod p_r
        const uint64_t qi = dnh::powmod(mod_p(1 - r) % p, p - 2, p);
        T->crt_k[0][r] = (uint32_t)(ni * qi % p);
        T->crt_k[1][r] = (uint32_t)((uint64_t)T->ipsi[r][1] * T->crt_k[0][r] % p);
        for (int m = 0; m < 2; ++m) T->crt_kq[m][r] = shoup(T->crt_k[m][r], p);
        uint64_t x = 1;
        for (int k = 0; k < M2N; ++k) {
            T->psipow[r][k] = (uint32_t)x;
            T->psipowq[r][k] = shoup(x, p);
            x = x * psi % p;
        }
    }
    // slot exponents: forward NTT of X gives psi^e_s in slot s
    {
        std::vector<uint64_t> a(N, 0);
        a[1] = 1;
        host_ntt_fwd(a, P1, T->psi[0]);
        std::map<uint64_t, int> dlog;
        for (int k = 0; k < M2N; ++k) dlog[T->psipow[0][k]] = k;
        for (int k = 0; k < M2N; ++k) T->exp_slot[k] = 0xFFFF;
        for (int s = 0; s < N; ++s) {
            auto it = dlog.find(a[s]);
            if (it == dlog.end() || (it->second & 1) == 0) return DNSPIR_E_HIP;
            T->slot_exp[s] = (uint16_t)it->second;
            T->exp_slot[it->second] = (uint16_t)s;
        }
        // the same exponents mod p2 (both moduli share the slot order)
        std::vector<uint64_t> b(N, 0);
        b[1] = 1;
        host_ntt_fwd(b, P2, T->psi[1]);
        for (int s = 0; s < N; ++s)
            if (b[s] != T->psipow[1][T->slot_exp[s]]) return DNSPIR_E_HIP;
    }
    uint64_t f = 1, fi = 1;
    const uint64_t inv5 = [] {
        for (uint64_t x = 1; x < (uint64_t)M2N; x += 2)
            if ((5 * x) % M2N == 1) return x;
        return (uint64_t)0;
    }();
    for (int i = 0; i < HALFN; ++i) {
        T->pow5[i] = (uint16_t)f;
        T->inv5pow[i] = (uint16_t)fi;
        f = f * 5 % M2N;
        fi = fi * inv5 % M2N;
    }
    for (int s = 0; s < N; ++s) T->sigma5[s] = T->exp_slot[(5 * T->slot_exp[s]) % M2N];
    // the four-CU chain keeps each half of the slots on its own blocks: sigma_5
    // must map each half to itself (the halves are exponents 1 and 3 mod 4)
    for (int s = 0; s < N; ++s)
        if ((T->sigma5[s] < HALFN) != (s < HALFN) || (T->slot_exp[s] & 3) != (s < HALFN ? 1 : 3)) return DNSPIR_E_HIP;
    for (int gp = 0; gp < N; ++gp) {
        int e = gp < HALFN ? T->pow5[gp] : (M2N - T->pow5[gp - HALFN]) % M2N;
        T->gpos_slot[gp] = T->exp_slot[e];
        T->slot_gpos[T->exp_slot[e]] = (uint16_t)gp;
    }
    std::vector<uint64_t> c(N), v(N);
    for (int sd = 0; sd < 6; ++sd) {
        dnh::expand(dnh::SEEDS[sd], c.data());
        for (int r = 0; r < 2; ++r) {
            const uint64_t p = mod_p(r);
            for (int i = 0; i < N; ++i) v[i] = c[i] % p;
            host_ntt_fwd(v, p, T->psi[r]);
            uint32_t* dst = sd < 2 ? T->gka[sd][r] : T->qa[sd - 2][r];
            for (int i = 0; i < N; ++i) dst[i] = (uint32_t)v[i];
        }
    }
    return 0;
}

}  // namespace

int dne::ctx_init(dnspir_ctx* c, hipStream_t s) {
    if (s) c->stream = s;
    else {
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own = true;
    }
    // the events time phases and tell the host that the GPU got past a point;
    // no host read depends on them (results come back by copies on the
    // stream), so they skip the system-scope cache write-back a default event
    // does when it is reached (about 5 us of idle GPU per event record)
    for (auto& e : c->ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    HIP_OK(hipHostMalloc((void**)&c->h_err, sizeof(unsigned), hipHostMallocDefault));
    *c->h_err = 0;
    return 0;
}

int dne::ctx_presize(dnspir_ctx* c, int max_nq) {
    int rc;
    if ((rc = c->ptrs.ensure(PTR_SLOTS * sizeof(void*)))) return rc;
    if ((rc = c->xch.ensure((chain_pair_xbuf_words(PAIR_MAX_CHAINS) + 64) * 4))) return rc;
    const size_t cap = c->idx.n;
    if ((rc = c->idx.ensure((size_t)3 * max_nq * (sizeof(long long) + 2 * sizeof(int)) + 64))) return rc;
    if (c->idx.n != cap) c->idx_uploaded = nullptr;  // the lane lists are uploaded again
    return c->keyp.ensure((size_t)2 * max_nq * sizeof(void*));
}

// once per GPU, with that GPU current: the device address of `scratch` and
// the context of the no-allocation entry point
static int noalloc_setup(int device) {
    if (device >= Global::MAX_DEVICES) return 0;  // that entry point then reports DNSPIR_E_NOINIT
    Global::NoAlloc& na = g.na[device];
    if (na.base) return 0;
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p || ((uintptr_t)p & 255)) return DNSPIR_E_ALIGN;
    dnspir_ctx* c = new (std::nothrow) dnspir_ctx;
    if (!c) return DNSPIR_E_OOM;
    const int rc = ctx_init(c, nullptr);
    if (rc) {
        delete c;
        return rc;
    }
    // one operation of each kind the entry point issues (kernel launch,
    // copies from pageable and to pinned host memory, event records) now, on
    // its stream, so that whatever the runtime sets up on first use of a
    // stream happens here and not inside a call
    {
        uint32_t host_words[64] = {0};
        HIP_OK(hipEventRecord(c->ev[0], c->stream));
        launch_fill_u32(static_cast<uint32_t*>(p), 64, 0, c->stream);
        HIP_OK(hipMemcpyAsync(p, host_words, sizeof(host_words), hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(c->h_err, p, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipEventRecord(c->ev[1], c->stream));
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(c->stream));
        *c->h_err = 0;
    }
    na.arena.base = static_cast<char*>(p);
    na.arena.cap = DNSPIR_NOALLOC_SCRATCH_BYTES;
    c->bind_all(&na.arena);  // from now on every buffer of c lives in scratch
    na.ctx = c;
    na.base = static_cast<char*>(p);
    return 0;
}

int dne::ensure_init() {
    if (g.device < 0 || !g.tables) return DNSPIR_E_NOINIT;
    return hipSetDevice(g.device) == hipSuccess ? 0 : DNSPIR_E_HIP;
}

const DevTables* dne::tables() { return g.tables; }

namespace {

dnspir_ctx* default_ctx() { return g.default_ctx; }

// rows per K1 block: enough blocks to fill the chip, few atomics per row
int k1_rows(int L, int nsh, int nst) {
    const int per_chunk_blocks = (N / 128) * nsh * 2 * nst;
    int chunks = (2048 + per_chunk_blocks - 1) / per_chunk_blocks;
    chunks = std::max(1, std::min(chunks, L / 16 > 0 ? L / 16 : 1));
    return (L + chunks - 1) / chunks;
}

// uploads the fleet's slab and a17 pointer lists (cached per context)
int upload_fleet(dnspir_ctx* c, int slot, const dnspir_shard* const* fleet, size_t len, const uint32_t* const** slabs_dev,
                 const uint32_t* const** a17_dev) {
    int rc = c->ptrs.ensure(PTR_SLOTS * sizeof(void*));
    if (rc) return rc;
    const uint32_t** base = c->ptrs.as<const uint32_t*>() + slot * 512;
    *slabs_dev = base;
    *a17_dev = base + 256;
    // the device list is compared by the pointers it holds, not by shard
    // handles: a handle freed and allocated again may come back at the same
    // address for a shard in other device memory
    std::vector<const uint32_t*> h(512, nullptr);
    for (size_t i = 0; i < len; ++i) {
        h[i] = fleet[i]->slab;
        h[256 + i] = fleet[i]->a17;
    }
    if (c->cached_base[slot] != (const void*)base || h != c->cached[slot]) {
        c->cached[slot].clear();  // a failed copy leaves nothing trusted
        HIP_OK(hipMemcpyAsync(base, h.data(), 512 * sizeof(void*), hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->cached[slot].swap(h);
        c->cached_base[slot] = base;
    }
    return 0;
}

int check_fleet(const dnspir_shard* const* fleet, size_t len, int l_p) {
    if (!fleet || len == 0 || len > (size_t)(N / LANES)) return DNSPIR_E_SHAPE;
    for (size_t i = 0; i < len; ++i) {
        if (!fleet[i] || !fleet[i]->has_db || fleet[i]->seeds != 0) return DNSPIR_E_ARG;
        if (fleet[i]->l != l_p) return DNSPIR_E_CONJ;  // double_pir.rs:160 (same shape throughout)
    }
    return 0;
}

// Lane lists of the secondary phase: lane index li -> entry offset, number
// of nonzero columns, query slot. Uploaded as [ent_off (ll) | ncols (int) | slot (int)].
int upload_lanes(dnspir_ctx* c, const std::vector<long long>& off, const std::vector<int>& ncols,
                 const std::vector<int>& slot, const long long** d_off, const int** d_ncols, const int** d_slot) {
    const size_t n = off.size();
    const size_t cap = c->idx.n;
    int rc = c->idx.ensure(n * (sizeof(long long) + 2 * sizeof(int)) + 64);
    if (rc) return rc;
    if (c->idx.n != cap) c->idx_uploaded = nullptr;  // reallocated (possibly at the same address)
    std::vector<long long> h(2 * n, 0);
    memcpy(h.data(), off.data(), n * sizeof(long long));
    int* ip = reinterpret_cast<int*>(h.data() + n);
    memcpy(ip, ncols.data(), n * sizeof(int));
    memcpy(ip + n, slot.data(), n * sizeof(int));
    // the same lists already in this buffer (repeated calls on one fleet): no copy
    if (c->idx.p != c->idx_uploaded || h != c->h_idx) {
        c->h_idx.assign(h.begin(), h.end());
        HIP_OK(hipMemcpyAsync(c->idx.p, c->h_idx.data(), 2 * n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
        c->idx_uploaded = c->idx.p;
    }
    *d_off = c->idx.as<long long>();
    *d_ncols = reinterpret_cast<const int*>(c->idx.as<long long>() + n);
    *d_slot = *d_ncols + n;
    return 0;
}

// Primary phase of one fleet for nq queries (qp [q][4][2][N]); leaves the
// per-(query, shard) results for the epilogue in c->outb (limb-block shards)
// or c->acc (rotated shards, nq = 1).
int run_primary(dnspir_ctx* c, const uint32_t* const* slabs, int len, int l_p, const uint32_t* qp, int nq, bool* used_outb,
                hipEvent_t k1_start, uint64_t* bytes = nullptr) {
    const int L_p = loop_len(l_p), nst_p = nstreams(l_p);
    const hipStream_t s = c->stream;
    int rc;
    if (bytes) *bytes = (uint64_t)len * slab_bytes(l_p);
    if (L_p == N / 2) {
        if ((rc = c->outb.ensure((size_t)nq * len * 2 * N * LANES * 4))) return rc;
        if (use_k1w(c, l_p, (size_t)len, nq)) {
            if ((rc = c->wscan.ensure(nq * k1w_scan_bytes())) || (rc = c->wtile.ensure(nq * k1w_tile_bytes(nst_p))))
                return rc;
            launch_k1w(slabs, len, qp, nq, c->wscan.as<uint32_t>(), c->wtile.as<uint32_t>(), c->outb.as<uint32_t>(), nst_p,
                       g.tables, s, k1_start);
            if (bytes) *bytes = *bytes / 3 * 2;  // the TE and IPGK1 streams
        } else {
            if ((rc = c->ql.ensure(qlimb_bytes(nq, nst_p)))) return rc;
            launch_k1b(slabs, len, qp, nq, c->ql.as<uint32_t>(), c->outb.as<uint32_t>(), nst_p, s, k1_start);
        }
        *used_outb = true;
    } else {
        if (nq != 1) return DNSPIR_E_SHAPE;
        if ((rc = c->acc.ensure((size_t)len * 2 * N * LANES * 8))) return rc;
        unsigned long long* acc = c->acc.as<unsigned long long>();
        launch_fill_u64(acc, (size_t)len * 2 * N * LANES, 0, s);
        if (k1_start) HIP_OK(hipEventRecord(k1_start, s));
        launch_k1(slabs, len, qp, acc, L_p, nst_p, k1_rows(L_p, len, nst_p), s);
        *used_outb = false;
    }
    return 0;
}

}  // namespace

// queues the copy of the two-CU chain's timeout flag to pinned host memory
// (outside any graph capture: a copy to host memory cannot be captured)
// the multi-CU chain's timeout flag reaches c->h_err through chain_out_kernel;
// check_pair reads it after the call's sync
static int copy_pair_flag(dnspir_ctx* c) {
    c->pair_pending = true;
    return 0;
}

// ---- the device's budget of multi-CU chain blocks -----------------------------
// Each block of chain_pair_kernel / chain_quad_kernel spins on its partner
// blocks, so a wait can end only once the partners are resident. One such
// block takes more than 80 KB of LDS, so at most one fits a CU (160 KB). If
// the multi-CU chain blocks in flight on the device never outnumber its CUs,
// every waiting block's partners find a CU as the blocks of other kernels
// (which never wait) retire. g.pair_used counts the blocks reserved by every
// context of the process; a call that would exceed the CU count runs its
// chains on fewer CUs each (four -> two -> one).
namespace {

void pair_release(dnspir_ctx* c) {
    if (!c->pair_held) return;
    std::lock_guard<std::mutex> lk(g.pair_mu);
    g.pair_used -= c->pair_held;
    c->pair_held = 0;
}

// a reservation left by a call that returned early (its kernels may still run)
void pair_settle(dnspir_ctx* c) {
    if (!c->pair_held) return;
    (void)hipStreamSynchronize(c->stream);
    pair_release(c);
}

// CUs per chain for nchains chains on this context: 4 up to QUAD_MAX_CHAINS
// chains, 2 up to PAIR_MAX_CHAINS, else 1, capped by the context's setting;
// reserves the blocks of a multi-CU launch (the result falls to fewer CUs
// when the device's budget has no room)
int chain_acquire(dnspir_ctx* c, int nchains) {
    pair_settle(c);
    int want = nchains <= QUAD_MAX_CHAINS ? 4 : nchains <= PAIR_MAX_CHAINS ? 2 : 1;
    if (c->chain_cus) want = std::min(want, c->chain_cus);
    if (want == 1) return 1;
    std::lock_guard<std::mutex> lk(g.pair_mu);
    for (int cus = want; cus > 1; cus /= 2) {
        const int b = chain_multi_blocks(nchains, cus == 4);
        if (g.pair_used + b <= g.pair_cap) {
            g.pair_used += b;
            c->pair_held = b;
            ++c->stats.pair_runs;
            if (cus == 4) ++c->stats.quad_runs;
            return cus;
        }
    }
    ++c->stats.pair_refused;
    return 1;
}

// records a reply encoding that follows the secondary phase (repeated by a
// retry of the phase)
void note_encode(dnspir_ctx* c, int lane0, int nq, uint8_t* reply) {
    dnspir_ctx::SecondaryCall& sc = c->last_sec;
    if (sc.n_enc >= 2) return;
    sc.enc_lane0[sc.n_enc] = lane0;
    sc.enc_nq[sc.n_enc] = nq;
    sc.enc_reply[sc.n_enc] = reply;
    ++sc.n_enc;
}

// the arguments of a secondary phase, for check_pair's retry
void note_secondary(dnspir_ctx* c, const uint16_t* entries, long long col_stride, const long long* d_off,
                    const int* d_ncols, const int* d_slot, int nlanes, int l_s) {
    dnspir_ctx::SecondaryCall& sc = c->last_sec;
    sc.entries = entries;
    sc.col_stride = col_stride;
    sc.d_off = d_off;
    sc.d_ncols = d_ncols;
    sc.d_slot = d_slot;
    sc.nlanes = nlanes;
    sc.l_s = l_s;
    sc.n_enc = 0;
}

// Secondary phase for `nlanes` lanes whose entries are described by the lane
// lists; the reply (a, b per lane) lands in c->acoef/bcoef/a17o/b12o.
// cus: CUs per chain. 2 or 4: chain_pair_kernel / chain_quad_kernel, lower
// latency for a few chains; the caller holds their chain_acquire reservation,
// queues copy_pair_flag and calls check_pair after its sync. The arguments
// are kept in c->last_sec so that check_pair can run the phase again on one
// CU per chain.
int run_secondary(dnspir_ctx* c, const uint16_t* entries, long long col_stride, const long long* d_off, const int* d_ncols,
                  const int* d_slot, int nlanes, int l_s, int cus = 1) {
    const int L_s = loop_len(l_s), nst_s = nstreams(l_s);
    const hipStream_t s = c->stream;
    int rc;
    if ((rc = c->M.ensure((size_t)nlanes * l_s * N * 4))) return rc;
    if ((rc = c->te.ensure((size_t)nst_s * nlanes * L_s * 2 * N * 4))) return rc;
    if ((rc = c->aout.ensure((size_t)nst_s * nlanes * 2 * N * 4))) return rc;
    if ((rc = c->cout.ensure((size_t)nst_s * nlanes * 2 * N * 4))) return rc;
    if ((rc = c->acoef.ensure((size_t)nlanes * N * 8))) return rc;
    if ((rc = c->bcoef.ensure((size_t)nlanes * N * 8))) return rc;
    if ((rc = c->a17o.ensure((size_t)nlanes * N * 4))) return rc;
    if ((rc = c->b12o.ensure((size_t)nlanes * N * 4))) return rc;
    note_secondary(c, entries, col_stride, d_off, d_ncols, d_slot, nlanes, l_s);
    // xch: the timeout flag in word 0, the exchange words from word 64, zeroed by K5
    const size_t xwords = cus > 1 ? chain_pair_xbuf_words(nst_s * nlanes) + 64 : 0;
    if (cus > 1 && (rc = c->xch.ensure(xwords * 4))) return rc;
    launch_k5(entries, d_off, col_stride, c->M.as<uint32_t>(), c->te.as<uint32_t>(), l_s, d_ncols, nlanes, g.tables, s,
              cus > 1 ? c->xch.as<uint32_t>() : nullptr, xwords);
    if (cus > 1) {
        launch_chain_pair(c->te.as<uint32_t>(), c->aout.as<uint32_t>(), c->cout.as<uint32_t>(), c->qs.as<uint32_t>(), d_slot,
                          L_s, nst_s, nlanes, c->xch.as<uint32_t>(), c->exchange_polls, g.tables, s, cus == 4, true);
    } else {
        launch_chain_sec(c->te.as<uint32_t>(), c->aout.as<uint32_t>(), c->cout.as<uint32_t>(), c->qs.as<uint32_t>(), d_slot,
                         L_s, nst_s, nlanes, g.tables, s);
    }
    launch_chain_out(c->aout.as<uint32_t>(), c->cout.as<uint32_t>(), nst_s, nlanes, c->acoef.as<uint64_t>(),
                     c->bcoef.as<uint64_t>(), c->a17o.as<uint32_t>(), c->b12o.as<uint32_t>(), g.tables, s,
                     cus > 1 ? c->xch.as<unsigned>() : nullptr, cus > 1 ? c->h_err : nullptr);
    return 0;
}

}  // namespace

// After the sync that follows a call: if that call's two-CU chain timed out
// waiting for a partner block, its outputs are invalid; the same secondary
// phase (and the reply encodings noted for it) runs again with one CU per
// chain, which never waits, and the stream is synchronised again.
int dne::check_pair(dnspir_ctx* c, bool* retried) {
    if (retried) *retried = false;
    const bool pending = c->pair_pending;
    c->pair_pending = false;
    const bool timed_out = pending && *c->h_err != 0;
    pair_release(c);
    if (!timed_out) return 0;
    ++c->stats.pair_retries;
    const dnspir_ctx::SecondaryCall sc = c->last_sec;
    int rc = run_secondary(c, sc.entries, sc.col_stride, sc.d_off, sc.d_ncols, sc.d_slot, sc.nlanes, sc.l_s, 1);
    if (rc) return rc;
    for (int i = 0; i < sc.n_enc; ++i)
        launch_encode_reply(c->a17o.as<uint32_t>(), c->b12o.as<uint32_t>(), sc.enc_lane0[i], sc.enc_nq[i], sc.enc_reply[i],
                            c->stream);
    c->last_sec = sc;
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    if (retried) *retried = true;
    return 0;
}

namespace {

// perform_batched_double_pir on ring-element queries already in c->qcoef
// ([2][6][N]); leaves the reply in c->acoef/bcoef/a17o/b12o with lane index
// li = (#present slots before s)*3 + part.
int run_double(dnspir_ctx* c, const dnspir_shard* const* fleets[2], const size_t lens[2], const bool present[2],
               dnspir_timing* tm) {
    const hipStream_t s = c->stream;
    const int l_p = fleets[0][0]->l;
    const int l_s = dnspir_choose_index_group_secondary(lens[0]);
    int rc;
    if ((rc = c->qs.ensure((size_t)2 * 6 * 2 * N * 4))) return rc;
    if ((rc = c->qp.ensure((size_t)2 * 4 * 2 * N * 4))) return rc;
    if ((rc = c->se.ensure((size_t)l_s * LANES * N * 2))) return rc;
    std::vector<long long> off;
    std::vector<int> ncols, slot;
    for (int sl = 0; sl < 2; ++sl)
        if (present[sl])
            for (int p = 0; p < 3; ++p) {
                off.push_back((long long)(3 * sl + p) * N);
                ncols.push_back((int)(LANES * lens[sl]));
                slot.push_back(sl);
            }
    const long long* d_off;
    const int *d_ncols, *d_slot;
    if ((rc = upload_lanes(c, off, ncols, slot, &d_off, &d_ncols, &d_slot))) return rc;
    const uint32_t* const* slabs[2] = {nullptr, nullptr};
    const uint32_t* const* a17s[2] = {nullptr, nullptr};
    for (int sl = 0; sl < 2; ++sl)
        if (present[sl] && (rc = upload_fleet(c, sl, fleets[sl], lens[sl], &slabs[sl], &a17s[sl]))) return rc;

    HIP_OK(hipEventRecord(c->ev[0], s));
    launch_qprep(c->qcoef.as<uint64_t>(), c->qs.as<uint32_t>(), c->qp.as<uint32_t>(), 2, g.tables, s);
    uint64_t streamed = 0, ref_streamed = 0;
    c->k1_launches = 0;
    c->k1_timed = true;
    // both slots on the same fleet: one two-query pass over the shards (the
    // two-stream inner product takes two queries; qp holds them back to back)
    const bool joint = present[0] && present[1] && lens[0] == lens[1] &&
                       std::equal(fleets[0], fleets[0] + lens[0], fleets[1]) && use_k1w(c, l_p, lens[0], 2);
    for (int sl = 0; sl < 2; ++sl) {
        if (!present[sl]) continue;
        const int len = (int)lens[sl];
        bool ob = true;
        uint64_t bytes = 0;
        if (!joint || sl == 0) {
            if ((rc = run_primary(c, slabs[sl], len, l_p, c->qp.as<uint32_t>() + (size_t)sl * 4 * 2 * N, joint ? 2 : 1, &ob,
                                  c->k1_launches == 0 ? c->ev[3] : nullptr, &bytes)))
                return rc;
            if (c->k1_launches == 0) HIP_OK(hipEventRecord(c->ev[4], s));
            ++c->k1_launches;
        }
        EpiLaunch e;
        if (ob) e.outb = c->outb.as<uint32_t>() + (joint ? (size_t)sl * len * 2 * N * LANES : 0);
        else e.acc = c->acc.as<unsigned long long>();
        e.nshards = len;
        e.nq = 1;
        e.a17s = a17s[sl];
        e.se = c->se.as<uint16_t>() + (size_t)3 * sl * N;  // entry lanes 3*slot + part
        e.lpr = LANES;
        e.lane0_mult = 0;
        launch_epi(e, g.tables, s);
        streamed += bytes;
        ref_streamed += (uint64_t)len * ref_slab_bytes(l_p);
    }
    HIP_OK(hipEventRecord(c->ev[1], s));
    const int cus = chain_acquire(c, nstreams(l_s) * (int)off.size());
    if ((rc = run_secondary(c, c->se.as<uint16_t>(), (long long)LANES * N, d_off, d_ncols, d_slot, (int)off.size(), l_s,
                            cus)))
        return rc;
    if (cus > 1 && (rc = copy_pair_flag(c))) return rc;
    HIP_OK(hipEventRecord(c->ev[2], s));
    if (tm) {
        tm->bytes_streamed = streamed;
        tm->ref_bytes = ref_streamed;
    }
    return 0;
}

void fill_timing(dnspir_ctx* c, dnspir_timing* tm, double wall_us) {
    if (!tm) return;
    float a = 0, b = 0, k = 0;
    bool err = hipEventElapsedTime(&a, c->ev[0], c->ev[1]) != hipSuccess;
    err |= hipEventElapsedTime(&b, c->ev[1], c->ev[2]) != hipSuccess;
    if (c->k1_timed) err |= hipEventElapsedTime(&k, c->ev[3], c->ev[4]) != hipSuccess;
    if (err) (void)hipGetLastError();  // a timing gap must not fail the next call
    tm->primary_us = a * 1000.0;
    tm->secondary_us = b * 1000.0;
    tm->total_us = wall_us;
    // the first inner-product launch of the call
    tm->k1_us = k * 1000.0;
    tm->k1_launches = c->k1_launches;
}

}  // namespace

double dne::now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t dne::query_layout(int l_p, int l_s, int offs[6], int bits[6]) {
    const int b[6] = {GK_BITS, GK_BITS, QRY_BITS, QRY_BITS, QRY_BITS, QRY_BITS};
    const int pres[6] = {1, 1, 1, l_p == N, 1, l_s == N};
    size_t off = 0;
    for (int v = 0; v < 6; ++v) {
        bits[v] = b[v];
        offs[v] = pres[v] ? (int)off : -1;
        if (pres[v]) off += dnh::condensed_size(b[v], N);
    }
    return off;
}

extern "C" {

const char* dnspir_last_error(void) { return g_last_error; }

const char* dnspir_strerror(int code) {
    switch (code) {
        case DNSPIR_OK: return "ok";
        case DNSPIR_E_SHORT_QUERY: return "not enough data";
        case DNSPIR_E_CONJ: return "conjugate query presence does not match the index group";
        case DNSPIR_E_COUNT: return "entry count does not match the index group";
        case DNSPIR_E_SHAPE: return "invalid index group or fleet shape";
        case DNSPIR_E_ARG: return "invalid argument";
        case DNSPIR_E_ALIGN: return "device memory misaligned or too small";
        case DNSPIR_E_OOM: return "device allocation failed";
        case DNSPIR_E_HIP: return "HIP runtime error";
        case DNSPIR_E_DECODE: return "reply digits out of range";
        case DNSPIR_E_NOINIT: return "dnspir_init has not been called";
        case DNSPIR_E_NOKEY: return "keyless query, but the client's Galois keys are not cached";
        case DNSPIR_E_CLOSED: return "the server is shutting down";
        case DNSPIR_E_IO: return "a shard file could not be opened, written or closed";
        case DNSPIR_E_FORMAT: return "not a shard file of this build's layout, or corrupt or truncated";
        case DNSPIR_E_SCRATCH: return "the call needs more device scratch than the no-allocation entry point holds";
        case DNSPIR_E_FRAME: return "frame too large, empty request or unknown command";
        default: return "unknown error";
    }
}

int dnspir_init(int device) {
    std::lock_guard<std::mutex> lk(g.mu);
    if (g.tables && g.device == device) return 0;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return DNSPIR_E_NOINIT;
    HIP_OK(hipSetDevice(device));
    DevTables* h = new (std::nothrow) DevTables;
    if (!h) return DNSPIR_E_OOM;
    int rc = build_tables(h);
    if (rc) {
        delete h;
        return rc;
    }
    // what a persisted shard's bytes depend on besides its entries: the slot
    // order of the NTT (Galois positions -> slots), the roots, the residue
    // packing of the slab
    uint64_t fp = 0xcbf29ce484222325ull ^ 0x2625u;
    for (int i = 0; i < N; ++i) fp = (fp ^ h->gpos_slot[i]) * 0x100000001b3ull;
    for (int r = 0; r < 2; ++r) fp = (fp ^ h->psi[r][1]) * 0x100000001b3ull;
    fp = (fp ^ (MB_BLOCK_BYTES_P1 * 65536ull + MB_BLOCK_BYTES_P2)) * 0x100000001b3ull;
    DevTables* d = nullptr;
    if (hipMalloc(&d, sizeof(DevTables)) != hipSuccess) {
        delete h;
        return DNSPIR_E_OOM;
    }
    if (hipMemcpy(d, h, sizeof(DevTables), hipMemcpyHostToDevice) != hipSuccess) {
        delete h;
        return DNSPIR_E_HIP;
    }
    delete h;
    g.tables = d;
    g.layout = fp;
    g.device = device;
    {
        hipDeviceProp_t prop;
        HIP_OK(hipGetDeviceProperties(&prop, device));
        std::lock_guard<std::mutex> pk(g.pair_mu);
        g.pair_cap = prop.multiProcessorCount;
        g.pair_used = 0;
    }
    HIP_OK(hipStreamCreateWithFlags(&g.pre_stream, hipStreamNonBlocking));
    g.default_ctx = new dnspir_ctx;
    if ((rc = ctx_init(g.default_ctx, nullptr))) return rc;
    return noalloc_setup(device);
}

size_t dnspir_required_bytes(int group_order) {
    if (!valid_group(group_order)) return 0;
    return slab_bytes(group_order) + EXTRA_BYTES;
}

int dnspir_shard_create(int group_order, int used_seeds, void* dev_mem, size_t bytes, dnspir_shard** out) {
    if (!out || !dev_mem) return DNSPIR_E_ARG;
    if (!valid_group(group_order)) return DNSPIR_E_SHAPE;
    if (used_seeds != 0 && used_seeds != 1) return DNSPIR_E_ARG;
    if (((uintptr_t)dev_mem & 255) != 0 || bytes < dnspir_required_bytes(group_order)) return DNSPIR_E_ALIGN;
    dnspir_shard* s = new (std::nothrow) dnspir_shard;
    if (!s) return DNSPIR_E_OOM;
    s->l = group_order;
    s->L = loop_len(group_order);
    s->nst = nstreams(group_order);
    s->seeds = used_seeds;
    s->slab = (uint32_t*)dev_mem;
    s->a_coef = (uint64_t*)((char*)dev_mem + slab_bytes(group_order));
    s->a17 = (uint32_t*)((char*)s->a_coef + (size_t)LANES * N * 8);
    s->has_db = false;
    *out = s;
    return 0;
}

int dnspir_shard_destroy(dnspir_shard* s) {
    delete s;
    return 0;
}

int dnspir_shard_group_order(const dnspir_shard* s) { return s ? s->l : 0; }

// ---- persisted shards: a preprocessed shard's device state in a file --------
// 64-byte header, then the dnspir_required_bytes(l) bytes of the shard's
// device memory (slab, a-part, its 17-bit form) as they lie in HBM.
namespace {
constexpr char SHARD_MAGIC[8] = {'D', 'N', 'S', 'P', 'I', 'R', 'S', 'H'};
constexpr uint32_t SHARD_FILE_VERSION = 1;
struct ShardFileHeader {
    char magic[8];
    uint32_t version, n;
    int32_t group_order, used_seeds;
    uint64_t layout;         // fingerprint of the device layout that wrote it
    uint64_t payload_bytes;  // dnspir_required_bytes(group_order)
    uint64_t checksum;       // of the payload, over 64-bit words
    uint64_t reserved[2];
};
static_assert(sizeof(ShardFileHeader) == 64, "shard file header");
constexpr size_t SHARD_FILE_CHUNK = (size_t)32 << 20;

struct PayloadHash {
    uint64_t h = 0x84222325cbf29ce4ull;
    void add(const void* p, size_t n) {  // n is a multiple of 8
        const uint64_t* w = static_cast<const uint64_t*>(p);
        for (size_t i = 0; i < n / 8; ++i) {
            h = (h ^ w[i]) * 0x100000001b3ull;
            h ^= h >> 29;
        }
    }
};

int read_header(FILE* f, ShardFileHeader* hd) {
    if (fread(hd, sizeof(*hd), 1, f) != 1) return DNSPIR_E_FORMAT;
    if (memcmp(hd->magic, SHARD_MAGIC, 8) != 0 || hd->version != SHARD_FILE_VERSION || hd->n != (uint32_t)N ||
        !valid_group(hd->group_order) || (hd->used_seeds != 0 && hd->used_seeds != 1) ||
        hd->payload_bytes != dnspir_required_bytes(hd->group_order))
        return DNSPIR_E_FORMAT;
    return 0;
}

struct PinnedChunk {
    void* p = nullptr;
    ~PinnedChunk() {
        if (p) (void)hipHostFree(p);
    }
};
}  // namespace

int dnspir_shard_save(const dnspir_shard* sh, const char* path) {
    if (!sh || !path) return DNSPIR_E_ARG;
    int rc = ensure_init();
    if (rc) return rc;
    if (!sh->has_db) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(g.mu);
    HIP_OK(hipStreamSynchronize(g.pre_stream));
    const size_t total = dnspir_required_bytes(sh->l);
    PinnedChunk buf;
    HIP_OK(hipHostMalloc(&buf.p, std::min(total, SHARD_FILE_CHUNK), hipHostMallocDefault));
    FILE* f = fopen(path, "wb");
    if (!f) return DNSPIR_E_IO;
    ShardFileHeader hd;
    memset(&hd, 0, sizeof(hd));
    memcpy(hd.magic, SHARD_MAGIC, 8);
    hd.version = SHARD_FILE_VERSION;
    hd.n = N;
    hd.group_order = sh->l;
    hd.used_seeds = sh->seeds;
    hd.layout = g.layout;
    hd.payload_bytes = total;
    bool ok = fwrite(&hd, sizeof(hd), 1, f) == 1;
    PayloadHash hash;
    for (size_t off = 0; ok && off < total; off += SHARD_FILE_CHUNK) {
        const size_t n = std::min(SHARD_FILE_CHUNK, total - off);
        if (hipMemcpy(buf.p, reinterpret_cast<const char*>(sh->slab) + off, n, hipMemcpyDeviceToHost) != hipSuccess) {
            fclose(f);
            HIP_OK(hipGetLastError());
            return DNSPIR_E_HIP;
        }
        hash.add(buf.p, n);
        ok = fwrite(buf.p, 1, n, f) == n;
    }
    hd.checksum = hash.h;
    ok = ok && fseek(f, 0, SEEK_SET) == 0 && fwrite(&hd, sizeof(hd), 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : DNSPIR_E_IO;
}

int dnspir_shard_file_info(const char* path, int* group_order, int* used_seeds, size_t* bytes) {
    if (!path) return DNSPIR_E_ARG;
    FILE* f = fopen(path, "rb");
    if (!f) return DNSPIR_E_IO;
    ShardFileHeader hd;
    const int rc = read_header(f, &hd);
    fclose(f);
    if (rc) return rc;
    if (group_order) *group_order = hd.group_order;
    if (used_seeds) *used_seeds = hd.used_seeds;
    if (bytes) *bytes = (size_t)hd.payload_bytes;
    return 0;
}

int dnspir_shard_load(const char* path, void* dev_mem, size_t bytes, dnspir_shard** out) {
    if (!path || !dev_mem || !out) return DNSPIR_E_ARG;
    int rc = ensure_init();
    if (rc) return rc;
    FILE* f = fopen(path, "rb");
    if (!f) return DNSPIR_E_IO;
    ShardFileHeader hd;
    if ((rc = read_header(f, &hd)) || hd.layout != g.layout) {
        fclose(f);
        return rc ? rc : DNSPIR_E_FORMAT;
    }
    dnspir_shard* s = nullptr;
    if ((rc = dnspir_shard_create(hd.group_order, hd.used_seeds, dev_mem, bytes, &s))) {
        fclose(f);
        return rc;
    }
    const size_t total = hd.payload_bytes;
    PinnedChunk buf;
    if (hipHostMalloc(&buf.p, std::min(total, SHARD_FILE_CHUNK), hipHostMallocDefault) != hipSuccess) {
        fclose(f);
        delete s;
        HIP_OK(hipGetLastError());
        return DNSPIR_E_OOM;
    }
    PayloadHash hash;
    rc = 0;
    for (size_t off = 0; !rc && off < total; off += SHARD_FILE_CHUNK) {
        const size_t n = std::min(SHARD_FILE_CHUNK, total - off);
        if (fread(buf.p, 1, n, f) != n) rc = DNSPIR_E_FORMAT;  // short file
        else {
            hash.add(buf.p, n);
            if (hipMemcpy(static_cast<char*>(dev_mem) + off, buf.p, n, hipMemcpyHostToDevice) != hipSuccess) rc = DNSPIR_E_HIP;
        }
    }
    if (!rc && (hash.h != hd.checksum || fgetc(f) != EOF)) rc = DNSPIR_E_FORMAT;
    fclose(f);
    if (rc) {
        delete s;
        if (rc == DNSPIR_E_HIP) HIP_OK(hipGetLastError());
        return rc;
    }
    s->has_db = true;
    *out = s;
    return 0;
}

// where set_db's entries come from
enum EntrySource { HOST_COEFFS, DEVICE_COEFFS, HOST_BUCKETS };

static int set_db_impl(dnspir_shard* sh, const void* src, size_t n, EntrySource kind) {
    if (!sh || !src) return DNSPIR_E_ARG;
    int rc = ensure_init();
    if (rc) return rc;
    const size_t n_entries = (size_t)LANES * sh->l;
    // base_pir.rs:693 asks for exactly 8*l entries; buckets may fill a
    // shard only partly (the last shard of a database, server.rs:664-739)
    if (kind == HOST_BUCKETS ? n > n_entries : n != n_entries) return DNSPIR_E_COUNT;
    std::lock_guard<std::mutex> lk(g.mu);
    const int l = sh->l, L = sh->L, nst = sh->nst;
    const hipStream_t s = g.pre_stream;
    const bool on_device = kind == DEVICE_COEFFS;
    if (!on_device && (rc = g.pre_entries.ensure(n_entries * N * 2))) return rc;
    if (kind == HOST_BUCKETS && (rc = g.pre_buckets.ensure(std::max<size_t>(n, 1) * BUCKET_BYTES))) return rc;
    if ((rc = g.pre_M.ensure((size_t)LANES * l * N * 4))) return rc;
    if ((rc = g.pre_te.ensure((size_t)nst * LANES * L * 2 * N * 4))) return rc;
    if ((rc = g.pre_ip.ensure((size_t)nst * LANES * L * 6 * N * 4))) return rc;
    if ((rc = g.pre_aout.ensure((size_t)nst * LANES * 2 * N * 4))) return rc;
    if ((rc = g.pre_idx.ensure(8 * sizeof(long long) + 8 * sizeof(int)))) return rc;
    long long h_idx[12];
    int* h_nc = reinterpret_cast<int*>(h_idx + 8);
    for (int i = 0; i < 8; ++i) {
        h_idx[i] = (long long)i * N;
        h_nc[i] = l;
    }
    HIP_OK(hipMemcpyAsync(g.pre_idx.p, h_idx, sizeof(h_idx), hipMemcpyHostToDevice, s));
    const uint16_t* ent = static_cast<const uint16_t*>(src);
    if (kind == HOST_COEFFS) {
        HIP_OK(hipMemcpyAsync(g.pre_entries.p, src, n_entries * N * 2, hipMemcpyHostToDevice, s));
        ent = g.pre_entries.as<uint16_t>();
    } else if (kind == HOST_BUCKETS) {
        if (n) HIP_OK(hipMemcpyAsync(g.pre_buckets.p, src, n * BUCKET_BYTES, hipMemcpyHostToDevice, s));
        launch_pack_buckets(g.pre_buckets.as<uint8_t>(), n, n_entries, g.pre_entries.as<uint16_t>(), s);
        ent = g.pre_entries.as<uint16_t>();
    }
    launch_k5(ent, g.pre_idx.as<long long>(), (long long)LANES * N, g.pre_M.as<uint32_t>(), g.pre_te.as<uint32_t>(), l,
              reinterpret_cast<const int*>(g.pre_idx.as<long long>() + 8), LANES, g.tables, s);
    launch_chain_pre(g.pre_te.as<uint32_t>(), g.pre_ip.as<uint32_t>(), g.pre_aout.as<uint32_t>(), L, nst,
                     sh->seeds == 0 ? 0 : 2, g.tables, s);
    launch_relayout(g.pre_ip.as<uint32_t>(), sh->slab, L, nst, s);
    launch_chain_out(g.pre_aout.as<uint32_t>(), nullptr, nst, LANES, sh->a_coef, nullptr, sh->a17, nullptr, g.tables, s);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s));
    sh->has_db = true;
    return 0;
}

int dnspir_shard_set_db(dnspir_shard* sh, const uint16_t* entries, size_t n_entries) {
    return set_db_impl(sh, entries, n_entries, HOST_COEFFS);
}

int dnspir_shard_set_db_device(dnspir_shard* sh, const uint16_t* entries_dev, size_t n_entries) {
    return set_db_impl(sh, entries_dev, n_entries, DEVICE_COEFFS);
}

int dnspir_shard_set_db_buckets(dnspir_shard* sh, const uint8_t* buckets, size_t n_buckets) {
    static const uint8_t none = 0;
    return set_db_impl(sh, n_buckets ? (const void*)buckets : (const void*)&none, n_buckets, HOST_BUCKETS);
}

int dnspir_pack_bytes_into_coeffs(const uint8_t* buckets, size_t n, uint16_t* coeffs) {
    if ((!buckets || !coeffs) && n) return DNSPIR_E_ARG;
    for (size_t e = 0; e < n; ++e)
        for (int gi = 0; gi < N / 4; ++gi) {
            const uint8_t* p = buckets + e * BUCKET_BYTES + gi * 5;
            uint64_t v = 0;
            for (int b = 0; b < 5; ++b) v |= (uint64_t)p[b] << (8 * b);
            for (int k = 0; k < 4; ++k) coeffs[e * N + gi * 4 + k] = (uint16_t)((v >> (10 * k)) & 1023);
        }
    return 0;
}

int dnspir_shard_get_a(const dnspir_shard* sh, uint64_t* out) {
    if (!sh || !out) return DNSPIR_E_ARG;
    int rc = ensure_init();
    if (rc) return rc;
    HIP_OK(hipMemcpy(out, sh->a_coef, (size_t)LANES * N * 8, hipMemcpyDeviceToHost));
    return 0;
}

static int perform_common(const dnspir_shard* sh, const uint64_t* qb, const uint64_t* qbc, const uint64_t* gkb,
                          uint64_t* out, bool split) {
    if (!sh || !qb || !gkb || !out) return DNSPIR_E_ARG;
    int rc = ensure_init();
    if (rc) return rc;
    if (!sh->has_db) return DNSPIR_E_ARG;
    if (sh->nst == 2 && !qbc) return DNSPIR_E_CONJ;  // base_pir.rs:965
    dnspir_ctx* c = default_ctx();
    std::lock_guard<std::mutex> lk(c->mu);
    const hipStream_t s = c->stream;
    const int nq = split ? LANES : 1;
    if ((rc = c->qcoef.ensure((size_t)LANES * 6 * N * 8))) return rc;
    if ((rc = c->qs.ensure((size_t)LANES * 6 * 2 * N * 4))) return rc;
    if ((rc = c->qp.ensure((size_t)LANES * 4 * 2 * N * 4))) return rc;
    if ((rc = c->acc.ensure((size_t)2 * N * LANES * 8))) return rc;
    if ((rc = c->bcoef.ensure((size_t)LANES * N * 8))) return rc;
    if ((rc = c->ptrs.ensure(PTR_SLOTS * sizeof(void*)))) return rc;
    c->stage.assign((size_t)nq * 6 * N, 0);
    for (int q = 0; q < nq; ++q) {
        uint64_t* blk = c->stage.data() + (size_t)q * 6 * N;
        memcpy(blk, gkb + (size_t)q * 2 * N, 2 * N * 8);
        memcpy(blk + 2 * N, qb + (size_t)q * N, N * 8);
        if (sh->nst == 2) memcpy(blk + 3 * N, qbc + (size_t)q * N, N * 8);
    }
    HIP_OK(hipMemcpyAsync(c->qcoef.p, c->stage.data(), c->stage.size() * 8, hipMemcpyHostToDevice, s));
    const uint32_t* slab = sh->slab;
    const uint32_t** pslab = c->ptrs.as<const uint32_t*>() + (PTR_SLOTS - 1);  // scratch slot outside the fleet cache
    HIP_OK(hipMemcpyAsync((void*)pslab, &slab, sizeof(slab), hipMemcpyHostToDevice, s));
    launch_qprep(c->qcoef.as<uint64_t>(), c->qs.as<uint32_t>(), c->qp.as<uint32_t>(), nq, g.tables, s);
    EpiLaunch e;
    e.nshards = 1;
    e.bcoef = c->bcoef.as<uint64_t>();
    if (sh->L == N / 2) {
        if ((rc = c->ql.ensure(qlimb_bytes(nq, sh->nst)))) return rc;
        if ((rc = c->outb.ensure((size_t)nq * 2 * N * LANES * 4))) return rc;
        launch_k1b(pslab, 1, c->qp.as<uint32_t>(), nq, c->ql.as<uint32_t>(), c->outb.as<uint32_t>(), sh->nst, s);
        e.outb = c->outb.as<uint32_t>();
        e.nq = nq;
        e.diag = split ? 1 : 0;
    } else {
        unsigned long long* acc = c->acc.as<unsigned long long>();
        launch_fill_u64(acc, (size_t)2 * N * LANES, 0, s);
        if (split) launch_k1_split(pslab, 1, c->qp.as<uint32_t>(), acc, sh->L, sh->nst, k1_rows(sh->L, 1, sh->nst), s);
        else launch_k1(pslab, 1, c->qp.as<uint32_t>(), acc, sh->L, sh->nst, k1_rows(sh->L, 1, sh->nst), s);
        e.acc = acc;
    }
    launch_epi(e, g.tables, s);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, c->bcoef.p, (size_t)LANES * N * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return 0;
}

int dnspir_shard_perform_pir(const dnspir_shard* sh, const uint64_t* qb, const uint64_t* qbc, const uint64_t* gkb,
                             uint64_t* out) {
    return perform_common(sh, qb, qbc, gkb, out, false);
}

int dnspir_shard_perform_pir_split(const dnspir_shard* sh, const uint64_t* qb, const uint64_t* qbc, const uint64_t* gkb,
                                   uint64_t* out) {
    return perform_common(sh, qb, qbc, gkb, out, true);
}

int dnspir_ctx_create(void* stream, dnspir_ctx** out) {
    if (!out) return DNSPIR_E_ARG;
    int rc = ensure_init();
    if (rc) return rc;
    dnspir_ctx* c = new (std::nothrow) dnspir_ctx;
    if (!c) return DNSPIR_E_OOM;
    rc = ctx_init(c, (hipStream_t)stream);
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

int dnspir_ctx_destroy(dnspir_ctx* c) {
    if (!c) return DNSPIR_E_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pair_release(c);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->own && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

void* dnspir_ctx_stream(dnspir_ctx* c) { return c ? (void*)c->stream : nullptr; }

int dnspir_ctx_set_graphs(dnspir_ctx* c, int on) {
    if (!c) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->graphs = on != 0;
    return c->graphs_broken ? 1 : 0;
}

int dnspir_ctx_set_exchange_polls(dnspir_ctx* c, int max_polls) {
    if (!c) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->drop_graphs();  // the bound is baked into captured launches
    c->exchange_polls = max_polls;
    return 0;
}

int dnspir_ctx_set_chain_cus(dnspir_ctx* c, int cus) {
    if (!c || (cus != 0 && cus != 1 && cus != 2 && cus != 4)) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->drop_graphs();  // the kernel is baked into captured launches
    c->chain_cus = cus;
    return 0;
}

int dnspir_ctx_get_stats(dnspir_ctx* c, dnspir_ctx_stats* out) {
    if (!out) return DNSPIR_E_ARG;
    if (!c) {
        int rc = ensure_init();
        if (rc) return rc;
        c = default_ctx();
    }
    std::lock_guard<std::mutex> lk(c->mu);
    *out = c->stats;
    return 0;
}

int dnspir_ctx_set_k1_mode(dnspir_ctx* c, int mode) {
    if (!c || (mode != 0 && mode != 2 && mode != 3)) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->k1_mode = mode;
    return 0;
}

int dnspir_choose_index_group_secondary(size_t num_shards) {
    if (num_shards > (size_t)(N / LANES)) return DNSPIR_E_SHAPE;  // pir_wrapper.rs:168
    int log2l = 0;
    while (((size_t)1 << log2l) < num_shards) ++log2l;
    return LANES << log2l;
}

int dnspir_database_shape(size_t n, int* lp, int* ls) {
    if (n < 1 || n > (size_t)N * N) return DNSPIR_E_SHAPE;  // pir_wrapper.rs:221-222
    const int l_p = n <= (size_t)N * N / 16 ? N / 2 : N;  // pir_wrapper.rs:188, 223-227
    const size_t shards = (n + (size_t)LANES * l_p - 1) / ((size_t)LANES * l_p);
    if (lp) *lp = l_p;
    if (ls) *ls = dnspir_choose_index_group_secondary(shards);
    return 0;
}

size_t dnspir_num_databases(size_t n) {
    int lp = 0;
    if (dnspir_database_shape(n, &lp, nullptr)) return 0;
    return (n + (size_t)LANES * lp - 1) / ((size_t)LANES * lp);
}

size_t dnspir_query_bytes(int lp, int ls) {
    size_t b = 2 * (N / 8) * GK_BITS + 2 * (N / 8) * QRY_BITS;
    if (lp == N) b += (N / 8) * QRY_BITS;
    if (ls == N) b += (N / 8) * QRY_BITS;
    return b;
}

size_t dnspir_reply_bytes(void) { return REPLY_BYTES; }

int dnspir_double_pir(dnspir_ctx* c, const dnspir_shard* const* f0, size_t len0, const dnspir_shard* const* f1,
                      size_t len1, const uint64_t* q0, const uint64_t* q1, uint64_t* out0, uint64_t* out1,
                      dnspir_timing* tm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!c) c = default_ctx();
    if (!f0 || len0 == 0) return DNSPIR_E_SHAPE;
    const int l_p = f0[0] ? f0[0]->l : 0;
    const bool present[2] = {q0 != nullptr, q1 != nullptr};
    const dnspir_shard* const* fleets[2] = {f0, f1};
    const size_t lens[2] = {len0, len1};
    if (len0 < len1) return DNSPIR_E_SHAPE;  // pir_wrapper.rs:362
    for (int sl = 0; sl < 2; ++sl)
        if (present[sl] && (rc = check_fleet(fleets[sl], lens[sl], l_p))) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    const double t0 = now_us();
    if ((rc = c->qcoef.ensure((size_t)LANES * 6 * N * 8))) return rc;
    const uint64_t* qs[2] = {q0, q1};
    for (int sl = 0; sl < 2; ++sl)
        if (present[sl])
            HIP_OK(hipMemcpyAsync(c->qcoef.as<uint64_t>() + (size_t)sl * 6 * N, qs[sl], (size_t)6 * N * 8,
                                  hipMemcpyHostToDevice, c->stream));
    if ((rc = run_double(c, fleets, lens, present, tm))) return rc;
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    if ((rc = check_pair(c))) return rc;
    const size_t nl = 3 * ((present[0] ? 1 : 0) + (present[1] ? 1 : 0));  // lanes carrying replies
    std::vector<uint64_t> a(nl * N), b(nl * N);
    HIP_OK(hipMemcpyAsync(a.data(), c->acoef.p, a.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(b.data(), c->bcoef.p, b.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    uint64_t* outs[2] = {out0, out1};
    int li = 0;
    for (int sl = 0; sl < 2; ++sl) {
        if (!present[sl]) continue;
        for (int p = 0; p < 3; ++p, ++li) {
            if (!outs[sl]) continue;
            memcpy(outs[sl] + ((size_t)p * 2 + 0) * N, a.data() + (size_t)li * N, N * 8);
            memcpy(outs[sl] + ((size_t)p * 2 + 1) * N, b.data() + (size_t)li * N, N * 8);
        }
    }
    fill_timing(c, tm, now_us() - t0);
    return 0;
}

// hr0 / hr1: pinned host memory that receives the replies r0 / r1 (copies on
// the stream, before the call's one synchronisation), or null
static int process_device_impl(dnspir_ctx* c, const dnspir_shard* const* f0, size_t len0, const dnspir_shard* const* f1,
                               size_t len1, const uint8_t* q0, size_t ql0, const uint8_t* q1, size_t ql1, uint8_t* r0,
                               uint8_t* r1, dnspir_timing* tm, double t0, uint8_t* hr0 = nullptr, uint8_t* hr1 = nullptr) {
    int rc;
    if (!f0 || len0 == 0 || !f0[0]) return DNSPIR_E_SHAPE;
    if (len0 < len1) return DNSPIR_E_SHAPE;
    const int l_p = f0[0]->l;
    const bool present[2] = {q0 != nullptr, q1 != nullptr};
    const dnspir_shard* const* fleets[2] = {f0, f1};
    const size_t lens[2] = {len0, len1};
    const uint8_t* qb[2] = {q0, q1};
    const size_t ql[2] = {ql0, ql1};
    for (int sl = 0; sl < 2; ++sl)
        if (present[sl] && (rc = check_fleet(fleets[sl], lens[sl], l_p))) return rc;
    const int l_s = dnspir_choose_index_group_secondary(len0);
    if (l_s < 0) return l_s;
    if ((rc = c->qcoef.ensure((size_t)LANES * 6 * N * 8))) return rc;
    const int conj_p = l_p == N, conj_s = l_s == N;
    for (int sl = 0; sl < 2; ++sl) {
        if (!present[sl]) continue;
        // pir_wrapper.rs:370-384: gk0 gk1 (50 bits) || primary (34) [|| conj] || secondary (34) [|| conj]
        int offs[6], bits[6] = {GK_BITS, GK_BITS, QRY_BITS, QRY_BITS, QRY_BITS, QRY_BITS};
        int pres[6] = {1, 1, 1, conj_p, 1, conj_s};
        size_t off = 0;
        for (int v = 0; v < 6; ++v) {
            if (!pres[v]) {
                offs[v] = -1;
                continue;
            }
            offs[v] = (int)off;
            off += dnh::condensed_size(bits[v], N);
        }
        if (ql[sl] < off) return DNSPIR_E_SHORT_QUERY;
        launch_decode_query(qb[sl], 0, 1, offs, bits, c->qcoef.as<uint64_t>() + (size_t)sl * 6 * N, c->stream);
    }
    if ((rc = run_double(c, fleets, lens, present, tm))) return rc;
    uint8_t* rs[2] = {r0, r1};
    int li = 0;
    for (int sl = 0; sl < 2; ++sl) {
        if (!present[sl]) continue;
        if (rs[sl]) {
            launch_encode_reply(c->a17o.as<uint32_t>(), c->b12o.as<uint32_t>(), li, 1, rs[sl], c->stream);
            note_encode(c, li, 1, rs[sl]);
        }
        li += 3;
    }
    uint8_t* hr[2] = {hr0, hr1};
    auto copy_back = [&]() -> int {
        for (int sl = 0; sl < 2; ++sl)
            if (present[sl] && rs[sl] && hr[sl])
                HIP_OK(hipMemcpyAsync(hr[sl], rs[sl], REPLY_BYTES, hipMemcpyDeviceToHost, c->stream));
        return 0;
    };
    if ((rc = copy_back())) return rc;
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    bool retried = false;
    if ((rc = check_pair(c, &retried))) return rc;
    if (retried) {  // the replies were encoded again: copy them again
        if ((rc = copy_back())) return rc;
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    fill_timing(c, tm, now_us() - t0);
    return 0;
}

int dnspir_process_queries_device(dnspir_ctx* c, const dnspir_shard* const* f0, size_t len0,
                                  const dnspir_shard* const* f1, size_t len1, const uint8_t* q0, size_t ql0,
                                  const uint8_t* q1, size_t ql1, uint8_t* r0, uint8_t* r1, dnspir_timing* tm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!c) c = default_ctx();
    std::lock_guard<std::mutex> lk(c->mu);
    return process_device_impl(c, f0, len0, f1, len1, q0, ql0, q1, ql1, r0, r1, tm, now_us());
}

}  // extern "C"

namespace {

// Carves, in the no-allocation context's arena, every buffer that
// process_device_impl takes for these fleets: first what lives through the
// whole call, then the primary phase's buffers, then -- over the same bytes,
// which are dead once the primary epilogue has written the secondary entries
// (one stream orders it) -- the secondary phase's. The call's own ensure()s
// are then no-ops. DNSPIR_E_SCRATCH if it does not fit.
int noalloc_reserve(dnspir_ctx* c, Arena* ar, int l_p, const size_t lens[2], const bool present[2], bool joint) {
    const int l_s = dnspir_choose_index_group_secondary(lens[0]);
    if (l_s < 0) return l_s;
    const int L_s = loop_len(l_s), nst_s = nstreams(l_s), nst_p = nstreams(l_p);
    const int nl = 3 * ((present[0] ? 1 : 0) + (present[1] ? 1 : 0));
    bool two = false, three = false;  // some slot takes the two-array / three-array inner product
    for (int sl = 0; sl < 2; ++sl)
        if (present[sl]) (use_k1w(c, l_p, lens[sl], joint ? 2 : 1) ? two : three) = true;
    const size_t whole[] = {
        PTR_SLOTS * sizeof(void*),                                 // ptrs
        (size_t)nl * (sizeof(long long) + 2 * sizeof(int)) + 64,   // idx
        (chain_pair_xbuf_words(nst_s * nl) + 64) * 4,              // xch
        (size_t)LANES * 6 * N * 8,                                 // qcoef
        (size_t)2 * 6 * 2 * N * 4,                                 // qs
        (size_t)2 * 4 * 2 * N * 4,                                 // qp
        (size_t)l_s * LANES * N * 2,                               // se
    };
    DevBuf* wb[] = {&c->ptrs, &c->idx, &c->xch, &c->qcoef, &c->qs, &c->qp, &c->se};
    int rc;
    for (size_t i = 0; i < sizeof(whole) / sizeof(whole[0]); ++i)
        if ((rc = wb[i]->ensure(whole[i]))) return DNSPIR_E_SCRATCH;
    const size_t mark = ar->used;
    if (loop_len(l_p) == N / 2) {
        const int nq = joint ? 2 : 1;
        if ((rc = c->outb.ensure((size_t)nq * lens[0] * 2 * N * LANES * 4))) return DNSPIR_E_SCRATCH;
        if (two && ((rc = c->wscan.ensure(nq * k1w_scan_bytes())) || (rc = c->wtile.ensure(nq * k1w_tile_bytes(nst_p)))))
            return DNSPIR_E_SCRATCH;
        if (three && (rc = c->ql.ensure(qlimb_bytes(1, nst_p)))) return DNSPIR_E_SCRATCH;
    } else if ((rc = c->acc.ensure((size_t)lens[0] * 2 * N * LANES * 8))) {
        return DNSPIR_E_SCRATCH;
    }
    const size_t primary_end = ar->used;
    ar->used = mark;
    const size_t sec[] = {
        (size_t)nl * l_s * N * 4,                  // M
        (size_t)nst_s * nl * L_s * 2 * N * 4,      // te
        (size_t)nst_s * nl * 2 * N * 4,            // aout
        (size_t)nst_s * nl * 2 * N * 4,            // cout
        (size_t)nl * N * 8,                        // acoef
        (size_t)nl * N * 8,                        // bcoef
        (size_t)nl * N * 4,                        // a17o
        (size_t)nl * N * 4,                        // b12o
    };
    DevBuf* sb[] = {&c->M, &c->te, &c->aout, &c->cout, &c->acoef, &c->bcoef, &c->a17o, &c->b12o};
    for (size_t i = 0; i < sizeof(sec) / sizeof(sec[0]); ++i)
        if ((rc = sb[i]->ensure(sec[i]))) return DNSPIR_E_SCRATCH;
    ar->used = std::max(ar->used, primary_end);
    return 0;
}

}  // namespace

extern "C" {

int dnspir_process_queries_noalloc(const dnspir_shard* const* f0, size_t len0, const dnspir_shard* const* f1, size_t len1,
                                   const uint8_t* q0, size_t ql0, const uint8_t* q1, size_t ql1, uint8_t* r0, uint8_t* r1,
                                   dnspir_timing* tm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (g.device >= Global::MAX_DEVICES || !g.na[g.device].ctx) return DNSPIR_E_NOINIT;
    Global::NoAlloc& na = g.na[g.device];
    dnspir_ctx* c = na.ctx;
    if (!f0 || len0 == 0 || !f0[0] || len0 < len1) return DNSPIR_E_SHAPE;
    const int l_p = f0[0]->l;
    const bool present[2] = {q0 != nullptr, q1 != nullptr};
    const dnspir_shard* const* fleets[2] = {f0, f1};
    const size_t lens[2] = {len0, len1};
    for (int sl = 0; sl < 2; ++sl)
        if (present[sl] && (rc = check_fleet(fleets[sl], lens[sl], l_p))) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    const bool joint = present[0] && present[1] && len0 == len1 && std::equal(f0, f0 + len0, f1) && use_k1w(c, l_p, len0, 2);
    const double t0 = now_us();
    // the carving starts over at every call; the array keeps the last call's
    // contents, and nothing here relies on them
    na.arena.used = 0;
    c->bind_all(&na.arena);
    if ((rc = noalloc_reserve(c, &na.arena, l_p, lens, present, joint))) return rc;
    rc = process_device_impl(c, f0, len0, f1, len1, q0, ql0, q1, ql1, r0, r1, tm, t0);
    return rc == DNSPIR_E_OOM ? DNSPIR_E_SCRATCH : rc;
}

int dnspir_process_queries(dnspir_ctx* c, const dnspir_shard* const* f0, size_t len0, const dnspir_shard* const* f1,
                           size_t len1, const uint8_t* q0, size_t ql0, const uint8_t* q1, size_t ql1, uint8_t* r0,
                           uint8_t* r1, dnspir_timing* tm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!c) c = default_ctx();
    std::lock_guard<std::mutex> lk(c->mu);
    const double t0 = now_us();
    const size_t qcap = dnspir_ctx::HOST_QCAP;
    if ((rc = c->qbytes.ensure(2 * qcap))) return rc;
    if ((rc = c->rbytes.ensure(2 * REPLY_BYTES))) return rc;
    // queries and replies travel through the context's pinned host buffers
    // (one DMA each way, the reply copies ordered before the call's one sync)
    if (!c->h_q) HIP_OK(hipHostMalloc((void**)&c->h_q, 2 * qcap, hipHostMallocDefault));
    if (!c->h_r) HIP_OK(hipHostMalloc((void**)&c->h_r, 2 * REPLY_BYTES, hipHostMallocDefault));
    const uint8_t* qh[2] = {q0, q1};
    const size_t ql[2] = {ql0, ql1};
    const uint8_t* qd[2] = {nullptr, nullptr};
    for (int sl = 0; sl < 2; ++sl) {
        if (!qh[sl]) continue;
        const size_t n = std::min(ql[sl], qcap);
        uint8_t* dst = c->qbytes.as<uint8_t>() + sl * qcap;
        if (n) {
            memcpy(c->h_q + sl * qcap, qh[sl], n);
            HIP_OK(hipMemcpyAsync(dst, c->h_q + sl * qcap, n, hipMemcpyHostToDevice, c->stream));
        }
        qd[sl] = dst;
    }
    uint8_t* rd[2] = {r0 ? c->rbytes.as<uint8_t>() : nullptr, r1 ? c->rbytes.as<uint8_t>() + REPLY_BYTES : nullptr};
    rc = process_device_impl(c, f0, len0, f1, len1, qd[0], std::min(ql0, qcap), qd[1], std::min(ql1, qcap), rd[0], rd[1],
                             tm, t0, c->h_r, c->h_r + REPLY_BYTES);
    if (rc) return rc;
    if (r0 && q0) memcpy(r0, c->h_r, REPLY_BYTES);
    if (r1 && q1) memcpy(r1, c->h_r + REPLY_BYTES, REPLY_BYTES);
    fill_timing(c, tm, now_us() - t0);  // bytes_streamed / ref_bytes were set by run_double
    return 0;
}

}  // extern "C"

size_t dne::batch_scratch_bytes(int l_p, size_t len, int nq, size_t qlen) {
    dnspir_ctx probe;  // never initialised: only its buffers' sizes are used
    Arena count;
    count.cap = ~(size_t)0 >> 1;
    probe.bind_arena(&count);
    if (batch_reserve(&probe, l_p, len, nq, qlen)) return 0;
    const size_t used = count.used;
    probe.bind_arena(nullptr);
    return used;
}

int dne::batch_reserve(dnspir_ctx* c, int l_p, size_t len, int nq, size_t qlen) {
    const int l_s = dnspir_choose_index_group_secondary(len);
    const int L_s = loop_len(l_s), nst_s = nstreams(l_s), nst_p = nstreams(l_p);
    const int nqc = std::min(64, nq), nl = 3 * nq;
    int rc;
    if (qlen) {
        if ((rc = c->qbytes.ensure((size_t)nq * qlen))) return rc;
        if ((rc = c->rbytes.ensure((size_t)nq * REPLY_BYTES))) return rc;
    }
    const size_t sizes[] = {
        (size_t)nq * 6 * N * 8,                      // qcoef
        (size_t)nq * 6 * 2 * N * 4,                  // qs
        (size_t)nq * 4 * 2 * N * 4,                  // qp
        (size_t)nq * l_s * 3 * N * 2,                // se
        qlimb_bytes(nqc, nst_p),                     // ql
        (size_t)nqc * len * 2 * N * LANES * 4,       // outb
        (size_t)nl * l_s * N * 4,                    // M
        (size_t)nst_s * nl * L_s * 2 * N * 4,        // te
        (size_t)nst_s * nl * 2 * N * 4,              // aout
        (size_t)nst_s * nl * 2 * N * 4,              // cout
        (size_t)nl * N * 8,                          // acoef
        (size_t)nl * N * 8,                          // bcoef
        (size_t)nl * N * 4,                          // a17o
        (size_t)nl * N * 4,                          // b12o
        use_k1w(c, l_p, len, nq) ? nq * k1w_scan_bytes() : 0,         // wscan
        use_k1w(c, l_p, len, nq) ? nq * k1w_tile_bytes(nst_p) : 0,    // wtile
    };
    DevBuf* bufs[] = {&c->qcoef, &c->qs, &c->qp, &c->se, &c->ql, &c->outb, &c->M, &c->te,
                      &c->aout, &c->cout, &c->acoef, &c->bcoef, &c->a17o, &c->b12o, &c->wscan, &c->wtile};
    for (size_t i = 0; i < sizeof(sizes) / sizeof(sizes[0]); ++i)
        if ((rc = bufs[i]->ensure(sizes[i]))) return rc;
    return 0;
}

// ---- batched queries -----------------------------------------------------------
// nq independent queries against one fleet, each answered as process_query
// (pir_wrapper.rs:328-337) but sharing one pass over the shards: the primary
// inner product runs as one MFMA contraction per 64 queries; every query then
// gets its own secondary database (3 lanes each) and all of those run
// together. Query q's reply bytes equal its single-query reply.
int dne::batch_impl(dnspir_ctx* c, const dnspir_shard* const* fleet, size_t len, const uint8_t* qdev, size_t qlen, int nq,
                    uint8_t* rdev, dnspir_timing* tm, double t0, const BatchKeys* keys, bool sync, bool graph) {
    int rc;
    if (!fleet || len == 0 || !fleet[0] || nq <= 0 || !qdev || !rdev) return DNSPIR_E_ARG;
    const int l_p = fleet[0]->l;
    if (loop_len(l_p) != N / 2) return DNSPIR_E_SHAPE;
    if ((rc = check_fleet(fleet, len, l_p))) return rc;
    const int l_s = dnspir_choose_index_group_secondary(len);
    int offs[6], bits[6];
    if (qlen < query_layout(l_p, l_s, offs, bits)) return DNSPIR_E_SHORT_QUERY;
    const hipStream_t s = c->stream;
    if ((rc = batch_reserve(c, l_p, len, nq, 0))) return rc;
    // few chains run on two CUs each; their exchange buffer is sized once for
    // the largest such batch (no allocation inside a graph capture)
    const int nchains = 3 * nq * nstreams(l_s);
    if (nchains <= PAIR_MAX_CHAINS && (rc = c->xch.ensure((chain_pair_xbuf_words(PAIR_MAX_CHAINS) + 64) * 4))) return rc;
    const uint32_t* const* slabs;
    const uint32_t* const* a17s;
    if ((rc = upload_fleet(c, 0, fleet, len, &slabs, &a17s))) return rc;
    std::vector<long long> loff;
    std::vector<int> ncols, slot;
    for (int q = 0; q < nq; ++q)
        for (int p = 0; p < 3; ++p) {
            loff.push_back(((long long)q * l_s * 3 + p) * N);
            ncols.push_back((int)(LANES * len));
            slot.push_back(q);
        }
    const long long* d_off;
    const int *d_ncols, *d_slot;
    if ((rc = upload_lanes(c, loff, ncols, slot, &d_off, &d_ncols, &d_slot))) return rc;
    // the pair form only while the device's pair budget has room (taken last,
    // after every early return above)
    const int cus = chain_acquire(c, nchains);
    const bool pair = cus > 1;
    const bool use_keys = keys && (keys->src || keys->dst);
    if (use_keys) {
        // cached keys replace the decoded ones; fresh ones are stored
        if ((rc = c->keyp.ensure((size_t)2 * nq * sizeof(void*)))) return rc;
        c->h_keyp.assign((size_t)2 * nq, nullptr);
        for (int q = 0; q < nq; ++q) {
            if (keys->src) c->h_keyp[q] = keys->src[q];
            if (keys->dst) c->h_keyp[nq + q] = keys->dst[q];
        }
        HIP_OK(hipMemcpyAsync(c->keyp.p, c->h_keyp.data(), (size_t)2 * nq * sizeof(void*), hipMemcpyHostToDevice, s));
    }
    // the device work of the batch, as two phases of kernels (timing events
    // only when launched eagerly)
    auto enqueue_primary = [&](bool events) -> int {
        launch_decode_query(qdev, qlen, nq, offs, bits, c->qcoef.as<uint64_t>(), s);
        launch_qprep(c->qcoef.as<uint64_t>(), c->qs.as<uint32_t>(), c->qp.as<uint32_t>(), nq, g.tables, s);
        if (use_keys)
            launch_key_cache(c->keyp.as<const uint32_t*>(), c->keyp.as<uint32_t*>() + nq, nq, c->qs.as<uint32_t>(),
                             c->qp.as<uint32_t>(), s);
        int r2;
        for (int q0 = 0; q0 < nq; q0 += 64) {
            const int nqc = std::min(64, nq - q0);
            bool ob = false;
            if ((r2 = run_primary(c, slabs, (int)len, l_p, c->qp.as<uint32_t>() + (size_t)q0 * 4 * 2 * N, nqc, &ob,
                                  events && q0 == 0 ? c->ev[3] : nullptr)))
                return r2;
            if (events && q0 == 0) HIP_OK(hipEventRecord(c->ev[4], s));
            EpiLaunch e;
            e.outb = c->outb.as<uint32_t>();
            e.nshards = (int)len;
            e.nq = nqc;
            e.a17s = a17s;
            e.se = c->se.as<uint16_t>() + (size_t)q0 * l_s * 3 * N;
            e.se_qstride = (size_t)l_s * 3 * N;
            e.lpr = 3;
            e.lane0_mult = 0;
            launch_epi(e, g.tables, s);
        }
        return 0;
    };
    auto enqueue_secondary = [&]() -> int {
        int r2;
        if ((r2 = run_secondary(c, c->se.as<uint16_t>(), 3LL * N, d_off, d_ncols, d_slot, 3 * nq, l_s, cus))) return r2;
        launch_encode_reply(c->a17o.as<uint32_t>(), c->b12o.as<uint32_t>(), 0, nq, rdev, s);
        note_encode(c, 0, nq, rdev);
        return 0;
    };
    c->k1_launches = (nq + 63) / 64;
    bool graphed = false;
    if (graph && c->graphs && !c->graphs_broken) {
        // everything the captured launches bake in
        std::vector<uintptr_t> key = {(uintptr_t)qdev, qlen, (uintptr_t)nq, (uintptr_t)rdev, len, (uintptr_t)l_p,
                                      (uintptr_t)use_keys, (uintptr_t)c->keyp.p, (uintptr_t)c->ptrs.p, (uintptr_t)c->idx.p,
                                      (uintptr_t)k1b_get_choice(1, nstreams(l_p)), (uintptr_t)k1b_get_choice(2, nstreams(l_p)),
                                      (uintptr_t)c->xch.p, (uintptr_t)c->h_err, (uintptr_t)use_k1w(c, l_p, len, nq),
                                      (uintptr_t)cus, (uintptr_t)c->exchange_polls};
        for (auto* b : c->scratch_bufs) key.push_back((uintptr_t)b->p);
        if (!c->graph_exec[0] || key != c->graph_key) {
            c->drop_graphs();
            bool ok = true;
            for (int ph = 0; ph < 2 && ok; ++ph) {
                hipGraph_t gr = nullptr;
                ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
                const int r2 = ok ? (ph == 0 ? enqueue_primary(false) : enqueue_secondary()) : 0;
                ok = (hipStreamEndCapture(s, &gr) == hipSuccess) && ok && r2 == 0 && gr;
                if (ok) ok = hipGraphInstantiate(&c->graph_exec[ph], gr, nullptr, nullptr, 0) == hipSuccess;
                if (gr) (void)hipGraphDestroy(gr);
            }
            if (ok) {
                c->graph_key = key;
            } else {
                (void)hipGetLastError();
                c->drop_graphs();
                c->graphs_broken = true;  // capture is not available here: launch eagerly from now on
            }
        }
        if (c->graph_exec[0] && c->graph_exec[1]) {
            // what the captured secondary phase does, for check_pair's retry
            note_secondary(c, c->se.as<uint16_t>(), 3LL * N, d_off, d_ncols, d_slot, 3 * nq, l_s);
            note_encode(c, 0, nq, rdev);
            HIP_OK(hipEventRecord(c->ev[0], s));
            HIP_OK(hipGraphLaunch(c->graph_exec[0], s));
            HIP_OK(hipEventRecord(c->ev[1], s));
            HIP_OK(hipGraphLaunch(c->graph_exec[1], s));
            HIP_OK(hipEventRecord(c->ev[2], s));
            graphed = true;
        }
    }
    if (!graphed) {
        HIP_OK(hipEventRecord(c->ev[0], s));
        if ((rc = enqueue_primary(true))) return rc;
        HIP_OK(hipEventRecord(c->ev[1], s));
        if ((rc = enqueue_secondary())) return rc;
        HIP_OK(hipEventRecord(c->ev[2], s));
    }
    c->k1_timed = !graphed;
    c->pair_pending = false;
    if (pair && (rc = copy_pair_flag(c))) return rc;
    HIP_OK(hipGetLastError());
    if (!sync) return 0;
    HIP_OK(hipStreamSynchronize(s));
    if ((rc = check_pair(c))) return rc;
    fill_timing(c, tm, now_us() - t0);
    if (tm) {
        tm->bytes_streamed = (uint64_t)((nq + 63) / 64) * len * slab_bytes(l_p);
        tm->ref_bytes = (uint64_t)((nq + 63) / 64) * len * ref_slab_bytes(l_p);
    }
    return 0;
}

extern "C" {

int dnspir_process_batch_device(dnspir_ctx* c, const dnspir_shard* const* fleet, size_t len, const uint8_t* queries_dev,
                                           size_t qlen, int nq, uint8_t* replies_dev, dnspir_timing* tm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!c) c = default_ctx();
    std::lock_guard<std::mutex> lk(c->mu);
    return batch_impl(c, fleet, len, queries_dev, qlen, nq, replies_dev, tm, now_us(), nullptr, true, true);
}

int dnspir_process_batch(dnspir_ctx* c, const dnspir_shard* const* fleet, size_t len, const uint8_t* queries,
                                    size_t qlen, int nq, uint8_t* replies, dnspir_timing* tm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!c) c = default_ctx();
    if (!queries || !replies || nq <= 0) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    const double t0 = now_us();
    if ((rc = c->qbytes.ensure((size_t)nq * qlen + 16))) return rc;
    if ((rc = c->rbytes.ensure((size_t)nq * REPLY_BYTES))) return rc;
    HIP_OK(hipMemcpyAsync(c->qbytes.p, queries, (size_t)nq * qlen, hipMemcpyHostToDevice, c->stream));
    rc = batch_impl(c, fleet, len, c->qbytes.as<uint8_t>(), qlen, nq, c->rbytes.as<uint8_t>(), nullptr, t0, nullptr, true, true);
    if (rc) return rc;
    HIP_OK(hipMemcpy(replies, c->rbytes.p, (size_t)nq * REPLY_BYTES, hipMemcpyDeviceToHost));
    fill_timing(c, tm, now_us() - t0);
    if (tm) {
        tm->bytes_streamed = (uint64_t)((nq + 63) / 64) * len * slab_bytes(fleet[0]->l);
        tm->ref_bytes = (uint64_t)((nq + 63) / 64) * len * ref_slab_bytes(fleet[0]->l);
    }
    return 0;
}

// ---- autotuner of the batched inner product --------------------------------------
int dnspir_autotune_batch(dnspir_ctx* c, const dnspir_shard* const* fleet, size_t len, int nq, dnspir_tune_entry* entries,
                          int max_entries, int* n_entries) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!c) c = default_ctx();
    if (!fleet || len == 0 || !fleet[0] || nq < 3 || nq > 64) return DNSPIR_E_ARG;
    const int l_p = fleet[0]->l;
    if (loop_len(l_p) != N / 2) return DNSPIR_E_SHAPE;
    if ((rc = check_fleet(fleet, len, l_p))) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    const hipStream_t s = c->stream;
    const int nst = nstreams(l_p), ct = nq > 32 ? 2 : 1;
    const size_t qlb = qlimb_bytes(nq, nst), ob = (size_t)nq * len * 2 * N * LANES * 4;
    if ((rc = c->ql.ensure(qlb)) || (rc = c->outb.ensure(ob)) || (rc = c->acc.ensure(ob + 256))) return rc;
    const uint32_t* const* slabs;
    const uint32_t* const* a17s;
    if ((rc = upload_fleet(c, 0, fleet, len, &slabs, &a17s))) return rc;
    // random int8 digits are valid operands (|digit| <= 128 bounds the sums)
    launch_fill_hash(c->ql.as<uint32_t>(), qlb / 4, 12345u, s);
    uint32_t* ref = c->acc.as<uint32_t>();
    unsigned* flag = reinterpret_cast<unsigned*>(c->acc.as<char>() + ob);
    hipEvent_t e0 = c->ev[3], e1 = c->ev[4];
    int n = 0, first = -1, best = -1;
    double best_ms = 1e30;
    for (int v = 0; v < k1b_variant_count(); ++v) {
        if (k1b_variant_ct(v) != ct) continue;
        uint32_t* out = first < 0 ? ref : c->outb.as<uint32_t>();
        launch_k1b_variant(v, slabs, (int)len, c->ql.as<uint32_t>(), out, nst, nq, s);  // warm-up
        HIP_OK(hipEventRecord(e0, s));
        const int reps = 3;
        for (int k = 0; k < reps; ++k) launch_k1b_variant(v, slabs, (int)len, c->ql.as<uint32_t>(), out, nst, nq, s);
        HIP_OK(hipEventRecord(e1, s));
        unsigned h_flag = 0;
        if (first >= 0) {
            launch_fill_u32(flag, 1, 0, s);
            launch_diff(out, ref, ob / 4, flag, s);
            HIP_OK(hipMemcpyAsync(&h_flag, flag, 4, hipMemcpyDeviceToHost, s));
        }
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(s));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        const double per = ms / reps;
        const bool exact = h_flag == 0;
        if (first < 0) first = v;
        if (exact && per < best_ms) {
            best_ms = per;
            best = v;
        }
        if (entries && n < max_entries) {
            dnspir_tune_entry& t = entries[n];
            memset(&t, 0, sizeof(t));
            snprintf(t.name, sizeof(t.name), "%s", k1b_variant_name(v));
            t.column_tiles = ct;
            t.ms = per;
            t.exact = exact ? 1 : 0;
        }
        ++n;
    }
    if (best < 0) return DNSPIR_E_HIP;
    k1b_set_choice(ct, nst, best);
    if (entries) {
        int k = 0;
        for (int v = 0; v < k1b_variant_count() && k < max_entries; ++v)
            if (k1b_variant_ct(v) == ct) entries[k++].chosen = v == best ? 1 : 0;
    }
    if (n_entries) *n_entries = std::min(n, max_entries);
    return 0;
}

int dnspir_probe_mfma_rate(dnspir_ctx* c, int zero_operands, int reps, double* tops) {
    if (!tops || reps < 1 || (zero_operands != 0 && zero_operands != 1)) return DNSPIR_E_ARG;
    int rc = ensure_init();
    if (rc) return rc;
    if (!c) c = default_ctx();
    std::lock_guard<std::mutex> lk(c->mu);
    const hipStream_t s = c->stream;
    int* out = nullptr;
    HIP_OK(hipMalloc(&out, mfma_rate_out_words() * sizeof(int)));
    hipEvent_t e0 = c->ev[3], e1 = c->ev[4];
    launch_mfma_rate(zero_operands, out, 1u, s);  // warm-up
    hipError_t e = hipEventRecord(e0, s);
    for (int k = 0; k < reps && e == hipSuccess; ++k) launch_mfma_rate(zero_operands, out, 2u + k, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipFree(out);
    HIP_OK(e);
    *tops = ms > 0 ? mfma_rate_ops() * reps / (ms * 1e-3) / 1e12 : 0;
    return 0;
}

int dnspir_probe_read_rate(dnspir_ctx* c, const dnspir_shard* const* fleet, size_t len, int reps, double* gbs) {
    if (!gbs || reps < 1 || !fleet || len == 0 || len > 65535) return DNSPIR_E_ARG;
    int rc = ensure_init();
    if (rc) return rc;
    if (!c) c = default_ctx();
    for (size_t k = 0; k < len; ++k)
        if (!fleet[k] || !fleet[k]->has_db) return DNSPIR_E_ARG;
    std::vector<const char*> hp(len);
    std::vector<size_t> hb(len);
    double total = 0;
    for (size_t k = 0; k < len; ++k) {
        hp[k] = reinterpret_cast<const char*>(fleet[k]->slab);
        hb[k] = slab_bytes(fleet[k]->l);
        total += (double)hb[k];
    }
    std::lock_guard<std::mutex> lk(c->mu);
    const hipStream_t s = c->stream;
    char* dev = nullptr;
    const size_t outw = read_rate_out_words((int)len);
    HIP_OK(hipMalloc(&dev, len * (sizeof(char*) + sizeof(size_t)) + outw * 4));
    const char** dptr = reinterpret_cast<const char**>(dev);
    size_t* dbytes = reinterpret_cast<size_t*>(dev + len * sizeof(char*));
    uint32_t* out = reinterpret_cast<uint32_t*>(dev + len * (sizeof(char*) + sizeof(size_t)));
    hipEvent_t e0 = c->ev[3], e1 = c->ev[4];
    hipError_t e = hipMemcpyAsync(dptr, hp.data(), len * sizeof(char*), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dbytes, hb.data(), len * sizeof(size_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) launch_read_rate(dptr, dbytes, (int)len, out, s);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int k = 0; k < reps && e == hipSuccess; ++k) launch_read_rate(dptr, dbytes, (int)len, out, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipFree(dev);
    HIP_OK(e);
    *gbs = ms > 0 ? total * reps / (ms * 1e-3) / 1e9 : 0;
    return 0;
}

// ---- client ------------------------------------------------------------------
struct dnspir_rng {
    dnh::ChaCha12 r;
    explicit dnspir_rng(const uint8_t* s) : r(s) {}
};

dnspir_rng* dnspir_rng_new(const uint8_t seed[32]) { return seed ? new (std::nothrow) dnspir_rng(seed) : nullptr; }
void dnspir_rng_free(dnspir_rng* r) { delete r; }

}  // extern "C"

namespace {

void client_sk(const uint8_t sk_seed[32], std::vector<int64_t>& sk) {
    sk.assign(N, 0);
    dnh::ChaCha12 skr(sk_seed);
    dnh::gen_sk(skr, 128, sk.data());
}

size_t emit_condensed(const uint64_t* x, int bits, uint8_t* out) {
    std::vector<uint64_t> tmp(N);
    for (int k = 0; k < N; ++k) tmp[k] = dnh::mod_switch_encode(x[k], bits);
    return dnh::condense(bits, tmp.data(), N, out);
}

// prepare_query_general (pir_wrapper.rs:276-317), key part: sk seed drawn
// from the rng (:289), Galois keys gk0 || gk1 (:291-293, :305-306)
int client_keys(dnspir_rng* rng, uint8_t sk_seed[32], uint8_t* keys, size_t* n) {
    for (int i = 0; i < 32; ++i) sk_seed[i] = (uint8_t)(rng->r.next_u32() & 0xFFFF);
    std::vector<int64_t> sk;
    client_sk(sk_seed, sk);
    std::vector<uint64_t> gk(2 * N);
    dnh::gen_gk_b(rng->r, sk.data(), gk.data());
    *n = emit_condensed(gk.data(), GK_BITS, keys);
    *n += emit_condensed(gk.data() + N, GK_BITS, keys + *n);
    return 0;
}

// query part: primary X^(i*kp) [and conj], secondary X^(j*ks) [and conj]
// (double_pir.rs:245-317, pir_wrapper.rs:294-315)
int client_query(dnspir_rng* rng, const uint8_t sk_seed[32], size_t pidx, size_t sidx, int lp, int ls, uint8_t* query,
                 size_t* qlen) {
    std::vector<int64_t> sk;
    client_sk(sk_seed, sk);
    const int kp = N / 2 / loop_len(lp), ks = N / 2 / loop_len(ls);
    const int64_t i = (int64_t)(pidx / LANES);
    const int64_t j = (int64_t)(sidx * LANES + pidx % LANES);
    std::vector<uint64_t> ct(N);
    std::vector<uint16_t> mono(N);
    auto set_mono = [&](int64_t e) {
        std::fill(mono.begin(), mono.end(), 0);
        e %= 2 * N;
        if (e < 0) e += 2 * N;
        if (e < N) mono[e] = 1;
        else mono[e - N] = (uint16_t)(dnh::T - 1);
    };
    set_mono(i * kp);
    dnh::enc_sym_b(dnh::SEEDS[2], rng->r, sk.data(), mono.data(), ct.data());
    std::vector<uint64_t> pq = ct, pqc, sq, sqc;
    if (lp == N) {
        set_mono(2 * N - i * kp);
        dnh::enc_sym_b(dnh::SEEDS[3], rng->r, sk.data(), mono.data(), ct.data());
        pqc = ct;
    }
    set_mono(j * ks);
    dnh::enc_sym_b(dnh::SEEDS[4], rng->r, sk.data(), mono.data(), ct.data());
    sq = ct;
    if (ls == N) {
        set_mono(2 * N - j * ks);
        dnh::enc_sym_b(dnh::SEEDS[5], rng->r, sk.data(), mono.data(), ct.data());
        sqc = ct;
    }
    size_t n = 0;
    n += emit_condensed(pq.data(), QRY_BITS, query + n);
    if (lp == N) n += emit_condensed(pqc.data(), QRY_BITS, query + n);
    n += emit_condensed(sq.data(), QRY_BITS, query + n);
    if (ls == N) n += emit_condensed(sqc.data(), QRY_BITS, query + n);
    *qlen = n;
    return 0;
}

}  // namespace

extern "C" {

int dnspir_client_keys(dnspir_rng* rng, uint8_t sk_seed[32], uint8_t* keys) {
    if (!rng || !sk_seed || !keys) return DNSPIR_E_ARG;
    size_t n = 0;
    return client_keys(rng, sk_seed, keys, &n);
}

int dnspir_client_query(dnspir_rng* rng, const uint8_t sk_seed[32], size_t pidx, size_t sidx, int lp, int ls,
                        uint8_t* query, size_t* qlen) {
    if (!rng || !sk_seed || !query || !qlen) return DNSPIR_E_ARG;
    if (!valid_group(lp) || !valid_group(ls)) return DNSPIR_E_SHAPE;
    if (pidx >= (size_t)lp * LANES || sidx >= (size_t)ls / LANES) return DNSPIR_E_ARG;
    return client_query(rng, sk_seed, pidx, sidx, lp, ls, query, qlen);
}

int dnspir_prepare_query(dnspir_rng* rng, size_t pidx, size_t sidx, int lp, int ls, uint8_t* query, size_t* qlen,
                         uint8_t sk_seed[32]) {
    if (!rng || !query || !qlen || !sk_seed) return DNSPIR_E_ARG;
    if (!valid_group(lp) || !valid_group(ls)) return DNSPIR_E_SHAPE;
    if (pidx >= (size_t)lp * LANES || sidx >= (size_t)ls / LANES) return DNSPIR_E_ARG;
    size_t nk = 0, nq = 0;
    client_keys(rng, sk_seed, query, &nk);
    client_query(rng, sk_seed, pidx, sidx, lp, ls, query + nk, &nq);
    *qlen = nk + nq;
    return 0;
}

int dnspir_process_reply(const uint8_t* reply, size_t rlen, const uint8_t sk_seed[32], uint16_t* out) {
    if (!reply || !sk_seed || !out) return DNSPIR_E_ARG;
    std::vector<int64_t> sk(N);
    {
        dnh::ChaCha12 skr(sk_seed);
        dnh::gen_sk(skr, 128, sk.data());
    }
    std::vector<uint64_t> tmp(N), a(N), b(N);
    std::vector<uint16_t> parts(3 * N);
    size_t off = 0;
    for (int p = 0; p < 3; ++p) {
        if (!dnh::uncondense(A_BITS, N, reply + off, rlen - off, tmp.data())) return DNSPIR_E_SHORT_QUERY;
        off += dnh::condensed_size(A_BITS, N);
        for (int k = 0; k < N; ++k) a[k] = dnh::mod_switch_decode(tmp[k], A_BITS);
        if (!dnh::uncondense(B_BITS, N, reply + off, rlen - off, tmp.data())) return DNSPIR_E_SHORT_QUERY;
        off += dnh::condensed_size(B_BITS, N);
        for (int k = 0; k < N; ++k) b[k] = dnh::mod_switch_decode(tmp[k], B_BITS);
        dnh::dec(sk.data(), a.data(), b.data(), parts.data() + (size_t)p * N);
    }
    std::vector<uint64_t> av(N), bv(N);
    for (int k = 0; k < N; ++k) {
        uint64_t v = 0;
        for (int p = 2; p >= 0; --p) v = v * 1025 + parts[(size_t)p * N + k];
        if ((v >> B_BITS) >= (1ull << A_BITS)) return DNSPIR_E_DECODE;  // double_pir.rs:88
        av[k] = dnh::mod_switch_decode(v >> B_BITS, A_BITS);
        bv[k] = dnh::mod_switch_decode(v & ((1ull << B_BITS) - 1), B_BITS);
    }
    dnh::dec(sk.data(), av.data(), bv.data(), out);
    return 0;
}

}  // extern "C"
