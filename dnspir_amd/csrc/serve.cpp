// Serving loop: continuous batching of PIR queries against several resident
// databases on one GPU.
//
// The reference answers each query on its connection's thread
// (crates/dns-over-pir/src/dns/server.rs:747-807 -> process_query,
// pir_wrapper.rs:328-337) and batches at most two queries per call
// (BATCH_COUNT, double_pir.rs:37). Here requests queue per database and a set
// of slots (one worker thread, query context and HIP stream each) drains
// them: whenever a slot is free it takes up to max_batch waiting requests of
// the database whose oldest request is oldest, if the batch's device scratch
// fits in what the scratch arena has free (admission by HBM budget). Batches
// of different slots run on the device at the same time. Each request is
// answered by batch_impl, whose replies do not depend on the other queries of
// the batch (lanes never interact), so a reply equals the one the query gets
// alone.
//
// Per-client Galois keys stay in HBM in an LRU cache: a client that sent its
// keys once may send keyless queries (the wire query without its 25,600-byte
// key prefix) until its entry is evicted, after which a keyless query fails
// with DNSPIR_E_NOKEY and the client sends the full query again.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

using namespace dne;

namespace {

constexpr int PENDING = 1;

struct Request {
    uint64_t ticket = 0;
    int db = 0;
    uint64_t client = 0;
    bool keyless = false;
    std::vector<uint8_t> q;
    int key_src = -1;  // cache entry the keys are read from
    int key_dst = -1;  // cache entry the decoded keys are written to
    int status = PENDING;
    int batch = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    std::vector<uint8_t> reply;
};

struct Database {
    std::vector<const dnspir_shard*> fleet;
    int l_p = 0, l_s = 0;
    size_t qlen = 0;           // full wire length
    int max_batch = 0;         // largest batch whose scratch fits the budget
    std::vector<size_t> need;  // need[nb]: scratch bytes of a batch of nb
    bool streaming = false;    // a batch is in its primary phase (reading the shards)
    double pass_us = 0;        // running mean of the primary phase's duration (0 = none yet)
    double last_arrival = 0;   // submit time of the latest request
    double gap_us = 0;         // running mean of the time between its requests (0 = none yet)
};

struct KeyEntry {
    uint64_t client = 0;
    bool mapped = false;
    bool ready = false;
    int pins = 0;  // in-flight or queued requests using the entry
    std::list<int>::iterator lru;
};

// first-fit allocator over the scratch arena's offsets
struct FreeList {
    std::map<size_t, size_t> free;  // offset -> length
    size_t in_use = 0;
    void init(size_t cap) {
        free.clear();
        free[0] = cap;
        in_use = 0;
    }
    size_t largest() const {
        size_t m = 0;
        for (auto& kv : free) m = std::max(m, kv.second);
        return m;
    }
    bool alloc(size_t n, size_t* off) {
        n = (n + 4095) & ~(size_t)4095;
        for (auto it = free.begin(); it != free.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const size_t rest = it->second - n;
            const size_t at = it->first + n;
            free.erase(it);
            if (rest) free[at] = rest;
            in_use += n;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = (n + 4095) & ~(size_t)4095;
        in_use -= n;
        auto it = free.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free.erase(nx);
        }
        if (it != free.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free.erase(it);
            }
        }
    }
};

struct Batch {
    int db = -1;
    std::vector<Request*> reqs;
    size_t off = 0, bytes = 0;
};

struct Slot {
    dnspir_ctx ctx;
    std::thread th;
    uint8_t* hq = nullptr;  // pinned query staging
    uint8_t* hr = nullptr;  // pinned reply staging
    uint64_t refused_seen = 0, retries_seen = 0;  // ctx.stats already added to the server's
};

}  // namespace

struct dnspir_server {
    dnspir_server_config cfg{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work;  // slots wait for formable batches
    std::condition_variable cv_done;  // callers wait for their requests
    bool stop = false;
    std::vector<Database> dbs;
    std::list<Request*> pending;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_ticket = 1;
    char* arena = nullptr;
    FreeList fl;
    uint32_t* key_mem = nullptr;
    std::vector<KeyEntry> keys;
    std::unordered_map<uint64_t, int> key_of;
    std::list<int> lru;  // front = most recently used
    std::vector<std::unique_ptr<Slot>> slots;
    size_t max_qlen = 0;
    int inflight = 0;
    bool waiting_for_memory = false;
    double wake_at = 0;  // earliest time a held-back batch becomes due (0 = none)
    dnspir_server_stats st{};

    // ---- key cache (caller holds mu) ----
    void touch(int e) {
        lru.erase(keys[e].lru);
        lru.push_front(e);
        keys[e].lru = lru.begin();
    }
    int lookup(uint64_t client) {
        auto it = key_of.find(client);
        return it == key_of.end() ? -1 : it->second;
    }
    // an entry for a new client: a free one, else the least recently used unpinned one
    int claim(uint64_t client) {
        for (auto it = lru.rbegin(); it != lru.rend(); ++it) {
            KeyEntry& k = keys[*it];
            if (k.pins) continue;
            const int e = *it;
            if (k.mapped) {
                key_of.erase(k.client);
                ++st.key_evictions;
            }
            k.client = client;
            k.mapped = true;
            k.ready = false;
            key_of[client] = e;
            touch(e);
            return e;
        }
        return -1;
    }
    void unmap(int e) {
        if (keys[e].mapped) key_of.erase(keys[e].client);
        keys[e].mapped = false;
        keys[e].ready = false;
    }

    // ---- batch formation (caller holds mu) ----
    void finish(Request* r, int status) {
        r->status = status;
        r->t_done = now_us();
        if (status) ++st.failed;
        else ++st.completed;
    }
    // false: the request waits (its cached keys are still being written)
    bool dispatchable(Request* r) {
        if (r->key_src < 0) return true;
        return keys[r->key_src].ready;
    }
    bool form(Batch& b) {
        std::vector<char> tried(dbs.size(), 0);
        bool blocked = false;
        for (auto it = pending.begin(); it != pending.end();) {
            Request* r = *it;
            if (r->key_src >= 0 && !keys[r->key_src].mapped) {
                // the write that would have provided the keys failed
                --keys[r->key_src].pins;
                r->key_src = -1;
                it = pending.erase(it);
                ++st.key_misses;
                finish(r, DNSPIR_E_NOKEY);
                cv_done.notify_all();
                continue;
            }
            if (tried[r->db] || dbs[r->db].streaming || !dispatchable(r)) {
                ++it;
                continue;
            }
            tried[r->db] = 1;
            Database& d = dbs[r->db];
            std::vector<std::list<Request*>::iterator> pick;
            for (auto jt = it; jt != pending.end() && (int)pick.size() < d.max_batch; ++jt)
                if ((*jt)->db == r->db && dispatchable(*jt)) pick.push_back(jt);
            // batching window: while other batches run, a database whose pass
            // streams for long lets its oldest request wait up to window x
            // pass time for more requests (each pass costs the same HBM bytes
            // whatever its batch size) if at least one more request is
            // expected in that time at its recent arrival rate; an idle GPU
            // starts at once
            const double hold = cfg.batch_window * d.pass_us;
            if (inflight > 0 && hold > 0 && d.gap_us > 0 && hold >= d.gap_us && (int)pick.size() < d.max_batch) {
                const double due = r->t_submit + hold;
                if (now_us() < due) {
                    wake_at = wake_at == 0 ? due : std::min(wake_at, due);
                    ++it;
                    continue;
                }
            }
            const size_t room = fl.largest();
            int nb = (int)pick.size();
            while (nb > 0 && d.need[nb] > room) --nb;
            if (nb == 0) {
                blocked = true;
                ++it;
                continue;
            }
            if (!fl.alloc(d.need[nb], &b.off)) {
                blocked = true;
                ++it;
                continue;
            }
            b.db = r->db;
            b.bytes = d.need[nb];
            d.streaming = true;
            b.reqs.clear();
            const double t = now_us();
            for (int i = 0; i < nb; ++i) {
                Request* x = *pick[i];
                x->t_start = t;
                x->batch = nb;
                b.reqs.push_back(x);
                pending.erase(pick[i]);
            }
            st.scratch_in_use_peak = std::max(st.scratch_in_use_peak, fl.in_use);
            waiting_for_memory = false;
            return true;
        }
        if (blocked && !waiting_for_memory) {
            ++st.admission_waits;
            waiting_for_memory = true;
        }
        return false;
    }

    int run(Slot* sl, Batch& b) {
        Database& d = dbs[b.db];
        const int nb = (int)b.reqs.size();
        const double t_dispatch = now_us();
        Arena ar;
        ar.base = arena + b.off;
        ar.cap = b.bytes;
        dnspir_ctx* c = &sl->ctx;
        c->bind_arena(&ar);
        int rc = batch_reserve(c, d.l_p, d.fleet.size(), nb, d.qlen);
        std::vector<const uint32_t*> src(nb, nullptr);
        std::vector<uint32_t*> dst(nb, nullptr);
        bool any_key = false;
        if (!rc) {
            for (int i = 0; i < nb; ++i) {
                Request* r = b.reqs[i];
                uint8_t* q = sl->hq + (size_t)i * d.qlen;
                if (r->keyless) {
                    memset(q, 0, KEY_BYTES);
                    memcpy(q + KEY_BYTES, r->q.data(), d.qlen - KEY_BYTES);
                } else {
                    memcpy(q, r->q.data(), d.qlen);
                }
                if (r->key_src >= 0) src[i] = key_mem + (size_t)r->key_src * KEY_ENTRY_WORDS;
                if (r->key_dst >= 0) dst[i] = key_mem + (size_t)r->key_dst * KEY_ENTRY_WORDS;
                any_key |= r->key_src >= 0 || r->key_dst >= 0;
            }
            if (hipMemcpyAsync(c->qbytes.p, sl->hq, (size_t)nb * d.qlen, hipMemcpyHostToDevice, c->stream) != hipSuccess)
                rc = DNSPIR_E_HIP;
        }
        if (!rc) {
            BatchKeys k;
            k.src = src.data();
            k.dst = dst.data();
            rc = batch_impl(c, d.fleet.data(), d.fleet.size(), c->qbytes.as<uint8_t>(), d.qlen, nb, c->rbytes.as<uint8_t>(),
                            nullptr, now_us(), any_key ? &k : nullptr, false);
        }
        // the shards are free for the next batch of this database once the
        // primary phase has read them; the secondary phase overlaps it
        if (!rc && hipEventSynchronize(c->ev[1]) != hipSuccess) rc = DNSPIR_E_HIP;
        {
            std::lock_guard<std::mutex> lk(mu);
            d.streaming = false;
            if (!rc) {
                const double dt = now_us() - t_dispatch;
                d.pass_us = d.pass_us == 0 ? dt : 0.8 * d.pass_us + 0.2 * dt;
            }
        }
        cv_work.notify_all();
        auto fetch = [&] {
            return hipMemcpyAsync(sl->hr, c->rbytes.p, (size_t)nb * REPLY_BYTES, hipMemcpyDeviceToHost, c->stream) ==
                           hipSuccess &&
                       hipStreamSynchronize(c->stream) == hipSuccess
                       ? 0
                       : DNSPIR_E_HIP;
        };
        if (!rc) rc = fetch();
        bool retried = false;
        if (!rc) rc = check_pair(c, &retried);
        if (!rc && retried) rc = fetch();  // the replies were encoded again
        {
            std::lock_guard<std::mutex> lk(mu);
            st.pair_refused += c->stats.pair_refused - sl->refused_seen;
            st.pair_retries += c->stats.pair_retries - sl->retries_seen;
            sl->refused_seen = c->stats.pair_refused;
            sl->retries_seen = c->stats.pair_retries;
        }
        if (rc == DNSPIR_E_HIP) fprintf(stderr, "dnspir server: %s\n", dnspir_last_error());
        if (!rc)
            for (int i = 0; i < nb; ++i)
                b.reqs[i]->reply.assign(sl->hr + (size_t)i * REPLY_BYTES, sl->hr + (size_t)(i + 1) * REPLY_BYTES);
        c->bind_arena(nullptr);
        return rc;
    }

    void complete(Batch& b, int rc) {
        fl.release(b.off, b.bytes);
        ++st.batches;
        for (Request* r : b.reqs) {
            if (r->key_dst >= 0) {
                KeyEntry& k = keys[r->key_dst];
                --k.pins;
                if (rc) unmap(r->key_dst);
                else k.ready = true;
            }
            if (r->key_src >= 0) --keys[r->key_src].pins;
            finish(r, rc);
        }
    }

    void worker(Slot* sl) {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            Batch b;
            for (;;) {
                wake_at = 0;
                if (stop || form(b)) break;
                if (wake_at > 0) {
                    const auto until = std::chrono::steady_clock::time_point(std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                        std::chrono::duration<double, std::micro>(wake_at)));
                    cv_work.wait_until(lk, until);
                } else {
                    cv_work.wait(lk);
                }
            }
            if (stop) break;
            ++inflight;
            st.inflight_peak = std::max(st.inflight_peak, inflight);
            lk.unlock();
            const int rc = run(sl, b);
            lk.lock();
            complete(b, rc);
            --inflight;
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    int submit(int db, uint64_t client, const uint8_t* query, size_t qlen, int flags, uint64_t* ticket) {
        if (db < 0 || db >= (int)dbs.size() || !query || !ticket) return DNSPIR_E_ARG;
        Database& d = dbs[db];
        const bool keyless = (flags & DNSPIR_QUERY_KEYLESS) != 0;
        const size_t want = keyless ? d.qlen - KEY_BYTES : d.qlen;
        if (qlen < want) return DNSPIR_E_SHORT_QUERY;
        std::unique_ptr<Request> r(new (std::nothrow) Request);
        if (!r) return DNSPIR_E_OOM;
        r->db = db;
        r->client = client;
        r->keyless = keyless;
        r->q.assign(query, query + want);
        std::lock_guard<std::mutex> lk(mu);
        if (stop) return DNSPIR_E_CLOSED;
        const int e = lookup(client);
        if (keyless) {
            if (e < 0) {
                ++st.key_misses;
                return DNSPIR_E_NOKEY;
            }
            ++st.key_hits;
            ++keys[e].pins;
            touch(e);
            r->key_src = e;
        } else if (!keys.empty()) {
            int w = e;
            if (w >= 0 && keys[w].pins) w = -1;  // in use: leave it, the query carries its own keys
            else if (w < 0) w = claim(client);
            else touch(w);
            if (w >= 0) {
                keys[w].ready = false;
                ++keys[w].pins;
                r->key_dst = w;
                ++st.key_inserts;
            }
        }
        r->ticket = next_ticket++;
        r->t_submit = now_us();
        if (d.last_arrival > 0) {  // arrival rate of this database (batching window)
            const double gap = r->t_submit - d.last_arrival;
            d.gap_us = d.gap_us == 0 ? gap : 0.9 * d.gap_us + 0.1 * gap;
        }
        d.last_arrival = r->t_submit;
        Request* raw = r.release();
        reqs[raw->ticket] = raw;
        pending.push_back(raw);
        *ticket = raw->ticket;
        cv_work.notify_one();
        return 0;
    }

    int wait(uint64_t ticket, uint8_t* reply, dnspir_request_info* info) {
        std::unique_lock<std::mutex> lk(mu);
        auto it = reqs.find(ticket);
        if (it == reqs.end()) return DNSPIR_E_ARG;
        Request* r = it->second;
        cv_done.wait(lk, [&] { return r->status != PENDING; });
        reqs.erase(it);
        lk.unlock();
        const int rc = r->status;
        if (!rc && reply) memcpy(reply, r->reply.data(), REPLY_BYTES);
        if (info) {
            info->status = rc;
            info->db = r->db;
            info->batch_size = r->batch;
            info->key_cached = r->key_src >= 0 ? 1 : 0;
            info->submit_us = r->t_submit;
            info->start_us = r->t_start;
            info->done_us = r->t_done;
        }
        delete r;
        return rc;
    }
};

extern "C" {

size_t dnspir_key_bytes(void) { return KEY_BYTES; }

size_t dnspir_batch_scratch_bytes(int l_primary, size_t num_shards, int nq) {
    if (loop_len(l_primary) != N / 2 || num_shards == 0 || num_shards > (size_t)(N / LANES) || nq <= 0) return 0;
    const int l_s = dnspir_choose_index_group_secondary(num_shards);
    int offs[6], bits[6];
    return batch_scratch_bytes(l_primary, num_shards, nq, query_layout(l_primary, l_s, offs, bits));
}

int dnspir_server_create(const dnspir_server_config* cfg, dnspir_server** out) {
    if (!cfg || !out) return DNSPIR_E_ARG;
    if (cfg->slots < 1 || cfg->slots > 8 || cfg->max_batch < 1 || cfg->max_batch > 128 || cfg->key_cache_entries < 0)
        return DNSPIR_E_ARG;
    if (!(cfg->batch_window >= 0 && cfg->batch_window <= 1000)) return DNSPIR_E_ARG;
    int rc = ensure_init();
    if (rc) return rc;
    std::unique_ptr<dnspir_server> s(new (std::nothrow) dnspir_server);
    if (!s) return DNSPIR_E_OOM;
    s->cfg = *cfg;
    s->cfg.scratch_budget &= ~(size_t)4095;  // the free list hands out whole 4 KiB pages
    if (s->cfg.scratch_budget == 0) return DNSPIR_E_ARG;
    (void)hipGetDevice(&s->device);
    if (hipMalloc(&s->arena, s->cfg.scratch_budget) != hipSuccess) return DNSPIR_E_OOM;
    s->fl.init(s->cfg.scratch_budget);
    if (cfg->key_cache_entries) {
        if (hipMalloc(&s->key_mem, (size_t)cfg->key_cache_entries * KEY_ENTRY_WORDS * 4) != hipSuccess) {
            (void)hipFree(s->arena);
            return DNSPIR_E_OOM;
        }
        s->keys.resize(cfg->key_cache_entries);
        for (int e = 0; e < cfg->key_cache_entries; ++e) {
            s->lru.push_back(e);
            s->keys[e].lru = std::prev(s->lru.end());
        }
    }
    *out = s.release();
    return 0;
}

int dnspir_server_add_database(dnspir_server* s, const dnspir_shard* const* fleet, size_t len, int* db) {
    if (!s || !fleet || len == 0 || !db) return DNSPIR_E_ARG;
    if (len > (size_t)(N / LANES) || !fleet[0]) return DNSPIR_E_SHAPE;
    const int l_p = fleet[0]->l;
    if (loop_len(l_p) != N / 2) return DNSPIR_E_SHAPE;
    for (size_t i = 0; i < len; ++i)
        if (!fleet[i] || !fleet[i]->has_db || fleet[i]->l != l_p || fleet[i]->seeds != 0) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->slots.empty()) return DNSPIR_E_ARG;  // databases are fixed once serving starts
    Database d;
    d.fleet.assign(fleet, fleet + len);
    d.l_p = l_p;
    d.l_s = dnspir_choose_index_group_secondary(len);
    int offs[6], bits[6];
    d.qlen = query_layout(l_p, d.l_s, offs, bits);
    d.need.assign(s->cfg.max_batch + 1, 0);
    for (int nb = 1; nb <= s->cfg.max_batch; ++nb) {
        d.need[nb] = batch_scratch_bytes(l_p, len, nb, d.qlen);
        if (d.need[nb] == 0 || d.need[nb] > s->cfg.scratch_budget) break;
        d.max_batch = nb;
    }
    if (d.max_batch == 0) return DNSPIR_E_OOM;  // not even one query fits the budget
    s->max_qlen = std::max(s->max_qlen, d.qlen);
    *db = (int)s->dbs.size();
    s->dbs.push_back(std::move(d));
    return 0;
}

// releases what slot_init set up (no worker thread runs on it)
static void slot_free(Slot* sl) {
    if (sl->ctx.stream) (void)hipStreamSynchronize(sl->ctx.stream);
    for (auto& e : sl->ctx.ev)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    if (sl->ctx.own && sl->ctx.stream) (void)hipStreamDestroy(sl->ctx.stream);
    sl->ctx.stream = nullptr;
    if (sl->hq) (void)hipHostFree(sl->hq);
    if (sl->hr) (void)hipHostFree(sl->hr);
    sl->hq = sl->hr = nullptr;
}

static int slot_init(dnspir_server* s, Slot* sl) {
    int rc = ctx_init(&sl->ctx, nullptr);
    if (rc) return rc;
    if ((rc = ctx_presize(&sl->ctx, s->cfg.max_batch))) return rc;
    HIP_OK(hipHostMalloc((void**)&sl->hq, (size_t)s->cfg.max_batch * s->max_qlen, hipHostMallocDefault));
    HIP_OK(hipHostMalloc((void**)&sl->hr, (size_t)s->cfg.max_batch * REPLY_BYTES, hipHostMallocDefault));
    return 0;
}

// starts the slot workers on first use (after the databases are known); the
// slots become visible only once every one of them is set up, so a failure
// leaves the server without slots (the next submit tries again)
static int start_slots(dnspir_server* s) {
    if (!s->slots.empty()) return 0;
    if (s->dbs.empty()) return DNSPIR_E_ARG;
    std::vector<std::unique_ptr<Slot>> made;
    for (int i = 0; i < s->cfg.slots; ++i) {
        std::unique_ptr<Slot> sl(new (std::nothrow) Slot);
        int rc = sl ? slot_init(s, sl.get()) : DNSPIR_E_OOM;
        if (sl) made.push_back(std::move(sl));
        if (rc) {
            for (auto& m : made) slot_free(m.get());
            return rc;
        }
    }
    s->slots = std::move(made);
    for (auto& sl : s->slots) {
        Slot* p = sl.get();
        p->th = std::thread([s, p] { s->worker(p); });
    }
    return 0;
}

int dnspir_server_submit(dnspir_server* s, int db, uint64_t client, const uint8_t* query, size_t qlen, int flags,
                         uint64_t* ticket) {
    if (!s) return DNSPIR_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        int rc = start_slots(s);
        if (rc) return rc;
    }
    return s->submit(db, client, query, qlen, flags, ticket);
}

int dnspir_server_wait(dnspir_server* s, uint64_t ticket, uint8_t* reply, dnspir_request_info* info) {
    if (!s) return DNSPIR_E_ARG;
    if (info) memset(info, 0, sizeof(*info));
    return s->wait(ticket, reply, info);
}

int dnspir_server_replay(dnspir_server* s, const dnspir_arrival* arr, size_t n, uint8_t* replies, dnspir_request_info* infos) {
    if (!s || (!arr && n)) return DNSPIR_E_ARG;
    using clk = std::chrono::steady_clock;
    std::vector<uint64_t> tickets(n, 0);
    std::vector<int> sub_rc(n, 0), resent(n, 0);
    std::vector<uint8_t> full;
    const auto t0 = clk::now();
    for (size_t i = 0; i < n; ++i) {
        const auto due = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double, std::micro>(arr[i].arrival_us));
        if (due > clk::now() + std::chrono::microseconds(200)) std::this_thread::sleep_until(due - std::chrono::microseconds(100));
        while (clk::now() < due) {
        }
        int rc = dnspir_server_submit(s, arr[i].db, arr[i].client, arr[i].query, arr[i].qlen, DNSPIR_QUERY_KEYLESS,
                                      &tickets[i]);
        if (rc == DNSPIR_E_NOKEY && arr[i].keys) {
            // the client sends its keys with the query (the reference wire format)
            full.assign(arr[i].keys, arr[i].keys + KEY_BYTES);
            full.insert(full.end(), arr[i].query, arr[i].query + arr[i].qlen);
            rc = dnspir_server_submit(s, arr[i].db, arr[i].client, full.data(), full.size(), 0, &tickets[i]);
            resent[i] = 1;
        }
        sub_rc[i] = rc;
    }
    int first = 0;
    for (size_t i = 0; i < n; ++i) {
        dnspir_request_info info;
        memset(&info, 0, sizeof(info));
        int rc = sub_rc[i];
        if (!rc) rc = dnspir_server_wait(s, tickets[i], replies ? replies + i * REPLY_BYTES : nullptr, &info);
        info.status = rc;
        info.resent = resent[i];
        if (infos) infos[i] = info;
        if (rc && !first) first = rc;
    }
    return first;
}

}  // extern "C"

int dne::server_db_shape(dnspir_server* s, int db, int* l_p, size_t* len) {
    if (!s) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (db < 0 || db >= (int)s->dbs.size()) return DNSPIR_E_ARG;
    *l_p = s->dbs[db].l_p;
    *len = s->dbs[db].fleet.size();
    return 0;
}

extern "C" {

int dnspir_server_get_stats(dnspir_server* s, dnspir_server_stats* out) {
    if (!s || !out) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    return 0;
}

int dnspir_server_max_batch(dnspir_server* s, int db) {
    if (!s) return DNSPIR_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (db < 0 || db >= (int)s->dbs.size()) return DNSPIR_E_ARG;
    return s->dbs[db].max_batch;
}

int dnspir_server_destroy(dnspir_server* s) {
    if (!s) return DNSPIR_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Request* r : s->pending) s->finish(r, DNSPIR_E_CLOSED);
        s->pending.clear();
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    for (auto& sl : s->slots)
        if (sl->th.joinable()) sl->th.join();
    for (auto& sl : s->slots) slot_free(sl.get());
    for (auto& kv : s->reqs) delete kv.second;
    if (s->arena) (void)hipFree(s->arena);
    if (s->key_mem) (void)hipFree(s->key_mem);
    delete s;
    return 0;
}

}  // extern "C"
