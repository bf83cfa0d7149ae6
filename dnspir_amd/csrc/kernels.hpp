// Launchers of the HIP kernels in kernels.hip (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dn_common.hpp"

namespace dn {

void launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s);
void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
// K5: entries [l][8][N] u16 -> te_tmp [st][lane_idx][row][r][N]; M scratch [nlanes][l][N]
// entry (column j, lane index li) at entries + ent_off[li] + j * col_stride;
// ncols_nonzero[li]: columns of that lane that may be nonzero (device arrays);
// clear (optional): clear_words words zeroed on the way (a multi-CU chain's
// exchange buffer, so that its launch needs no fill of its own)
void launch_k5(const uint16_t* entries, const long long* ent_off, long long col_stride, uint32_t* M, uint32_t* te_tmp, int l,
               const int* ncols_nonzero, int nlanes, const DevTables* T, hipStream_t s, uint32_t* clear = nullptr,
               size_t clear_words = 0);
void launch_chain_pre(const uint32_t* te_tmp, uint32_t* ip_tmp, uint32_t* a_out, int L, int nst, int qa_base,
                      const DevTables* T, hipStream_t s);
void launch_chain_sec(const uint32_t* te_tmp, uint32_t* a_out, uint32_t* c_out, const uint32_t* qs, const int* lane_slot,
                      int L, int nst, int nlanes, const DevTables* T, hipStream_t s);
// the same on two CUs per chain (one per modulus) with a per-step exchange
// through xch: word 0 the timeout flag, then 63 words unused, the exchange
// words (chain_pair_xbuf_words) from word 64; all cleared here. Every wait
// ends after spin_limit polls (<= 0: at once); xch[0] != 0 after the launch
// means a wait timed out (the outputs are then invalid). cleared: xch was
// zeroed already on this stream (launch_k5's clear), no fill here
// quad: four CUs per chain (modulus x half of the slots, chain_quad_kernel)
// instead of two; the same xbuf size. chain_multi_blocks: the launch's grid
size_t chain_pair_xbuf_words(int nchains);
int chain_multi_blocks(int nchains, bool quad);
void launch_chain_pair(const uint32_t* te_tmp, uint32_t* a_out, uint32_t* c_out, const uint32_t* qs, const int* lane_slot,
                       int L, int nst, int nlanes, uint32_t* xch, int spin_limit, const DevTables* T, hipStream_t s,
                       bool quad = false, bool cleared = false);
// rows_tmp [st][lane][row][arr][r][gpos] (written by launch_chain_pre) -> slab;
// L = N/2 gives the limb-block layout read by launch_k1b, else the rotated one (launch_k1)
void launch_relayout(const uint32_t* rows_tmp, uint32_t* slab, int L, int nst, hipStream_t s);
// bytes of the query-limb scratch `ql` for nq queries
size_t qlimb_bytes(int nq, int nst);
// K1 on limb-block shards for nq <= 64 queries (qp [q][4][2][N] Galois order):
// outb [q][shard][r][lane][gpos] reduced residues
// main_start (optional) is recorded right before the inner-product launch
void launch_k1b(const uint32_t* const* slabs, int nshards, const uint32_t* qp, int nq, uint32_t* ql, uint32_t* outb, int nst,
                hipStream_t s, hipEvent_t main_start = nullptr);
// Two-stream K1 for one or two queries on limb-block shards (the outb of
// launch_k1b): per-query weights folded from the IPGK0 stream, then an inner
// product over the TE and IPGK1 streams only. scan: nq * k1w_scan_bytes(),
// tiles: nq * k1w_tile_bytes(nst).
__host__ __device__ constexpr size_t k1w_scan_bytes() { return (size_t)4 * dn::HALFN * dn::HALFN * 4; }
__host__ __device__ constexpr size_t k1w_tile_bytes(int nst) {
    return (size_t)4 * (dn::HALFN / 4) * 2 * nst * 32 * 512 + (size_t)4 * dn::HALFN * 4;
}
void launch_k1w(const uint32_t* const* slabs, int nshards, const uint32_t* qp, int nq, uint32_t* scan, uint32_t* tiles,
                uint32_t* outb, int nst, const DevTables* T, hipStream_t s, hipEvent_t main_start = nullptr);
// batched inner-product candidates (all exact), chosen per (column tiles, streams)
int k1b_variant_count();
const char* k1b_variant_name(int i);
int k1b_variant_ct(int i);
void k1b_set_choice(int ct, int nst, int i);
int k1b_get_choice(int ct, int nst);
// variant i on query limbs already in ql (qlimb layout for column tiles k1b_variant_ct(i))
void launch_k1b_variant(int i, const uint32_t* const* slabs, int nshards, const uint32_t* ql, uint32_t* outb, int nst, int nq,
                        hipStream_t s);
void launch_fill_hash(uint32_t* p, size_t n, uint32_t seed, hipStream_t s);
// the batched inner product's MFMA mix on register operands (zero = 1: all-zero
// digits, else hashed digits changing every step): out needs mfma_rate_out_words()
// words; one launch performs mfma_rate_ops() int8 operations
size_t mfma_rate_out_words();
double mfma_rate_ops();
void launch_mfma_rate(int zero, int* out, uint32_t seed, hipStream_t s);
// a streaming read of nranges device ranges (ptrs, bytes: device arrays of
// nranges entries; bytes multiples of 16); out needs read_rate_out_words(nranges) words
size_t read_rate_out_words(int nranges);
void launch_read_rate(const char* const* ptrs, const size_t* bytes, int nranges, uint32_t* out, hipStream_t s);
// n_total entries [n_total][N] u16 from n_valid buckets of 2560 bytes (the
// rest zero): pack_bytes_into_coeffs, dns/bucket.rs:289-303
void launch_pack_buckets(const uint8_t* buckets, size_t n_valid, size_t n_total, uint16_t* coeffs, hipStream_t s);
// *flag |= 1 where a and b differ
void launch_diff(const uint32_t* a, const uint32_t* b, size_t n, unsigned* flag, hipStream_t s);
// flag_dst (optional): receives *flag_src (a multi-CU chain's timeout flag),
// written to pinned host memory by the kernel
void launch_chain_out(const uint32_t* a_out, const uint32_t* c_out, int nst, int nlanes, uint64_t* a_coef, uint64_t* b_coef,
                      uint32_t* a17, uint32_t* b12, const DevTables* T, hipStream_t s, const unsigned* flag_src = nullptr,
                      unsigned* flag_dst = nullptr);
void launch_qprep(const uint64_t* qcoef, uint32_t* qs, uint32_t* qp, int nq, const DevTables* T, hipStream_t s);
// per-query Galois-key cache copies (src[q] -> key rows, else key rows -> dst[q]);
// src, dst: device arrays of nq entry pointers (null = none)
void launch_key_cache(const uint32_t* const* src, uint32_t* const* dst, int nq, uint32_t* qs, uint32_t* qp, hipStream_t s);
void launch_k1(const uint32_t* const* slabs, int nshards, const uint32_t* qp, unsigned long long* acc, int L, int nst,
               int rows_per_block, hipStream_t s);
void launch_k1_split(const uint32_t* const* slabs, int nshards, const uint32_t* qpl, unsigned long long* acc, int L, int nst,
                     int rows_per_block, hipStream_t s);
struct EpiLaunch {
    const uint32_t* outb = nullptr;            // k1b output, or
    const unsigned long long* acc = nullptr;   // k1 output (rotated shards)
    int nshards = 0, nq = 1;
    const uint32_t* const* a17s = nullptr;
    uint16_t* se = nullptr;                    // secondary entries (null = none)
    size_t se_qstride = 0;
    int lpr = LANES, lane0_mult = 3;
    uint64_t* bcoef = nullptr;                 // API output (null = none)
    int diag = 0;
};
void launch_epi(const EpiLaunch& e, const DevTables* T, hipStream_t s);

// offs[v] < 0: element absent (zero); bits[v]: wire width
// nq queries of qstride bytes each -> qcoef [nq][6][N]
void launch_decode_query(const uint8_t* qbytes, size_t qstride, int nq, const int offs[6], const int bits[6], uint64_t* qcoef,
                         hipStream_t s);
// nq replies: reply q from lanes lane0 + 3q .. lane0 + 3q + 2, at reply + q * 22272
void launch_encode_reply(const uint32_t* a17, const uint32_t* b12, int lane0, int nq, uint8_t* reply, hipStream_t s);

}  // namespace dn
