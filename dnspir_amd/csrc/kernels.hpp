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
// ncols_nonzero[li]: columns of lane `lanes[li]` that may be nonzero (device arrays)
void launch_k5(const uint16_t* entries, uint32_t* M, uint32_t* te_tmp, int l, const int* ncols_nonzero, const int* lanes,
               int nlanes, const DevTables* T, hipStream_t s);
void launch_chain_pre(const uint32_t* te_tmp, uint32_t* ip_tmp, uint32_t* a_out, int L, int nst, int qa_base,
                      const DevTables* T, hipStream_t s);
void launch_chain_sec(const uint32_t* te_tmp, uint32_t* a_out, uint32_t* c_out, const uint32_t* qs, const int* lane_slot,
                      int L, int nst, int nlanes, const DevTables* T, hipStream_t s);
void launch_relayout(const uint32_t* te_tmp, const uint32_t* ip_tmp, uint32_t* slab, int L, int nst, const DevTables* T,
                     hipStream_t s);
void launch_chain_out(const uint32_t* a_out, const uint32_t* c_out, int nst, int nlanes, uint64_t* a_coef, uint64_t* b_coef,
                      uint32_t* a17, uint32_t* b12, const DevTables* T, hipStream_t s);
void launch_qprep(const uint64_t* qcoef, uint32_t* qs, uint32_t* qp, int nq, const DevTables* T, hipStream_t s);
void launch_k1(const uint32_t* const* slabs, int nshards, const uint32_t* qp, unsigned long long* acc, int L, int nst,
               int rows_per_block, hipStream_t s);
void launch_k1_split(const uint32_t* const* slabs, int nshards, const uint32_t* qpl, unsigned long long* acc, int L, int nst,
                     int rows_per_block, hipStream_t s);
void launch_epi_primary(const unsigned long long* acc, int nshards, const uint32_t* const* a17s, uint16_t* se,
                        uint64_t* bcoef, int slot, const DevTables* T, hipStream_t s);

// offs[v] < 0: element absent (zero); bits[v]: wire width
void launch_decode_query(const uint8_t* qbytes, const int offs[6], const int bits[6], uint64_t* qcoef, hipStream_t s);
void launch_encode_reply(const uint32_t* a17, const uint32_t* b12, int lane0, uint8_t* reply, hipStream_t s);

}  // namespace dn
