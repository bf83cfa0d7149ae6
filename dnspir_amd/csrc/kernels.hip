// HIP kernels of the PIR response engine (gfx950).
//
// Data layouts (all residues u32 in [0, p_r)):
//   slot order   my negacyclic NTT's output order; slot s holds f(psi^e_s)
//   Galois order position g = sgn*N/2 + a holds exponent (-1)^sgn 5^a, so
//                sigma_5 is a cyclic shift by one inside each half
//   shard slab   [r][stream][row i][array (TE, IPGK0, IPGK1)][gpos][lane];
//                row i is stored pre-rotated by L-1-i inside each half so the
//                streaming inner product reads it at the output position
//                (the Galois scatter of base_pir.rs:1085-1247 made static)
//   lane-major temporaries [stream][lane][row][...][r][slot] during encoding
#include <atomic>
#include <type_traits>

#include "dn_common.hpp"
#include "kernels.hpp"

namespace dn {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// block-level transforms over LDS (radix 2, `nt` cooperating threads with
// index `tid`; every thread of the block runs the same barrier sequence)
// ---------------------------------------------------------------------------

// forward negacyclic NTT (Cooley-Tukey, psi powers in bit-reversed order):
// natural coefficients -> slot order
__device__ void ntt_fwd(uint32_t* a, const DevTables* __restrict__ T, int r, int tid, int nt) {
    const uint32_t p = mod_p(r);
    const uint32_t* psi = T->psi[r];
    const uint32_t* psiq = T->psiq[r];
#pragma unroll 1
    for (int s = 0; s < LOGN; ++s) {
        const int m = 1 << s;
        const int lt = LOGN - 1 - s;
        const int t = 1 << lt;
        for (int b = tid; b < N / 2; b += nt) {
            const int i = b >> lt;
            const int j = (i << (lt + 1)) + (b & (t - 1));
            const uint32_t U = a[j];
            const uint32_t V = mul_shoup(a[j + t], psi[m + i], psiq[m + i], p);
            a[j] = add_mod(U, V, p);
            a[j + t] = sub_mod(U, V, p);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int brv_bits(int v, int bits) {
    return bits == 0 ? 0 : (int)(__brev((unsigned)v) >> (32 - bits));
}

// ---------------------------------------------------------------------------
// register-radix NTT passes (radix 8, radix 4 for stages 9-10) over an LDS
// array; lazy values (forward: growing by 2p per stage, below 24p after a
// transform; inverse: Harvey's [0, 2p)), Shoup products by twiddles
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t red2p(uint32_t a, uint32_t p2) {  // [0, 4p) -> [0, 2p)
    return csub(a, p2);
}
__device__ __forceinline__ uint32_t shoup_lazy(uint32_t a, uint32_t w, uint32_t wq, uint32_t p) {  // -> [0, 2p)
    return a * w - __umulhi(a, wq) * p;
}
// Shoup companion floor(w 2^32 / p) of a constant w < p
__device__ __forceinline__ uint32_t shoup_companion(uint32_t w, uint32_t p) {
    return (uint32_t)(((uint64_t)w << 32) / p);
}
// [0, 8p) -> [0, p) (p < 2^26): three conditional subtractions
__device__ __forceinline__ uint32_t red8p(uint32_t x, uint32_t p) {
    return csub(csub(csub(x, 4 * p), 2 * p), p);
}
// The top balanced base-2^26 digit D1 of the centered lift of the coefficient
// whose CRT terms are e1 = c (q/p1)^-1 mod p1 and e2 = c (q/p2)^-1 mod p2 (both
// reduced), base_pir.rs:544-570: x = e1 p2 + e2 p1 mod q, centered; |x| < q/2 <
// 2^50, so the second digit is hi = (x + 2^25) >> 26 itself (the third is 0).
__device__ __forceinline__ int32_t crt_digit1(uint32_t e1, uint32_t e2) {
    uint64_t v = (uint64_t)e1 * P2 + (uint64_t)e2 * P1;
    v = v >= Q ? v - Q : v;
    const int64_t x = center_q(v);
    return (int32_t)((x + (1ll << (LOG2_B - 1))) >> LOG2_B);  // |d1| <= 2^24
}
// forward CT butterfly without reducing x: the Shoup product takes any 32-bit
// y to [0, 2p), so inputs below B give outputs below B + 2p. An 11-stage
// transform of inputs below 2p ends below 24p < 2^31 (p < 2^26); consumers
// either feed the outputs to Shoup products (any 32-bit input) or reduce
// them with red32.
__device__ __forceinline__ void bf_ct(uint32_t& x, uint32_t& y, uint32_t w, uint32_t wq, uint32_t p) {
    const uint32_t u = x;
    const uint32_t t = shoup_lazy(y, w, wq, p);
    x = u + t;
    y = u - t + 2 * p;
}
// any 32-bit x -> [0, p): a Shoup product by 1 (companion m = floor(2^32 / p))
// leaves [0, 2p), one conditional subtraction
constexpr uint32_t M32_P1 = (uint32_t)((1ull << 32) / P1), M32_P2 = (uint32_t)((1ull << 32) / P2);
__device__ __forceinline__ uint32_t red32(uint32_t x, uint32_t p, uint32_t m) { return csub(x - __umulhi(x, m) * p, p); }
// inverse GS butterfly, inputs/outputs in [0, 2p)
__device__ __forceinline__ void bf_gs(uint32_t& x, uint32_t& y, uint32_t w, uint32_t wq, uint32_t p) {
    const uint32_t p2 = 2 * p;
    const uint32_t u = x, v = y;
    x = red2p(u + v, p2);
    y = shoup_lazy(u - v + p2, w, wq, p);
}

// Barrier that orders LDS traffic only: global loads stay in flight across it
// (__syncthreads also waits for every outstanding global access). For loops
// whose threads share data through LDS alone.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders the LDS accesses of one wave: a wave's LDS instructions execute in
// issue order, so only the compiler has to be kept from moving loads and
// stores across this point. For passes whose data stays within one wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    asm volatile("" ::: "memory");
}

// LDS index of element j in a transform buffer. Padded buffers (N + N/8
// words) insert 4 words after every 32: the radix-8 pass at stride 4 then
// hits 32 distinct banks per half-wave instead of 4, and 4-aligned groups
// stay 16-byte aligned.
constexpr int NPAD = N + N / 8;
template <bool PAD>
__device__ __forceinline__ int lx(int j) {
    return PAD ? j + ((j >> 5) << 2) : j;
}

// one radix-8 forward pass (stages 3P..3P+2) for group g of 256
template <bool PAD = false>
__device__ __forceinline__ void fwd_pass8(uint32_t* buf, const uint32_t* tw, const uint32_t* twq, uint32_t p, int P, int g) {
    const int m = 1 << (3 * P);
    const int q = N >> (3 * P + 3);
    const int i = g / q, r = g % q;
    const int j0 = i * 8 * q + r;
    uint32_t v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = buf[lx<PAD>(j0 + c * q)];
    {
        const uint32_t w = tw[m + i], wq = twq[m + i];
#pragma unroll
        for (int c = 0; c < 4; ++c) bf_ct(v[c], v[c + 4], w, wq, p);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t w = tw[2 * m + 2 * i + h], wq = twq[2 * m + 2 * i + h];
        bf_ct(v[4 * h + 0], v[4 * h + 2], w, wq, p);
        bf_ct(v[4 * h + 1], v[4 * h + 3], w, wq, p);
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const uint32_t w = tw[4 * m + 4 * i + h], wq = twq[4 * m + 4 * i + h];
        bf_ct(v[2 * h], v[2 * h + 1], w, wq, p);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) buf[lx<PAD>(j0 + c * q)] = v[c];
}
// last forward pass: stages 9 and 10 for the two radix-4 groups g, g + 256
template <bool PAD = false>
__device__ __forceinline__ void fwd_pass4(uint32_t* buf, const uint32_t* tw, const uint32_t* twq, uint32_t p, int g) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = g + 256 * k;
        const int b4 = lx<PAD>(4 * i);
        uint32_t v0 = buf[b4], v1 = buf[b4 + 1], v2 = buf[b4 + 2], v3 = buf[b4 + 3];
        const uint32_t w = tw[512 + i], wq = twq[512 + i];
        bf_ct(v0, v2, w, wq, p);
        bf_ct(v1, v3, w, wq, p);
        bf_ct(v0, v1, tw[1024 + 2 * i], twq[1024 + 2 * i], p);
        bf_ct(v2, v3, tw[1025 + 2 * i], twq[1025 + 2 * i], p);
        buf[b4] = v0; buf[b4 + 1] = v1; buf[b4 + 2] = v2; buf[b4 + 3] = v3;
    }
}
// one radix-8 inverse pass (stages 3P+2..3P)
template <bool PAD = false>
__device__ __forceinline__ void inv_pass8(uint32_t* buf, const uint32_t* tw, const uint32_t* twq, uint32_t p, int P, int g) {
    const int m = 1 << (3 * P);
    const int q = N >> (3 * P + 3);
    const int i = g / q, r = g % q;
    const int j0 = i * 8 * q + r;
    uint32_t v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = buf[lx<PAD>(j0 + c * q)];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const uint32_t w = tw[4 * m + 4 * i + h], wq = twq[4 * m + 4 * i + h];
        bf_gs(v[2 * h], v[2 * h + 1], w, wq, p);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t w = tw[2 * m + 2 * i + h], wq = twq[2 * m + 2 * i + h];
        bf_gs(v[4 * h + 0], v[4 * h + 2], w, wq, p);
        bf_gs(v[4 * h + 1], v[4 * h + 3], w, wq, p);
    }
    {
        const uint32_t w = tw[m + i], wq = twq[m + i];
#pragma unroll
        for (int c = 0; c < 4; ++c) bf_gs(v[c], v[c + 4], w, wq, p);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) buf[lx<PAD>(j0 + c * q)] = v[c];
}

// first inverse pass: stages 10 and 9 for the radix-4 groups g, g + 256
template <bool PAD = false>
__device__ __forceinline__ void inv_pass4(uint32_t* buf, const uint32_t* tw, const uint32_t* twq, uint32_t p, int g) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = g + 256 * k;
        const int b4 = lx<PAD>(4 * i);
        uint32_t v0 = buf[b4], v1 = buf[b4 + 1], v2 = buf[b4 + 2], v3 = buf[b4 + 3];
        bf_gs(v0, v1, tw[1024 + 2 * i], twq[1024 + 2 * i], p);
        bf_gs(v2, v3, tw[1025 + 2 * i], twq[1025 + 2 * i], p);
        const uint32_t w = tw[512 + i], wq = twq[512 + i];
        bf_gs(v0, v2, w, wq, p);
        bf_gs(v1, v3, w, wq, p);
        buf[b4] = v0; buf[b4 + 1] = v1; buf[b4 + 2] = v2; buf[b4 + 3] = v3;
    }
}

// Radix-4 passes for 512 threads per transform (group g < 512): stages t, t+1
// (forward) or t+1, t (inverse) with 2^t groups at stage t; element stride
// q = N / 2^(t+2) inside a group of 4. Each twiddle sits beside its Shoup
// companion (tw2[k] = (w_k, wq_k), 16-byte aligned): one 8-byte LDS read for
// the first stage's twiddle, one 16-byte read for the second stage's adjacent
// pair, so a group issues 2 twiddle reads instead of 6.
template <bool PAD = false>
__device__ __forceinline__ void fwd_pass4gi(uint32_t* buf, const uint2* tw2, uint32_t p, int t, int g) {
    const int m = 1 << t, lq = LOGN - t - 2, q = 1 << lq;
    const int i = g >> lq, j0 = (i << (lq + 2)) + (g & (q - 1));
    uint32_t v0 = buf[lx<PAD>(j0)], v1 = buf[lx<PAD>(j0 + q)], v2 = buf[lx<PAD>(j0 + 2 * q)], v3 = buf[lx<PAD>(j0 + 3 * q)];
    const uint2 wa = tw2[m + i];
    bf_ct(v0, v2, wa.x, wa.y, p);
    bf_ct(v1, v3, wa.x, wa.y, p);
    const uint4 wb = *reinterpret_cast<const uint4*>(tw2 + 2 * m + 2 * i);
    bf_ct(v0, v1, wb.x, wb.y, p);
    bf_ct(v2, v3, wb.z, wb.w, p);
    buf[lx<PAD>(j0)] = v0; buf[lx<PAD>(j0 + q)] = v1; buf[lx<PAD>(j0 + 2 * q)] = v2; buf[lx<PAD>(j0 + 3 * q)] = v3;
}
template <bool PAD = false>
__device__ __forceinline__ void inv_pass4gi(uint32_t* buf, const uint2* tw2, uint32_t p, int t, int g) {
    const int m = 1 << t, lq = LOGN - t - 2, q = 1 << lq;
    const int i = g >> lq, j0 = (i << (lq + 2)) + (g & (q - 1));
    uint32_t v0 = buf[lx<PAD>(j0)], v1 = buf[lx<PAD>(j0 + q)], v2 = buf[lx<PAD>(j0 + 2 * q)], v3 = buf[lx<PAD>(j0 + 3 * q)];
    const uint4 wb = *reinterpret_cast<const uint4*>(tw2 + 2 * m + 2 * i);
    bf_gs(v0, v1, wb.x, wb.y, p);
    bf_gs(v2, v3, wb.z, wb.w, p);
    const uint2 wa = tw2[m + i];
    bf_gs(v0, v2, wa.x, wa.y, p);
    bf_gs(v1, v3, wa.x, wa.y, p);
    buf[lx<PAD>(j0)] = v0; buf[lx<PAD>(j0 + q)] = v1; buf[lx<PAD>(j0 + 2 * q)] = v2; buf[lx<PAD>(j0 + 3 * q)] = v3;
}
// radix-8 forward pass (stages 3P..3P+2) and the last radix-4 forward pass
// (stages 9, 10 for groups g, g + 256) on the interleaved twiddle table
template <bool PAD = false>
__device__ __forceinline__ void fwd_pass8i(uint32_t* buf, const uint2* tw2, uint32_t p, int P, int g) {
    const int m = 1 << (3 * P);
    const int q = N >> (3 * P + 3);
    const int i = g / q, r = g % q;
    const int j0 = i * 8 * q + r;
    uint32_t v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = buf[lx<PAD>(j0 + c * q)];
    const uint2 w1 = tw2[m + i];
#pragma unroll
    for (int c = 0; c < 4; ++c) bf_ct(v[c], v[c + 4], w1.x, w1.y, p);
    const uint4 w2 = *reinterpret_cast<const uint4*>(tw2 + 2 * m + 2 * i);
    bf_ct(v[0], v[2], w2.x, w2.y, p);
    bf_ct(v[1], v[3], w2.x, w2.y, p);
    bf_ct(v[4], v[6], w2.z, w2.w, p);
    bf_ct(v[5], v[7], w2.z, w2.w, p);
    const uint4 w3a = *reinterpret_cast<const uint4*>(tw2 + 4 * m + 4 * i);
    const uint4 w3b = *reinterpret_cast<const uint4*>(tw2 + 4 * m + 4 * i + 2);
    bf_ct(v[0], v[1], w3a.x, w3a.y, p);
    bf_ct(v[2], v[3], w3a.z, w3a.w, p);
    bf_ct(v[4], v[5], w3b.x, w3b.y, p);
    bf_ct(v[6], v[7], w3b.z, w3b.w, p);
#pragma unroll
    for (int c = 0; c < 8; ++c) buf[lx<PAD>(j0 + c * q)] = v[c];
}
// fwd_pass4i with the groups of wave w (g = thread index of 256) in elements
// [512w, 512w + 512), where the radix-8 passes at strides 32 and 4 leave them
template <bool PAD = false>
__device__ __forceinline__ void fwd_pass4iw(uint32_t* buf, const uint2* tw2, uint32_t p, int g) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = ((g >> 6) << 7) + (g & 63) + 64 * k;
        const int b4 = lx<PAD>(4 * i);
        uint32_t v0 = buf[b4], v1 = buf[b4 + 1], v2 = buf[b4 + 2], v3 = buf[b4 + 3];
        const uint2 wa = tw2[512 + i];
        bf_ct(v0, v2, wa.x, wa.y, p);
        bf_ct(v1, v3, wa.x, wa.y, p);
        const uint4 wb = *reinterpret_cast<const uint4*>(tw2 + 1024 + 2 * i);
        bf_ct(v0, v1, wb.x, wb.y, p);
        bf_ct(v2, v3, wb.z, wb.w, p);
        buf[b4] = v0; buf[b4 + 1] = v1; buf[b4 + 2] = v2; buf[b4 + 3] = v3;
    }
}
// the last forward pass with the groups of wave w in elements [256w, 256w + 256)
// (g = thread index of 512): after the stride-128 stage the passes of a wave
// touch only its own block, so they need no block barrier
template <bool PAD = false>
__device__ __forceinline__ void fwd_pass2_lastw(uint32_t* buf, const uint2* tw2, uint32_t p, int g) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = ((g >> 6) << 7) + (g & 63) + 64 * k;
        uint32_t x = buf[lx<PAD>(2 * i)], y = buf[lx<PAD>(2 * i + 1)];
        const uint2 w = tw2[1024 + i];
        bf_ct(x, y, w.x, w.y, p);
        buf[lx<PAD>(2 * i)] = x;
        buf[lx<PAD>(2 * i + 1)] = y;
    }
}
// whole-block forward transform by 256 threads (g = thread index) on an LDS
// array, inputs in [0, 2p), outputs left below 24p (callers reduce them as
// they write them out); the inverse leaves the N factor in, outputs in [0, 2p)
template <bool PAD = false>
__device__ __forceinline__ void ntt256_fwd_lazy(uint32_t* buf, const uint32_t* tw, const uint32_t* twq, uint32_t p, int g) {
#pragma unroll
    for (int P = 0; P < 3; ++P) {
        fwd_pass8<PAD>(buf, tw, twq, p, P, g);
        __syncthreads();
    }
    fwd_pass4<PAD>(buf, tw, twq, p, g);
    __syncthreads();
}
template <bool PAD = false>
__device__ __forceinline__ void ntt256_inv_unscaled(uint32_t* buf, const uint32_t* tw, const uint32_t* twq, uint32_t p, int g) {
    inv_pass4<PAD>(buf, tw, twq, p, g);
    __syncthreads();
#pragma unroll
    for (int P = 2; P >= 0; --P) {
        inv_pass8<PAD>(buf, tw, twq, p, P, g);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// K5: database encoding (base_pir.rs:662-772), modulus p1 for steps 1-5
// ---------------------------------------------------------------------------

// steps 1-2: column j of lane index li -> NTT of the entry's coefficient
// vector; the entry sits at entries + ent_off[li] + j * col_stride (u16 [N]).
// M [lane_idx][j][slot]
__global__ __launch_bounds__(256) void k5_col_ntt_kernel(const uint16_t* __restrict__ entries, const long long* __restrict__ ent_off,
                                                         long long col_stride, uint32_t* __restrict__ M, int l,
                                                         const int* __restrict__ ncols_nonzero, const DevTables* __restrict__ T,
                                                         uint32_t* __restrict__ clear, int clear_per_block, size_t clear_words) {
    __shared__ uint32_t a[NPAD];
    const int j = blockIdx.x, li = blockIdx.y;
    // side task: zero a slice of `clear` (the chain's exchange words, which the
    // chain launched after this phase needs cleared)
    if (clear) {
        const size_t c0 = ((size_t)li * gridDim.x + j) * clear_per_block;
        for (int k = threadIdx.x; k < clear_per_block; k += blockDim.x)
            if (c0 + k < clear_words) clear[c0 + k] = 0;
    }
    uint32_t* out = M + ((size_t)li * l + j) * N;
    if (j >= ncols_nonzero[li]) {
        for (int s = threadIdx.x; s < N; s += blockDim.x) out[s] = 0;
        return;
    }
    const uint16_t* src = entries + ent_off[li] + (long long)j * col_stride;
    for (int s = threadIdx.x; s < N; s += blockDim.x) a[lx<true>(s)] = src[s];
    __syncthreads();
    ntt256_fwd_lazy<true>(a, T->psi[0], T->psiq[0], P1, threadIdx.x);
    for (int s = threadIdx.x; s < N; s += blockDim.x) out[s] = red32(a[lx<true>(s)], P1, M32_P1);
}

// step 3 (coset_ntt, base_pir.rs:397-444) for a tile of TS slots of one lane,
// in place in M. Row s: twist by psi^(-coset j), DFT over j, then
// new[i] = evaluation at psi^(-z 5^i) (main half) / psi^(z 5^(i-N/2)) (conj
// half), z = e_s * k mod 2N.
// Stages top, top-1, ..., top-R+1 of the DIF transform over the l rows of a
// tile [l][TS] (TS interleaved columns), in registers per group of 2^R
// elements; lazy values in [0, 2p), Shoup products by psi^-(2N/2^(lh+1) j),
// read from the LDS table dtw[2^lh + j] (companions in dtwq).
template <int TS, int R, int NT>
__device__ __forceinline__ void k5_dif_pass(uint32_t* tile, int logl, int top, const uint32_t* dtw, const uint32_t* dtwq, int tid) {
    constexpr int G = 1 << R;
    const int sh = top - R + 1;  // group stride 2^sh in rows
    const int ngroups = (1 << logl) / G * TS;
    for (int g = tid; g < ngroups; g += NT) {
        const int d = g % TS, gi = g / TS;
        const int jlow = gi & ((1 << sh) - 1), blk = gi >> sh;
        const int base = (blk << (top + 1)) + jlow;
        uint32_t v[G];
#pragma unroll
        for (int k = 0; k < G; ++k) v[k] = tile[(base + (k << sh)) * TS + d];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const int lh = top - t, hk = G >> (t + 1);
#pragma unroll
            for (int k = 0; k < G; ++k) {
                if (k & hk) continue;
                const int ti = (1 << lh) + ((base + (k << sh)) & ((1 << lh) - 1));
                bf_gs(v[k], v[k + hk], dtw[ti], dtwq[ti], P1);
            }
        }
#pragma unroll
        for (int k = 0; k < G; ++k) tile[(base + (k << sh)) * TS + d] = v[k];
    }
}

// K5R_NT threads per block: the 64 KiB tile allows two blocks per CU, so a
// block brings 16 waves to hide the LDS and strided-load latency
constexpr int K5R_NT = 1024;
template <int TS>
__global__ __launch_bounds__(K5R_NT) void k5_rows_kernel(uint32_t* __restrict__ M, int l, int logl, int kq, int conj,
                                                         const DevTables* __restrict__ T) {
    extern __shared__ uint32_t tile[];  // [l][TS], l * TS <= 16384
    __shared__ uint32_t dtw[16384 / TS], dtwq[16384 / TS];  // DIF twiddles, l entries
    const int s0 = blockIdx.x * TS, li = blockIdx.y;
    uint32_t* base = M + (size_t)li * l * N;
    const int tid = threadIdx.x;
    const int lstep = LOGN + 1 - logl;  // 2N / l = 2^lstep
    for (int t = tid + 1; t < l; t += K5R_NT) {  // stage lh, position j: t = 2^lh + j
        const int lh = 31 - __clz(t), j = t - (1 << lh);
        const int e = (M2N - (((M2N >> (lh + 1)) * j) & (M2N - 1))) & (M2N - 1);
        dtw[t] = T->psipow[0][e];
        dtwq[t] = T->psipowq[0][e];
    }
    // each thread keeps one slot of the tile (K5R_NT is a multiple of TS)
    const int d = tid % TS, s = s0 + d;
    const int z = (T->slot_exp[s] * kq) & (M2N - 1);
    const int coset = z & ((1 << lstep) - 1);
    for (int j = tid / TS; j < l; j += K5R_NT / TS) {
        const int e = (M2N - ((coset * j) & (M2N - 1))) & (M2N - 1);
        tile[j * TS + d] = mul_shoup(base[(size_t)j * N + s], T->psipow[0][e], T->psipowq[0][e], P1);
    }
    __syncthreads();
    int top = logl - 1;
    while (top >= 2) {
        k5_dif_pass<TS, 3, K5R_NT>(tile, logl, top, dtw, dtwq, tid);
        __syncthreads();
        top -= 3;
    }
    if (top == 1) k5_dif_pass<TS, 2, K5R_NT>(tile, logl, top, dtw, dtwq, tid);
    else if (top == 0) k5_dif_pass<TS, 1, K5R_NT>(tile, logl, top, dtw, dtwq, tid);
    if (top >= 0) __syncthreads();
    for (int i = tid / TS; i < l; i += K5R_NT / TS) {
        int y;
        if (conj && i >= N / 2) y = (z * T->pow5[i - N / 2]) & (M2N - 1);
        else y = ((M2N - z) * T->pow5[i]) & (M2N - 1);
        const int u = (((M2N - y) & (M2N - 1)) - coset) >> lstep;  // natural DFT index
        const uint32_t v = tile[brv_bits(u, logl) * TS + d];
        base[(size_t)i * N + s] = csub(v, P1);
    }
}

// k5_rows_kernel for l <= 256 with one wave per 64-slot row of the tile: the
// DIF groups of a wave share their row indices, so every twiddle is
// wave-uniform and comes from the scalar cache (psipow and its Shoup
// companion), leaving LDS to the data. Same arithmetic as k5_rows_kernel.
__global__ __launch_bounds__(K5R_NT) void k5_rows_w_kernel(uint32_t* __restrict__ M, int l, int logl, int kq, int conj,
                                                           const DevTables* __restrict__ T) {
    constexpr int TS = 64, NW = K5R_NT / 64;
    extern __shared__ uint32_t tile[];  // [l][64]
    const int s0 = blockIdx.x * TS, li = blockIdx.y;
    uint32_t* base = M + (size_t)li * l * N;
    const int tid = threadIdx.x, d = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lstep = LOGN + 1 - logl;  // 2N / l = 2^lstep
    const int s = s0 + d;
    const int z = (T->slot_exp[s] * kq) & (M2N - 1);
    const int coset = z & ((1 << lstep) - 1);
    for (int j = wave; j < l; j += NW) {
        const int e = (M2N - ((coset * j) & (M2N - 1))) & (M2N - 1);
        tile[j * TS + d] = mul_shoup(base[(size_t)j * N + s], T->psipow[0][e], T->psipowq[0][e], P1);
    }
    __syncthreads();
    // DIF stages top .. top-R+1 for the groups of 2^R rows this wave owns
    auto pass = [&](auto RR, int top) {
        constexpr int R = decltype(RR)::value, G = 1 << R;
        const int sh = top - R + 1;
        for (int gi = wave; gi < (l >> R); gi += NW) {
            const int jlow = gi & ((1 << sh) - 1), blk = gi >> sh;
            const int b0 = (blk << (top + 1)) + jlow;
            uint32_t v[G];
#pragma unroll
            for (int k = 0; k < G; ++k) v[k] = tile[(b0 + (k << sh)) * TS + d];
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const int lh = top - t, hk = G >> (t + 1);
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    if (k & hk) continue;
                    // twiddle psi^-(2N / 2^(lh+1) * position) (the table dtw of k5_rows_kernel)
                    const int pos = (b0 + (k << sh)) & ((1 << lh) - 1);
                    const int e = (M2N - (((M2N >> (lh + 1)) * pos) & (M2N - 1))) & (M2N - 1);
                    bf_gs(v[k], v[k + hk], T->psipow[0][e], T->psipowq[0][e], P1);
                }
            }
#pragma unroll
            for (int k = 0; k < G; ++k) tile[(b0 + (k << sh)) * TS + d] = v[k];
        }
    };
    int top = logl - 1;
    while (top >= 2) {
        pass(std::integral_constant<int, 3>(), top);
        __syncthreads();
        top -= 3;
    }
    if (top == 1) pass(std::integral_constant<int, 2>(), top);
    else if (top == 0) pass(std::integral_constant<int, 1>(), top);
    if (top >= 0) __syncthreads();
    for (int i = wave; i < l; i += NW) {
        int y;
        if (conj && i >= N / 2) y = (z * T->pow5[i - N / 2]) & (M2N - 1);
        else y = ((M2N - z) * T->pow5[i]) & (M2N - 1);
        const int u = (((M2N - y) & (M2N - 1)) - coset) >> lstep;  // natural DFT index
        const uint32_t v = tile[brv_bits(u, logl) * TS + d];
        base[(size_t)i * N + s] = csub(v, P1);
    }
}

// steps 4-6: final column c takes column j = rev(c) (step 5), gathers its
// slots by sigma_{5^-j} (step 4), then inverse NTT, centered lift, times
// l^-1 mod t, centered lift, embedding into p1 and p2 and forward NTTs.
// te_tmp [st][lane_idx][row][r][slot]
__global__ __launch_bounds__(256) void k5_cols_final_kernel(const uint32_t* __restrict__ M, uint32_t* __restrict__ te_tmp,
                                                            int l, int L, int conj, int nlanes, int scale,
                                                            const DevTables* __restrict__ T) {
    __shared__ uint32_t a[2][NPAD];
    const int c = blockIdx.x, li = blockIdx.y;
    const int tid = threadIdx.x, nt = blockDim.x;
    int j;
    if (conj) j = c < N / 2 ? N / 2 - 1 - c : N / 2 + (N - 1 - c);
    else j = l - 1 - c;
    const uint32_t* col = M + ((size_t)li * l + j) * N;
    const int au = T->inv5pow[j & (HALFN - 1)];
    for (int s = tid; s < N; s += nt) a[0][lx<true>(s)] = col[T->exp_slot[(au * T->slot_exp[s]) & (M2N - 1)]];
    __syncthreads();
    ntt256_inv_unscaled<true>(a[0], T->ipsi[0], T->ipsiq[0], P1, tid);
    for (int s = tid; s < N; s += nt) {
        const int xs = lx<true>(s);
        uint32_t u = a[0][xs];
        u = csub(u, P1);
        int32_t x = (int32_t)mul_shoup(u, T->ninv[0], T->ninvq[0], P1);
        if (x > (int32_t)(P1 / 2)) x -= (int32_t)P1;
        int32_t xt = x % (int32_t)T_PLAIN;
        if (xt < 0) xt += T_PLAIN;
        int32_t y = (int32_t)(((uint32_t)xt * (uint32_t)scale) % T_PLAIN);  // < 1025^2
        if (y > (int32_t)(T_PLAIN / 2)) y -= (int32_t)T_PLAIN;
        a[0][xs] = y < 0 ? (uint32_t)(y + (int32_t)P1) : (uint32_t)y;
        a[1][xs] = y < 0 ? (uint32_t)(y + (int32_t)P2) : (uint32_t)y;
    }
    __syncthreads();
    ntt256_fwd_lazy<true>(a[0], T->psi[0], T->psiq[0], P1, tid);
    ntt256_fwd_lazy<true>(a[1], T->psi[1], T->psiq[1], P2, tid);
    const int st = c < L ? 0 : 1;
    const int row = c < L ? c : c - L;
    uint32_t* dst = te_tmp + (((size_t)(st * nlanes + li) * L + row) * 2) * N;
    for (int s = tid; s < N; s += nt) {  // below 24p -> [0, p)
        dst[s] = red32(a[0][lx<true>(s)], P1, M32_P1);
        dst[N + s] = red32(a[1][lx<true>(s)], P2, M32_P2);
    }
}

// ---------------------------------------------------------------------------
// K4/K6: the sequential a-chain (preprocess_loop, base_pir.rs:819-862, with
// partial_gal :496-594). One block of 1024 threads per (lane, stream). Step i:
//   A  gather sigma_5(a) and inverse NTT in p1 and p2     (512 threads each)
//   B  CRT, centered lift, top balanced base-2^26 digit D1  (all threads)
//   C  forward NTTs of D1 in p1 and p2                      (256 threads each)
//   D  a <- D0 gka0 + D1 gka1 + qa TE_i (per slot), with NTT(D0) =
//      sigma_5(a) - 2^26 NTT(D1): a <- sigma_5(a) gka0 + D1 (gka1 - 2^26 gka0) + qa TE_i
// SEC additionally runs the secondary b-recurrence in gathered form,
// c <- sigma_5(c) + D0 gkb0 + D1 gkb1 + TE_i qb (main_pir_loop,
// base_pir.rs:1085-1247), so the digits never leave the chip; PRE stores the
// digits as the shard's IPGK rows.
// NTT passes are radix-4 (inverse) and radix-8 (forward) over registers with
// LDS exchanges; values stay in [0, 2p) between inverse butterflies (Harvey's
// lazy reduction) and grow below 24p through a forward transform, with Shoup
// multiplications by twiddles; twiddles live in LDS. Passes
// whose data stays within one wave order LDS per wave, not per block.
// ---------------------------------------------------------------------------
constexpr uint32_t NINV_Q2INV_P1 = 67055624u;  // N^-1 p2^-1 mod p1
constexpr uint32_t NINV_Q1INV_P2 = 8189u;      // N^-1 p1^-1 mod p2
// their Shoup companions floor(w 2^32 / p)
constexpr uint32_t NINV_Q2INV_P1_SH = (uint32_t)(((uint64_t)NINV_Q2INV_P1 << 32) / P1);
constexpr uint32_t NINV_Q1INV_P2_SH = (uint32_t)(((uint64_t)NINV_Q1INV_P2 << 32) / P2);

template <bool SEC>
__global__ __launch_bounds__(1024) void chain_kernel(const uint32_t* __restrict__ te_tmp, uint32_t* __restrict__ ip_tmp,
                                                     uint32_t* __restrict__ a_out, uint32_t* __restrict__ c_out,
                                                     const uint32_t* __restrict__ qs, const int* __restrict__ lane_slot,
                                                     int L, int nlanes, int qa_base, const DevTables* __restrict__ T) {
    __shared__ __attribute__((aligned(16))) uint2 tw2[2][2][N];  // [fwd/inv][r][k] = (twiddle, Shoup companion)
    __shared__ uint32_t sa[2][N];
    __shared__ uint32_t sc[2][2][N];                // [buffer][r][slot]
    __shared__ uint32_t xb[2][NPAD];  // transform buffers [r], padded index lx<true>
    const int li = blockIdx.x, st = blockIdx.y;
    const int tid = threadIdx.x;
    for (int k = tid; k < N; k += 1024) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            tw2[0][r][k] = make_uint2(T->psi[r][k], T->psiq[r][k]);
            tw2[1][r][k] = make_uint2(T->ipsi[r][k], T->ipsiq[r][k]);
            sa[r][k] = 0;
            sc[0][r][k] = 0;
        }
    }
    // per-thread constants for the update phase: slots tid, tid + 1024, both moduli
    uint32_t cg0[4], cg1[4], cqa[4], cb0[4], cb1[4], cqb[4];
    int sig_upd[2];
    const uint32_t* qa = T->qa[qa_base + st][0];
    const uint32_t* qv = SEC ? qs + (size_t)lane_slot[li] * 6 * 2 * N : nullptr;  // [v][r][slot]
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int s = tid + 1024 * (k >> 1), r = k & 1;
        cg0[k] = T->gka[0][r][s];
        cg1[k] = T->gka[1][r][s];
        cqa[k] = qa[r * N + s];
        if (SEC) {
            cb0[k] = qv[(0 * 2 + r) * N + s];
            cb1[k] = qv[(1 * 2 + r) * N + s];
            cqb[k] = qv[((4 + st) * 2 + r) * N + s];
        }
        // NTT(D0) = sigma_5(a) - 2^26 NTT(D1) (x = D0 + 2^26 D1 and NTT(x mod p) =
        // sigma_5(a)): fold -2^26 into the D1 constants, keep the D0 ones for sigma_5(a)
        const uint32_t p = mod_p(r);
        const uint64_t t26 = (1ull << LOG2_B) % p;
        cg1[k] = (uint32_t)(((uint64_t)cg1[k] + p - t26 * cg0[k] % p) % p);
        if (SEC) cb1[k] = (uint32_t)(((uint64_t)cb1[k] + p - t26 * cb0[k] % p) % p);
    }
    // the update multiplies by these constants every step: Shoup companions
    uint32_t cg0q[4], cg1q[4], cqaq[4], cb0q[4], cb1q[4], cqbq[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t p = mod_p(k & 1);
        cg0q[k] = shoup_companion(cg0[k], p);
        cg1q[k] = shoup_companion(cg1[k], p);
        cqaq[k] = shoup_companion(cqa[k], p);
        if (SEC) {
            cb0q[k] = shoup_companion(cb0[k], p);
            cb1q[k] = shoup_companion(cb1[k], p);
            cqbq[k] = shoup_companion(cqb[k], p);
        }
    }
    sig_upd[0] = T->sigma5[tid];
    sig_upd[1] = T->sigma5[tid + 1024];
    int gpos_upd[2] = {T->slot_gpos[tid], T->slot_gpos[tid + 1024]};

    // phase A: 512 threads per modulus rA, radix-4 group gA first (stages 10, 9)
    const int rA = tid >> 9, gA = tid & 511;
    int sigA[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) sigA[c] = T->sigma5[4 * gA + c];
    const int q4 = tid >> 8, rC = q4 & 1, gC = tid & 255;
    lds_barrier();

    const uint32_t* te_base = te_tmp + ((size_t)(st * nlanes + li) * L) * 2 * N;
    int cur = 0;
#pragma unroll 1
    for (int i = 0; i < L; ++i) {
        // ---- A: inverse NTTs of sigma_5(a) (512 threads per modulus, radix 4) ----
        {
            const uint32_t p = mod_p(rA);
            const uint2* itw2 = tw2[1][rA];
            const int ii = gA;
            uint32_t v0 = sa[rA][sigA[0]], v1 = sa[rA][sigA[1]], v2 = sa[rA][sigA[2]], v3 = sa[rA][sigA[3]];
            const uint4 wb = *reinterpret_cast<const uint4*>(itw2 + 1024 + 2 * ii);
            bf_gs(v0, v1, wb.x, wb.y, p);
            bf_gs(v2, v3, wb.z, wb.w, p);
            const uint2 wa = itw2[512 + ii];
            bf_gs(v0, v2, wa.x, wa.y, p);
            bf_gs(v1, v3, wa.x, wa.y, p);
            uint32_t* b = xb[rA];
            const int b4 = lx<true>(4 * ii);
            b[b4] = v0; b[b4 + 1] = v1; b[b4 + 2] = v2; b[b4 + 3] = v3;
        }
        {
            // the gather and the stride-4..128 passes of a wave stay in its 256
            // elements; the stride-256/512 and stride-1024 passes cross waves
            int gAx = gA;
            asm volatile("" : "+v"(gAx));  // pass addresses recomputed per step (VGPRs)
#pragma unroll
            for (int t = 7; t >= 3; t -= 2) {
                wave_lds_sync();
                inv_pass4gi<true>(xb[rA], tw2[1][rA], mod_p(rA), t, gAx);
            }
            lds_barrier();
            inv_pass4gi<true>(xb[rA], tw2[1][rA], mod_p(rA), 1, gAx);
            lds_barrier();
        }
        // ---- B: the last inverse stage, CRT, centered lift, digit D1
        // (base_pir.rs:544-570) for the coefficient pair (tid, tid + 1024). The
        // stage's butterfly (x, y) -> (x + y, (x - y) psi^-1) and the CRT factors
        // N^-1 (q/p_r)^-1 fold into one Shoup product per coefficient. ----
        uint32_t ecrt[2][2];  // [k: lo, hi][r]
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const uint32_t p = mod_p(r);
            const uint32_t x = xb[r][lx<true>(tid)], y = xb[r][lx<true>(tid + 1024)];  // [0, 2p)
            ecrt[0][r] = mul_shoup(x + y, T->crt_k[0][r], T->crt_kq[0][r], p);
            ecrt[1][r] = mul_shoup(x - y + 2 * p, T->crt_k[1][r], T->crt_kq[1][r], p);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int j = tid + 1024 * k;
            const int xj = lx<true>(j);
            const int32_t d1 = crt_digit1(ecrt[k][0], ecrt[k][1]);
            xb[0][xj] = (uint32_t)(d1 + (int32_t)P1);  // [0, 2p): |d1| <= 2^24 < p
            xb[1][xj] = (uint32_t)(d1 + (int32_t)P2);
        }
        lds_barrier();
        // sigma_5(a) of this step (sa is rewritten in D)
        uint32_t sga[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) sga[k] = sa[k & 1][sig_upd[k >> 1]];
        // TE row of this step (used in phase D; loaded here so phase C hides it)
        uint32_t te[4];
        const uint32_t* te_row = te_base + (size_t)i * 2 * N;
#pragma unroll
        for (int k = 0; k < 4; ++k) te[k] = te_row[(k & 1) * N + tid + 1024 * (k >> 1)];
        // ---- C: forward NTTs of D1 mod p1 and mod p2 (threads 0..511) ----
        {
            const uint32_t p = mod_p(rC);
            int gCx = gC;
            asm volatile("" : "+v"(gCx));  // pass addresses recomputed per step (VGPRs)
            // after the stride-256 pass a wave's passes stay in its 512 elements
            if (tid < 512) fwd_pass8i<true>(xb[rC], tw2[0][rC], p, 0, gCx);
            lds_barrier();
            if (tid < 512) {
                fwd_pass8i<true>(xb[rC], tw2[0][rC], p, 1, gCx);
                wave_lds_sync();
                fwd_pass8i<true>(xb[rC], tw2[0][rC], p, 2, gCx);
                wave_lds_sync();
                fwd_pass4iw<true>(xb[rC], tw2[0][rC], p, gCx);
            }
        }
        lds_barrier();
        // ---- D: update a (and c) per slot ----
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int s = tid + 1024 * (k >> 1), r = k & 1;
            const uint32_t p = mod_p(r);
            const int K = mod_k(r);
            const uint32_t d1 = xb[r][lx<true>(s)];  // NTT(D1), below 24p
            // three Shoup products in [0, 2p) each
            sa[r][s] = red8p(shoup_lazy(sga[k], cg0[k], cg0q[k], p) + shoup_lazy(d1, cg1[k], cg1q[k], p) +
                                 shoup_lazy(te[k], cqa[k], cqaq[k], p),
                             p);
            if (SEC) {
                const uint32_t cv = sc[cur][r][sig_upd[k >> 1]] + shoup_lazy(sga[k], cb0[k], cb0q[k], p) +
                                    shoup_lazy(d1, cb1[k], cb1q[k], p) + shoup_lazy(te[k], cqb[k], cqbq[k], p);
                sc[cur ^ 1][r][s] = red8p(cv, p);  // [0, 7p) -> [0, p)
            } else {
                // IPGK rows and the TE row, Galois order: [arr][r][gpos];
                // NTT(D0) = sigma_5(a) - 2^26 NTT(D1)
                uint32_t* dst = ip_tmp + (((size_t)(st * nlanes + li) * L + i) * 6) * N;
                const int gp = gpos_upd[k >> 1];
                dst[(0 * 2 + r) * N + gp] = te[k];
                const uint32_t e1 = red32(d1, p, r == 0 ? M32_P1 : M32_P2);
                const uint32_t m = mul_mod(e1, (uint32_t)((1ull << LOG2_B) % p), p, K);
                dst[(1 * 2 + r) * N + gp] = sga[k] >= m ? sga[k] - m : sga[k] + p - m;
                dst[(2 * 2 + r) * N + gp] = e1;
            }
        }
        cur ^= 1;
        lds_barrier();
    }
    uint32_t* ao = a_out + ((size_t)(st * nlanes + li) * 2) * N;
    for (int s = tid; s < N; s += 1024) {
        ao[s] = sa[0][s];
        ao[N + s] = sa[1][s];
    }
    if (SEC) {
        uint32_t* co = c_out + ((size_t)(st * nlanes + li) * 2) * N;
        for (int s = tid; s < N; s += 1024) {
            co[s] = sc[cur][0][s];
            co[N + s] = sc[cur][1][s];
        }
    }
}

// Shard rows (Galois order, lane-major temporaries rows[st][lane][i][arr][r][gpos]
// written by the preprocessing chain) -> the slab.
//
// Rotated layout (any L): [r][st][i][arr][gpos][lane], row i rotated by
// L-1-i inside each half; used by k1_kernel.
__global__ __launch_bounds__(256) void relayout_rot_kernel(const uint32_t* __restrict__ rows_tmp, uint32_t* __restrict__ slab,
                                                           int L, int nst) {
    __shared__ uint32_t rows[LANES][N];
    const int i = blockIdx.x, arr = blockIdx.y;
    const int r = blockIdx.z / nst, st = blockIdx.z % nst;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int idx = tid; idx < LANES * N; idx += nt) {
        const int lane = idx / N, g = idx % N;
        rows[lane][g] = rows_tmp[(((size_t)(st * LANES + lane) * L + i) * 6 + arr * 2 + r) * N + g];
    }
    __syncthreads();
    uint32_t* out = slab + ((((size_t)(r * nst + st) * L + i) * 3 + arr) * N) * LANES;
    for (int o = tid; o < N * LANES; o += nt) {
        const int g = o >> 3, lane = o & 7;
        const int hb = g & ~(HALFN - 1);
        out[o] = rows[lane][hb + (((g & (HALFN - 1)) + L - 1 - i) & (HALFN - 1))];
    }
}

// Output-major limb-block layout (L = N/2), read by the k1b/k1d kernels.
// Logically M2[r][half][a][s][j][lane] = row i = (a - j + L - 1) mod N/2 at
// Galois position (half, j): the closed form of main_pir_loop as a plain
// matrix, out[a] = sum_(s,j) M2[a][s][j] Q_s[j] (s = stream*3 + array).
//
// Each residue x in [0, p) is stored as 26 (p1) or 25 (p2) bits: the
// representative x' = x, or x - p when x + 0x808080 >= 2^25, written as
// balanced base-256 digits x' = d0 + 2^8 d1 + 2^16 d2 + 2^24 t with d0..d2 in
// [-128, 127] and t in {-2, -1, 0, 1} mod p1 (p1 < 2^26), t in {0, 1} mod p2.
// Mod p1, t is kept biased as u = t + 2 in 2 bits, so x = d0 + 2^8 d1 +
// 2^16 d2 + 2^24 u - 2^25 (mod p1) and the constant -2^25 sum_k Q[k] of every
// output is added per query at recombination; mod p2, u = t in 1 bit.
//
// Stored as blocks of MB_BLOCK_BYTES_P1 / _P2 bytes (all p1 blocks, then all
// p2 blocks; mb_offset), one per (r, half, a-group g of 4, s, j-chunk c of
// 32): byte v*1024 + l*16 + kk = digit d_v (v < 3) of the value at lane
// l = h*32 + (a%4)*8 + lane8, j = 32c + 16h + kk (exactly the A operand a
// 64-lane wave feeds to v_mfma_i32_32x32x32_i8), then the lane's 16 top
// digits u: mod p1 one word at byte 3072 + 4l, value kk at bit
// 8*(kk%4) + 2*(kk/4), so (word >> 2m) & 0x03030303 is the int8 vector of
// kk = 4m..4m+3; mod p2 one half-word at byte 3072 + 2l, value kk = 4m + b
// at bit m + {0, 8, 4, 12}[b] (top_digits1 unpacks it).
__device__ __forceinline__ uint32_t limbs8(uint32_t x) { return (x + 0x00808080u) ^ 0x00808080u; }

// the representative whose top balanced digit lies in [-2, 1], as its 4
// balanced digits (two's complement bytes)
__device__ __forceinline__ uint32_t limbs8_top2(uint32_t x, uint32_t p) {
    x = x >= p ? x - p : x;
    return limbs8(x + 0x00808080u >= (1u << 25) ? x - p : x);
}

__global__ __launch_bounds__(256) void relayout_mb_kernel(const uint32_t* __restrict__ rows_tmp, uint32_t* __restrict__ slab,
                                                          int nst) {
    // tile of 63 rows x 8 lanes x 32 positions covering a in [a0, a0+32), j in [j0, j0+32)
    __shared__ uint32_t tile[63][LANES][32];
    const int S = 3 * nst;
    const int c = blockIdx.x & 31, at = blockIdx.x >> 5;  // j-chunk, a-tile (32 a's)
    const int half = blockIdx.y & 1, r = blockIdx.y >> 1;
    const int s = blockIdx.z;
    const int st = s / 3, arr = s % 3;
    const int a0 = at * 32, j0 = c * 32;
    const int ibase = (a0 - j0 - 31 + (HALFN - 1)) & (HALFN - 1);  // row of (a0, j0 + 31)
    const int tid = threadIdx.x;
    const uint32_t p = mod_p(r);
    for (int idx = tid; idx < 63 * LANES * 32; idx += 256) {
        const int jl = idx & 31, lane = (idx >> 5) & 7, ri = idx >> 8;
        const int i = (ibase + ri) & (HALFN - 1);
        tile[ri][lane][jl] = rows_tmp[(((size_t)(st * LANES + lane) * HALFN + i) * 6 + arr * 2 + r) * N + half * HALFN + j0 + jl];
    }
    __syncthreads();
    // 8 blocks (a-groups) x 64 lanes; each item writes the 3 digit chunks and
    // the top-digit word of one lane
    for (int it = tid; it < 8 * 64; it += 256) {
        const int gi = it >> 6, l = it & 63;
        const int h = l >> 5, ai = (l >> 3) & 3, lane8 = l & 7;
        const int a = a0 + gi * 4 + ai;
        uint32_t w[16];
        uint32_t top = 0;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const int jl = 16 * h + kk;
            // row i = a - j + L-1  ->  tile row (a - a0) - jl + 31
            w[kk] = limbs8_top2(tile[(a - a0) - jl + 31][lane8][jl], p);
            if (r == 0) top |= (((w[kk] >> 24) + 2u) & 3u) << (8 * (kk & 3) + 2 * (kk >> 2));
            else top |= ((w[kk] >> 24) & 1u) << ((kk >> 2) + (0xC480u >> (4 * (kk & 3)) & 15u));
        }
        const int g = a0 / 4 + gi;
        uint32_t* blk = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(slab) +
                                                    mb_offset(r, (((size_t)half * (HALFN / 4) + g) * S + s) * 32 + c, S));
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            u32x4 o;
            o.x = ((w[0] >> (8 * v)) & 255) | (((w[1] >> (8 * v)) & 255) << 8) | (((w[2] >> (8 * v)) & 255) << 16) | (((w[3] >> (8 * v)) & 255) << 24);
            o.y = ((w[4] >> (8 * v)) & 255) | (((w[5] >> (8 * v)) & 255) << 8) | (((w[6] >> (8 * v)) & 255) << 16) | (((w[7] >> (8 * v)) & 255) << 24);
            o.z = ((w[8] >> (8 * v)) & 255) | (((w[9] >> (8 * v)) & 255) << 8) | (((w[10] >> (8 * v)) & 255) << 16) | (((w[11] >> (8 * v)) & 255) << 24);
            o.w = ((w[12] >> (8 * v)) & 255) | (((w[13] >> (8 * v)) & 255) << 8) | (((w[14] >> (8 * v)) & 255) << 16) | (((w[15] >> (8 * v)) & 255) << 24);
            *reinterpret_cast<u32x4*>(blk + v * 256 + l * 4) = o;
        }
        if (r == 0) blk[768 + l] = top;
        else reinterpret_cast<uint16_t*>(blk + 768)[l] = (uint16_t)top;
    }
}

// Query vectors in Galois order (qp [q][4][r][gpos]) -> B operands of the
// k1b kernels: ql[r][half][s][c][ct][v][l] (16 bytes each): limb v of
// Q_b[s][j], b = 32 ct + l % 32, j = 32c + 16 (l / 32) + kk; zero for b >= nq.
__global__ __launch_bounds__(256) void qlimb_kernel(const uint32_t* __restrict__ qp, uint32_t* __restrict__ ql, int nq, int nst,
                                                    int CT) {
    const int S = 3 * nst;
    const int c = blockIdx.x, sidx = blockIdx.y;  // sidx = (r*2 + half)*S + s
    const int s = sidx % S, rh = sidx / S;
    const int half = rh & 1, r = rh >> 1;
    const int st = s / 3, arr = s % 3;
    const int vq = arr == 0 ? 2 + st : arr - 1;
    for (int it = threadIdx.x; it < CT * 64; it += blockDim.x) {
        const int ct = it >> 6, l = it & 63;
        const int b = 32 * ct + (l & 31), h = l >> 5;
        uint32_t w[16];
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const int j = 32 * c + 16 * h + kk;
            w[kk] = b < nq ? limbs8(qp[(((size_t)b * 4 + vq) * 2 + r) * N + half * HALFN + j]) : 0;
        }
        uint32_t* dst = ql + ((((size_t)sidx * 32 + c) * CT + ct) * 4) * 256;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            u32x4 o;
            o.x = ((w[0] >> (8 * v)) & 255) | (((w[1] >> (8 * v)) & 255) << 8) | (((w[2] >> (8 * v)) & 255) << 16) | (((w[3] >> (8 * v)) & 255) << 24);
            o.y = ((w[4] >> (8 * v)) & 255) | (((w[5] >> (8 * v)) & 255) << 8) | (((w[6] >> (8 * v)) & 255) << 16) | (((w[7] >> (8 * v)) & 255) << 24);
            o.z = ((w[8] >> (8 * v)) & 255) | (((w[9] >> (8 * v)) & 255) << 8) | (((w[10] >> (8 * v)) & 255) << 16) | (((w[11] >> (8 * v)) & 255) << 24);
            o.w = ((w[12] >> (8 * v)) & 255) | (((w[13] >> (8 * v)) & 255) << 8) | (((w[14] >> (8 * v)) & 255) << 16) | (((w[15] >> (8 * v)) & 255) << 24);
            *reinterpret_cast<u32x4*>(dst + v * 256 + l * 4) = o;
        }
    }
}

// The per-query constant of the biased top digits: corr[(r*2+half)][b] =
// -2^25 sum_(s,j) Q_b[s][j] mod p_r over the contraction of (r, half)
// (zero for b >= nq). Written behind the limb tiles in the ql buffer.
__global__ __launch_bounds__(256) void qcorr_kernel(const uint32_t* __restrict__ qp, uint32_t* __restrict__ corr, int nq, int nst,
                                                    int ncols) {
    __shared__ unsigned long long part[256];
    const int rh = blockIdx.x, b = blockIdx.y;
    const int half = rh & 1, r = rh >> 1;
    const int tid = threadIdx.x;
    unsigned long long sum = 0;
    if (b < nq) {
        for (int s = 0; s < 3 * nst; ++s) {
            const int st = s / 3, arr = s % 3;
            const int vq = arr == 0 ? 2 + st : arr - 1;
            const uint32_t* v = qp + (((size_t)b * 4 + vq) * 2 + r) * N + half * HALFN;
            for (int j = tid; j < HALFN; j += 256) sum += v[j];
        }
    }
    part[tid] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) part[tid] += part[tid + w];
        __syncthreads();
    }
    if (tid == 0) {  // the p2 top digits are unbiased
        const uint64_t p = mod_p(r);
        const uint64_t m = r == 0 ? ((1ull << 25) % p) * (part[0] % p) % p : 0;
        corr[(size_t)rh * ncols + b] = (uint32_t)((p - m) % p);
    }
}

// ---------------------------------------------------------------------------
// K1 (single, split and batched queries) on the limb-block layout: for each
// (shard, r, half) the product OUT[(a, lane)][b] = sum_(s,j) M2[(a,lane)][(s,j)] Q_b[(s,j)]
// on int8 MFMA with exact limb decomposition: the shard operand is split into
// 3 signed base-256 digits and the top digit u (biased, in [0, 3], mod p1; 0
// or 1 mod p2), the query into 4 signed digits; the 16 digit products are summed in int32 per weight
// w = u + v (|sum| <= 4 * 6144 * 2^14 < 2^29), and
// OUT = sum_w 2^(8w) ACC_w + corr_b mod p. One wave per a-group (32 output
// rows), CT column tiles of 32 queries. Writes outb[q][shard][r][lane][gpos]
// (lane-major, so the epilogue of one (query, shard, lane) reads 8 KiB rows).
// ---------------------------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// The shard operand through global-address-space pointers: the loads are
// global (not flat) loads, so waiting for them never waits for LDS traffic.
typedef __attribute__((address_space(1))) const v4i gv4i;
typedef __attribute__((address_space(1))) const uint32_t gu32;

typedef __attribute__((address_space(1))) const uint16_t gu16;

// one K-step of the shard operand for one lane: digit planes 0..2 and the
// packed top digits
struct ATile {
    v4i d[3];
    uint32_t top;
};

// the lane's blocks of one (r, half, a-group) row: first block's address,
// block size in 16-byte units
struct ARow {
    const char* base;
    int bsz16;
};
__device__ __forceinline__ ARow a_row(const uint32_t* slab, int r, int half, int g, int S) {
    ARow a;
    a.base = reinterpret_cast<const char*>(slab) + mb_offset(r, (((size_t)half * (HALFN / 4) + g) * S) * 32, S);
    a.bsz16 = (r == 0 ? MB_BLOCK_BYTES_P1 : MB_BLOCK_BYTES_P2) / 16;
    return a;
}
__device__ __forceinline__ void load_a(ATile& t, const ARow& a, int r, int l, int k) {
    gv4i* A = (gv4i*)a.base + (size_t)k * a.bsz16 + l;
    t.d[0] = __builtin_nontemporal_load(A);
    t.d[1] = __builtin_nontemporal_load(A + 64);
    t.d[2] = __builtin_nontemporal_load(A + 128);
    if (r == 0) t.top = __builtin_nontemporal_load((gu32*)(A - l + 192) + l);
    else t.top = __builtin_nontemporal_load((gu16*)(A - l + 192) + l);
}

// the four int8 vectors of a lane's top digits (2 bits each, mod p1)
__device__ __forceinline__ v4i top_digits(uint32_t w) {
    v4i o;
    o.x = (int)(w & 0x03030303u);
    o.y = (int)((w >> 2) & 0x03030303u);
    o.z = (int)((w >> 4) & 0x03030303u);
    o.w = (int)((w >> 6) & 0x03030303u);
    return o;
}
// the same for 1-bit top digits (mod p2): value 4m + b at bit m + {0, 8, 4, 12}[b]
__device__ __forceinline__ int top1_vec(uint32_t w, int m) {
    const uint32_t t = (w >> m) & 0x1111u;
    return (int)((t | (t << 12)) & 0x01010101u);
}
__device__ __forceinline__ v4i top_digits1(uint32_t w) {
    v4i o;
    o.x = top1_vec(w, 0);
    o.y = top1_vec(w, 1);
    o.z = top1_vec(w, 2);
    o.w = top1_vec(w, 3);
    return o;
}

// 2^(8w) mod p for the 7 digit weights
__device__ __forceinline__ void weight_consts(uint32_t p, uint32_t cw[7]) {
    uint64_t x = 1;
#pragma unroll
    for (int w = 0; w < 7; ++w) {
        cw[w] = (uint32_t)(x % p);
        x = (x % p) * 256;
    }
}

// the per-query corrections behind the limb tiles of a ql buffer built for
// CT column tiles and S = 3*nst arrays
__device__ __forceinline__ const uint32_t* qcorr_ptr(const void* ql, int S, int CT) {
    return reinterpret_cast<const uint32_t*>(ql) + (size_t)4 * S * 32 * CT * 1024;
}

// k1b epilogue: the seven weight accumulators of each of the lane's 16 rows
// per column tile combined with the digit weights and the query's top-digit
// correction, reduced mod p and stored as the 32-bit per-shard result.
template <int CT>
__device__ __forceinline__ void k1b_store(const v16i (&acc)[7][CT], const void* ql, uint32_t* __restrict__ outb, int S, int r,
                                          int half, int sh, int g, int l, int nq, size_t q_stride) {
    const uint32_t p = mod_p(r);
    const int K = mod_k(r);
    uint32_t cw[7];
    weight_consts(p, cw);
    const uint32_t* corr = qcorr_ptr(ql, S, CT) + (size_t)(r * 2 + half) * CT * 32;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const int q = ct * 32 + (l & 31);
        if (q >= nq) continue;
        uint32_t* o = outb + (size_t)q * q_stride + ((size_t)sh * 2 + r) * N * LANES;
        const int64_t cq = corr[q];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            int64_t t = cq;
#pragma unroll
            for (int w = 0; w < 7; ++w) t += (int64_t)acc[w][ct][e] * cw[w];
            const int row = 8 * (e >> 2) + 4 * (l >> 5) + (e & 3);
            const int a = g * 4 + (row >> 3), lane8 = row & 7;
            o[(size_t)lane8 * N + half * HALFN + a] = from_i64(t, p, K);
        }
    }
}

// k1b: one wave per a-group, CT column tiles per wave. The shard tiles are
// prefetched PD K-steps ahead into a register ring; the query tiles are
// loaded PD steps ahead into a register ring too and written to a 2-stage LDS
// buffer one step before use, so waiting for a query tile never waits for a
// younger shard prefetch (vector memory counters retire in order). Loads are
// unconditional (clamped index). U steps are unrolled per loop iteration
// (U a multiple of PD; a long body keeps the compiler's wait counts exact).
template <int CT, int PD, int U>
__global__ __launch_bounds__(256) void k1b_kernel(const uint32_t* const* __restrict__ slabs, const v4i* __restrict__ ql,
                                                  uint32_t* __restrict__ outb, int nst, int nq, size_t q_stride) {
    static_assert(U % PD == 0, "PD divides U");
    __shared__ v4i Bt[2][CT][4][64];
    const int S = 3 * nst;
    const int half = blockIdx.y;
    const int r = blockIdx.z & 1, sh = blockIdx.z >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const int g = blockIdx.x * 4 + wave;
    const ARow arow = a_row(slabs[sh], r, half, g, S);
    gv4i* Q = (gv4i*)ql + ((size_t)(r * 2 + half) * S) * 32 * CT * 256 + tid;
    const int nk = S * 32;
    v16i acc[7][CT];
#pragma unroll
    for (int w = 0; w < 7; ++w)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[w][ct][e] = 0;
    ATile abuf[PD];
    v4i bq[PD][CT];
    // step 0's query tile straight to LDS; steps 1..PD in the register ring
#pragma unroll
    for (int i = 0; i < CT; ++i) (&Bt[0][0][0][0])[tid + 256 * i] = Q[256 * i];
#pragma unroll
    for (int d = 0; d < PD; ++d) {
        load_a(abuf[d], arow, r, l, d);
        const int kb = min(d + 1, nk - 1);
#pragma unroll
        for (int i = 0; i < CT; ++i) bq[d][i] = Q[(size_t)kb * CT * 256 + 256 * i];
    }
    for (int kk0 = 0; kk0 < nk; kk0 += U) {
#pragma unroll
        for (int dd = 0; dd < U; ++dd) {
            const int d = dd % PD;
            const int kk = kk0 + dd;
            __syncthreads();  // tile kk visible; tile kk-1's buffer free
            // tile kk+1 (loaded PD steps ago) -> LDS
#pragma unroll
            for (int i = 0; i < CT; ++i) (&Bt[(kk + 1) & 1][0][0][0])[tid + 256 * i] = bq[d][i];
            v4i a[4];
            a[0] = abuf[d].d[0];
            a[1] = abuf[d].d[1];
            a[2] = abuf[d].d[2];
            a[3] = r == 0 ? top_digits(abuf[d].top) : top_digits1(abuf[d].top);
            v4i b[CT][4];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v) b[ct][v] = Bt[kk & 1][ct][v][l];
            const int ka = min(kk + PD, nk - 1), kb = min(kk + PD + 1, nk - 1);
            load_a(abuf[d], arow, r, l, ka);
#pragma unroll
            for (int i = 0; i < CT; ++i) bq[d][i] = Q[(size_t)kb * CT * 256 + 256 * i];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        acc[u + v][ct] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[u], b[ct][v], acc[u + v][ct], 0, 0, 0);
        }
    }
    k1b_store<CT>(acc, ql, outb, S, r, half, sh, g, l, nq, q_stride);
}

// 3-slot LDS ring: the query tile of step k+1 is read from LDS into a second
// register set before step k's MFMAs issue, so the matrix pipe never waits on
// an LDS read after the step's barrier. At step k: slot (k+1)%3 was written
// before the previous barrier; slot (k+2)%3, which receives tile k+2 from the
// register ring, held tile k-1, read at step k-2, and the barrier after that
// step (each barrier waits for the wave's LDS reads) separates the read from
// this write. U is a multiple of 2, 3 and PD, so every slot index and register
// set is a compile-time constant (the K extent 96 * nst is a multiple of U).
template <int CT, int PD, int U>
__global__ __launch_bounds__(256) void k1b_ring_kernel(const uint32_t* const* __restrict__ slabs, const v4i* __restrict__ ql,
                                                       uint32_t* __restrict__ outb, int nst, int nq, size_t q_stride) {
    static_assert(U % PD == 0 && U % 6 == 0, "U a multiple of PD, 2 and 3");
    __shared__ v4i Bt[3][CT][4][64];
    const int S = 3 * nst;
    const int half = blockIdx.y;
    const int r = blockIdx.z & 1, sh = blockIdx.z >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const int g = blockIdx.x * 4 + wave;
    const ARow arow = a_row(slabs[sh], r, half, g, S);
    gv4i* Q = (gv4i*)ql + ((size_t)(r * 2 + half) * S) * 32 * CT * 256 + tid;
    const int nk = S * 32;
    v16i acc[7][CT];
#pragma unroll
    for (int w = 0; w < 7; ++w)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[w][ct][e] = 0;
    ATile abuf[PD];
    v4i bq[PD][CT];
    // tiles 0 and 1 straight to LDS; tiles 2..PD+1 in the register ring
#pragma unroll
    for (int i = 0; i < CT; ++i) {
        (&Bt[0][0][0][0])[tid + 256 * i] = Q[256 * i];
        (&Bt[1][0][0][0])[tid + 256 * i] = Q[(size_t)CT * 256 + 256 * i];
    }
#pragma unroll
    for (int d = 0; d < PD; ++d) {
        load_a(abuf[d], arow, r, l, d);
        const int kb = min(d + 2, nk - 1);
#pragma unroll
        for (int i = 0; i < CT; ++i) bq[d][i] = Q[(size_t)kb * CT * 256 + 256 * i];
    }
    __syncthreads();
    v4i b[2][CT][4];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) b[0][ct][v] = Bt[0][ct][v][l];
    for (int kk0 = 0; kk0 < nk; kk0 += U) {
#pragma unroll
        for (int dd = 0; dd < U; ++dd) {
            const int d = dd % PD, cur = dd & 1, nxt = cur ^ 1;
            const int kk = kk0 + dd;
            // tile kk+1 -> registers (slot (dd+1)%3), tile kk+2 -> slot (dd+2)%3
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v) b[nxt][ct][v] = Bt[(dd + 1) % 3][ct][v][l];
#pragma unroll
            for (int i = 0; i < CT; ++i) (&Bt[(dd + 2) % 3][0][0][0])[tid + 256 * i] = bq[d][i];
            v4i a[4];
            a[0] = abuf[d].d[0];
            a[1] = abuf[d].d[1];
            a[2] = abuf[d].d[2];
            a[3] = r == 0 ? top_digits(abuf[d].top) : top_digits1(abuf[d].top);
            const int ka = min(kk + PD, nk - 1), kb = min(kk + PD + 2, nk - 1);
            load_a(abuf[d], arow, r, l, ka);
#pragma unroll
            for (int i = 0; i < CT; ++i) bq[d][i] = Q[(size_t)kb * CT * 256 + 256 * i];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        acc[u + v][ct] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[u], b[cur][ct][v], acc[u + v][ct], 0, 0, 0);
            __syncthreads();
        }
    }
    k1b_store<CT>(acc, ql, outb, S, r, half, sh, g, l, nq, q_stride);
}

// No LDS and no barriers: every wave loads the query tiles itself, PB steps
// ahead in a register ring (the four waves of a block read the same tile, so
// three of the four reads hit the CU's vector cache). The waves of a block
// run free of each other.
template <int CT, int PD, int PB, int U>
__global__ __launch_bounds__(256) void k1b_direct_kernel(const uint32_t* const* __restrict__ slabs, const v4i* __restrict__ ql,
                                                         uint32_t* __restrict__ outb, int nst, int nq, size_t q_stride) {
    static_assert(U % PD == 0 && U % PB == 0, "U a multiple of PD and PB");
    const int S = 3 * nst;
    const int half = blockIdx.y;
    const int r = blockIdx.z & 1, sh = blockIdx.z >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const int g = blockIdx.x * 4 + wave;
    const ARow arow = a_row(slabs[sh], r, half, g, S);
    gv4i* Q = (gv4i*)ql + ((size_t)(r * 2 + half) * S) * 32 * CT * 256 + l;
    const int nk = S * 32;
    v16i acc[7][CT];
#pragma unroll
    for (int w = 0; w < 7; ++w)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[w][ct][e] = 0;
    ATile abuf[PD];
    v4i bq[PB][CT][4];
#pragma unroll
    for (int d = 0; d < PD; ++d) load_a(abuf[d], arow, r, l, d);
#pragma unroll
    for (int d = 0; d < PB; ++d)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) bq[d][ct][v] = Q[(size_t)d * CT * 256 + (ct * 4 + v) * 64];
    for (int kk0 = 0; kk0 < nk; kk0 += U) {
#pragma unroll
        for (int dd = 0; dd < U; ++dd) {
            const int da = dd % PD, db = dd % PB;
            const int kk = kk0 + dd;
            v4i a[4];
            a[0] = abuf[da].d[0];
            a[1] = abuf[da].d[1];
            a[2] = abuf[da].d[2];
            a[3] = r == 0 ? top_digits(abuf[da].top) : top_digits1(abuf[da].top);
            load_a(abuf[da], arow, r, l, min(kk + PD, nk - 1));
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        acc[u + v][ct] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[u], bq[db][ct][v], acc[u + v][ct], 0, 0, 0);
            const int kb = min(kk + PB, nk - 1);
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v) bq[db][ct][v] = Q[(size_t)kb * CT * 256 + (ct * 4 + v) * 64];
        }
    }
    k1b_store<CT>(acc, ql, outb, S, r, half, sh, g, l, nq, q_stride);
}

// Diagnostic: the batched inner product's MFMA instruction mix with the
// operands held in registers (no memory traffic): per wave 7 x 2 accumulator
// tiles, per K-step 32 v_mfma_i32_32x32x32_i8 in k1b's (u, v, column tile)
// order, 96 K-steps, one 256-thread block per (a-group quad, half, shard,
// modulus) of a 32-shard half-size fleet. ZERO = 0 draws every operand digit
// from a hash and changes it each step (what shards and queries look like to
// the matrix pipe); ZERO = 1 keeps every digit zero. The clock this GPU holds
// depends on how much the operand bits switch, so the two bound what the k1b
// kernels can reach on real data.
template <int ZERO>
__global__ __launch_bounds__(256) void mfma_rate_kernel(int* __restrict__ out, int nk, uint32_t seed) {
    uint32_t h = (blockIdx.x * 256u + threadIdx.x) * 2654435761u ^ seed;
    v4i a[4], b[2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        h = h * 1664525u + 1013904223u;
        a[i] = ZERO ? (v4i){0, 0, 0, 0} : (v4i){(int)h, (int)(h * 3u), (int)(h * 5u), (int)(h * 7u)};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            h = h * 1664525u + 1013904223u;
            b[c][i] = ZERO ? (v4i){0, 0, 0, 0} : (v4i){(int)h, (int)(h * 11u), (int)(h * 13u), (int)(h * 17u)};
        }
    }
    v16i acc[7][2];
#pragma unroll
    for (int w = 0; w < 7; ++w)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[w][c][e] = 0;
    for (int k = 0; k < nk; ++k) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    acc[u + v][ct] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[u], b[ct][v], acc[u + v][ct], 0, 0, 0);
        if (!ZERO) {
            const v4i kv = {k, k, k, k};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] ^= kv;
                b[0][i] += kv;
                b[1][i] ^= kv;
            }
        }
    }
    int sum = 0;
#pragma unroll
    for (int w = 0; w < 7; ++w)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int e = 0; e < 16; ++e) sum += acc[w][c][e];
    out[blockIdx.x * 256 + threadIdx.x] = sum;
}

// Diagnostic: a plain streaming read of device ranges (16-byte nontemporal
// loads, as the K1 kernels issue them), for the HBM rate the inner products
// are compared with. Block (x, y) reads chunk x of range y; each thread folds
// what it reads into one word so the loads are kept.
__global__ __launch_bounds__(256) void read_rate_kernel(const char* const* __restrict__ ptrs, const size_t* __restrict__ bytes,
                                                        uint32_t* __restrict__ out) {
    const size_t n16 = bytes[blockIdx.y] / 16;
    const gv4i* src = (const gv4i*)ptrs[blockIdx.y];
    const size_t per = (n16 + gridDim.x - 1) / gridDim.x;
    const size_t b = (size_t)blockIdx.x * per, e = min(n16, b + per);
    v4i acc = {0, 0, 0, 0};
    size_t i = b + threadIdx.x;
    for (; i + 3 * 256 < e; i += 4 * 256) {
        const v4i x0 = __builtin_nontemporal_load(src + i), x1 = __builtin_nontemporal_load(src + i + 256);
        const v4i x2 = __builtin_nontemporal_load(src + i + 512), x3 = __builtin_nontemporal_load(src + i + 768);
        acc ^= x0 ^ x1 ^ x2 ^ x3;
    }
    for (; i < e; i += 256) acc ^= __builtin_nontemporal_load(src + i);
    out[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 256 + threadIdx.x] = (uint32_t)(acc.x ^ acc.y ^ acc.z ^ acc.w);
}
constexpr int READ_RATE_CHUNKS = 512;
size_t read_rate_out_words(int nranges) { return (size_t)nranges * READ_RATE_CHUNKS * 256; }
void launch_read_rate(const char* const* ptrs, const size_t* bytes, int nranges, uint32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(read_rate_kernel, dim3(READ_RATE_CHUNKS, nranges), dim3(256), 0, s, ptrs, bytes, out);
}

constexpr int MFMA_RATE_BLOCKS = 8192, MFMA_RATE_STEPS = 96;
size_t mfma_rate_out_words() { return (size_t)MFMA_RATE_BLOCKS * 256; }
double mfma_rate_ops() { return (double)MFMA_RATE_BLOCKS * 4 * MFMA_RATE_STEPS * 32 * 65536; }
void launch_mfma_rate(int zero, int* out, uint32_t seed, hipStream_t s) {
    if (zero) hipLaunchKernelGGL(mfma_rate_kernel<1>, dim3(MFMA_RATE_BLOCKS), dim3(256), 0, s, out, MFMA_RATE_STEPS, seed);
    else hipLaunchKernelGGL(mfma_rate_kernel<0>, dim3(MFMA_RATE_BLOCKS), dim3(256), 0, s, out, MFMA_RATE_STEPS, seed);
}

// Candidates for the batched inner product, per column-tile count. All give
// identical results (exact integer arithmetic); the autotuner times them.
typedef void (*K1bFn)(const uint32_t* const*, const v4i*, uint32_t*, int, int, size_t);
struct K1bVariant {
    const char* name;
    int ct;
    K1bFn fn;
    int threads = 256;
};
static const K1bVariant k_k1b_variants[] = {
    {"prefetch 2", 1, k1b_kernel<1, 2, 2>},
    {"prefetch 3", 1, k1b_kernel<1, 3, 3>},
    {"prefetch 4", 1, k1b_kernel<1, 4, 4>},
    {"prefetch 4, unroll 32", 1, k1b_kernel<1, 4, 32>},
    {"prefetch 2", 2, k1b_kernel<2, 2, 2>},
    {"prefetch 3", 2, k1b_kernel<2, 3, 3>},
    {"prefetch 4", 2, k1b_kernel<2, 4, 4>},
    {"prefetch 4, unroll 8", 2, k1b_kernel<2, 4, 8>},
    {"prefetch 4, unroll 32", 2, k1b_kernel<2, 4, 32>},
    {"LDS ring 3, prefetch 2", 1, k1b_ring_kernel<1, 2, 6>},
    {"LDS ring 3, prefetch 3", 1, k1b_ring_kernel<1, 3, 6>},
    {"LDS ring 3, prefetch 2", 2, k1b_ring_kernel<2, 2, 6>},
    {"LDS ring 3, prefetch 3", 2, k1b_ring_kernel<2, 3, 6>},
    {"no LDS, prefetch 3/2", 1, k1b_direct_kernel<1, 3, 2, 6>},
    {"no LDS, prefetch 3/3", 1, k1b_direct_kernel<1, 3, 3, 6>},
    {"no LDS, prefetch 2/2", 2, k1b_direct_kernel<2, 2, 2, 2>},
    {"no LDS, prefetch 3/2", 2, k1b_direct_kernel<2, 3, 2, 6>},
    {"no LDS, prefetch 4/3", 2, k1b_direct_kernel<2, 4, 3, 12>},
};
constexpr int K1B_NVAR = sizeof(k_k1b_variants) / sizeof(k_k1b_variants[0]);
// chosen variant per (column tiles, streams); -1 = default (register ring, prefetch 4)
static std::atomic<int> g_k1b_choice[3][3] = {{{-1}, {-1}, {-1}}, {{-1}, {-1}, {-1}}, {{-1}, {-1}, {-1}}};

int k1b_variant_count() { return K1B_NVAR; }
const char* k1b_variant_name(int i) { return i >= 0 && i < K1B_NVAR ? k_k1b_variants[i].name : ""; }
int k1b_variant_ct(int i) { return i >= 0 && i < K1B_NVAR ? k_k1b_variants[i].ct : 0; }
void k1b_set_choice(int ct, int nst, int i) {
    if (ct >= 1 && ct <= 2 && nst >= 1 && nst <= 2) g_k1b_choice[ct][nst] = i;
}
int k1b_get_choice(int ct, int nst) {
    int i = (ct >= 1 && ct <= 2 && nst >= 1 && nst <= 2) ? g_k1b_choice[ct][nst].load() : -1;
    if (i >= 0 && i < K1B_NVAR && k_k1b_variants[i].ct == ct) return i;
    for (int k = 0; k < K1B_NVAR; ++k)
        if (k_k1b_variants[k].ct == ct && k_k1b_variants[k].fn == (ct == 1 ? (K1bFn)k1b_kernel<1, 4, 4> : (K1bFn)k1b_kernel<2, 4, 4>)) return k;
    return 0;
}

// sum of the chains' a (and c) over the streams -> inverse NTT -> CRT ->
// coefficients mod q and their 17-bit (a) / 12-bit (b) mod switch.
// One block per lane index and ring element (a, c). a_out/c_out [st][li][r][slot].
__global__ __launch_bounds__(512) void chain_out_kernel(const uint32_t* __restrict__ a_out, const uint32_t* __restrict__ c_out,
                                                       int nst, int nlanes, uint64_t* __restrict__ a_coef,
                                                       uint64_t* __restrict__ b_coef, uint32_t* __restrict__ a17,
                                                       uint32_t* __restrict__ b12, const unsigned* __restrict__ flag_src,
                                                       unsigned* flag_dst, const DevTables* __restrict__ T) {
    __shared__ uint32_t w[2][NPAD];
    const int li = blockIdx.x, part = blockIdx.y;  // part 0: a, 1: c (the b of the reply)
    const int tid = threadIdx.x;
    // the chains' timeout flag to pinned host memory (read after the call's sync)
    if (flag_dst && li == 0 && part == 0 && tid == 0)
        __hip_atomic_store(flag_dst, *flag_src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint32_t* src = part == 0 ? a_out : c_out;
    if (!src) return;
    for (int s = tid; s < N; s += 512) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const uint32_t p = mod_p(r);
            uint32_t v = 0;
            for (int st = 0; st < nst; ++st) v = add_mod(v, src[((size_t)(st * nlanes + li) * 2 + r) * N + s], p);
            w[r][lx<true>(s)] = v;
        }
    }
    __syncthreads();
    {
        // one modulus per 256 threads; N^-1 is folded into the CRT below
        const int r = tid >> 8;
        ntt256_inv_unscaled<true>(w[r], T->ipsi[r], T->ipsiq[r], mod_p(r), tid & 255);
    }
    for (int s = tid; s < N; s += 512) {
        const int x = lx<true>(s);
        const uint32_t e1 = mul_shoup(w[0][x], NINV_Q2INV_P1, NINV_Q2INV_P1_SH, P1);
        const uint32_t e2 = mul_shoup(w[1][x], NINV_Q1INV_P2, NINV_Q1INV_P2_SH, P2);
        uint64_t v = (uint64_t)e1 * P2 + (uint64_t)e2 * P1;
        v = v >= Q ? v - Q : v;
        if (part == 0) {
            if (a_coef) a_coef[(size_t)li * N + s] = v;
            if (a17) a17[(size_t)li * N + s] = mod_switch_enc(v, 17);
        } else {
            if (b_coef) b_coef[(size_t)li * N + s] = v;
            if (b12) b12[(size_t)li * N + s] = mod_switch_enc(v, 12);
        }
    }
}

// query ring elements (coefficients mod q) -> NTT vectors mod p1 and p2.
// qcoef [nq][6][N] (gk0, gk1, primary, primary conj, secondary, secondary conj)
// qs [nq][6][2][slot]; qp [nq][4][2][gpos]
__global__ __launch_bounds__(256) void qprep_kernel(const uint64_t* __restrict__ qcoef, uint32_t* __restrict__ qs,
                                                    uint32_t* __restrict__ qp, const DevTables* __restrict__ T) {
    __shared__ uint32_t a[N];
    const int v = blockIdx.x, r = blockIdx.y, q = blockIdx.z;
    const int tid = threadIdx.x, nt = blockDim.x;
    const uint32_t p = mod_p(r);
    const int K = mod_k(r);
    const uint64_t* src = qcoef + ((size_t)q * 6 + v) * N;
    for (int s = tid; s < N; s += nt) a[s] = red64(src[s], p, K);
    __syncthreads();
    ntt_fwd(a, T, r, tid, nt);
    uint32_t* o = qs + (((size_t)q * 6 + v) * 2 + r) * N;
    for (int s = tid; s < N; s += nt) o[s] = a[s];
    if (v < 4) {
        uint32_t* g = qp + (((size_t)q * 4 + v) * 2 + r) * N;
        for (int s = tid; s < N; s += nt) g[s] = a[T->gpos_slot[s]];
    }
}

// Device cache of per-client Galois keys. Query q's key rows are rows 0-1
// (gk0, gk1, both moduli) of its qs and qp blocks, 4N words each. An entry
// holds [qs rows 0-1 | qp rows 0-1]. src[q] set: the entry replaces the rows
// decoded from the query bytes; else dst[q] set: the rows are stored into it.
__global__ __launch_bounds__(256) void key_cache_kernel(const uint32_t* const* __restrict__ src, uint32_t* const* __restrict__ dst,
                                                        uint32_t* __restrict__ qs, uint32_t* __restrict__ qp) {
    const int q = blockIdx.y, part = blockIdx.z;
    u32x4* rows = reinterpret_cast<u32x4*>(part == 0 ? qs + (size_t)q * 6 * 2 * N : qp + (size_t)q * 4 * 2 * N);
    const int i = blockIdx.x * 256 + threadIdx.x;  // one of N 16-byte words
    const u32x4* s = reinterpret_cast<const u32x4*>(src[q]);
    u32x4* d = reinterpret_cast<u32x4*>(dst[q]);
    if (s) rows[i] = s[part * N + i];
    else if (d) d[part * N + i] = rows[i];
}

// ---------------------------------------------------------------------------
// K1: the primary inner product (main_pir_loop, base_pir.rs:1085-1247) in
// closed form  c[g] = sum_i sum_arr DBrot_arr[i][g] * Q_arr[rot_i(g)].
// Block: 128 Galois positions x 2 lane halves, a chunk of rows of one
// (shard, r, stream). Partial sums are reduced mod p and added atomically
// into acc[shard][r][gpos][lane] (u64, at most 2 * 1024 partials < 2^37).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k1_kernel(const uint32_t* const* __restrict__ slabs, const uint32_t* __restrict__ qp,
                                                 unsigned long long* __restrict__ acc, int L, int nst, int rows_per_block) {
    __shared__ uint32_t qv[3][N];
    const int z = blockIdx.z;
    const int st = z % nst;
    const int r = (z / nst) & 1;
    const int sh = z / (2 * nst);
    const int tid = threadIdx.x;
    for (int s = tid; s < N; s += 256) {
        qv[0][s] = qp[((2 + st) * 2 + r) * N + s];
        qv[1][s] = qp[(0 * 2 + r) * N + s];
        qv[2][s] = qp[(1 * 2 + r) * N + s];
    }
    __syncthreads();
    const int g = blockIdx.x * 128 + (tid >> 1);
    const int half = tid & 1;
    const int hb = g & ~(HALFN - 1);
    const int ga = g & (HALFN - 1);
    const uint32_t* base = slabs[sh] + ((size_t)(r * nst + st) * L) * 3 * N * LANES;
    const int i0 = blockIdx.y * rows_per_block;
    const int i1 = min(L, i0 + rows_per_block);
    unsigned long long a0 = 0, a1 = 0, a2 = 0, a3 = 0;
#pragma unroll 4
    for (int i = i0; i < i1; ++i) {
        const int src = hb + ((ga + L - 1 - i) & (HALFN - 1));
        const uint32_t qt = qv[0][src], g0 = qv[1][src], g1 = qv[2][src];
        const u32x4* rowp = reinterpret_cast<const u32x4*>(base + (size_t)i * 3 * N * LANES);
        const u32x4 te = __builtin_nontemporal_load(rowp + g * 2 + half);
        const u32x4 x0 = __builtin_nontemporal_load(rowp + (N + g) * 2 + half);
        const u32x4 x1 = __builtin_nontemporal_load(rowp + (2 * N + g) * 2 + half);
        a0 += (unsigned long long)te.x * qt + (unsigned long long)x0.x * g0 + (unsigned long long)x1.x * g1;
        a1 += (unsigned long long)te.y * qt + (unsigned long long)x0.y * g0 + (unsigned long long)x1.y * g1;
        a2 += (unsigned long long)te.z * qt + (unsigned long long)x0.z * g0 + (unsigned long long)x1.z * g1;
        a3 += (unsigned long long)te.w * qt + (unsigned long long)x0.w * g0 + (unsigned long long)x1.w * g1;
    }
    const uint32_t p = mod_p(r);
    const int K = mod_k(r);
    unsigned long long* o = acc + (((size_t)sh * 2 + r) * N + g) * LANES + half * 4;
    atomicAdd(o + 0, (unsigned long long)red64(a0, p, K));
    atomicAdd(o + 1, (unsigned long long)red64(a1, p, K));
    atomicAdd(o + 2, (unsigned long long)red64(a2, p, K));
    atomicAdd(o + 3, (unsigned long long)red64(a3, p, K));
}

// Per-lane query vectors (perform_pir_split): same as k1_kernel but Q is read
// per lane from global memory: qpl [lane][v][r][gpos].
__global__ __launch_bounds__(256) void k1_split_kernel(const uint32_t* const* __restrict__ slabs, const uint32_t* __restrict__ qpl,
                                                       unsigned long long* __restrict__ acc, int L, int nst, int rows_per_block) {
    const int z = blockIdx.z;
    const int st = z % nst;
    const int r = (z / nst) & 1;
    const int sh = z / (2 * nst);
    const int tid = threadIdx.x;
    const int g = blockIdx.x * 128 + (tid >> 1);
    const int half = tid & 1;
    const int hb = g & ~(HALFN - 1);
    const int ga = g & (HALFN - 1);
    const uint32_t* base = slabs[sh] + ((size_t)(r * nst + st) * L) * 3 * N * LANES;
    const int i0 = blockIdx.y * rows_per_block;
    const int i1 = min(L, i0 + rows_per_block);
    unsigned long long a[4] = {0, 0, 0, 0};
    for (int i = i0; i < i1; ++i) {
        const int src = hb + ((ga + L - 1 - i) & (HALFN - 1));
        const uint32_t* rowp = base + (size_t)i * 3 * N * LANES;
        for (int k = 0; k < 4; ++k) {
            const int lane = half * 4 + k;
            const uint32_t* ql = qpl + (size_t)lane * 4 * 2 * N;
            a[k] += (unsigned long long)rowp[g * LANES + lane] * ql[((2 + st) * 2 + r) * N + src] +
                    (unsigned long long)rowp[(N + g) * LANES + lane] * ql[(0 * 2 + r) * N + src] +
                    (unsigned long long)rowp[(2 * N + g) * LANES + lane] * ql[(1 * 2 + r) * N + src];
        }
    }
    const uint32_t p = mod_p(r);
    const int K = mod_k(r);
    unsigned long long* o = acc + (((size_t)sh * 2 + r) * N + g) * LANES + half * 4;
    for (int k = 0; k < 4; ++k) atomicAdd(o + k, (unsigned long long)red64(a[k], p, K));
}

// primary epilogue per (query, shard, lane): reduce, Galois order -> slot
// order, inverse NTT, CRT (from_simd_element, base_pir.rs:905-935); then
// either the b coefficients (perform_pir) or, with the shard's 17-bit a-part,
// the three base-1025 digits of a*2^12 + b written as secondary entries
// (double_pir.rs:157-177, 70-81): column shard*8 + lane, entry lane
// lane0_mult*q + part, at se + q*se_qstride + (col*lpr + lane)*N.
struct EpiArgs {
    const uint32_t* outb;             // [q][shard][r][lane][gpos] reduced (k1b, k1d, k1w), or null
    const unsigned long long* acc;    // [shard][r][gpos][lane] partial sums (k1), or null
    size_t q_stride;                  // elements per query in outb
    const uint32_t* const* a17s;      // per shard: [lane][N]
    uint16_t* se;                     // secondary entries, or null
    size_t se_qstride;
    int lpr, lane0_mult;
    uint64_t* bcoef;                  // [shard][lane][N] (query index = lane when diag), or null
    int diag;                         // query q = lane (perform_pir_split)
};
__global__ __launch_bounds__(512) void epi_primary_kernel(EpiArgs A, const DevTables* __restrict__ T) {
    __shared__ uint32_t w[2][NPAD];
    const int sh = blockIdx.x, lane = blockIdx.y;
    const int q = A.diag ? lane : blockIdx.z;
    const int tid = threadIdx.x;
    for (int s = tid; s < N; s += 512) {
        const int g = T->slot_gpos[s];
        const int x = lx<true>(s);
        if (A.outb) {
            const uint32_t* o = A.outb + (size_t)q * A.q_stride + (size_t)sh * 2 * N * LANES;
            w[0][x] = o[(size_t)lane * N + g];
            w[1][x] = o[((size_t)LANES + lane) * N + g];
        } else {
            w[0][x] = red64(A.acc[(((size_t)sh * 2 + 0) * N + g) * LANES + lane], P1, 26);
            w[1][x] = red64(A.acc[(((size_t)sh * 2 + 1) * N + g) * LANES + lane], P2, 25);
        }
    }
    __syncthreads();
    {
        // one modulus per 256 threads; N^-1 is folded into the CRT below
        const int r = tid >> 8;
        ntt256_inv_unscaled<true>(w[r], T->ipsi[r], T->ipsiq[r], mod_p(r), tid & 255);
    }
    for (int s = tid; s < N; s += 512) {
        const int x = lx<true>(s);
        // [0, 2p) -> CRT terms in [0, p) (Shoup products by the constants)
        const uint32_t e1 = mul_shoup(w[0][x], NINV_Q2INV_P1, NINV_Q2INV_P1_SH, P1);
        const uint32_t e2 = mul_shoup(w[1][x], NINV_Q1INV_P2, NINV_Q1INV_P2_SH, P2);
        uint64_t b = (uint64_t)e1 * P2 + (uint64_t)e2 * P1;
        b = b >= Q ? b - Q : b;
        if (A.bcoef) A.bcoef[((size_t)sh * LANES + lane) * N + s] = b;
        if (A.se) {
            const uint32_t b12 = mod_switch_enc(b, 12);
            const uint32_t a17 = A.a17s[sh][(size_t)lane * N + s];
            uint32_t v = (a17 << 12) + b12;
            const int col = sh * LANES + lane;
            uint16_t* dst = A.se + (size_t)q * A.se_qstride + ((size_t)col * A.lpr + A.lane0_mult * q) * N + s;
            dst[0] = (uint16_t)(v % T_PLAIN);
            v /= T_PLAIN;
            dst[N] = (uint16_t)(v % T_PLAIN);
            dst[2 * N] = (uint16_t)(v / T_PLAIN);
        }
    }
}

__global__ void fill_u64_kernel(unsigned long long* p, size_t n, unsigned long long v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s) {
    if (!n) return;
    int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(blocks), dim3(256), 0, s, p, n, v);
}
void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    if (!n) return;
    int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, p, n, v);
}

void launch_k5(const uint16_t* entries, const long long* ent_off, long long col_stride, uint32_t* M, uint32_t* te_tmp, int l,
               const int* ncols_nonzero, int nlanes, const DevTables* T, hipStream_t s, uint32_t* clear, size_t clear_words) {
    const int L = l < N ? l : N / 2;
    const int conj = l == N;
    const int kq = N / 2 / L;
    int logl = 0;
    while ((1 << logl) < l) ++logl;
    int scale = 1;
    for (int x = 1; x < (int)T_PLAIN; ++x)
        if ((x * l) % (int)T_PLAIN == 1) { scale = x; break; }
    const int nblocks = l * nlanes;
    const int clear_per_block = clear ? (int)((clear_words + nblocks - 1) / nblocks) : 0;
    hipLaunchKernelGGL(k5_col_ntt_kernel, dim3(l, nlanes), dim3(256), 0, s, entries, ent_off, col_stride, M, l, ncols_nonzero, T,
                       clear, clear_per_block, clear_words);
    // tiles of 64 KiB: TS slots x l rows, TS * 4 bytes contiguous per row
    if (l <= 256)
        hipLaunchKernelGGL(k5_rows_w_kernel, dim3(N / 64, nlanes), dim3(K5R_NT), (size_t)l * 64 * 4, s, M, l, logl, kq, conj, T);
    else if (l <= 512)
        hipLaunchKernelGGL(k5_rows_kernel<32>, dim3(N / 32, nlanes), dim3(K5R_NT), (size_t)l * 32 * 4, s, M, l, logl, kq, conj, T);
    else if (l == 1024)
        hipLaunchKernelGGL(k5_rows_kernel<16>, dim3(N / 16, nlanes), dim3(K5R_NT), (size_t)l * 16 * 4, s, M, l, logl, kq, conj, T);
    else
        hipLaunchKernelGGL(k5_rows_kernel<8>, dim3(N / 8, nlanes), dim3(K5R_NT), (size_t)l * 8 * 4, s, M, l, logl, kq, conj, T);
    hipLaunchKernelGGL(k5_cols_final_kernel, dim3(l, nlanes), dim3(256), 0, s, M, te_tmp, l, L, conj, nlanes, scale, T);
}

void launch_chain_pre(const uint32_t* te_tmp, uint32_t* ip_tmp, uint32_t* a_out, int L, int nst, int qa_base,
                      const DevTables* T, hipStream_t s) {
    hipLaunchKernelGGL(chain_kernel<false>, dim3(LANES, nst), dim3(1024), 0, s, te_tmp, ip_tmp, a_out,
                       (uint32_t*)nullptr, (const uint32_t*)nullptr, (const int*)nullptr, L, LANES, qa_base, T);
}

void launch_chain_sec(const uint32_t* te_tmp, uint32_t* a_out, uint32_t* c_out, const uint32_t* qs, const int* lane_slot,
                      int L, int nst, int nlanes, const DevTables* T, hipStream_t s) {
    hipLaunchKernelGGL(chain_kernel<true>, dim3(nlanes, nst), dim3(1024), 0, s, te_tmp, (uint32_t*)nullptr, a_out, c_out,
                       qs, lane_slot, L, nlanes, 2, T);
}

// c <- sigma_5(c) + sigma_5(a) cb0 + NTT(D1) cb1' + TE cqb mod p for slots
// tid, tid + 1024 (the b-recurrence of the fused secondary inner product,
// with NTT(D0) = sigma_5(a) - 2^26 NTT(D1) folded into cb1' = cb1 - 2^26 cb0)
// (cq: the Shoup companions of the three constants [cb0, cb1, cqb][k])
__device__ __forceinline__ void c_update(const uint32_t* __restrict__ c_in, uint32_t* __restrict__ c_out,
                                         const uint32_t* d1t, const uint32_t sga[2], const uint32_t te[2],
                                         const uint32_t cb0[2], const uint32_t cb1[2], const uint32_t cqb[2],
                                         const uint32_t cq[3][2], const int sig[2], uint32_t p, int tid) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int s = tid + 1024 * k;
        const uint32_t d1 = d1t[lx<true>(s)];  // below 24p
        const uint32_t cv = c_in[sig[k]] + shoup_lazy(sga[k], cb0[k], cq[0][k], p) + shoup_lazy(d1, cb1[k], cq[1][k], p) +
                            shoup_lazy(te[k], cqb[k], cq[2][k], p);
        c_out[s] = red8p(cv, p);  // [0, 7p) -> [0, p)
    }
}

// The secondary chain on two CUs per chain (latency form for a few queries):
// block (chain, r) keeps a and c mod p_r. Each step it inverse-transforms
// sigma_5(a) mod p_r down to stage 1, swaps those 2048 words with its partner
// block (the other modulus) through xbuf, applies the last inverse stage of
// both moduli inside the CRT, computes the top gadget digit D1 of every
// coefficient, forward-transforms it mod p_r (NTT(D0) = sigma_5(a) - 2^26
// NTT(D1) needs no transform) and updates a and c mod p_r: half the
// transforms of chain_kernel<true> per block, plus one exchange. Partners are blocks b and b ^ 8 (one XCD under round-robin
// dispatch). Each exchanged word carries a step tag in bits 27-31 (values
// are < 2p < 2^27), so a block waits for exactly its partner's words of this
// step and needs no fence; xbuf starts zeroed (tag 0 is never used). Every
// wait is bounded by spin_limit polls; a timeout sets *err and the block
// runs on, so the kernel always completes.
__global__ __launch_bounds__(1024) void chain_pair_kernel(const uint32_t* __restrict__ te_tmp, uint32_t* __restrict__ a_out,
                                                          uint32_t* __restrict__ c_out, const uint32_t* __restrict__ qs,
                                                          const int* __restrict__ lane_slot, int L, int nlanes, int nchains,
                                                          uint32_t* xbuf, unsigned* err, int spin_limit,
                                                          const DevTables* __restrict__ T) {
    __shared__ __attribute__((aligned(16))) uint2 tw2[2][N];  // [fwd/inv][k] = (twiddle, Shoup companion), modulus r
    __shared__ uint32_t sa[N];
    __shared__ uint32_t sc[2][N];      // [buffer][slot]
    __shared__ uint32_t xa[NPAD];      // inverse transform of sigma_5(a)
    __shared__ uint32_t xd[2][NPAD];   // NTT(D1) in xd[0]; xd[1] only keeps LDS > 80 KB (one block per CU)
    const int b = blockIdx.x;
    const int chain = (b >> 4) * 8 + (b & 7), r = (b >> 3) & 1;
    if (chain >= nchains) return;
    const int st = chain / nlanes, li = chain % nlanes;
    const int tid = threadIdx.x;
    const uint32_t p = mod_p(r);
    for (int k = tid; k < N; k += 1024) {
        tw2[0][k] = make_uint2(T->psi[r][k], T->psiq[r][k]);
        tw2[1][k] = make_uint2(T->ipsi[r][k], T->ipsiq[r][k]);
        sa[k] = 0;
        sc[0][k] = 0;
    }
    // per-thread constants for the update phase: slots tid, tid + 1024
    uint32_t cg0[2], cg1[2], cqa[2], cb0[2], cb1[2], cqb[2];
    int sig_upd[2];
    const uint32_t* qa = T->qa[2 + st][0];
    const uint32_t* qv = qs + (size_t)lane_slot[li] * 6 * 2 * N;  // [v][r][slot]
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int s = tid + 1024 * k;
        cg0[k] = T->gka[0][r][s];
        cg1[k] = T->gka[1][r][s];
        cqa[k] = qa[r * N + s];
        cb0[k] = qv[(0 * 2 + r) * N + s];
        cb1[k] = qv[(1 * 2 + r) * N + s];
        cqb[k] = qv[((4 + st) * 2 + r) * N + s];
        sig_upd[k] = T->sigma5[s];
        // NTT(D0) = sigma_5(a) - 2^26 NTT(D1) (x = D0 + 2^26 D1 and NTT(x mod p) =
        // sigma_5(a)): fold -2^26 into the D1 constants, keep the D0 ones for sigma_5(a)
        const uint64_t t26 = (1ull << LOG2_B) % p;
        cg1[k] = (uint32_t)(((uint64_t)cg1[k] + p - t26 * cg0[k] % p) % p);
        cb1[k] = (uint32_t)(((uint64_t)cb1[k] + p - t26 * cb0[k] % p) % p);
    }
    // Shoup companions of the update constants
    uint32_t cg0q[2], cg1q[2], cqaq[2], cq[3][2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        cg0q[k] = shoup_companion(cg0[k], p);
        cg1q[k] = shoup_companion(cg1[k], p);
        cqaq[k] = shoup_companion(cqa[k], p);
        cq[0][k] = shoup_companion(cb0[k], p);
        cq[1][k] = shoup_companion(cb1[k], p);
        cq[2][k] = shoup_companion(cqb[k], p);
    }
    // phase A: threads 0..511, radix-4 group g = tid first (stages 10, 9)
    const int gA = tid & 511;
    int sigA[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) sigA[c] = T->sigma5[4 * gA + c];
    lds_barrier();

    uint32_t* xo = xbuf + (size_t)chain * 4 * N;  // [parity][r][N]
    const uint32_t* te_base = te_tmp + ((size_t)(st * nlanes + li) * L) * 2 * N;
    int cur = 0;
    uint32_t te[2] = {0, 0};   // TE row of the current (then previous) step
    uint32_t sga[2] = {0, 0};  // sigma_5(a) at slots tid, tid + 1024, same steps
#pragma unroll 1
    for (int i = 0; i < L; ++i) {
        // this step's TE row, in flight through A
        uint32_t tn[2];
        {
            const uint32_t* te_row = te_base + (size_t)i * 2 * N + (size_t)r * N;
#pragma unroll
            for (int k = 0; k < 2; ++k) tn[k] = te_row[tid + 1024 * k];
        }
        // ---- A: inverse NTT of sigma_5(a) mod p_r (threads 0..511, radix 4) ----
        // The gather and the stride-4..128 passes of wave w stay in elements
        // [256w, 256w + 256), so only the last two passes wait at block barriers
        // (the forward transform below likewise from its stride-32 pass on).
        if (tid < 512) {
            const int ii = gA;
            uint32_t v0 = sa[sigA[0]], v1 = sa[sigA[1]], v2 = sa[sigA[2]], v3 = sa[sigA[3]];
            const uint4 wb = *reinterpret_cast<const uint4*>(&tw2[1][1024 + 2 * ii]);
            bf_gs(v0, v1, wb.x, wb.y, p);
            bf_gs(v2, v3, wb.z, wb.w, p);
            const uint2 wa = tw2[1][512 + ii];
            bf_gs(v0, v2, wa.x, wa.y, p);
            bf_gs(v1, v3, wa.x, wa.y, p);
            const int b4 = lx<true>(4 * ii);
            xa[b4] = v0; xa[b4 + 1] = v1; xa[b4 + 2] = v2; xa[b4 + 3] = v3;
            wave_lds_sync();
            inv_pass4gi<true>(xa, tw2[1], p, 7, gA);
            wave_lds_sync();
            inv_pass4gi<true>(xa, tw2[1], p, 5, gA);
            wave_lds_sync();
            inv_pass4gi<true>(xa, tw2[1], p, 3, gA);
        }
        lds_barrier();
        if (tid < 512) inv_pass4gi<true>(xa, tw2[1], p, 1, gA);
        lds_barrier();
        // (the last inverse stage, pairs (j, j + 1024), folds into phase B)
        // ---- exchange with the partner (other modulus) ----
        const uint32_t tag = (uint32_t)(i % 31 + 1) << 27;
        // words j = tid and tid + 1024 side by side: one 8-byte store and load per thread
        uint64_t* mine = reinterpret_cast<uint64_t*>(xo + ((size_t)(i & 1) * 2 + r) * N);
        const uint64_t* theirs = reinterpret_cast<const uint64_t*>(xo + ((size_t)(i & 1) * 2 + (r ^ 1)) * N);
        uint32_t own[2], oth[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) own[k] = xa[lx<true>(tid + 1024 * k)];  // [0, 2p)
        __hip_atomic_store(mine + tid, (uint64_t)(tag | own[0]) | ((uint64_t)(tag | own[1]) << 32), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        // while the partner's words travel: the c-update of the previous step
        // (its digit transforms are still in xd, its TE row in te)
        if (i > 0) {
            c_update(sc[cur], sc[cur ^ 1], xd[0], sga, te, cb0, cb1, cqb, cq, sig_upd, p, tid);
            cur ^= 1;
        }
        // this step's sigma_5(a) (sa is rewritten in D) and TE row, and the two
        // products of the a-update that do not need the digits
        uint32_t pa[2];  // [0, 4p)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            te[k] = tn[k];
            sga[k] = sa[sig_upd[k]];
            pa[k] = shoup_lazy(sga[k], cg0[k], cg0q[k], p) + shoup_lazy(te[k], cqa[k], cqaq[k], p);
        }
        // both words in one round trip, polled again until both carry this step's tag
        const uint64_t tag2 = (uint64_t)tag | ((uint64_t)tag << 32);
        uint64_t u2 = __hip_atomic_load(theirs + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (spin_limit <= 0) {  // no waiting at all (exercises the timeout path)
            atomicOr(err, 1u);
            u2 = tag2;
        }
        for (int n = 0; (u2 & 0xF8000000F8000000ull) != tag2;) {
            // give up at the limit, or soon after any wait has (later steps then pass quickly)
            if ((++n & 255) == 0 &&
                (n >= spin_limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            u2 = __hip_atomic_load(theirs + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        oth[0] = (uint32_t)u2 & 0x07FFFFFFu;
        oth[1] = (uint32_t)(u2 >> 32) & 0x07FFFFFFu;
        // ---- B: the last inverse stage of both moduli for the pair (tid,
        // tid + 1024), CRT, centered lift, digits (base_pir.rs:544-570): D1 mod
        // p_r. The stage's butterfly and the CRT factors fold into one Shoup
        // product per coefficient (as in chain_kernel). ----
        uint32_t ecrt[2][2];  // [k: lo, hi][modulus]
        {
            const uint32_t x1 = r == 0 ? own[0] : oth[0], y1 = r == 0 ? own[1] : oth[1];  // [0, 2p1)
            const uint32_t x2 = r == 0 ? oth[0] : own[0], y2 = r == 0 ? oth[1] : own[1];  // [0, 2p2)
            ecrt[0][0] = mul_shoup(x1 + y1, T->crt_k[0][0], T->crt_kq[0][0], P1);
            ecrt[1][0] = mul_shoup(x1 - y1 + 2 * P1, T->crt_k[1][0], T->crt_kq[1][0], P1);
            ecrt[0][1] = mul_shoup(x2 + y2, T->crt_k[0][1], T->crt_kq[0][1], P2);
            ecrt[1][1] = mul_shoup(x2 - y2 + 2 * P2, T->crt_k[1][1], T->crt_kq[1][1], P2);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int j = tid + 1024 * k;
            const int xj = lx<true>(j);
            const int32_t d1 = crt_digit1(ecrt[k][0], ecrt[k][1]);
            xd[0][xj] = (uint32_t)(d1 + (int32_t)p);  // [0, 2p): |d1| <= 2^24 < p
        }
        lds_barrier();
        // ---- C: forward NTT of D1 mod p_r (threads 0..511, radix 4) ----
        if (tid < 512) fwd_pass4gi<true>(xd[0], tw2[0], p, 0, gA);
        lds_barrier();
        if (tid < 512) fwd_pass4gi<true>(xd[0], tw2[0], p, 2, gA);
        lds_barrier();
        if (tid < 512) {
            fwd_pass4gi<true>(xd[0], tw2[0], p, 4, gA);
            wave_lds_sync();
            fwd_pass4gi<true>(xd[0], tw2[0], p, 6, gA);
            wave_lds_sync();
            fwd_pass4gi<true>(xd[0], tw2[0], p, 8, gA);
            wave_lds_sync();
            fwd_pass2_lastw<true>(xd[0], tw2[0], p, gA);
        }
        lds_barrier();
        // ---- D: update a mod p_r (c follows in the next step's exchange) ----
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int s = tid + 1024 * k;
            const uint32_t d1 = xd[0][lx<true>(s)];  // NTT(D1), below 24p
            sa[s] = red8p(pa[k] + shoup_lazy(d1, cg1[k], cg1q[k], p), p);  // [0, 6p) -> [0, p)
        }
        lds_barrier();
    }
    c_update(sc[cur], sc[cur ^ 1], xd[0], sga, te, cb0, cb1, cqb, cq, sig_upd, p, tid);
    cur ^= 1;
    lds_barrier();
    uint32_t* ao = a_out + ((size_t)chain * 2 + r) * N;
    uint32_t* co = c_out + ((size_t)chain * 2 + r) * N;
    for (int s = tid; s < N; s += 1024) {
        ao[s] = sa[s];
        co[s] = sc[cur][s];
    }
}

// The secondary chain on four CUs per chain (the lowest latency for a few
// chains): block (chain, r, h) keeps a and c mod p_r on the slots of half h.
// sigma_5 maps each half of the slots to itself (slot s holds f(psi^e), the
// halves are e = 1 and e = 3 mod 4 -- the two factors X^(N/2) -+ psi^(N/2) of
// the first NTT stage -- and 5e = e mod 4), and below its first stage the
// negacyclic transform splits into two independent half transforms. So each
// step a block gathers sigma_5(a) on its half and runs the inverse transform
// down to stage 1 (1024 points), all four blocks of the chain swap those 1024
// words, and every block finishes the last inverse stage of both moduli for
// all 2048 coefficients, takes their CRT and top gadget digit D1 (the work
// the two-CU chain also repeats in both blocks), applies the first forward
// stage for its half and transforms it mod p_r (stages 1..10, 1024 points),
// then updates a and c on its half. The arithmetic is chain_pair_kernel's,
// stage by stage, so the outputs are identical; each block runs half the
// transform work of that kernel and the same one exchange per step. The
// blocks of a chain are b, b + 8, b + 16, b + 24 (one XCD under round-robin
// dispatch); tags, bounded waits and the timeout flag work as in
// chain_pair_kernel. LDS above 80 KB keeps one block per CU.
constexpr int QUAD_NT = 512;
__device__ __forceinline__ void c_update_half(const uint32_t* __restrict__ c_in, uint32_t* __restrict__ c_out,
                                              const uint32_t* d1t, int off, const uint32_t sga[2], const uint32_t te[2],
                                              const uint32_t cb0[2], const uint32_t cb1[2], const uint32_t cqb[2],
                                              const uint32_t cq[3][2], const int sig[2], uint32_t p, int tid) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int sl = tid + QUAD_NT * k;
        const uint32_t d1 = d1t[lx<true>(off + sl)];  // below 24p
        const uint32_t cv = c_in[sig[k]] + shoup_lazy(sga[k], cb0[k], cq[0][k], p) + shoup_lazy(d1, cb1[k], cq[1][k], p) +
                            shoup_lazy(te[k], cqb[k], cq[2][k], p);
        c_out[sl] = red8p(cv, p);  // [0, 7p) -> [0, p)
    }
}

// crt_digit1 mod p, in [0, 2p) (|d1| <= 2^24 < p)
__device__ __forceinline__ uint32_t digit1_mod(uint32_t e1, uint32_t e2, uint32_t p) {
    return (uint32_t)(crt_digit1(e1, e2) + (int32_t)p);
}

__global__ __launch_bounds__(QUAD_NT) void chain_quad_kernel(const uint32_t* __restrict__ te_tmp, uint32_t* __restrict__ a_out,
                                                             uint32_t* __restrict__ c_out, const uint32_t* __restrict__ qs,
                                                             const int* __restrict__ lane_slot, int L, int nlanes, int nchains,
                                                             uint32_t* xbuf, unsigned* err, int spin_limit,
                                                             const DevTables* __restrict__ T) {
    __shared__ __attribute__((aligned(16))) uint2 tw2[2][N];  // [fwd/inv][k] = (twiddle, Shoup companion), modulus r
    __shared__ uint32_t sa[HALFN];     // a on this half's slots
    __shared__ uint32_t sc[2][HALFN];  // c, [buffer][slot]
    __shared__ uint32_t xa[NPAD];      // inverse transform of this half (elements of the half only)
    __shared__ uint32_t xd[4][NPAD];   // forward transform of D1 on this half in xd[0]; xd[1..3] only keep LDS > 80 KB
    const int b = blockIdx.x;
    const int chain = (b >> 5) * 8 + (b & 7), role = (b >> 3) & 3;
    const int r = role & 1, h = role >> 1;
    if (chain >= nchains) return;
    const int st = chain / nlanes, li = chain % nlanes;
    const int tid = threadIdx.x;
    const uint32_t p = mod_p(r);
    const int off = h * HALFN;  // first slot (and transform element) of the half
    for (int k = tid; k < N; k += QUAD_NT) {
        tw2[0][k] = make_uint2(T->psi[r][k], T->psiq[r][k]);
        tw2[1][k] = make_uint2(T->ipsi[r][k], T->ipsiq[r][k]);
    }
    for (int k = tid; k < HALFN; k += QUAD_NT) {
        sa[k] = 0;
        sc[0][k] = 0;
    }
    // stage 0: the last inverse stage of both moduli, the first forward stage mod p_r
    // The inverse stage's butterfly (x, y) -> (x + y, (x - y) psi^-1) and the
    // CRT factors N^-1 (q/p_r)^-1 fold into one Shoup product per output:
    // e_lo = (x + y) k_r, e_hi = (x - y) psi_r^-1 k_r (mod p_r).
    const uint32_t kc[2] = {T->crt_k[0][0], T->crt_k[0][1]}, kcq[2] = {T->crt_kq[0][0], T->crt_kq[0][1]};
    const uint32_t wk[2] = {T->crt_k[1][0], T->crt_k[1][1]}, wkq[2] = {T->crt_kq[1][0], T->crt_kq[1][1]};
    const uint32_t wf = T->psi[r][1], wfq = T->psiq[r][1];
    // per-thread constants of the c-update: local slots tid, tid + 512
    uint32_t cb0[2], cb1[2], cqb[2];
    int sig_upd[2];
    const uint64_t t26 = (1ull << LOG2_B) % p;
    const uint32_t* qa = T->qa[2 + st][0];
    const uint32_t* qv = qs + (size_t)lane_slot[li] * 6 * 2 * N;  // [v][r][slot]
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int s = off + tid + QUAD_NT * k;
        cb0[k] = qv[(0 * 2 + r) * N + s];
        cb1[k] = qv[(1 * 2 + r) * N + s];
        cqb[k] = qv[((4 + st) * 2 + r) * N + s];
        sig_upd[k] = T->sigma5[s] - off;  // within the half
        cb1[k] = (uint32_t)(((uint64_t)cb1[k] + p - t26 * cb0[k] % p) % p);
    }
    uint32_t cq[3][2];  // Shoup companions of cb0, cb1, cqb
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        cq[0][k] = shoup_companion(cb0[k], p);
        cq[1][k] = shoup_companion(cb1[k], p);
        cq[2][k] = shoup_companion(cqb[k], p);
    }
    // the transform passes: threads 0..255, radix-4 group g = 256 h + tid of the
    // full-length passes (the groups of half h)
    const int gq = off / 4 + (tid & 255);
    // the a-update runs inside the last forward pass: thread tid < 256 updates
    // local slots 4 tid .. 4 tid + 3, the elements of its last radix-4 group
    const int sd0 = 4 * (tid & 255);
    uint32_t cg0[4], cg1[4], cqa[4], cg0q[4], cg1q[4], cqaq[4];
    int sig_d[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int s = off + sd0 + c;
        cg0[c] = T->gka[0][r][s];
        cg1[c] = (uint32_t)(((uint64_t)T->gka[1][r][s] + p - t26 * cg0[c] % p) % p);
        cqa[c] = qa[r * N + s];
        sig_d[c] = T->sigma5[s] - off;
        cg0q[c] = shoup_companion(cg0[c], p);
        cg1q[c] = shoup_companion(cg1[c], p);
        cqaq[c] = shoup_companion(cqa[c], p);
    }
    int sigA[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) sigA[c] = T->sigma5[4 * gq + c] - off;
    lds_barrier();

    uint32_t* xo = xbuf + (size_t)chain * 4 * N;  // [parity][role][HALFN]
    const uint32_t* te_base = te_tmp + ((size_t)(st * nlanes + li) * L) * 2 * N;
    int cur = 0;
    uint32_t te[2] = {0, 0};
    uint32_t sga[2] = {0, 0};
#pragma unroll 1
    for (int i = 0; i < L; ++i) {
        uint32_t tn[2];
        uint4 ted;  // TE at the a-update slots
        {
            const uint32_t* te_row = te_base + (size_t)i * 2 * N + (size_t)r * N + off;
#pragma unroll
            for (int k = 0; k < 2; ++k) tn[k] = te_row[tid + QUAD_NT * k];
            ted = *reinterpret_cast<const uint4*>(te_row + sd0);
        }
        // ---- A: sigma_5 gather and inverse stages 10..1 on this half ----
        if (tid < 256) {
            uint32_t v0 = sa[sigA[0]], v1 = sa[sigA[1]], v2 = sa[sigA[2]], v3 = sa[sigA[3]];
            const uint4 wb = *reinterpret_cast<const uint4*>(&tw2[1][1024 + 2 * gq]);
            bf_gs(v0, v1, wb.x, wb.y, p);
            bf_gs(v2, v3, wb.z, wb.w, p);
            const uint2 wa = tw2[1][512 + gq];
            bf_gs(v0, v2, wa.x, wa.y, p);
            bf_gs(v1, v3, wa.x, wa.y, p);
            const int b4 = lx<true>(4 * gq);
            xa[b4] = v0; xa[b4 + 1] = v1; xa[b4 + 2] = v2; xa[b4 + 3] = v3;
            wave_lds_sync();
            inv_pass4gi<true>(xa, tw2[1], p, 7, gq);
            wave_lds_sync();
            inv_pass4gi<true>(xa, tw2[1], p, 5, gq);
            wave_lds_sync();
            inv_pass4gi<true>(xa, tw2[1], p, 3, gq);
        }
        lds_barrier();
        // ---- exchange: the four halves of stage 1's output ----
        // The last inverse pass (stages 2, 1; group g = 256 h + tid holds local
        // elements tid + 256 c) sends its words as soon as they are computed,
        // each tagged with the step, and keeps them in xa for this block's own use.
        const uint32_t tag = (uint32_t)(i % 31 + 1) << 27;
        uint32_t* base = xo + (size_t)(i & 1) * 4 * HALFN;
        if (tid < 256) {
            uint32_t v[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = xa[lx<true>(off + tid + 256 * c)];
            const uint4 wb = *reinterpret_cast<const uint4*>(&tw2[1][4 + 2 * h]);
            bf_gs(v[0], v[1], wb.x, wb.y, p);
            bf_gs(v[2], v[3], wb.z, wb.w, p);
            const uint2 wa = tw2[1][2 + h];
            bf_gs(v[0], v[2], wa.x, wa.y, p);
            bf_gs(v[1], v[3], wa.x, wa.y, p);
#pragma unroll
            for (int c = 0; c < 4; ++c) xa[lx<true>(off + tid + 256 * c)] = v[c];  // [0, 2p)
            // words j and j + 512 side by side: pairs (tid, tid + 512), (tid + 256, tid + 768)
            uint64_t* mine = reinterpret_cast<uint64_t*>(base + role * HALFN);
            __hip_atomic_store(mine + tid, (uint64_t)(tag | v[0]) | ((uint64_t)(tag | v[2]) << 32), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(mine + tid + 256, (uint64_t)(tag | v[1]) | ((uint64_t)(tag | v[3]) << 32), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        lds_barrier();
        uint32_t own[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) own[k] = xa[lx<true>(off + tid + QUAD_NT * k)];
        if (i > 0) {
            c_update_half(sc[cur], sc[cur ^ 1], xd[0], off, sga, te, cb0, cb1, cqb, cq, sig_upd, p, tid);
            cur ^= 1;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            te[k] = tn[k];
            sga[k] = sa[sig_upd[k]];
        }
        // the two products of the a-update that do not need the digits (only
        // the transform threads update a)
        uint32_t pa[4] = {0, 0, 0, 0};  // [0, 4p)
        if (tid < 256) {
            const uint32_t tv[4] = {ted.x, ted.y, ted.z, ted.w};
#pragma unroll
            for (int c = 0; c < 4; ++c)
                pa[c] = shoup_lazy(sa[sig_d[c]], cg0[c], cg0q[c], p) + shoup_lazy(tv[c], cqa[c], cqaq[c], p);
        }
        // the other three roles' words for pairs j = tid, tid + 512 (one 8-byte
        // load per role): every poll reloads all pairs not yet tagged in one round trip
        const uint64_t tag2 = (uint64_t)tag | ((uint64_t)tag << 32);
        uint64_t w2[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            w2[q] = q == role ? (tag2 | own[0] | ((uint64_t)own[1] << 32))
                              : __hip_atomic_load(reinterpret_cast<const uint64_t*>(base + q * HALFN) + tid, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        if (spin_limit <= 0) {  // no waiting at all (exercises the timeout path)
            atomicOr(err, 1u);
#pragma unroll
            for (int q = 0; q < 4; ++q) w2[q] = tag2;
        }
        for (int n = 0;;) {
            bool ready = true;
#pragma unroll
            for (int q = 0; q < 4; ++q) ready = ready && (w2[q] & 0xF8000000F8000000ull) == tag2;
            if (ready) break;
            // give up at the limit, or soon after any wait has (later steps then pass quickly)
            if ((++n & 255) == 0 &&
                (n >= spin_limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if ((w2[q] & 0xF8000000F8000000ull) != tag2)
                    w2[q] = __hip_atomic_load(reinterpret_cast<const uint64_t*>(base + q * HALFN) + tid, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
        }
        uint32_t w[2][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            w[0][q] = (uint32_t)w2[q];
            w[1][q] = (uint32_t)(w2[q] >> 32);
        }
        // ---- B: stage 0 inverse (both moduli), CRT, D1, stage 0 forward (this half) ----
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int j = tid + QUAD_NT * k;
            // roles: 0 = (p1, half 0), 1 = (p2, half 0), 2 = (p1, half 1), 3 = (p2, half 1)
            const uint32_t x1lo = w[k][0] & 0x07FFFFFFu, x1hi = w[k][2] & 0x07FFFFFFu;  // [0, 2p1)
            const uint32_t x2lo = w[k][1] & 0x07FFFFFFu, x2hi = w[k][3] & 0x07FFFFFFu;  // [0, 2p2)
            const uint32_t e1lo = mul_shoup(x1lo + x1hi, kc[0], kcq[0], P1);
            const uint32_t e1hi = mul_shoup(x1lo - x1hi + 2 * P1, wk[0], wkq[0], P1);
            const uint32_t e2lo = mul_shoup(x2lo + x2hi, kc[1], kcq[1], P2);
            const uint32_t e2hi = mul_shoup(x2lo - x2hi + 2 * P2, wk[1], wkq[1], P2);
            const uint32_t dlo = digit1_mod(e1lo, e2lo, p), dhi = digit1_mod(e1hi, e2hi, p);
            const uint32_t t = shoup_lazy(dhi, wf, wfq, p);
            xd[0][lx<true>(off + j)] = h == 0 ? dlo + t : dlo - t + 2 * p;  // [0, 4p)
        }
        lds_barrier();
        // ---- C: forward stages 1..10 on this half ----
        if (tid < 256) fwd_pass4gi<true>(xd[0], tw2[0], p, 1, gq);
        lds_barrier();
        if (tid < 256) {
            fwd_pass4gi<true>(xd[0], tw2[0], p, 3, gq);
            wave_lds_sync();
            fwd_pass4gi<true>(xd[0], tw2[0], p, 5, gq);
            wave_lds_sync();
            fwd_pass4gi<true>(xd[0], tw2[0], p, 7, gq);
            wave_lds_sync();
            // ---- D: the last pass (stages 9, 10) and the a-update of its four
            // slots (c follows in the next step's exchange) ----
            const int b4 = lx<true>(4 * gq);
            uint32_t v0 = xd[0][b4], v1 = xd[0][b4 + 1], v2 = xd[0][b4 + 2], v3 = xd[0][b4 + 3];
            const uint2 wa = tw2[0][512 + gq];
            bf_ct(v0, v2, wa.x, wa.y, p);
            bf_ct(v1, v3, wa.x, wa.y, p);
            const uint4 wb = *reinterpret_cast<const uint4*>(&tw2[0][1024 + 2 * gq]);
            bf_ct(v0, v1, wb.x, wb.y, p);
            bf_ct(v2, v3, wb.z, wb.w, p);
            xd[0][b4] = v0; xd[0][b4 + 1] = v1; xd[0][b4 + 2] = v2; xd[0][b4 + 3] = v3;  // NTT(D1), below 24p
            // [0, 6p) -> [0, p)
            sa[sd0] = red8p(pa[0] + shoup_lazy(v0, cg1[0], cg1q[0], p), p);
            sa[sd0 + 1] = red8p(pa[1] + shoup_lazy(v1, cg1[1], cg1q[1], p), p);
            sa[sd0 + 2] = red8p(pa[2] + shoup_lazy(v2, cg1[2], cg1q[2], p), p);
            sa[sd0 + 3] = red8p(pa[3] + shoup_lazy(v3, cg1[3], cg1q[3], p), p);
        }
        lds_barrier();
    }
    c_update_half(sc[cur], sc[cur ^ 1], xd[0], off, sga, te, cb0, cb1, cqb, cq, sig_upd, p, tid);
    cur ^= 1;
    lds_barrier();
    uint32_t* ao = a_out + ((size_t)chain * 2 + r) * N + off;
    uint32_t* co = c_out + ((size_t)chain * 2 + r) * N + off;
    for (int sl = tid; sl < HALFN; sl += QUAD_NT) {
        ao[sl] = sa[sl];
        co[sl] = sc[cur][sl];
    }
}

// exchange words per chain of chain_pair_kernel
size_t chain_pair_xbuf_words(int nchains) { return (size_t)nchains * 4 * N; }

void launch_chain_pair(const uint32_t* te_tmp, uint32_t* a_out, uint32_t* c_out, const uint32_t* qs, const int* lane_slot,
                       int L, int nst, int nlanes, uint32_t* xch, int spin_limit, const DevTables* T, hipStream_t s,
                       bool quad, bool cleared) {
    const int nchains = nst * nlanes;
    if (!cleared) launch_fill_u32(xch, 64 + chain_pair_xbuf_words(nchains), 0, s);
    uint32_t* xbuf = xch + 64;
    unsigned* err = reinterpret_cast<unsigned*>(xch);
    if (quad)
        hipLaunchKernelGGL(chain_quad_kernel, dim3(chain_multi_blocks(nchains, true)), dim3(QUAD_NT), 0, s, te_tmp, a_out, c_out,
                           qs, lane_slot, L, nlanes, nchains, xbuf, err, spin_limit, T);
    else
        hipLaunchKernelGGL(chain_pair_kernel, dim3(chain_multi_blocks(nchains, false)), dim3(1024), 0, s, te_tmp, a_out, c_out,
                           qs, lane_slot, L, nlanes, nchains, xbuf, err, spin_limit, T);
}

int chain_multi_blocks(int nchains, bool quad) { return (quad ? 32 : 16) * ((nchains + 7) / 8); }

// K1 for one or two queries on the limb-block layout: the same exact limb
// arithmetic as k1b_kernel, on the vector ALU (v_dot4_i32_i8 over 4 digit
// pairs at a time) so that many waves per CU keep enough loads in flight to
// stream at HBM rate. Lane l of a wave owns output row l % 32 and k-half l / 32
// of its a-group; the two halves are summed at the end.
template <int NQ, int PD>
__global__ __launch_bounds__(256) void k1d_kernel(const uint32_t* const* __restrict__ slabs, const uint32_t* __restrict__ ql,
                                                  uint32_t* __restrict__ outb, int nst, size_t q_stride) {
    // query digits for this (r, half): [s*32 + c][q][v][h] x 16 bytes
    __shared__ u32x4 qs[6 * 32][NQ][4][2];
    const int S = 3 * nst;
    const int half = blockIdx.y;
    const int r = blockIdx.z & 1, sh = blockIdx.z >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, h = l >> 5;
    const int CT = 1;  // ql built for <= 32 queries
    const u32x4* Q = reinterpret_cast<const u32x4*>(ql) + ((size_t)(r * 2 + half) * S) * 32 * CT * 256;
    for (int it = tid; it < S * 32 * NQ * 4 * 2; it += 256) {
        const int hh = it & 1, v = (it >> 1) & 3, q = (it >> 3) % NQ, kc = it / (8 * NQ);
        qs[kc][q][v][hh] = Q[(size_t)kc * CT * 256 + v * 64 + hh * 32 + q];
    }
    __syncthreads();
    const int g = blockIdx.x * 4 + wave;
    const ARow arow = a_row(slabs[sh], r, half, g, S);
    const int nk = S * 32;
    int acc[NQ][7];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int w = 0; w < 7; ++w) acc[q][w] = 0;
    ATile abuf[PD];
#pragma unroll
    for (int d = 0; d < PD; ++d) load_a(abuf[d], arow, r, l, d);
    for (int kk0 = 0; kk0 < nk; kk0 += PD) {
#pragma unroll
        for (int d = 0; d < PD; ++d) {
            const int kk = kk0 + d;
            v4i a[4];
            a[0] = abuf[d].d[0];
            a[1] = abuf[d].d[1];
            a[2] = abuf[d].d[2];
            a[3] = r == 0 ? top_digits(abuf[d].top) : top_digits1(abuf[d].top);
            if (kk + PD < nk) load_a(abuf[d], arow, r, l, kk + PD);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                u32x4 b[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) b[v] = qs[kk][q][v][h];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        int t = acc[q][u + v];
                        t = __builtin_amdgcn_sdot4(a[u].x, (int)b[v].x, t, false);
                        t = __builtin_amdgcn_sdot4(a[u].y, (int)b[v].y, t, false);
                        t = __builtin_amdgcn_sdot4(a[u].z, (int)b[v].z, t, false);
                        t = __builtin_amdgcn_sdot4(a[u].w, (int)b[v].w, t, false);
                        acc[q][u + v] = t;
                    }
            }
        }
    }
    const uint32_t p = mod_p(r);
    const int K = mod_k(r);
    uint32_t cw[7];
    weight_consts(p, cw);
    const uint32_t* corr = qcorr_ptr(ql, S, CT) + (size_t)(r * 2 + half) * CT * 32;
    const int row = l & 31;
    const int a = g * 4 + (row >> 3), lane8 = row & 7;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        int64_t t = corr[q];
#pragma unroll
        for (int w = 0; w < 7; ++w) {
            const int tot = acc[q][w] + __shfl_xor(acc[q][w], 32);
            t += (int64_t)tot * cw[w];
        }
        if (h == 0) outb[(size_t)q * q_stride + (((size_t)sh * 2 + r) * LANES + lane8) * N + half * HALFN + a] = from_i64(t, p, K);
    }
}

// k1d for a fleet of one shard: one block per (a-group, r, half) whose four
// waves each take a quarter of the K-steps, summed through LDS at the end.
// Four times the waves of k1d_kernel (4096 instead of 1024), so enough loads
// are in flight to stream at HBM rate; the integer sums are the same, so are
// the results.
template <int NQ, int PD>
__global__ __launch_bounds__(256) void k1d_split_kernel(const uint32_t* const* __restrict__ slabs, const uint32_t* __restrict__ ql,
                                                        uint32_t* __restrict__ outb, int nst, size_t q_stride) {
    __shared__ u32x4 qs[6 * 32][NQ][4][2];
    const int S = 3 * nst;
    const int half = blockIdx.y;
    const int r = blockIdx.z & 1, sh = blockIdx.z >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, h = l >> 5;
    const int CT = 1;
    const u32x4* Q = reinterpret_cast<const u32x4*>(ql) + ((size_t)(r * 2 + half) * S) * 32 * CT * 256;
    for (int it = tid; it < S * 32 * NQ * 4 * 2; it += 256) {
        const int hh = it & 1, v = (it >> 1) & 3, q = (it >> 3) % NQ, kc = it / (8 * NQ);
        qs[kc][q][v][hh] = Q[(size_t)kc * CT * 256 + v * 64 + hh * 32 + q];
    }
    __syncthreads();
    const int g = blockIdx.x;
    const ARow arow = a_row(slabs[sh], r, half, g, S);
    const int nk = S * 32, kq = nk / 4;  // S * 32 is a multiple of 4
    const int k0 = wave * kq, k1 = k0 + kq;
    int acc[NQ][7];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int w = 0; w < 7; ++w) acc[q][w] = 0;
    ATile abuf[PD];
#pragma unroll
    for (int d = 0; d < PD; ++d) load_a(abuf[d], arow, r, l, k0 + d);
    for (int kk0 = k0; kk0 < k1; kk0 += PD) {
#pragma unroll
        for (int d = 0; d < PD; ++d) {
            const int kk = kk0 + d;
            v4i a[4];
            a[0] = abuf[d].d[0];
            a[1] = abuf[d].d[1];
            a[2] = abuf[d].d[2];
            a[3] = r == 0 ? top_digits(abuf[d].top) : top_digits1(abuf[d].top);
            if (kk + PD < k1) load_a(abuf[d], arow, r, l, kk + PD);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                u32x4 b[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) b[v] = qs[kk][q][v][h];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        int t = acc[q][u + v];
                        t = __builtin_amdgcn_sdot4(a[u].x, (int)b[v].x, t, false);
                        t = __builtin_amdgcn_sdot4(a[u].y, (int)b[v].y, t, false);
                        t = __builtin_amdgcn_sdot4(a[u].z, (int)b[v].z, t, false);
                        t = __builtin_amdgcn_sdot4(a[u].w, (int)b[v].w, t, false);
                        acc[q][u + v] = t;
                    }
            }
        }
    }
    // waves 1..3 hand their partial sums to wave 0 (|sum| < 2^29 in total)
    __syncthreads();  // every wave is done reading qs
    int* part = reinterpret_cast<int*>(&qs[0][0][0][0]);  // [3][NQ][7][64]
    if (wave > 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int w = 0; w < 7; ++w) part[(((wave - 1) * NQ + q) * 7 + w) * 64 + l] = acc[q][w];
    }
    __syncthreads();
    if (wave > 0) return;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int w = 0; w < 7; ++w)
            acc[q][w] += part[(q * 7 + w) * 64 + l] + part[((NQ + q) * 7 + w) * 64 + l] + part[((2 * NQ + q) * 7 + w) * 64 + l];
    const uint32_t p = mod_p(r);
    const int K = mod_k(r);
    uint32_t cw[7];
    weight_consts(p, cw);
    const uint32_t* corr = qcorr_ptr(ql, S, CT) + (size_t)(r * 2 + half) * CT * 32;
    const int row = l & 31;
    const int a = g * 4 + (row >> 3), lane8 = row & 7;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        int64_t t = corr[q];
#pragma unroll
        for (int w = 0; w < 7; ++w) {
            const int tot = acc[q][w] + __shfl_xor(acc[q][w], 32);
            t += (int64_t)tot * cw[w];
        }
        if (h == 0) outb[(size_t)q * q_stride + (((size_t)sh * 2 + r) * LANES + lane8) * N + half * HALFN + a] = from_i64(t, p, K);
    }
}

// ---------------------------------------------------------------------------
// Two-stream K1 for one or two queries (limb-block shards). The key-switch
// digits satisfy NTT(D0_i) = sigma(a_i) - 2^26 NTT(D1_i) mod p_r
// (base_pir.rs:544-579: D0 + 2^26 D1 is the exact lift of sigma(a_i)), and the
// chain a_(i+1) = sigma(a_i) gka0 + D1_i (gka1 - 2^26 gka0) + qry_a TE_i
// (base_pir.rs:838-861, a_0 = 0) is linear in the stored rows. Unrolled, the
// IPGK0 term of main_pir_loop (base_pir.rs:1153-1208) becomes a row-dependent
// weight on the IPGK1 and TE rows, so K1 reads two of the three streams:
//   out[a] = sum_(st,j) IPGK1[a][j] G[a][j] + TE_st[a][j] Q_st[a][j]
//   G[a][j]    = gkb1'[j] + gka1'[j] C[(j - a) mod L][j]
//   Q_st[a][j] = qb_st[j] + qa_st[j] C[(j - a) mod L][j]
//   C[M][j]    = sum_(n=1..M) gkb0[j-n] prod_(m=1..n-1) gka0[j-m]
// in Galois positions (j, j-n within one half, mod L = N/2), x' = x - 2^26
// x0. The weights depend on the query and the seeds only, so all shards of a
// fleet share them.
// ---------------------------------------------------------------------------

// C[q][r][h][M][j] (u32). The recurrence over M (c += gkb0[j-M] P,
// P *= gka0[j-M]) is affine in (c, P), so it is split into 16 segments of 64
// steps: each thread (j, segment) first forms its segment's own product and
// sum from (0, 1), the segments' starting values come from a prefix over the
// 16 segments, then each thread runs its segment again from the true start
// and writes it. Block: 64 positions j x 16 segments.
__global__ __launch_bounds__(256) void k1w_scan_kernel(const uint32_t* __restrict__ qp, uint32_t* __restrict__ C,
                                                        const DevTables* __restrict__ T) {
    // 16 positions x 16 segments per block: 4 x HALFN / 16 blocks per query
    // spread the sequential steps over every CU
    constexpr int SEG = 16, LEN = HALFN / SEG, JB = 16;
    __shared__ uint32_t g0s[HALFN], gas[HALFN];
    __shared__ uint32_t segc[SEG][JB], segp[SEG][JB];
    const int h = blockIdx.y, r = blockIdx.z & 1, q = blockIdx.z >> 1;
    const uint32_t p = mod_p(r);
    const int K = mod_k(r);
    const int tid = threadIdx.x, jj = tid % JB, seg = tid / JB;
    const uint32_t* g0 = qp + ((size_t)q * 4 * 2 + r) * N + h * HALFN;  // vector 0 (gkb0) of query q
    for (int y = tid; y < HALFN; y += 256) {
        g0s[y] = g0[y];
        gas[y] = T->gka[0][r][T->gpos_slot[h * HALFN + y]];
    }
    __syncthreads();
    const int j = blockIdx.x * JB + jj;
    const int m0 = seg * LEN;  // steps M = m0 .. m0 + LEN - 1 (M = 0 contributes nothing)
    uint32_t c = 0, P = 1;
    for (int M = (seg == 0 ? 1 : m0); M < m0 + LEN; ++M) {
        const int y = (j - M) & (HALFN - 1);
        c = add_mod(c, mul_mod(g0s[y], P, p, K), p);
        P = mul_mod(P, gas[y], p, K);
    }
    segc[seg][jj] = c;
    segp[seg][jj] = P;
    __syncthreads();
    // start of this segment: fold the earlier segments, (c, P) -> (c + P c_k, P P_k)
    uint32_t c0 = 0, P0 = 1;
    for (int k = 0; k < seg; ++k) {
        c0 = add_mod(c0, mul_mod(P0, segc[k][jj], p, K), p);
        P0 = mul_mod(P0, segp[k][jj], p, K);
    }
    uint32_t* out = C + ((size_t)(q * 2 + r) * 2 + h) * HALFN * HALFN + j;
    if (seg == 0) out[0] = 0;
    c = c0;
    P = P0;
    for (int M = (seg == 0 ? 1 : m0); M < m0 + LEN; ++M) {
        const int y = (j - M) & (HALFN - 1);
        c = add_mod(c, mul_mod(g0s[y], P, p, K), p);
        P = mul_mod(P, gas[y], p, K);
        out[(size_t)M * HALFN] = c;
    }
}

// Weight tiles in the layout k1w_kernel reads: tile (r, h, g, s', c) holds at
// byte v*128 + (ai*2 + hh)*16 + kk digit v of the weight of a = 4g + ai,
// j = 32c + 16hh + kk (balanced digits, as the query limbs of the batched
// kernels). corr[a] = -2^25 sum_(s',j) W[a][j] mod p1 for the biased top
// digits (zero mod p2). One block per (16 a's, h, r, q).
// 1024 threads: the 99 KiB of LDS holds one block per CU, so the block brings
// all 16 of its waves
constexpr int K1W_TILES_NT = 1024;
__global__ __launch_bounds__(K1W_TILES_NT) void k1w_tiles_kernel(const uint32_t* __restrict__ qp, const uint32_t* __restrict__ C,
                                                        uint32_t* __restrict__ W, int nst, const DevTables* __restrict__ T) {
    __shared__ uint32_t E[16][HALFN + 1];  // E[d][j] = C[(j - a0 - d) mod L][j]; +1 spreads the rows over banks
    __shared__ uint32_t base[4][HALFN], mult[4][HALFN];
    __shared__ unsigned long long csum[16];
    const int a0 = blockIdx.x * 16, h = blockIdx.y, r = blockIdx.z & 1, q = blockIdx.z >> 1;
    const int tid = threadIdx.x;
    const uint32_t p = mod_p(r);
    const int K = mod_k(r);
    const uint32_t two26 = (uint32_t)((1ull << 26) % p);
    const uint32_t* Cb = C + ((size_t)(q * 2 + r) * 2 + h) * HALFN * HALFN;
    for (int e = tid; e < 16 * HALFN; e += K1W_TILES_NT) {
        const int M = e >> 4, d = e & 15;
        const int j = (M + a0 + d) & (HALFN - 1);
        E[d][j] = Cb[(size_t)M * HALFN + j];
    }
    const uint32_t* qv = qp + ((size_t)q * 4 * 2 + r) * N + h * HALFN;  // vector v at qv + v * 2N
    for (int y = tid; y < HALFN; y += K1W_TILES_NT) {
        const int sl = T->gpos_slot[h * HALFN + y];
        const uint32_t g0 = qv[y], g1 = qv[2 * N + y];
        const uint32_t g1p = sub_mod(g1, mul_mod(two26, g0, p, K), p);
        const uint32_t ga1p = sub_mod(T->gka[1][r][sl], mul_mod(two26, T->gka[0][r][sl], p, K), p);
        for (int st = 0; st < nst; ++st) {
            base[2 * st][y] = qv[(2 + st) * 2 * N + y];
            mult[2 * st][y] = T->qa[st][r][sl];
            base[2 * st + 1][y] = g1p;
            mult[2 * st + 1][y] = ga1p;
        }
    }
    if (tid < 16) csum[tid] = 0;
    __syncthreads();
    // 32 (g, ai, hh) combinations x K1W_TILES_NT / 32 threads, each over (s', c) items
    constexpr int NPART = K1W_TILES_NT / 32;
    const int combo = tid & 31, part = tid >> 5;
    const int gl = combo >> 3, ai = (combo >> 1) & 3, hh = combo & 1;
    const int d = gl * 4 + ai;
    const int g = blockIdx.x * 4 + gl;
    const int nsp = 2 * nst;
    unsigned long long sum = 0;
    uint32_t* Wq = W + (size_t)q * (k1w_tile_bytes(nst) / 4);
    for (int it = part; it < nsp * 32; it += NPART) {
        const int sp = it >> 5, c = it & 31;
        uint32_t w[16];
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const int j = 32 * c + 16 * hh + kk;
            const uint32_t x = add_mod(base[sp][j], mul_mod(mult[sp][j], E[d][j], p, K), p);
            sum += x;
            w[kk] = limbs8(x);
        }
        uint32_t* dst = Wq + ((((size_t)(r * 2 + h) * (HALFN / 4) + g) * nsp + sp) * 32 + c) * 128 + (ai * 2 + hh) * 4;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            u32x4 o;
            o.x = ((w[0] >> (8 * v)) & 255) | (((w[1] >> (8 * v)) & 255) << 8) | (((w[2] >> (8 * v)) & 255) << 16) | (((w[3] >> (8 * v)) & 255) << 24);
            o.y = ((w[4] >> (8 * v)) & 255) | (((w[5] >> (8 * v)) & 255) << 8) | (((w[6] >> (8 * v)) & 255) << 16) | (((w[7] >> (8 * v)) & 255) << 24);
            o.z = ((w[8] >> (8 * v)) & 255) | (((w[9] >> (8 * v)) & 255) << 8) | (((w[10] >> (8 * v)) & 255) << 16) | (((w[11] >> (8 * v)) & 255) << 24);
            o.w = ((w[12] >> (8 * v)) & 255) | (((w[13] >> (8 * v)) & 255) << 8) | (((w[14] >> (8 * v)) & 255) << 16) | (((w[15] >> (8 * v)) & 255) << 24);
            *reinterpret_cast<u32x4*>(dst + v * 32) = o;
        }
    }
    atomicAdd(&csum[d], sum);
    __syncthreads();
    if (tid < 16) {
        uint32_t cv = 0;
        if (r == 0) {
            const uint64_t m = ((1ull << 25) % p) * (csum[tid] % p) % p;
            cv = (uint32_t)((p - m) % p);
        }
        Wq[(size_t)4 * (HALFN / 4) * nsp * 32 * 128 + (r * 2 + h) * HALFN + a0 + tid] = cv;
    }
}

// The two-stream inner product: the k1d arithmetic (v_dot4 over 4 digit
// pairs, 7 weight accumulators) on the TE and IPGK1 K-steps only. A block is
// one a-group (and r, half) that streams groups of 16 shards, one wave per
// shard (8 for two queries): the weights of that a-group for every K-step
// (nsp x 32 tiles of 512 B) are staged in LDS once and read like the query
// digits of k1d, so the loop has no barrier, and the block walks every shard
// group against them (the weight tiles are fetched once per launch). Grid x =
// a-group: the blocks in flight cover all a-groups of one (r, half) as the
// shard groups pass (measured: 86% of HBM peak, against 80-84% with the
// shard group fastest, which keeps pages of all 256 shards open at once).
template <int R, int PD, int U, int NQ, int SPB>
__device__ __forceinline__ void k1w_body(const uint32_t* const* __restrict__ slabs, int nshards, const uint32_t* __restrict__ W,
                                         uint32_t* __restrict__ outb, int nst, size_t q_stride, const u32x4* wl, int g, int sg0,
                                         int sg_step, int half) {
    const int S = 3 * nst, nsp = 2 * nst;
    // the wave index as a wave-uniform value (shard, rotation and row pointers in SGPRs)
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, h = l >> 5;
    const int ai = (l & 31) >> 3;
    const int nk = nsp * 32;
    // this wave's shards: wave w of shard groups sg0, sg0 + sg_step, ...
    int sh = sg0 * SPB + wave;
    if (sh >= nshards) return;
    const int a = g * 4 + ((l & 31) >> 3);
    // the top-digit correction of this lane's output position, per query (the
    // same for every shard)
    uint32_t corr[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        corr[q] = W[q * (k1w_tile_bytes(nst) / 4) + (size_t)4 * (HALFN / 4) * nsp * 32 * 128 + (R * 2 + half) * HALFN + a];
    // K-step q: s' = q / 32 -> slab stream s = 3 st + (0: TE, 2: IPGK1), chunk
    // q % 32, i.e. slab K-step q + 32 * ((s' + 1) / 2). Each shard walks the
    // K-steps from its own starting point (the integer sums do not depend on
    // the order), so the 16 shards of a block, which stream the same a-group,
    // read at different offsets. The last PD steps of a shard prefetch the
    // first PD steps of the wave's next shard, so the stream does not drain
    // at the shard's epilogue; after the wave's last shard they reload its last
    // step.
    ARow arow = a_row(slabs[sh], R, half, g, S);
    int rot = (sh * 37) & (nk - 1);
    ATile abuf[PD];
#pragma unroll
    for (int d = 0; d < PD; ++d) {
        const int q = (d + rot) & (nk - 1);
        load_a(abuf[d], arow, R, l, q + 32 * (((q >> 5) + 1) >> 1));
    }
    const int lb = ai * 2 + h;
    for (;;) {
        const int shn = sh + sg_step * SPB;
        const bool more = shn < nshards;
        const ARow nrow = a_row(slabs[more ? shn : sh], R, half, g, S);
        const int nrot = (shn * 37) & (nk - 1);
        int acc[NQ][7];
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int w = 0; w < 7; ++w) acc[q][w] = 0;
        // U steps per loop iteration (U a multiple of PD dividing nk): the body is
        // long enough that the compiler keeps the PD-step prefetch within it
        for (int kk0 = 0; kk0 < nk; kk0 += U) {
#pragma unroll
            for (int dd = 0; dd < U; ++dd) {
                const int d = dd % PD;
                const int kk = kk0 + dd;
                v4i av[4];
                av[0] = abuf[d].d[0];
                av[1] = abuf[d].d[1];
                av[2] = abuf[d].d[2];
                av[3] = R == 0 ? top_digits(abuf[d].top) : top_digits1(abuf[d].top);
                const int kn = kk + PD;
                const bool next = kn >= nk && more;
                const int qn = kn < nk ? (kn + rot) & (nk - 1) : next ? (kn - nk + nrot) & (nk - 1) : (nk - 1 + rot) & (nk - 1);
                load_a(abuf[d], next ? nrow : arow, R, l, qn + 32 * (((qn >> 5) + 1) >> 1));
                // keep the prefetch here: without this the scheduler moves each
                // load down to its use and only one step stays in flight
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    // one weight digit vector at a time (4 registers live, not 16)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const u32x4 b = wl[(q * nk + ((kk + rot) & (nk - 1))) * 32 + v * 8 + lb];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            int t = acc[q][u + v];
                            t = __builtin_amdgcn_sdot4(av[u].x, (int)b.x, t, false);
                            t = __builtin_amdgcn_sdot4(av[u].y, (int)b.y, t, false);
                            t = __builtin_amdgcn_sdot4(av[u].z, (int)b.z, t, false);
                            t = __builtin_amdgcn_sdot4(av[u].w, (int)b.w, t, false);
                            acc[q][u + v] = t;
                        }
                    }
                }
            }
        }
        const uint32_t p = mod_p(R);
        const int K = mod_k(R);
        uint32_t cw[7];
        weight_consts(p, cw);
        const int lane8 = l & 7;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            int64_t t = corr[q];
#pragma unroll
            for (int w = 0; w < 7; ++w) {
                const int tot = acc[q][w] + __shfl_xor(acc[q][w], 32);
                t += (int64_t)tot * cw[w];
            }
            if (h == 0)
                outb[(size_t)q * q_stride + (((size_t)sh * 2 + R) * LANES + lane8) * N + half * HALFN + a] = from_i64(t, p, K);
        }
        if (!more) break;
        sh = shn;
        arow = nrow;
        rot = nrot;
    }
}

template <int PD, int U, int NQ, int SPB>
__global__ __launch_bounds__(64 * SPB) void k1w_kernel(const uint32_t* const* __restrict__ slabs, int nshards,
                                                       const uint32_t* __restrict__ W, uint32_t* __restrict__ outb, int nst,
                                                       size_t q_stride) {
    extern __shared__ u32x4 wl[];  // [NQ][nsp * 32 K-steps][32]
    const int g = blockIdx.x, sg = blockIdx.y;
    const int half = blockIdx.z & 1, r = blockIdx.z >> 1;
    const int nk = 2 * nst * 32;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const u32x4* src = reinterpret_cast<const u32x4*>(W + q * (k1w_tile_bytes(nst) / 4)) +
                           ((size_t)(r * 2 + half) * (HALFN / 4) + g) * nk * 32;
        for (int i = threadIdx.x; i < nk * 32; i += 64 * SPB) wl[q * nk * 32 + i] = src[i];
    }
    __syncthreads();
    // shard groups sg, sg + gridDim.y, ...: the weights stay in LDS across them
    if (r == 0) k1w_body<0, PD, U, NQ, SPB>(slabs, nshards, W, outb, nst, q_stride, wl, g, sg, gridDim.y, half);
    else k1w_body<1, PD, U, NQ, SPB>(slabs, nshards, W, outb, nst, q_stride, wl, g, sg, gridDim.y, half);
}

// blocks per (a-group, r, half): one block fills the a-group's weights into
// LDS once and streams every shard group against them
constexpr int K1W_SG_BLOCKS = 1;
void launch_k1w(const uint32_t* const* slabs, int nshards, const uint32_t* qp, int nq, uint32_t* scan, uint32_t* tiles,
                uint32_t* outb, int nst, const DevTables* T, hipStream_t s, hipEvent_t main_start) {
    hipLaunchKernelGGL(k1w_scan_kernel, dim3(HALFN / 16, 2, 2 * nq), dim3(256), 0, s, qp, scan, T);
    hipLaunchKernelGGL(k1w_tiles_kernel, dim3(HALFN / 16, 2, 2 * nq), dim3(K1W_TILES_NT), 0, s, qp, scan, tiles, nst, T);
    if (main_start) (void)hipEventRecord(main_start, s);
    const size_t lds = (size_t)nq * 2 * nst * 32 * 512;
    const size_t qstride = (size_t)nshards * 2 * N * LANES;
    const int gy1 = std::min(K1W_SG_BLOCKS, (nshards + 15) / 16), gy2 = std::min(K1W_SG_BLOCKS, (nshards + 7) / 8);
    if (nq == 1)
        hipLaunchKernelGGL((k1w_kernel<4, 32, 1, 16>), dim3(HALFN / 4, gy1, 4), dim3(1024), lds, s, slabs,
                           nshards, tiles, outb, nst, qstride);
    else  // two queries: both weight sets in LDS (128 KiB at full group order), 8 shards per block
        hipLaunchKernelGGL((k1w_kernel<4, 32, 2, 8>), dim3(HALFN / 4, gy2, 4), dim3(512), lds, s, slabs,
                           nshards, tiles, outb, nst, qstride);
}

void launch_relayout(const uint32_t* rows_tmp, uint32_t* slab, int L, int nst, hipStream_t s) {
    if (L == HALFN)
        hipLaunchKernelGGL(relayout_mb_kernel, dim3(32 * 32, 4, 3 * nst), dim3(256), 0, s, rows_tmp, slab, nst);
    else
        hipLaunchKernelGGL(relayout_rot_kernel, dim3(L, 3, 2 * nst), dim3(256), 0, s, rows_tmp, slab, L, nst);
}

size_t qlimb_bytes(int nq, int nst) {
    const int CT = nq > 32 ? 2 : 1;
    // limb tiles, then the per-query corrections [r*2+half][CT*32]
    return (size_t)2 * 2 * 3 * nst * 32 * CT * 4 * 1024 + (size_t)4 * CT * 32 * 4;
}

// fleets of up to this many shards take k1d_split_kernel for one or two queries
// (one full shard: 105 instead of 160 us; from two shards on k1d_kernel has
// enough waves and is as fast or faster)
constexpr int K1D_SPLIT_MAX_SHARDS = 1;
void launch_k1b(const uint32_t* const* slabs, int nshards, const uint32_t* qp, int nq, uint32_t* ql, uint32_t* outb, int nst,
                hipStream_t s, hipEvent_t main_start) {
    const int CT = nq > 32 ? 2 : 1;
    const int S = 3 * nst;
    hipLaunchKernelGGL(qlimb_kernel, dim3(32, 4 * S), dim3(128), 0, s, qp, ql, nq, nst, CT);
    hipLaunchKernelGGL(qcorr_kernel, dim3(4, CT * 32), dim3(256), 0, s, qp, ql + (size_t)4 * S * 32 * CT * 1024, nq, nst, CT * 32);
    if (main_start) (void)hipEventRecord(main_start, s);
    const size_t qstride = (size_t)nshards * 2 * N * LANES;
    const bool split = nshards <= K1D_SPLIT_MAX_SHARDS;
    if (nq == 1 && split)
        hipLaunchKernelGGL((k1d_split_kernel<1, 2>), dim3(HALFN / 4, 2, nshards * 2), dim3(256), 0, s, slabs, ql, outb, nst, qstride);
    else if (nq == 2 && split)
        hipLaunchKernelGGL((k1d_split_kernel<2, 2>), dim3(HALFN / 4, 2, nshards * 2), dim3(256), 0, s, slabs, ql, outb, nst, qstride);
    else if (nq == 1)
        hipLaunchKernelGGL((k1d_kernel<1, 2>), dim3(HALFN / 16, 2, nshards * 2), dim3(256), 0, s, slabs, ql, outb, nst, qstride);
    else if (nq == 2)
        hipLaunchKernelGGL((k1d_kernel<2, 2>), dim3(HALFN / 16, 2, nshards * 2), dim3(256), 0, s, slabs, ql, outb, nst, qstride);
    else
        launch_k1b_variant(k1b_get_choice(CT, nst), slabs, nshards, ql, outb, nst, nq, s);
}

void launch_k1b_variant(int i, const uint32_t* const* slabs, int nshards, const uint32_t* ql, uint32_t* outb, int nst, int nq,
                        hipStream_t s) {
    const size_t qstride = (size_t)nshards * 2 * N * LANES;
    hipLaunchKernelGGL(k_k1b_variants[i].fn, dim3(HALFN / 16, 2, nshards * 2), dim3(k_k1b_variants[i].threads), 0, s, slabs,
                       (const v4i*)ql, outb, nst, nq, qstride);
}

__global__ void fill_hash_kernel(uint32_t* p, size_t n, uint32_t seed) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t x = (i + 1) * 0x9E3779B97F4A7C15ull ^ seed;
        x ^= x >> 31;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 27;
        p[i] = (uint32_t)(x >> 16);
    }
}

__global__ void diff_kernel(const uint32_t* a, const uint32_t* b, size_t n, unsigned* flag) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (a[i] != b[i]) {
            atomicOr(flag, 1u);
            return;
        }
}

// Database entries from 2560-byte buckets (pack_bytes_into_coeffs,
// dns/bucket.rs:289-303): one thread per 5-byte group, read as a
// little-endian 40-bit integer v, writes the four ten-bit coefficients
// v >> 0, >> 10, >> 20, >> 30. Entries from n_valid to n_total are zero
// (the padding of a database's last shard).
__global__ void pack_buckets_kernel(const uint8_t* __restrict__ b, size_t n_valid, size_t n_total, uint16_t* __restrict__ c) {
    const size_t groups = n_total * (N / 4);
    for (size_t gi = blockIdx.x * (size_t)blockDim.x + threadIdx.x; gi < groups; gi += (size_t)gridDim.x * blockDim.x) {
        uint64_t v = 0;
        if (gi < n_valid * (N / 4)) {
            const uint8_t* p = b + gi * 5;
            v = (uint64_t)p[0] | ((uint64_t)p[1] << 8) | ((uint64_t)p[2] << 16) | ((uint64_t)p[3] << 24) | ((uint64_t)p[4] << 32);
        }
        ushort4 o;
        o.x = (uint16_t)(v & 1023);
        o.y = (uint16_t)((v >> 10) & 1023);
        o.z = (uint16_t)((v >> 20) & 1023);
        o.w = (uint16_t)((v >> 30) & 1023);
        reinterpret_cast<ushort4*>(c)[gi] = o;
    }
}

void launch_pack_buckets(const uint8_t* buckets, size_t n_valid, size_t n_total, uint16_t* coeffs, hipStream_t s) {
    hipLaunchKernelGGL(pack_buckets_kernel, dim3(1024), dim3(256), 0, s, buckets, n_valid, n_total, coeffs);
}

void launch_fill_hash(uint32_t* p, size_t n, uint32_t seed, hipStream_t s) {
    hipLaunchKernelGGL(fill_hash_kernel, dim3(2048), dim3(256), 0, s, p, n, seed);
}
void launch_diff(const uint32_t* a, const uint32_t* b, size_t n, unsigned* flag, hipStream_t s) {
    hipLaunchKernelGGL(diff_kernel, dim3(2048), dim3(256), 0, s, a, b, n, flag);
}

void launch_chain_out(const uint32_t* a_out, const uint32_t* c_out, int nst, int nlanes, uint64_t* a_coef, uint64_t* b_coef,
                      uint32_t* a17, uint32_t* b12, const DevTables* T, hipStream_t s, const unsigned* flag_src,
                      unsigned* flag_dst) {
    // one block per lane and ring element (a, then c), one inverse transform
    // per modulus of 256 threads each
    hipLaunchKernelGGL(chain_out_kernel, dim3(nlanes, 2), dim3(512), 0, s, a_out, c_out, nst, nlanes, a_coef, b_coef, a17, b12,
                       flag_src, flag_dst, T);
}

void launch_qprep(const uint64_t* qcoef, uint32_t* qs, uint32_t* qp, int nq, const DevTables* T, hipStream_t s) {
    hipLaunchKernelGGL(qprep_kernel, dim3(6, 2, nq), dim3(256), 0, s, qcoef, qs, qp, T);
}

void launch_key_cache(const uint32_t* const* src, uint32_t* const* dst, int nq, uint32_t* qs, uint32_t* qp, hipStream_t s) {
    hipLaunchKernelGGL(key_cache_kernel, dim3(N / 256, nq, 2), dim3(256), 0, s, src, dst, qs, qp);
}

void launch_k1(const uint32_t* const* slabs, int nshards, const uint32_t* qp, unsigned long long* acc, int L, int nst,
               int rows_per_block, hipStream_t s) {
    const int chunks = (L + rows_per_block - 1) / rows_per_block;
    hipLaunchKernelGGL(k1_kernel, dim3(N / 128, chunks, nshards * 2 * nst), dim3(256), 0, s, slabs, qp, acc, L, nst,
                       rows_per_block);
}

void launch_k1_split(const uint32_t* const* slabs, int nshards, const uint32_t* qpl, unsigned long long* acc, int L, int nst,
                     int rows_per_block, hipStream_t s) {
    const int chunks = (L + rows_per_block - 1) / rows_per_block;
    hipLaunchKernelGGL(k1_split_kernel, dim3(N / 128, chunks, nshards * 2 * nst), dim3(256), 0, s, slabs, qpl, acc, L, nst,
                       rows_per_block);
}

void launch_epi(const EpiLaunch& e, const DevTables* T, hipStream_t s) {
    EpiArgs A;
    A.outb = e.outb;
    A.acc = e.acc;
    A.q_stride = (size_t)e.nshards * 2 * N * LANES;
    A.a17s = e.a17s;
    A.se = e.se;
    A.se_qstride = e.se_qstride;
    A.lpr = e.lpr;
    A.lane0_mult = e.lane0_mult;
    A.bcoef = e.bcoef;
    A.diag = e.diag;
    hipLaunchKernelGGL(epi_primary_kernel, dim3(e.nshards, LANES, e.diag ? 1 : e.nq), dim3(512), 0, s, A, T);
}

// ---------------------------------------------------------------------------
// wire codec on the device (pir_wrapper.rs:84-158 bit planes, bfv.rs:51-69)
// ---------------------------------------------------------------------------
// value j of a `bits`-bit stream of N values condensed at d
__device__ uint64_t uncondense_one(const uint8_t* __restrict__ d, int bits, int j) {
    uint64_t v = 0;
    int off = 0;
    for (int i = 0; i < bits / 8; ++i) {
        v |= (uint64_t)d[off + j] << (8 * i);
        off += N;
    }
    if (bits % 8 >= 4) {
        const uint8_t b = d[off + j / 2];
        v |= (uint64_t)((j & 1) ? (b & 0xF) : (b >> 4)) << ((bits / 8) * 8);
        off += N / 2;
    }
    if (bits % 4 >= 2) {
        const uint8_t b = d[off + j / 4];
        v |= (uint64_t)((b >> (6 - 2 * (j & 3))) & 3) << ((bits / 4) * 4);
        off += N / 4;
    }
    if (bits % 2 == 1) {
        const uint8_t b = d[off + j / 8];
        v |= (uint64_t)((b >> (7 - (j & 7))) & 1) << ((bits / 2) * 2);
    }
    return v;
}

// byte k of the condensed stream of N `bits`-bit values x[]
__device__ uint8_t condense_byte(const uint32_t* __restrict__ x, int bits, int k) {
    const int nbp = bits / 8;
    if (k < nbp * N) return (uint8_t)(x[k % N] >> (8 * (k / N)));
    k -= nbp * N;
    if (bits % 8 >= 4) {
        if (k < N / 2) {
            const int sh = nbp * 8;
            return (uint8_t)((((x[2 * k] >> sh) & 0xF) << 4) | ((x[2 * k + 1] >> sh) & 0xF));
        }
        k -= N / 2;
    }
    if (bits % 4 >= 2) {
        if (k < N / 4) {
            const int sh = (bits / 4) * 4;
            uint8_t v = 0;
            for (int c = 0; c < 4; ++c) v |= (uint8_t)(((x[4 * k + c] >> sh) & 3) << (6 - 2 * c));
            return v;
        }
        k -= N / 4;
    }
    const int sh = (bits / 2) * 2;
    uint8_t v = 0;
    for (int c = 0; c < 8; ++c) v |= (uint8_t)(((x[8 * k + c] >> sh) & 1) << (7 - c));
    return v;
}

// query bytes -> ring elements [6][N] (mod_switch_decode: floor(x q / 2^B)).
// offs[v] < 0 marks an absent element (written as zero).
struct QueryLayout {
    int offs[6];
    int bits[6];
};
__global__ __launch_bounds__(256) void decode_query_kernel(const uint8_t* __restrict__ qbytes, size_t qstride, QueryLayout lay,
                                                           uint64_t* __restrict__ qcoef) {
    const int v = blockIdx.y;
    qbytes += (size_t)blockIdx.z * qstride;
    qcoef += (size_t)blockIdx.z * 6 * N;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    uint64_t out = 0;
    if (lay.offs[v] >= 0) {
        const uint64_t x = uncondense_one(qbytes + lay.offs[v], lay.bits[v], j);
        out = (uint64_t)(((unsigned __int128)x * Q) >> lay.bits[v]);
    }
    qcoef[(size_t)v * N + j] = out;
}

// reply bytes: for part p: condense<17>(a17[lane0 + p]) || condense<12>(b12[lane0 + p])
__global__ __launch_bounds__(256) void encode_reply_kernel(const uint32_t* __restrict__ a17, const uint32_t* __restrict__ b12,
                                                           int lane0, uint8_t* __restrict__ reply) {
    constexpr int AB = N * 17 / 8, BB = N * 12 / 8;
    lane0 += 3 * blockIdx.y;
    reply += (size_t)blockIdx.y * 3 * (AB + BB);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3 * (AB + BB)) return;
    const int p = k / (AB + BB);
    const int o = k % (AB + BB);
    const int li = lane0 + p;
    reply[k] = o < AB ? condense_byte(a17 + (size_t)li * N, 17, o) : condense_byte(b12 + (size_t)li * N, 12, o - AB);
}

void launch_decode_query(const uint8_t* qbytes, size_t qstride, int nq, const int offs[6], const int bits[6], uint64_t* qcoef,
                         hipStream_t s) {
    QueryLayout lay;
    for (int i = 0; i < 6; ++i) {
        lay.offs[i] = offs[i];
        lay.bits[i] = bits[i];
    }
    hipLaunchKernelGGL(decode_query_kernel, dim3(N / 256, 6, nq), dim3(256), 0, s, qbytes, qstride, lay, qcoef);
}

void launch_encode_reply(const uint32_t* a17, const uint32_t* b12, int lane0, int nq, uint8_t* reply, hipStream_t s) {
    const int total = 3 * (N * 17 / 8 + N * 12 / 8);
    hipLaunchKernelGGL(encode_reply_kernel, dim3((total + 255) / 256, nq), dim3(256), 0, s, a17, b12, lane0, reply);
}

}  // namespace dn
