// irc_screen.hip -- MFMA distance screen of the kNN pass for gfx950.
//
// The screen replaces the per-thread binned-minimum search of search_min_drift.cu:24-94
// (and the host sort of its nThreads x nA output, matlab/irc2.m:3784-3787) by:
//   1. distance tiles as a dense contraction on MFMA: with candidate rows
//      c~ = [c/s, w, |c|^2/(s^2 w)] and query rows q~ = [-2q/s, |q|^2/(s^2 w), w], one 32x32 product
//      gives 32 candidates x 32 queries of (|q|^2 + |c|^2 - 2 q.c) / s^2 directly;
//        FP32 form: v_mfma_f32_32x32x2_f32                               (s = w = 1)
//        BF16 form: v_mfma_f32_32x32x16_bf16 on a hi/lo split of every operand
//                   (x = hi + lo + r, |r| <= 2^-16 |x|), products hi*hi + hi*lo + lo*hi   (s = w = 1)
//        FP8 form:  v_mfma_scale_f32_32x32x64_f8f6f4 on a three-term e4m3 split
//                   (x = h + m + l + r, |r| <= 2^-12 |x| + 1.07 * 2^-10), eight products (all but l*l);
//                   s is a per-site power of two (max |x| / s in (16, 32]), w = 256
//      all accumulated in FP32;
//   2. a per-query list of the K' smallest screened keys (32-bit keys: truncated
//      distance bits above the candidate's stream position), kept sorted in VGPRs and
//      refreshed from an LDS append buffer with sorting networks;
//   3. a rigorous bound on |screened - exact| certifies that every candidate outside the list is
//      farther than the k-th; the candidates that can still be in the top k are re-ranked with the
//      canonical FP64 distance and ordered by (d2, spike id).  Uncertified queries go to the exact
//      kernel.
// A wavefront owns 64 queries of one (site, chunk) block: two 32-query column tiles share each
// candidate tile; a lane permutation (v_permlane32_swap) gives every lane all 32 distances of its
// own query.  The MFMAs of the next candidate tile are issued before the current tile's distances
// are selected, so the matrix core and the vector ALU work at the same time within a wavefront.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "irc_kernels.h"
#include "irc_screen.h"

#pragma clang fp contract(off)

namespace irc {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef short short8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2bf_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// OCP e4m3 (bias 7, 3 mantissa bits, subnormals m * 2^-9, largest finite 448): the value nearest v,
// ties to even, for |v| <= 448.  v / q and rint(v / q) * q are exact (q a power of two).
__device__ __forceinline__ float e4m3_round(float v) {
    const int e = (int)((__float_as_uint(v) >> 23) & 0xff) - 127;
    const float q = __uint_as_float((uint32_t)((e < -6 ? -9 : e - 3) + 127) << 23);
    return rintf(v / q) * q;
}
// the code of an e4m3 value r (exactly representable)
__device__ __forceinline__ uint32_t e4m3_code(float r) {
    const uint32_t u = __float_as_uint(r);
    const uint32_t s = (u >> 24) & 0x80u;
    const uint32_t a = u & 0x7fffffffu;
    if (a == 0) return s;
    const int e = (int)(a >> 23) - 127;
    if (e < -6) return s | (uint32_t)(__uint_as_float(a) * 512.0f);   // m * 2^-9, m = 1..7
    return s | ((uint32_t)(e + 7) << 3) | ((a >> 20) & 7u);
}
// the three-term split x = h + m + l + r: term t's code
__device__ __forceinline__ uint32_t e4m3_term(float x, int t) {
    const float h = e4m3_round(x);
    if (t == 0) return e4m3_code(h);
    const float r1 = x - h;                 // exact
    const float m = e4m3_round(r1);
    if (t == 1) return e4m3_code(m);
    return e4m3_code(e4m3_round(r1 - m));   // r1 - m exact
}
// -2 x on e4m3 codes (no overflow for |x| <= 224): one exponent step up for normal codes, twice the
// code for subnormal ones (m * 2^-9 -> 2m * 2^-9, which is also the code of 2^-6 (1 + (2m-8)/8) for
// 2m >= 8), then the sign flipped
__device__ __forceinline__ uint32_t e4m3_neg2x4(uint32_t w) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t b = (w >> (8 * i)) & 0xffu;
        const uint32_t r = (b & 0x78u) ? b + 8u : ((b & 0x80u) | ((b & 7u) << 1));
        o |= ((r ^ 0x80u) & 0xffu) << (8 * i);
    }
    return o;
}

// candidate-form value at dimension d of a column: x_d / s below D, w at D, |x|^2 / (s^2 w) at D + 1
__device__ __forceinline__ float cform(const float* x, int d, int D, float inv_s, float w, float nrm_sw) {
    return d < D ? x[d] * inv_s : (d == D ? w : (d == D + 1 ? nrm_sw : 0.0f));
}

// A workgroup packs 64 columns: the contiguous span of their rows is staged in LDS as is (16-byte
// loads when it is aligned), one thread per column forms |x|^2 (and, for FP8, looks up its site's
// scale), then one thread per (column, piece) writes the 16-byte pieces of the candidate-form rows in
// the lane-half layout of irc_screen.h: consecutive threads take consecutive columns of one piece, so
// 32 threads store 512 contiguous bytes.  The query-form rows are not stored: the screen derives them
// per item.  FS <= IRC_DMAX (the batch passes D, the per-site forms DP), so the span fits.
constexpr int PACK_COLS = 64;

template <int PREC>
__global__ __launch_bounds__(256) void k_screen_pack(int64_t n_cols, int D, int FS, int np, const float* __restrict__ F,
                                                     uint4* __restrict__ rows, float* __restrict__ norm,
                                                     float* __restrict__ feat, int DP,
                                                     const int64_t* __restrict__ seg_ptr, int n_seg,
                                                     const float* __restrict__ seg_scale) {
    __shared__ __align__(16) float srow[PACK_COLS * IRC_DMAX];   // column c's values at c * FS (float4 stores)
    __shared__ float snorm[PACK_COLS];
    __shared__ float sscale[PACK_COLS];
    const int64_t col0 = (int64_t)blockIdx.x * PACK_COLS;
    const int ncol = n_cols - col0 < PACK_COLS ? (int)(n_cols - col0) : PACK_COLS;
    // rows col0 .. col0+ncol-1 of F: (ncol-1)*FS + D contiguous floats
    const int span = (ncol - 1) * FS + D;
    const float* src = F + col0 * FS;
    if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        const int n4 = span >> 2;
        const float4* s4 = reinterpret_cast<const float4*>(src);
        float4* d4 = reinterpret_cast<float4*>(srow);
        for (int e = threadIdx.x; e < n4; e += blockDim.x) d4[e] = s4[e];
        for (int e = 4 * n4 + threadIdx.x; e < span; e += blockDim.x) srow[e] = src[e];
    } else {
        for (int e = threadIdx.x; e < span; e += blockDim.x) srow[e] = src[e];
    }
    __syncthreads();
    if ((int)threadIdx.x < ncol) {
        const float* x = srow + threadIdx.x * FS;
        float n = 0.0f;
        for (int d = 0; d < D; ++d) n = fmaf(x[d], x[d], n);
        snorm[threadIdx.x] = n;
        norm[col0 + threadIdx.x] = n;
        float s = 1.0f;
        if (PREC == SCREEN_FP8) {
            // the segment (site) of the column: largest g with seg_ptr[g] <= col
            const int64_t col = col0 + threadIdx.x;
            int lo = 0, hi = n_seg - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (seg_ptr[mid] <= col) lo = mid; else hi = mid - 1;
            }
            s = seg_scale[lo];
        }
        sscale[threadIdx.x] = s;
    }
    __syncthreads();
    if (feat) {
        // the FP32 rows of these columns, zero padded to DP: contiguous float4 stores
        const int n4 = DP / 4;
        float4* f4 = reinterpret_cast<float4*>(feat + col0 * DP);
        for (int e = threadIdx.x; e < ncol * n4; e += blockDim.x) {
            const int cc = e / n4, d0 = 4 * (e - cc * n4);
            const float* x = srow + cc * FS;
            f4[e] = make_float4(d0 < D ? x[d0] : 0.0f, d0 + 1 < D ? x[d0 + 1] : 0.0f, d0 + 2 < D ? x[d0 + 2] : 0.0f,
                                d0 + 3 < D ? x[d0 + 3] : 0.0f);
        }
    }
    const int nph = np >> 1;
    for (int e = threadIdx.x; e < PACK_COLS * np; e += blockDim.x) {
        const int c = e & (PACK_COLS - 1), hj = e / PACK_COLS;
        if (c >= ncol) continue;
        const int h = hj / nph, j = hj - h * nph;
        const int64_t col = col0 + c;
        const float* x = srow + c * FS;
        uint4 out;
        if (PREC == SCREEN_FP32) {
            const int d0 = 8 * j + 4 * h;
            const float n = snorm[c];
            out = make_uint4(__float_as_uint(cform(x, d0, D, 1.0f, 1.0f, n)),
                             __float_as_uint(cform(x, d0 + 1, D, 1.0f, 1.0f, n)),
                             __float_as_uint(cform(x, d0 + 2, D, 1.0f, 1.0f, n)),
                             __float_as_uint(cform(x, d0 + 3, D, 1.0f, 1.0f, n)));
        } else if (PREC == SCREEN_BF16) {
            const int kb = j >> 1, t = j & 1, d0 = 16 * kb + 8 * h;
            const float n = snorm[c];
            uint32_t w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                uint32_t pr = 0;
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
                    const float v = cform(x, d0 + 2 * i + hh, D, 1.0f, 1.0f, n);
                    const uint16_t hi = f2bf_rne(v);
                    const uint16_t b = t == 0 ? hi : f2bf_rne(v - bf2f(hi));
                    pr |= (uint32_t)b << (16 * hh);
                }
                w[i] = pr;
            }
            out = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
            const int t = j >> 1, d0 = 32 * h + 16 * (j & 1);
            const float s = sscale[c], inv = 1.0f / s;   // powers of two: exact
            const float nsw = snorm[c] * (inv * inv) * (1.0f / 256.0f);
            uint32_t w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                uint32_t pr = 0;
#pragma unroll
                for (int bb = 0; bb < 4; ++bb)
                    pr |= e4m3_term(cform(x, d0 + 4 * i + bb, D, inv, 256.0f, nsw), t) << (8 * bb);
                w[i] = pr;
            }
            out = make_uint4(w[0], w[1], w[2], w[3]);
        }
        rows[screen_off(col, h, np) + 32 * j] = out;
    }
}

// One workgroup per segment (site): s = the smallest power of two with max |x| <= 32 s.
__global__ __launch_bounds__(256) void k_site_scale(const int64_t* __restrict__ seg_ptr, const float* __restrict__ F,
                                                    int FS, int D, float* __restrict__ seg_scale) {
    __shared__ float red[4];
    const int g = blockIdx.x;
    const int64_t c0 = seg_ptr[g], c1 = seg_ptr[g + 1];
    float m = 0.0f;
    for (int64_t e = c0 * D + threadIdx.x; e < c1 * D; e += blockDim.x) {
        const int64_t c = e / D;
        const int d = (int)(e - c * D);
        m = fmaxf(m, fabsf(F[c * FS + d]));
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        float s = 1.0f;
        if (m > 0.0f) {
            int e;
            const float fr = frexpf(m * (1.0f / 32.0f), &e);   // m / 32 = fr * 2^e, fr in [0.5, 1)
            s = ldexpf(1.0f, fr == 0.5f ? e - 1 : e);
        }
        seg_scale[g] = s;
    }
}

// Query-form pieces from the candidate-form pieces of the query columns: below D every value is
// scaled by -2 (exact in each form), at D and D+1 the query's norm term and w are put in.
// FP32 piece j of half h: dimensions 8 j + 4 h ..
__device__ __forceinline__ float4 qform_f32(float4 v, bool valid, int d0, int D, float nrm) {
    float c[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int d = d0 + t;
        const float r = d < D ? -2.0f * c[t] : (d == D ? nrm : (d == D + 1 ? 1.0f : 0.0f));
        c[t] = valid ? r : 0.0f;
    }
    return make_float4(c[0], c[1], c[2], c[3]);
}
// BF16 piece: dimensions d0 .. d0+7 of one part (at_D, at_D1: that part of |q|^2 and of 1)
__device__ __forceinline__ uint4 qform_bf16(uint4 v, bool valid, int d0, int D, uint16_t at_D, uint16_t at_D1) {
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        uint32_t o = 0;
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            const int d = d0 + 2 * t + hh;
            const uint16_t b = (uint16_t)(w[t] >> (16 * hh));
            const uint16_t r = d < D ? f2bf_rne(-2.0f * bf2f(b)) : (d == D ? at_D : (d == D + 1 ? at_D1 : 0));
            o |= (uint32_t)r << (16 * hh);
        }
        w[t] = valid ? o : 0u;
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}
// FP8 piece: dimensions d0 .. d0+15 of one term (at_D, at_D1: codes of that term of |q|^2/(s^2 w) and w)
__device__ __forceinline__ uint4 qform_fp8(uint4 v, bool valid, int d0, int D, uint32_t at_D, uint32_t at_D1) {
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        uint32_t o = e4m3_neg2x4(w[t]);
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int d = d0 + 4 * t + bb;
            if (d >= D) {
                const uint32_t r = d == D ? at_D : (d == D + 1 ? at_D1 : 0u);
                o = (o & ~(0xffu << (8 * bb))) | (r << (8 * bb));
            }
        }
        w[t] = valid ? o : 0u;
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// keys are signed: a negative screened distance (a candidate within the error bound of the
// query) keeps its sign bit and sorts first
template <int N>
__device__ __forceinline__ void cex(int32_t& a, int32_t& b, bool up) {
    const int32_t lo = a < b ? a : b;
    const int32_t hi = a < b ? b : a;
    a = up ? lo : hi;
    b = up ? hi : lo;
}

// Batcher's odd-even merge sort (ascending): 191 compare-exchanges for 32 keys (a bitonic
// sorting network needs 240)
template <int N>
__device__ __forceinline__ void oddeven_sort(int32_t* v) {
#pragma unroll
    for (int p = 1; p < N; p <<= 1) {
#pragma unroll
        for (int k = p; k >= 1; k >>= 1) {
#pragma unroll
            for (int j = k % p; j + k < N; j += 2 * k) {
#pragma unroll
                for (int i = 0; i < k; ++i) {
                    if (i + j + k < N && (i + j) / (2 * p) == (i + j + k) / (2 * p)) cex<N>(v[i + j], v[i + j + k], true);
                }
            }
        }
    }
}

template <int N>
__device__ __forceinline__ void bitonic_clean(int32_t* v) {
#pragma unroll
    for (int j = N >> 1; j > 0; j >>= 1) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const int l = i ^ j;
            if (l > i) cex<N>(v[i], v[l], true);
        }
    }
}

__device__ __forceinline__ int ceil_log2(int n) {
    int b = 0;
    while ((1 << b) < n) ++b;
    return b;
}

// the screened distance a key stands for, negative values read as 0 (a valid lower bound, and
// non-decreasing along a sorted list); the empty-slot key reads as NaN, which fails every test
__device__ __forceinline__ float key_dist(int32_t key, uint32_t hmask) {
    const float v = __uint_as_float((uint32_t)key & hmask);
    return v < 0.0f ? 0.0f : v;
}
constexpr int32_t KEY_EMPTY = 0x7fffffff;
constexpr int32_t KEY_NONE = (int32_t)0x80000000;   // threshold of lanes without a query

// Appends `value` of every lane with `pred` to list[0 .. *n) with one atomic per wavefront (the
// lanes' order within the wavefront is kept).  Called by all 64 lanes.
__device__ __forceinline__ void wave_append(int* list, int* n, bool pred, int value) {
    const uint64_t bal = __ballot(pred);
    if (bal == 0) return;
    int base = 0;
    if (threadIdx.x % 64 == (unsigned)__builtin_ctzll(bal)) base = atomicAdd(n, __builtin_popcountll(bal));
    base = __shfl(base, __builtin_ctzll(bal), 64);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
    if (pred) list[base + rank] = value;
}

// pieces per lane half of the operand rows
template <int PREC, int KB>
struct Form {
    static constexpr int NPH = PREC == SCREEN_FP8 ? 6 : 2 * KB;
    static constexpr int NP = 2 * NPH;
};

// HALF: the kernel of the items of at most 32 queries (run in halves, below); the items of
// segment x are items[seg_lo[x] .. seg_hi[x])
template <int PREC, int KB, int KPB, bool HALF>
__global__ __launch_bounds__(64, 2) void k_screen_knn(ScreenArgs a, int* __restrict__ seg_ctr,
                                                   const int* __restrict__ seg_lo, const int* __restrict__ seg_hi) {
    constexpr int NPH = Form<PREC, KB>::NPH;
    constexpr int NP = Form<PREC, KB>::NP;
    constexpr int LS = KPB <= 32 ? 32 : 64;   // list slots in registers
    __shared__ uint32_t sbuf[65 * 64];   // rows 0..63: append buffer / list; row 64: sink
    __shared__ int s_rb[IRC_SCREEN_MAXRUNS + 1];
    __shared__ int s_ro[IRC_SCREEN_MAXRUNS + 1];
    __shared__ int s_item;
    const DevPlan& p = a.p;
    const uint4* rows = reinterpret_cast<const uint4*>(a.rows);
    const int lane = threadIdx.x;
    const int h = lane >> 5, m = lane & 31;
    // workgroups are dispatched round-robin over the 8 XCDs: blockIdx & 7 names the XCD
    const int my_seg = blockIdx.x & 7;
    // claim of an item (lane 0): own XCD's segment first (its blocks share candidate chunks in
    // that XCD's L2), then the other segments; -1 once all are empty, which ends every wave
    auto claim = [&]() -> int {
        int got = -1;
        for (int t = 0; t < 8 && got < 0; ++t) {
            const int sg = (my_seg + t) & 7;
            if (seg_ctr[sg] >= seg_hi[sg] - seg_lo[sg]) continue;
            const int idx = atomicAdd(&seg_ctr[sg], 1);
            if (seg_lo[sg] + idx < seg_hi[sg]) got = seg_lo[sg] + idx;
        }
        return got;
    };
    // the next item is claimed while the current one is finalized
    int next = 0;
    if (lane == 0) next = claim();
    for (;;) {
        __syncthreads();
        if (lane == 0) s_item = next;
        __syncthreads();
        const int it = s_item;
        if (it < 0) break;
        const ScreenItem item = a.items_kind[it];
        const int s = item.blk / p.C, c = item.blk - s * p.C;
        // the block's candidate runs (linked chunks in link order, primary part then
        // secondary), one lane per (link, part): stream offsets by a wave prefix sum
        const int lp0 = p.link_ptr[c], ne = 2 * (p.link_ptr[c + 1] - lp0);
        int nr = 0, off = 0;
        float nmax_w = 0.0f;   // largest |c|^2 of the block's candidates
        for (int e0 = 0; e0 < ne; e0 += 64) {
            const int e = e0 + lane;
            int len = 0, beg = 0;
            float mx = 0.0f;
            if (e < ne) {
                const int idx = 2 * (s * p.C + p.link_idx[lp0 + (e >> 1)]) + (e & 1);
                len = p.run_len[idx];
                if (len > 0) {
                    beg = p.run_beg[idx];
                    mx = a.run_nmax[idx];
                }
            }
            const uint64_t bal = __ballot(len > 0);
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            int incl = len;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int t = __shfl_up(incl, d, 64);
                if (lane >= d) incl += t;
            }
            if (len > 0) {
                s_rb[nr + rank] = beg;
                s_ro[nr + rank] = off + incl - len;
            }
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) mx = fmaxf(mx, __shfl_xor(mx, d, 64));
            nmax_w = fmaxf(nmax_w, mx);
            off += __shfl(incl, 63, 64);
            nr += __popcll(bal);
        }
        if (lane == 0) s_ro[nr] = off;
        __syncthreads();
        const int ncand = item.ncand;
        // FP8: the site's scale (a power of two); distances are screened in units of s^2
        const float sc = PREC == SCREEN_FP8 ? a.site_scale[s] : 1.0f;

        // ---- query operands (B) for the two 32-query column tiles: query-form pieces derived from
        // the candidate-form pieces of the query columns
        uint4 Bq[2][NPH];
#pragma unroll
        for (int T = 0; T < (HALF ? 1 : 2); ++T) {
            const int qj = 32 * T + m;
            const bool qv = qj < item.nq;
            const int qc = p.q_col[item.q0 + (qv ? qj : 0)];
            const float qn = a.norm[qc];
            const int64_t o = screen_off(qc, h, NP);
            if constexpr (PREC == SCREEN_FP32) {
#pragma unroll
                for (int j = 0; j < NPH; ++j) {
                    const uint4 v = rows[o + 32 * j];
                    const float4 f = qform_f32(make_float4(__uint_as_float(v.x), __uint_as_float(v.y),
                                                           __uint_as_float(v.z), __uint_as_float(v.w)),
                                               qv, 8 * j + 4 * h, p.D, qn);
                    Bq[T][j] = make_uint4(__float_as_uint(f.x), __float_as_uint(f.y), __float_as_uint(f.z),
                                          __float_as_uint(f.w));
                }
            } else if constexpr (PREC == SCREEN_BF16) {
                const uint16_t nh = f2bf_rne(qn), nl = f2bf_rne(qn - bf2f(nh));
#pragma unroll
                for (int j = 0; j < NPH; ++j) {
                    const int t = j & 1;
                    Bq[T][j] = qform_bf16(rows[o + 32 * j], qv, 16 * (j >> 1) + 8 * h, p.D, t ? nl : nh,
                                          t ? 0 : f2bf_rne(1.0f));
                }
            } else {
                const float inv = 1.0f / sc;
                const float nsw = qn * (inv * inv) * (1.0f / 256.0f);
#pragma unroll
                for (int j = 0; j < NPH; ++j) {
                    const int t = j >> 1;
                    Bq[T][j] = qform_fp8(rows[o + 32 * j], qv, 32 * h + 16 * (j & 1), p.D, e4m3_term(nsw, t),
                                         t == 0 ? 0x78u : 0u);   // 0x78: 256
                }
            }
        }
        // after the lane swap, lane l owns query l of the item.  An item of at most 32 queries runs
        // in halves: both MFMAs of a step take query tile 0, against two consecutive candidate
        // tiles, so lane l < 32 screens its query over the even 32-candidate tiles and lane l + 32
        // the same query over the odd ones; the two lists merge before the finalize.
        constexpr bool half = HALF;
        const bool qmine = lane < item.nq;
        const bool qact0 = half ? (lane & 31) < item.nq : qmine;
        const int q_mine = item.q0 + (qmine ? lane : 0);
        const int qc_mine = p.q_col[q_mine];
        const float nq_mine = a.norm[qc_mine];
        // retire the item's own loads here: the stream loop then waits on operand loads only
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0), expcnt and lgkmcnt unconstrained

        // the item's list depth K' is a compile-time constant: slots past it hold KEY_EMPTY
        // between merges, so the merge networks lose their dead half
        auto body = [&](auto kp_c) {
        constexpr int kprime = decltype(kp_c)::value;
        int32_t L[LS];
#pragma unroll
        for (int i = 0; i < LS; ++i) L[i] = KEY_EMPTY;
        bool qact = qact0;   // the lane screens a query (in a half item: half of its stream)
        int32_t tau = qact ? KEY_EMPTY : KEY_NONE;
        // the lane's append position as a byte offset into sbuf: row cnt of the lane's column,
        // wb = 256 cnt + 4 lane
        const int lane4 = 4 * lane;
        int wb = lane4;
        const int ib = ceil_log2(ncand > 2 ? ncand : 2);
        const uint32_t pmask = (1u << ib) - 1u, hmask = ~pmask;
        // the position mask in a vector register, so the key is one v_bfi_b32 with the scalar
        // position as its only scalar operand
        uint32_t pmask_v;
        asm volatile("v_mov_b32 %0, %1" : "=v"(pmask_v) : "s"(pmask));
        int run = 0;

        auto kth = [&]() -> int32_t { return L[kprime - 1]; };
        // merge the oldest 32 buffered keys into L, move the rest (at most 32) down.  Buffer rows
        // of a lane past its count hold KEY_EMPTY or keys that failed the threshold when written
        // (larger than L[K'-1] then and ever after, so a merge never keeps them); rows already
        // merged are cleared, so the merge reads the rows unmasked.
        auto compact_half = [&]() {
            int32_t B[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) B[i] = (int32_t)sbuf[i * 64 + lane];
            oddeven_sort<32>(B);
#pragma unroll
            for (int i = 0; i < 32; ++i) {
                const int32_t v = B[31 - i];
                L[LS - 32 + i] = L[LS - 32 + i] < v ? L[LS - 32 + i] : v;
            }
            bitonic_clean<LS>(L);
#pragma unroll
            for (int i = kprime; i < LS; ++i) L[i] = KEY_EMPTY;
            if (__any(wb > lane4 + 32 * 256)) {
#pragma unroll
                for (int i = 0; i < 32; ++i) sbuf[i * 64 + lane] = sbuf[(32 + i) * 64 + lane];
#pragma unroll
                for (int i = 32; i < 64; ++i) sbuf[i * 64 + lane] = (uint32_t)KEY_EMPTY;
            } else {
#pragma unroll
                for (int i = 0; i < 32; ++i) sbuf[i * 64 + lane] = (uint32_t)KEY_EMPTY;
            }
            wb = wb > lane4 + 32 * 256 ? wb - 32 * 256 : lane4;
            tau = qact ? kth() : KEY_NONE;
        };

        uint4 An0[NPH], An1[NPH];   // operands of the next two tiles
        // tile [pos0, pos0 + 32): `run` is the run holding pos0, with its start column, stream
        // offset and the next run's offset kept wave-uniform (scalar registers); runs that start
        // inside the tile are walked with uniform branches, each lane keeping its column
        int run_rb = 0, run_ro = 0, run_next = 0;
        auto run_enter = [&](int rr) {
            run = rr;
            run_rb = __builtin_amdgcn_readfirstlane(s_rb[rr]);
            run_ro = __builtin_amdgcn_readfirstlane(s_ro[rr]);
            run_next = rr + 1 < nr ? __builtin_amdgcn_readfirstlane(s_ro[rr + 1]) : 0x7fffffff;
        };
        auto issue = [&](int pos0, uint4 (&An)[NPH]) {
            while (pos0 >= run_next) run_enter(run + 1);
            const int pos = pos0 + m;
            int col = run_rb + (pos - run_ro);
            if (run_next < pos0 + 32) {
                for (int rr = run + 1; rr < nr; ++rr) {
                    const int ro = __builtin_amdgcn_readfirstlane(s_ro[rr]);
                    if (ro >= pos0 + 32) break;
                    if (pos >= ro) col = __builtin_amdgcn_readfirstlane(s_rb[rr]) + (pos - ro);
                }
            }
            if (pos >= ncand) col = run_rb;
            // a lane's pieces of one column are 512 bytes apart: one address, immediate offsets
            const uint4* src = rows + screen_off(col, h, NP);
#pragma unroll
            for (int j = 0; j < NPH; ++j) An[j] = src[32 * j];
        };
        // the MFMAs of one candidate tile against both query tiles
        auto mma = [&](const uint4 (&Ac)[NPH], f32x16& c0, f32x16& c1) {
            c0 = f32x16{};
            c1 = f32x16{};
            if constexpr (PREC == SCREEN_BF16) {
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) {
                    const bf16x8 ah = __builtin_bit_cast(bf16x8, Ac[2 * kb]);
                    const bf16x8 al = __builtin_bit_cast(bf16x8, Ac[2 * kb + 1]);
                    const bf16x8 h0 = __builtin_bit_cast(bf16x8, Bq[0][2 * kb]);
                    const bf16x8 h1 = __builtin_bit_cast(bf16x8, Bq[HALF ? 0 : 1][2 * kb]);
                    const bf16x8 l0 = __builtin_bit_cast(bf16x8, Bq[0][2 * kb + 1]);
                    const bf16x8 l1 = __builtin_bit_cast(bf16x8, Bq[HALF ? 0 : 1][2 * kb + 1]);
                    c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, h0, c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, h1, c1, 0, 0, 0);
                    c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, l0, c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, l1, c1, 0, 0, 0);
                    c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, h0, c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, h1, c1, 0, 0, 0);
                }
            } else if constexpr (PREC == SCREEN_FP32) {
#pragma unroll
                for (int g = 0; g < NPH; ++g) {
                    const uint4 av = Ac[g], b0 = Bq[0][g], b1 = Bq[HALF ? 0 : 1][g];
                    c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.x), __uint_as_float(b0.x), c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.x), __uint_as_float(b1.x), c1, 0, 0, 0);
                    c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.y), __uint_as_float(b0.y), c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.y), __uint_as_float(b1.y), c1, 0, 0, 0);
                    c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.z), __uint_as_float(b0.z), c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.z), __uint_as_float(b1.z), c1, 0, 0, 0);
                    c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.w), __uint_as_float(b0.w), c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.w), __uint_as_float(b1.w), c1, 0, 0, 0);
                }
            } else {
                auto v8 = [](const uint4& x, const uint4& y) {
                    return i32x8{(int)x.x, (int)x.y, (int)x.z, (int)x.w, (int)y.x, (int)y.y, (int)y.z, (int)y.w};
                };
                const i32x8 Ah = v8(Ac[0], Ac[1]), Am = v8(Ac[2], Ac[3]), Al = v8(Ac[4], Ac[5]);
#define IRC_F8(A, B, C) C = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A, B, C, 0, 0, 0, 127, 0, 127)
#pragma unroll
                for (int T = 0; T < 2; ++T) {
                    const int TB = HALF ? 0 : T;
                    const i32x8 Bh = v8(Bq[TB][0], Bq[TB][1]), Bm = v8(Bq[TB][2], Bq[TB][3]), Bl = v8(Bq[TB][4], Bq[TB][5]);
                    f32x16& cc = T == 0 ? c0 : c1;
                    IRC_F8(Am, Bl, cc);
                    IRC_F8(Al, Bm, cc);
                    IRC_F8(Am, Bm, cc);
                    IRC_F8(Ah, Bl, cc);
                    IRC_F8(Al, Bh, cc);
                    IRC_F8(Ah, Bm, cc);
                    IRC_F8(Am, Bh, cc);
                    IRC_F8(Ah, Bh, cc);
                }
#undef IRC_F8
            }
        };
        // half items: query tile 0 against one candidate tile
        auto mma_one = [&](const uint4 (&Ac)[NPH], f32x16& cc) {
            cc = f32x16{};
            if constexpr (PREC == SCREEN_BF16) {
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) {
                    const bf16x8 ah = __builtin_bit_cast(bf16x8, Ac[2 * kb]);
                    const bf16x8 al = __builtin_bit_cast(bf16x8, Ac[2 * kb + 1]);
                    const bf16x8 h0 = __builtin_bit_cast(bf16x8, Bq[0][2 * kb]);
                    const bf16x8 l0 = __builtin_bit_cast(bf16x8, Bq[0][2 * kb + 1]);
                    cc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, h0, cc, 0, 0, 0);
                    cc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, l0, cc, 0, 0, 0);
                    cc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, h0, cc, 0, 0, 0);
                }
            } else if constexpr (PREC == SCREEN_FP32) {
#pragma unroll
                for (int g = 0; g < NPH; ++g) {
                    const uint4 av = Ac[g], b0 = Bq[0][g];
                    cc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.x), __uint_as_float(b0.x), cc, 0, 0, 0);
                    cc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.y), __uint_as_float(b0.y), cc, 0, 0, 0);
                    cc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.z), __uint_as_float(b0.z), cc, 0, 0, 0);
                    cc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av.w), __uint_as_float(b0.w), cc, 0, 0, 0);
                }
            } else {
                auto v8 = [](const uint4& x, const uint4& y) {
                    return i32x8{(int)x.x, (int)x.y, (int)x.z, (int)x.w, (int)y.x, (int)y.y, (int)y.z, (int)y.w};
                };
                const i32x8 Bh = v8(Bq[0][0], Bq[0][1]), Bm = v8(Bq[0][2], Bq[0][3]), Bl = v8(Bq[0][4], Bq[0][5]);
                const i32x8 Ah = v8(Ac[0], Ac[1]), Am = v8(Ac[2], Ac[3]), Al = v8(Ac[4], Ac[5]);
#define IRC_F8(A, B, C) C = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A, B, C, 0, 0, 0, 127, 0, 127)
                IRC_F8(Am, Bl, cc);
                IRC_F8(Al, Bm, cc);
                IRC_F8(Am, Bm, cc);
                IRC_F8(Ah, Bl, cc);
                IRC_F8(Al, Bh, cc);
                IRC_F8(Ah, Bm, cc);
                IRC_F8(Am, Bh, cc);
                IRC_F8(Ah, Bh, cc);
#undef IRC_F8
            }
        };
        // the 32 distances of the lane's query in tile [base, base + 32) from (c0, c1):
        // lane l < 32 gets query l of tile 0, lane l >= 32 query l-32 of tile 1
        auto swap_tile = [&](const f32x16& c0, const f32x16& c1, uint32_t (&u)[32]) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(c0[r]), __float_as_uint(c1[r]),
                                                                false, false);
                u[2 * r] = sw[0];
                u[2 * r + 1] = sw[1];
            }
        };
        auto select_tile = [&](const uint32_t (&u)[32], int base, auto&& consume) {
            auto tile_out = [&](auto full) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        const int cpos = base + 8 * (r >> 2) + 4 * hh + (r & 3);
                        consume((int)u[2 * r + hh], cpos, decltype(full)::value || cpos < ncand);
                    }
                }
            };
            if (base + 32 <= ncand) tile_out(std::true_type{});
            else tile_out(std::false_type{});
        };
        // half items: lane half h holds candidates base + 32 h + (0..31); the key carries the
        // position less 32 h (wave-uniform), put right when the halves merge
        auto select_half = [&](const uint32_t (&u)[32], int base, auto&& consume) {
            if (base + 64 <= ncand) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh)
                        consume((int)u[2 * r + hh], base + 8 * (r >> 2) + 4 * hh + (r & 3), true);
                }
            } else {
                const int lim = ncand - 32 * h;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        const int cpos = base + 8 * (r >> 2) + 4 * hh + (r & 3);
                        consume((int)u[2 * r + hh], cpos, cpos < lim);
                    }
                }
            }
        };
        // one pass over the candidate stream, two tiles per step: the MFMAs of a tile, the loads of
        // the tile two ahead (into the operand registers just consumed), then the tile's selection.
        // Two wavefronts per SIMD hide each other's MFMA and load latency.  A tile past the stream
        // end is multiplied (its rows are valid memory) but not selected.
        auto stream = [&](auto&& consume, auto&& after_tile) {
            run_enter(0);
            issue(0, An0);
            issue(32, An1);
            f32x16 c0, c1;
            if constexpr (half) {
                for (int base = 0; base < ncand; base += 64) {
                    mma_one(An0, c0);
                    issue(base + 64, An0);
                    mma_one(An1, c1);
                    issue(base + 96, An1);
                    uint32_t u[32];
                    swap_tile(c0, c1, u);
                    select_half(u, base, consume);
                    after_tile();
                }
                return;
            }
            auto tile = [&](int base, uint4 (&Ac)[NPH]) {
                mma(Ac, c0, c1);
                issue(base + 64, Ac);
                if (base >= ncand) return;
                uint32_t u[32];
                swap_tile(c0, c1, u);
                select_tile(u, base, consume);
                after_tile();
            };
            for (int base = 0; base < ncand; base += 64) {
                tile(base, An0);
                tile(base + 32, An1);
            }
        };

        // ---- keys (the screened distance's bits, sign kept, above the candidate's stream
        // position) below the threshold go to the lane's append buffer; full buffers merge into L
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 64; ++i) sbuf[i * 64 + lane] = (uint32_t)KEY_EMPTY;
        stream(
            [&](int sbits, int cpos, bool valid) {
                // key: one bit-field insert (low ib bits from the wave-uniform position, the rest
                // from the value); the lane's next buffer row (< 64: at most 32 buffered when a
                // tile starts) is scratch until counted: the key is stored unconditionally and the
                // position advances by one row when it passes the threshold.  The compare and the
                // select do not depend on the position, so each distance adds one dependent add.
                // (positions past the stream end -- wave-uniform -- write nothing: the buffer rows
                // past a lane's count must only hold keys that can never enter the list)
                if (!valid) return;
                int32_t key;
                asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(key) : "v"(pmask_v), "s"(cpos), "v"(sbits));
                *reinterpret_cast<int32_t*>(reinterpret_cast<char*>(sbuf) + wb) = key;
                wb += key < tau ? 256 : 0;
            },
            [&]() {
                if (__any(wb > lane4 + 32 * 256)) compact_half();
            });
        // flush the buffers; a half item then hands lane l + 32's list (positions put right) to
        // lane l, which merges it as kprime buffered keys, and flushes again
#pragma nounroll
        for (int pass = 0; pass < (half ? 2 : 1); ++pass) {
            if (pass == 1) {
                __syncthreads();
                if (h == 1) {
#pragma unroll
                    for (int i = 0; i < kprime; ++i)
                        sbuf[i * 64 + m] = (uint32_t)(L[i] == KEY_EMPTY ? KEY_EMPTY : L[i] + 32);
                }
                __syncthreads();
                qact = qmine;
                wb = h == 0 ? lane4 + 256 * kprime : lane4;
            }
            while (__any(wb > lane4)) compact_half();
        }
        if (lane == 0) next = claim();

        // ---- finalize: certification and the list of candidates the FP64 re-rank must see
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kprime; ++i) sbuf[i * 64 + lane] = (uint32_t)L[i];
        // rows kprime..63 hold a byte table: the run of every stream position (runs < 256), so a
        // kept candidate's column is two independent LDS reads away
        constexpr int TBL = (64 - kprime) * 64 * 4;
        uint8_t* const tbl = reinterpret_cast<uint8_t*>(sbuf + kprime * 64);
        const bool use_tbl = ncand <= TBL;
        auto run_of = [&](int pos) {
            int lo = 0, hi = nr - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_ro[mid] <= pos) lo = mid; else hi = mid - 1;
            }
            return lo;
        };
        if (use_tbl) {
            for (int rr = 0; rr < nr; ++rr) {
                const int e = s_ro[rr + 1];
                for (int t = s_ro[rr] + lane; t < e; t += 64) tbl[t] = (uint8_t)rr;
            }
        }
        __syncthreads();
        bool to_exact = false;
        if (qmine) {
            const int k = p.k;
            const int nl = ncand < kprime ? ncand : kprime;
            const int kk = ncand < k ? ncand : k;
            // the bound in real units, then in the screen's units (s^2; powers of two: exact)
            const double s2 = (double)sc * (double)sc;
            const double eps = a.eta * ((double)nq_mine + (double)nmax_w) +
                               a.eta2 * (double)sc * (sqrt((double)nq_mine) + sqrt((double)nmax_w)) +
                               a.eta3 * s2 + 1e-30;
            const double eps_s = eps / s2;
            const double qf = ldexp(1.0, -(22 - ib));
            const double dkt = key_dist((int32_t)sbuf[(kk - 1) * 64 + lane], hmask);
            // every one of the first kk kept candidates has d^2 <= U (screen units)
            const double U = dkt * (1.0 + qf) + eps_s;
            // every candidate outside L has a key at least L[K'-1]
            const int32_t kexcl = (int32_t)sbuf[(kprime - 1) * 64 + lane];
            const bool ok = ncand <= kprime || ((double)key_dist(kexcl, hmask) - eps_s > U);
            if (ok) {
                // the re-rank needs every kept candidate whose lower bound reaches U
                // (the list is sorted and its screened distances nondecreasing, so the test
                // holds for a prefix of the slots past kk: count it from the registers)
                int r = kk;
#pragma unroll
                for (int j = 0; j < kprime; ++j)
                    r += (j >= kk && j < nl && (double)key_dist(L[j], hmask) - eps_s <= U) ? 1 : 0;
                for (int j = 0; j < r; ++j) {
                    const int32_t key = (int32_t)sbuf[j * 64 + lane];
                    const int pos = (int)(key & pmask);
                    const int lo = use_tbl ? (int)tbl[pos] : run_of(pos);
                    a.cand[(int64_t)j * a.n_q + q_mine] =
                        make_int2(s_rb[lo] + (pos - s_ro[lo]), __float_as_int((float)((double)key_dist(key, hmask) * s2)));
                }
                a.cand_n[q_mine] = r | (kk << 16);
                a.cand_eps[q_mine] = eps;
            } else {
                a.cand_n[q_mine] = -1;
                to_exact = true;
            }
        }
        wave_append(a.fallback, a.n_fallback, to_exact, q_mine);
        };
        // long candidate streams keep a deeper list: their keys carry more position bits (coarser
        // distances) and their neighbour distances are closer together
        if constexpr (LS == 64 && KPB < 64) {
            if (ncand > a.large_ncand) {
                body(std::integral_constant<int, 64>{});
                continue;
            }
        }
        body(std::integral_constant<int, KPB>{});
    }
}

// FP64 re-rank of the certified lists.  One wavefront takes 64 consecutive queries (lane l =
// query l) and walks the list slots in ascending screened order.  For each slot the 64
// candidate rows are gathered cooperatively -- consecutive lanes fetch consecutive 16-byte
// pieces of a row, straight into LDS (global_load_lds), so a load instruction touches a few
// rows instead of 64 -- and each lane then reads its own row back and runs the canonical FP64
// chain against its query row, held in registers.  The next slot's gather is in flight while
// the chain runs.  Each lane keeps a window of W exact distances in (d^2, id) order and emits
// its head once the head is below the next slot's lower bound d_j - eps (no later slot can
// precede it).  A window overflow (more than W mutually ambiguous candidates) sends the query
// on: the first pass keeps a short window (its predicated insertion and shifts are most of the
// VALU work) and hands overflows to a second pass over that list with the long window; what
// overflows there goes to the exact kernel.  (The pass is bound by the gathers' path through
// the vector memory unit, not by the FP64 arithmetic: a first pass in FP32 with a certified
// margin halves the VALU instructions and leaves the time unchanged.)
constexpr int WIN1 = 8;    // first pass: every certified query
constexpr int WIN2 = 16;   // second pass: the first pass's overflows

// one group of up to 64 queries (lane l = query q, valid when qv); overflows are appended to spill
struct FinOut {
    float* rho;
    float* dk;
    int64_t* knn;
    double* d2head;   // [IRC_D2HEAD][n_q]
    int first;        // the first pass visits every query: it marks the ones it does not finish
    float* rho_col;   // rho by column, for the queries' own columns
};

template <int NV, int W>
__device__ __forceinline__ void rerank_group(const DevPlan& p, int64_t n_q, int64_t q, bool qv,
                                             const int* __restrict__ cand_n, const int2* __restrict__ cand,
                                             const double* __restrict__ cand_eps,
                                             int* __restrict__ knn_col, double* __restrict__ dk2,
                                             int* __restrict__ spill, int* __restrict__ n_spill, float4* stage,
                                             int* skey, int* obuf, const FinOut& fin) {
    const int l = threadIdx.x;
    const int rn = qv ? cand_n[q] : -1;
    const int r = rn < 0 ? 0 : (rn & 0xffff);
    int rmax = r;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) rmax = max(rmax, __shfl_xor(rmax, off, 64));
    if (rmax == 0) {
        if (fin.d2head && fin.first && qv)
            for (int j = 0; j < IRC_D2HEAD; ++j) fin.d2head[(int64_t)j * n_q + q] = __builtin_nan("");
        return;
    }
    const int qc = p.q_col[qv ? q : 0];
    float4 qr[NV];
    {
        const float4* q4 = reinterpret_cast<const float4*>(p.feat + (int64_t)qc * p.DP);
#pragma unroll
        for (int i = 0; i < NV; ++i) qr[i] = q4[i];
    }
    // slot j's (column, screened distance); lanes past their own list fetch their query's row (a
    // valid address) and ignore it
    auto slot = [&](int j) { return j < r ? cand[(int64_t)j * n_q + q] : make_int2(qc, 0); };
    // slot j's rows and ids, straight into LDS
    auto gather = [&](int colj) {
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(p.col_key + colj),
                                         (__attribute__((address_space(3))) void*)skey, 4, 0, 0);
        int crow[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) crow[i] = __shfl(colj, (i * 64 + l) / NV, 64);
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int piece = i * 64 + l;
            const int part = piece - (piece / NV) * NV;
            const float* src = p.feat + (int64_t)crow[i] * p.DP + 4 * part;
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                             (__attribute__((address_space(3))) void*)(stage + i * 64), 16, 0, 0);
        }
    };
    const int k = p.k;
    const int kk = rn >> 16;
    const double eps = qv && rn >= 0 ? cand_eps[q] : 0.0;
    double wt[W];
    int wk[W], wc[W];
    double dh[IRC_D2HEAD] = {};   // exact d^2 of the first emitted entries
    int wn = 0, out = 0;
    double last_t = 0.0;
    int last_c = -1;
    bool ok = rn >= 0;
    int colj = slot(0).x;
    int2 sn = slot(1);
    gather(colj);
    for (int j = 0; j < rmax; ++j) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // the gather of slot j has landed in LDS
        float4 b[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) b[i] = stage[l * NV + i];
        const int kc = skey[l];
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");   // LDS read before the next gather lands
        const bool act = ok && j < r && out < kk;
        const int coln = sn.x;
        const float dn = __int_as_float(sn.y);   // the next slot's screened distance
        sn = slot(j + 2);
        if (j + 1 < rmax) gather(coln);
        if (act) {
#pragma unroll
            for (int i = 0; i < NV; ++i)
                asm volatile("" : "+v"(qr[i].x), "+v"(qr[i].y), "+v"(qr[i].z), "+v"(qr[i].w));
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                double d;
                d = (double)qr[i].x - (double)b[i].x; t = fma(d, d, t);
                d = (double)qr[i].y - (double)b[i].y; t = fma(d, d, t);
                d = (double)qr[i].z - (double)b[i].z; t = fma(d, d, t);
                d = (double)qr[i].w - (double)b[i].w; t = fma(d, d, t);
            }
            const double lb = j + 1 < r ? (double)dn - eps : __builtin_inf();
            if (wn == W) {
                ok = false;
            } else {
                bool placed = false;
#pragma unroll
                for (int i = W - 1; i >= 0; --i) {
                    if (i <= wn && !placed) {
                        if (i > 0 && irc_less(t, kc, wt[i - 1], wk[i - 1])) {
                            wt[i] = wt[i - 1]; wk[i] = wk[i - 1]; wc[i] = wc[i - 1];
                        } else {
                            wt[i] = t; wk[i] = kc; wc[i] = colj;
                            placed = true;
                        }
                    }
                }
                ++wn;
                while (wn > 0 && out < kk && wt[0] < lb) {
#pragma unroll
                    for (int i = 0; i < IRC_D2HEAD; ++i) dh[i] = out == i ? wt[0] : dh[i];
                    if (k <= 32) obuf[out * 64 + l] = wc[0];
                    else knn_col[(int64_t)out * n_q + q] = wc[0];
                    last_t = wt[0];
                    last_c = wc[0];
                    ++out;
#pragma unroll
                    for (int i = 0; i < W - 1; ++i) {
                        if (i + 1 < wn) { wt[i] = wt[i + 1]; wk[i] = wk[i + 1]; wc[i] = wc[i + 1]; }
                    }
                    --wn;
                }
            }
        }
        colj = coln;
    }
    const bool done = qv && rn >= 0 && ok && out == kk;
    if (done) {
        for (int j = kk; j < k; ++j) {
            if (k <= 32) obuf[j * 64 + l] = last_c;
            else knn_col[(int64_t)j * n_q + q] = last_c;
        }
        dk2[q] = last_t;
    }
    wave_append(spill, n_spill, qv && rn >= 0 && !done, (int)q);
    // the buffered lists leave slot by slot: 64 queries per store
    if (k <= 32) {
        for (int j = 0; j < k; ++j)
            if (done) knn_col[(int64_t)j * n_q + q] = obuf[j * 64 + l];
        // batch form: rho, d_k and the reported ids of the finished queries (the finish of
        // rho_paged_site_, matlab/irc2.m:3669-3676): the list's columns become spike ids in place,
        // then each finished query's row of k ids leaves contiguously (consecutive lanes store
        // consecutive entries of one row)
        if (fin.d2head) {
            // padded entries repeat the last one
            for (int j = 0; j < IRC_D2HEAD; ++j) {
                if (done) fin.d2head[(int64_t)j * n_q + q] = j < kk ? dh[j] : last_t;
                else if (fin.first && qv) fin.d2head[(int64_t)j * n_q + q] = __builtin_nan("");
            }
        }
        if (fin.rho || fin.knn) {
            const int slot = done ? p.q_out[q] : 0;
            if (done && fin.rho) fin.rho[slot] = (float)(1.0 / sqrt(last_t));
            if (done && fin.rho_col) fin.rho_col[qc] = (float)(1.0 / sqrt(last_t));
            if (done && fin.dk) fin.dk[slot] = (float)sqrt(last_t);
            if (fin.knn) {
                if (done)
                    for (int j = 0; j < k; ++j) obuf[j * 64 + l] = p.col_ret[obuf[j * 64 + l]];
                __syncthreads();
                const uint64_t dm = __ballot(done);
                for (int e = l; e < k * 64; e += 64) {
                    const int c = e / k, j = e - c * k;
                    const int sc = __shfl(slot, c, 64);
                    if ((dm >> c) & 1ull) fin.knn[(int64_t)sc * k + j] = obuf[j * 64 + c];
                }
            }
        }
        __syncthreads();   // obuf is reused by the next group
    }
}

template <int NV>
__global__ __launch_bounds__(64) void k_rerank(DevPlan p, int64_t n_q, const int* __restrict__ cand_n,
                                              const int2* __restrict__ cand,
                                              const double* __restrict__ cand_eps, int* __restrict__ knn_col,
                                              double* __restrict__ dk2, int* __restrict__ spill,
                                              int* __restrict__ n_spill, FinOut fin) {
    __shared__ float4 stage[64 * NV];   // [row][piece]: the slot's 64 candidate rows
    __shared__ int skey[64];            // their spike ids
    extern __shared__ int obuf[];       // [k][64] emitted columns when k <= 32 (else written directly)
    // workgroups are dispatched round-robin over the 8 XCDs: give each XCD a contiguous eighth
    // of the query groups, so the candidate rows its groups share stay in its L2
    const unsigned G = gridDim.x, x = blockIdx.x & 7, per = (G + 7) >> 3, lo8 = G & 7;
    const unsigned grp = (lo8 == 0 || x < lo8) ? x * per + (blockIdx.x >> 3)
                                               : lo8 * per + (x - lo8) * (per - 1) + (blockIdx.x >> 3);
    const int64_t q = (int64_t)grp * 64 + threadIdx.x;
    rerank_group<NV, WIN1>(p, n_q, q, q < n_q, cand_n, cand, cand_eps, knn_col, dk2, spill, n_spill, stage, skey,
                           obuf, fin);
}

// second pass: the queries of list[0 .. *n_list), 64 per group, groups strided over the grid
template <int NV>
__global__ __launch_bounds__(64) void k_rerank_list(DevPlan p, int64_t n_q, const int* __restrict__ list,
                                                   const int* __restrict__ n_list, const int* __restrict__ cand_n,
                                                   const int2* __restrict__ cand,
                                                   const double* __restrict__ cand_eps, int* __restrict__ knn_col,
                                                   double* __restrict__ dk2, int* __restrict__ fallback,
                                                   int* __restrict__ n_fallback, FinOut fin) {
    __shared__ float4 stage[64 * NV];
    __shared__ int skey[64];
    extern __shared__ int obuf[];
    const int64_t n = *n_list;
    for (int64_t g = blockIdx.x; g * 64 < n; g += gridDim.x) {
        const int64_t i = g * 64 + threadIdx.x;
        const bool qv = i < n;
        rerank_group<NV, WIN2>(p, n_q, qv ? (int64_t)list[i] : 0, qv, cand_n, cand, cand_eps, knn_col, dk2,
                               fallback, n_fallback, stage, skey, obuf, fin);
    }
}

// one wavefront per run
__global__ void k_run_nmax(int64_t n_runs, const int* __restrict__ run_beg, const int* __restrict__ run_len,
                           const float* __restrict__ norm, float* __restrict__ out) {
    const int64_t rr = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int l = threadIdx.x & 63;
    if (rr >= n_runs) return;
    const int beg = run_beg[rr], len = run_len[rr];
    float m = 0.0f;
    for (int i = l; i < len; i += 64) m = fmaxf(m, norm[beg + i]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (l == 0) out[rr] = m;
}

__global__ void k_fill_int(int* p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Lane layout probe of the MFMA forms and the lane swap the screen relies on.
// out[0..1023]   : 32x32x16 bf16 with A[m][k] = (k == 0) * (m + 1), B[k][n] = (k == 0) * 64 * (n + 1)
//                  -> acc[r] of lane l, stored at out[r * 64 + l]
// out[1024..2047]: 32x32x2 f32 with the same A, B
// out[2048..4095]: permlane32_swap(acc0, acc1) of the bf16 result with acc1 = acc0 + 0.5
//                  (first and second result at out[2048 + r*64 + l], out[3072 + r*64 + l])
// out[4096..5119]: 32x32x64 fp8 (e4m3, unit scales) with every byte of lane (m, 0) = code 4 m + 1 and of
//                  lane (n, 0) (B) = code (8 + n) | 0x80 * (n & 1); lanes of half 1 zero
// out[5120..6143]: the same with the halves exchanged (half 1 carries the codes, half 0 zero)
__global__ void k_mfma_probe(float* out) {
    const int l = threadIdx.x;
    const int h = l >> 5, m = l & 31;
    short8 va, vb;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = 8 * h + j;
        va[j] = (short)(__float_as_uint(k == 0 ? (float)(m + 1) : 0.0f) >> 16);
        vb[j] = (short)(__float_as_uint(k == 0 ? 64.0f * (m + 1) : 0.0f) >> 16);
    }
    f32x16 acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, va), __builtin_bit_cast(bf16x8, vb), acc,
                                                  0, 0, 0);
    f32x16 acc2 = {};
    acc2 = __builtin_amdgcn_mfma_f32_32x32x2f32(h == 0 ? (float)(m + 1) : 0.0f, h == 0 ? 64.0f * (m + 1) : 0.0f,
                                                acc2, 0, 0, 0);
    f32x16 acc3 = {}, acc4 = {};
    {
        const uint32_t ca = (uint32_t)(4 * m + 1), cb = (uint32_t)((8 + m) | (m & 1 ? 0x80 : 0));
        const int wa = (int)(ca * 0x01010101u), wb = (int)(cb * 0x01010101u);
        const i32x8 A0 = h == 0 ? i32x8{wa, wa, wa, wa, wa, wa, wa, wa} : i32x8{};
        const i32x8 B0 = h == 0 ? i32x8{wb, wb, wb, wb, wb, wb, wb, wb} : i32x8{};
        const i32x8 A1 = h == 1 ? i32x8{wa, wa, wa, wa, wa, wa, wa, wa} : i32x8{};
        const i32x8 B1 = h == 1 ? i32x8{wb, wb, wb, wb, wb, wb, wb, wb} : i32x8{};
        acc3 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A0, B0, acc3, 0, 0, 0, 127, 0, 127);
        acc4 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A1, B1, acc4, 0, 0, 0, 127, 0, 127);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        out[r * 64 + l] = acc[r];
        out[1024 + r * 64 + l] = acc2[r];
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[r]), __float_as_uint(acc[r] + 0.5f),
                                                        false, false);
        out[2048 + r * 64 + l] = __uint_as_float(sw[0]);
        out[3072 + r * 64 + l] = __uint_as_float(sw[1]);
        out[4096 + r * 64 + l] = acc3[r];
        out[5120 + r * 64 + l] = acc4[r];
    }
}

// Stable split of every XCD segment [seg_beg[x], seg_beg[x + 1]) of items into its items of more
// than 32 queries, then those of at most 32: out[seg_beg[x] .. seg_half[x]) and
// out[seg_half[x] .. seg_beg[x + 1]).  One workgroup per segment.
__global__ __launch_bounds__(1024) void k_split_items(const ScreenItem* __restrict__ in, ScreenItem* __restrict__ out,
                                                     const int* __restrict__ seg_beg, int* __restrict__ seg_half) {
    __shared__ int s_wave[16];
    __shared__ int s_tot;
    const int x = blockIdx.x, t = threadIdx.x, w = t >> 6, l = t & 63;
    const int b = seg_beg[x], e = seg_beg[x + 1];
    // pass 1: items of the full kind
    int nf = 0;
    for (int i = b + t; i < e; i += 1024) nf += in[i].nq > 32;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) nf += __shfl_xor(nf, off, 64);
    if (l == 0) s_wave[w] = nf;
    __syncthreads();
    if (t == 0) {
        int tot = 0;
        for (int i = 0; i < 16; ++i) tot += s_wave[i];
        s_tot = tot;
        seg_half[x] = b + tot;
    }
    __syncthreads();
    const int half0 = b + s_tot;
    // pass 2: ranks within each kind, chunk by chunk
    int done_f = 0, done_h = 0;
    for (int c0 = b; c0 < e; c0 += 1024) {
        const int i = c0 + t;
        const bool v = i < e;
        ScreenItem it{};
        if (v) it = in[i];
        const bool full = v && it.nq > 32;
        const uint64_t bf = __ballot(full), bv = __ballot(v);
        const int rf = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bf >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bf, 0u));
        const int rv = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bv >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bv, 0u));
        __syncthreads();
        if (l == 0) s_wave[w] = __popcll(bf) | (__popcll(bv) << 16);
        __syncthreads();
        int pf = 0, pv = 0;
        for (int j = 0; j < w; ++j) {
            pf += s_wave[j] & 0xffff;
            pv += s_wave[j] >> 16;
        }
        int cf = 0, cv = 0;
        for (int j = 0; j < 16; ++j) {
            cf += s_wave[j] & 0xffff;
            cv += s_wave[j] >> 16;
        }
        if (v) {
            const int kf = pf + rf, kh = (pv + rv) - kf;   // full / half items before this one
            if (full) out[b + done_f + kf] = it;
            else out[half0 + done_h + kh] = it;
        }
        done_f += cf;
        done_h += cv - cf;
    }
    if (x == 7 && t == 0) seg_half[8] = e;
}

template <int PREC, int KB>
hipError_t launch_kb(hipStream_t st, const ScreenArgs& a, int grid, int* ctr, const int* seg_beg) {
    // ctr: [0, 8) counters of the full kernel, [8, 16) of the half kernel, [16, 25) split offsets
    int* seg_half = ctr + 16;
    k_zero_ints<<<1, 64, 0, st>>>(ctr, 16);
    k_split_items<<<8, 1024, 0, st>>>(a.items, a.items_kind, seg_beg, seg_half);
#define IRC_SK(KPB)                                                                                         \
    k_screen_knn<PREC, KB, KPB, false><<<grid, 64, 0, st>>>(a, ctr, seg_beg, seg_half);                     \
    k_screen_knn<PREC, KB, KPB, true><<<grid, 64, 0, st>>>(a, ctr + 8, seg_half, seg_beg + 1)
    switch (a.kprime) {
        case 32: IRC_SK(32); break;
        case 40: IRC_SK(40); break;
        case 48: IRC_SK(48); break;
        case 64: IRC_SK(64); break;
        default: return hipErrorInvalidValue;
    }
#undef IRC_SK
    return hipGetLastError();
}

}  // namespace

int screen_kp(int D) { return (D + 2 + 15) / 16 * 16; }

void screen_eta(int prec, int D, int KP, double* eta, double* eta2, double* eta3) {
    // unit roundoff of the FP32 accumulation taken as 2^-23 (twice FP32's) so the bound also covers
    // truncating accumulation inside the matrix core
    const double u = ldexp(1.0, -23);
    auto gamma = [&](double n) { return n * u / (1.0 - n * u); };
    const double gD = gamma(D);
    double e1, e2 = 0.0, e3 = 0.0;
    if (prec == SCREEN_FP32) {
        e1 = 2.0 * (1.0 + gD) * gamma(KP + 2) + gD;
    } else if (prec == SCREEN_BF16) {
        e1 = (3.02 * ldexp(1.0, -16) + 1.02 * gamma(3 * KP + 2)) * 2.0 * (1.0 + gD) + gD;
    } else {
        // Three-term e4m3 split of every scaled operand value a: each term is an e4m3 rounding of
        // the previous remainder, |v - rnd(v)| <= 2^-4 |v| + 2^-10 (subnormal spacing 2^-9), so the
        // remainder r after three terms has |r| <= 2^-12 a + 1.07 phi (phi = 2^-10) and the third term
        // |l| <= 0.00415 a + 2.2 phi.  The eight kept products miss r_q c + q~ r_c + l_q l_c:
        // per position at most 5.06e-4 a_q a_c + 1.08 phi (a_q + a_c) + 6 phi^2.  Summed over the row
        // with sum a_q a_c <= 2 (|q|^2 + |c|^2) / s^2 (features, -2x and the norm entries), sum of the
        // feature a's <= 2 sqrt(D) |x| / s, the norm entries' floors 1.08 phi |x|^2/(256 s^2), and the
        // exact entry w = 256 meeting the other side's norm term (256 * 1.07 phi each way).  In real
        // units (times s^2):
        //   eta  = 2 * 5.06e-4 + 1.08 phi / 256 + accumulation (8 MFMAs of 64 products: gamma(520) * 2)
        //   eta2 = 1.08 phi * 2 sqrt(D)          (times s (|q| + |c|))
        //   eta3 = 2 * 256 * 1.07 phi + 64 * 6 phi^2   (times s^2)
        const double phi = ldexp(1.0, -10);
        e1 = 2.0 * 5.06e-4 + 1.08 * phi / 256.0 + 2.0 * gamma(8 * 65 + 2);
        e2 = 1.08 * phi * 2.0 * std::sqrt((double)D);
        e3 = 2.0 * 256.0 * 1.07 * phi + 64.0 * 6.0 * phi * phi;
    }
    *eta = 1.25 * e1 + 1e-12;
    *eta2 = 1.25 * e2;
    *eta3 = 1.25 * e3;
}

hipError_t launch_screen_pack(hipStream_t st, int prec, int64_t n_cols, int D, int FS, int KB, const float* F,
                              void* rows, float* norm, float* feat, int DP, const int64_t* seg_ptr, int n_seg,
                              const float* seg_scale) {
    if (feat && FS != D) return hipErrorInvalidValue;
    if (n_cols <= 0) return hipSuccess;
    const int64_t blocks = (n_cols + PACK_COLS - 1) / PACK_COLS;
    const int np = screen_np(prec, KB);
    uint4* r = reinterpret_cast<uint4*>(rows);
    if (prec == SCREEN_FP32)
        k_screen_pack<SCREEN_FP32><<<(unsigned)blocks, 256, 0, st>>>(n_cols, D, FS, np, F, r, norm, feat, DP, seg_ptr,
                                                                      n_seg, seg_scale);
    else if (prec == SCREEN_BF16)
        k_screen_pack<SCREEN_BF16><<<(unsigned)blocks, 256, 0, st>>>(n_cols, D, FS, np, F, r, norm, feat, DP, seg_ptr,
                                                                      n_seg, seg_scale);
    else
        k_screen_pack<SCREEN_FP8><<<(unsigned)blocks, 256, 0, st>>>(n_cols, D, FS, np, F, r, norm, feat, DP, seg_ptr,
                                                                     n_seg, seg_scale);
    return hipGetLastError();
}

hipError_t launch_site_scale(hipStream_t st, int n_seg, const int64_t* seg_ptr, const float* F, int FS, int D,
                             float* seg_scale) {
    if (n_seg <= 0) return hipSuccess;
    k_site_scale<<<n_seg, 256, 0, st>>>(seg_ptr, F, FS, D, seg_scale);
    return hipGetLastError();
}

hipError_t launch_screen_knn_q(hipStream_t st, int prec, int KB, const ScreenArgs& a, int grid, int* seg_ctr,
                               const int* seg_beg) {
    if (a.n_items <= 0) return hipSuccess;
    if (prec == SCREEN_FP8) return launch_kb<SCREEN_FP8, 4>(st, a, grid, seg_ctr, seg_beg);
    if (prec == SCREEN_BF16) {
        switch (KB) {
            case 1: return launch_kb<SCREEN_BF16, 1>(st, a, grid, seg_ctr, seg_beg);
            case 2: return launch_kb<SCREEN_BF16, 2>(st, a, grid, seg_ctr, seg_beg);
            case 3: return launch_kb<SCREEN_BF16, 3>(st, a, grid, seg_ctr, seg_beg);
            case 4: return launch_kb<SCREEN_BF16, 4>(st, a, grid, seg_ctr, seg_beg);
        }
    } else {
        switch (KB) {
            case 1: return launch_kb<SCREEN_FP32, 1>(st, a, grid, seg_ctr, seg_beg);
            case 2: return launch_kb<SCREEN_FP32, 2>(st, a, grid, seg_ctr, seg_beg);
            case 3: return launch_kb<SCREEN_FP32, 3>(st, a, grid, seg_ctr, seg_beg);
            case 4: return launch_kb<SCREEN_FP32, 4>(st, a, grid, seg_ctr, seg_beg);
        }
    }
    return hipErrorInvalidValue;
}

hipError_t launch_screen_rerank(hipStream_t st, const ScreenArgs& a, int* knn_col, double* dk2, int* spill,
                                int* n_spill) {
    if (a.n_q <= 0) return hipSuccess;
    k_zero_ints<<<1, 64, 0, st>>>(n_spill, 1);
    const unsigned grid = (unsigned)((a.n_q + 63) / 64);
    const unsigned grid2 = grid < 1024 ? grid : 1024;   // the second pass takes few queries
    const size_t obytes = a.p.k <= 32 ? sizeof(int) * 64 * (size_t)a.p.k : 0;
    FinOut fin{nullptr, nullptr, nullptr, nullptr, 1, nullptr};
    if (screen_rerank_finishes(a)) fin = FinOut{a.fin_rho, a.fin_dk, a.fin_knn, a.fin_d2head, 1, a.fin_rho_col};
    FinOut fin2 = fin;
    fin2.first = 0;
#define IRC_RR(NV)                                                                                         \
    case NV:                                                                                               \
        k_rerank<NV><<<grid, 64, obytes, st>>>(a.p, a.n_q, a.cand_n, a.cand, a.cand_eps, knn_col, dk2, spill, \
                                          n_spill, fin);                                                   \
        k_rerank_list<NV><<<grid2, 64, obytes, st>>>(a.p, a.n_q, spill, n_spill, a.cand_n, a.cand, a.cand_eps, \
                                                knn_col, dk2, a.fallback, a.n_fallback, fin2);             \
        break;
    switch (a.p.DP / 4) {
        IRC_RR(1) IRC_RR(2) IRC_RR(3) IRC_RR(4) IRC_RR(5) IRC_RR(6) IRC_RR(7) IRC_RR(8)
        IRC_RR(9) IRC_RR(10) IRC_RR(11) IRC_RR(12) IRC_RR(13) IRC_RR(14) IRC_RR(15) IRC_RR(16)
        default: return hipErrorInvalidValue;
    }
#undef IRC_RR
    return hipGetLastError();
}

hipError_t launch_run_nmax(hipStream_t st, int64_t n_runs, const int* run_beg, const int* run_len, const float* norm,
                           float* out) {
    if (n_runs <= 0) return hipSuccess;
    k_run_nmax<<<(unsigned)((n_runs + 3) / 4), 256, 0, st>>>(n_runs, run_beg, run_len, norm, out);
    return hipGetLastError();
}

hipError_t launch_fill_int(hipStream_t st, int* p, int64_t n, int v) {
    if (n <= 0) return hipSuccess;
    k_fill_int<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_mfma_layout_probe(hipStream_t st, float* out) {
    k_mfma_probe<<<1, 64, 0, st>>>(out);
    return hipGetLastError();
}

// Accumulation probe: tile t's A_t (32 x K, row-major) times B_t (K x 32, row-major) by one chain of
// the MFMAs the screen uses -- prec 1: v_mfma_f32_32x32x2_f32, two k per step; prec 2:
// v_mfma_f32_32x32x16_bf16, sixteen k per step, the operands being floats that bf16 represents
// exactly (their upper halves are passed) -- written as out[t][row][col] through the 32 x 32
// accumulator layout (register r of lane l: row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32).
__global__ __launch_bounds__(64) void k_mfma_accum_probe(int prec, int K, const float* __restrict__ A,
                                                         const float* __restrict__ B, float* __restrict__ out) {
    const int l = threadIdx.x, h = l >> 5, m = l & 31;
    const int64_t t = blockIdx.x;
    const float* a = A + t * 32 * K;
    const float* b = B + t * K * 32;
    f32x16 acc = {};
    if (prec == 1) {
        for (int k0 = 0; k0 < K; k0 += 2)
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[m * K + k0 + h], b[(k0 + h) * 32 + m], acc, 0, 0, 0);
    } else {
        for (int k0 = 0; k0 < K; k0 += 16) {
            short8 va, vb;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + 8 * h + j;
                va[j] = (short)(__float_as_uint(a[m * K + k]) >> 16);
                vb[j] = (short)(__float_as_uint(b[k * 32 + m]) >> 16);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, va),
                                                          __builtin_bit_cast(bf16x8, vb), acc, 0, 0, 0);
        }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) out[t * 1024 + (8 * (r >> 2) + 4 * h + (r & 3)) * 32 + m] = acc[r];
}

hipError_t launch_mfma_accum_probe(hipStream_t st, int prec, int K, int n_tiles, const float* A, const float* B,
                                   float* out) {
    k_mfma_accum_probe<<<n_tiles, 64, 0, st>>>(prec, K, A, B, out);
    return hipGetLastError();
}

}  // namespace irc
