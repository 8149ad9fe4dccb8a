// irc_screen.hip -- MFMA distance screen of the kNN pass for gfx950.
//
// The screen replaces the per-thread binned-minimum search of search_min_drift.cu:24-94
// (and the host sort of its nThreads x nA output, matlab/irc2.m:3784-3787) by:
//   1. distance tiles as a dense contraction on MFMA: with candidate rows
//      c~ = [c, 1, |c|^2] and query rows q~ = [-2q, |q|^2, 1], one 32x32 product gives
//      32 candidates x 32 queries of |q|^2 + |c|^2 - 2 q.c directly;
//        FP32 form: v_mfma_f32_32x32x2_f32
//        BF16 form: v_mfma_f32_32x32x16_bf16 on a hi/lo split of every operand
//                   (x = hi + lo + r, |r| <= 2^-16 |x|), products hi*hi + hi*lo + lo*hi,
//                   accumulated in FP32;
//   2. a per-query list of the K' smallest screened keys (32-bit keys: truncated
//      distance bits above the candidate's stream position), kept sorted in VGPRs and
//      refreshed from an LDS append buffer with bitonic networks;
//   3. a rigorous bound |screened - exact| <= eta * (|q|^2 + max|c|^2) certifies that
//      every candidate outside the list is farther than the k-th; the candidates that
//      can still be in the top k are re-ranked with the canonical FP64 distance and
//      ordered by (d2, spike id).  Uncertified queries go to the exact kernel.
// A wavefront owns 64 queries of one (site, chunk) block: two 32-query column tiles
// share each candidate tile; a lane permutation (v_permlane32_swap) gives every lane
// all 32 distances of its own query.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "irc_kernels.h"
#include "irc_screen.h"

#pragma clang fp contract(off)

namespace irc {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef short short8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2bf_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// A workgroup packs 32 columns: their rows are staged in LDS with coalesced loads, one thread
// per column forms |x|^2, then one thread per (column, 8-dimension chunk) writes coalesced
// 16-byte pieces of the candidate-form rows [x, 1, |x|^2, 0...] (bf16 hi/lo split, or fp32).
// The query-form rows [-2x, |x|^2, 1, 0...] are not stored: the screen derives them from the
// candidate-form rows and |x|^2 (qform_bf16 / qform_f32).
constexpr int PACK_COLS = 32;

__global__ void k_screen_pack(int prec, int64_t n_cols, int D, int FS, int KP, const float* __restrict__ F,
                              void* __restrict__ hi, void* __restrict__ lo, float* __restrict__ norm,
                              float* __restrict__ feat, int DP) {
    __shared__ float srow[PACK_COLS * IRC_DMAX];
    __shared__ float snorm[PACK_COLS];
    const int nch = KP / 8;
    const int64_t col0 = (int64_t)blockIdx.x * PACK_COLS;
    const int ncol = n_cols - col0 < PACK_COLS ? (int)(n_cols - col0) : PACK_COLS;
    // rows col0 .. col0+ncol-1 of F: (ncol-1)*FS + D contiguous floats
    const int span = (ncol - 1) * FS + D;
    const float* src = F + col0 * FS;
    for (int e = threadIdx.x; e < span; e += blockDim.x) {
        const int c = e / FS, d = e - c * FS;
        if (d < D) srow[c * IRC_DMAX + d] = src[e];
    }
    __syncthreads();
    if ((int)threadIdx.x < ncol) {
        const float* x = srow + threadIdx.x * IRC_DMAX;
        float n = 0.0f;
        for (int d = 0; d < D; ++d) n = fmaf(x[d], x[d], n);
        snorm[threadIdx.x] = n;
        norm[col0 + threadIdx.x] = n;
    }
    __syncthreads();
    if (feat) {
        // the FP32 rows of these columns, zero padded to DP: contiguous float4 stores
        const int n4 = DP / 4;
        float4* f4 = reinterpret_cast<float4*>(feat + col0 * DP);
        for (int e = threadIdx.x; e < ncol * n4; e += blockDim.x) {
            const int cc = e / n4, d0 = 4 * (e - cc * n4);
            const float* x = srow + cc * IRC_DMAX;
            f4[e] = make_float4(d0 < D ? x[d0] : 0.0f, d0 + 1 < D ? x[d0 + 1] : 0.0f, d0 + 2 < D ? x[d0 + 2] : 0.0f,
                                d0 + 3 < D ? x[d0 + 3] : 0.0f);
        }
    }
    const int c = threadIdx.x / nch, j = threadIdx.x - c * nch;
    if (c >= ncol) return;
    const int64_t col = col0 + c;
    const float* x = srow + c * IRC_DMAX;
    const float n = snorm[c];
    float cv[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int d = 8 * j + t;
        cv[t] = d < D ? x[d] : (d == D ? 1.0f : (d == D + 1 ? n : 0.0f));
    }
    if (prec == SCREEN_FP32) {
        // fp32 rows are KP / 4 pieces: this thread's 8 values are pieces 2 j and 2 j + 1
        float4* h4 = reinterpret_cast<float4*>(hi);
        const int np = KP / 4;
        h4[screen_piece(col, 2 * j, np)] = make_float4(cv[0], cv[1], cv[2], cv[3]);
        h4[screen_piece(col, 2 * j + 1, np)] = make_float4(cv[4], cv[5], cv[6], cv[7]);
    } else {
        uint32_t ph[4], pl[4];
#pragma unroll
        for (int t = 0; t < 8; t += 2) {
            uint16_t h0 = f2bf_rne(cv[t]), h1 = f2bf_rne(cv[t + 1]);
            uint16_t l0 = f2bf_rne(cv[t] - bf2f(h0)), l1 = f2bf_rne(cv[t + 1] - bf2f(h1));
            ph[t / 2] = (uint32_t)h0 | ((uint32_t)h1 << 16);
            pl[t / 2] = (uint32_t)l0 | ((uint32_t)l1 << 16);
        }
        // bf16 rows are KP / 8 pieces: this thread's 8 values are piece j
        const int64_t o = screen_piece(col, j, KP / 8);
        reinterpret_cast<uint4*>(hi)[o] = make_uint4(ph[0], ph[1], ph[2], ph[3]);
        reinterpret_cast<uint4*>(lo)[o] = make_uint4(pl[0], pl[1], pl[2], pl[3]);
    }
}

// Query-form piece [-2x, |x|^2, 1, 0...] of dimensions d0 .. d0+7 from the candidate-form piece
// [x, 1, |x|^2, 0...] of the same column: below D every part is scaled by -2 (exact in bf16),
// at D and D+1 the parts of |x|^2 and of 1 are put in (at_D, at_D1).
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

// the same for an fp32 piece (dimensions d0 .. d0+3)
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

__device__ __forceinline__ double exact_d2(const float* __restrict__ q, const float* __restrict__ x, int DP) {
    double acc = 0.0;
    const float4* q4 = reinterpret_cast<const float4*>(q);
    const float4* x4 = reinterpret_cast<const float4*>(x);
    for (int d4 = 0; d4 < DP / 4; ++d4) {
        const float4 a = q4[d4], b = x4[d4];
        double t;
        t = (double)a.x - (double)b.x; acc = fma(t, t, acc);
        t = (double)a.y - (double)b.y; acc = fma(t, t, acc);
        t = (double)a.z - (double)b.z; acc = fma(t, t, acc);
        t = (double)a.w - (double)b.w; acc = fma(t, t, acc);
    }
    return acc;
}

// the screened distance a key stands for, negative values read as 0 (a valid lower bound, and
// non-decreasing along a sorted list); the empty-slot key reads as NaN, which fails every test
__device__ __forceinline__ float key_dist(int32_t key, uint32_t hmask) {
    const float v = __uint_as_float((uint32_t)key & hmask);
    return v < 0.0f ? 0.0f : v;
}
constexpr int32_t KEY_EMPTY = 0x7fffffff;
constexpr int32_t KEY_NONE = (int32_t)0x80000000;   // threshold of lanes without a query

template <int PREC, int KB, int KPB, bool HIST>
__global__ __launch_bounds__(64, 2) void k_screen_knn(ScreenArgs a, int* __restrict__ seg_ctr,
                                                   const int* __restrict__ seg_beg) {
    constexpr int KP = 16 * KB;
    constexpr int LS = KPB <= 32 ? 32 : 64;   // list slots in registers
    __shared__ uint32_t sbuf[65 * 64];   // rows 0..63: append buffer / histogram / list; row 64: sink
    __shared__ int s_rb[IRC_SCREEN_MAXRUNS + 1];
    __shared__ int s_ro[IRC_SCREEN_MAXRUNS + 1];
    __shared__ int s_item;
    const DevPlan& p = a.p;
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
            if (seg_ctr[sg] >= seg_beg[sg + 1] - seg_beg[sg]) continue;
            const int idx = atomicAdd(&seg_ctr[sg], 1);
            if (seg_beg[sg] + idx < seg_beg[sg + 1]) got = seg_beg[sg] + idx;
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
        const ScreenItem item = a.items[it];
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

        // ---- query operands (B) for the two 32-query column tiles: query-form rows derived from
        // the candidate-form rows of the query columns
        bf16x8 bh[2][KB], bl[2][KB];
        float4 bq[2][2 * KB];
#pragma unroll
        for (int T = 0; T < 2; ++T) {
            const int qj = 32 * T + m;
            const bool qv = qj < item.nq;
            const int qc = p.q_col[item.q0 + (qv ? qj : 0)];
            const float qn = a.norm[qc];
            if constexpr (PREC == SCREEN_BF16) {
                const uint4* rh = reinterpret_cast<const uint4*>(a.hi);
                const uint4* rl = reinterpret_cast<const uint4*>(a.lo);
                const uint16_t nh = f2bf_rne(qn), nl = f2bf_rne(qn - bf2f(nh));
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) {
                    const int pc = 2 * kb + h;
                    const int64_t o = screen_piece(qc, pc, 2 * KB);
                    bh[T][kb] = __builtin_bit_cast(bf16x8, qform_bf16(rh[o], qv, 8 * pc, p.D, nh, f2bf_rne(1.0f)));
                    bl[T][kb] = __builtin_bit_cast(bf16x8, qform_bf16(rl[o], qv, 8 * pc, p.D, nl, 0));
                }
            } else {
                const float4* r = reinterpret_cast<const float4*>(a.hi);
#pragma unroll
                for (int g = 0; g < 2 * KB; ++g) {
                    const int pc = 2 * g + h;
                    bq[T][g] = qform_f32(r[screen_piece(qc, pc, 4 * KB)], qv, 4 * pc, p.D, qn);
                }
            }
        }
        // after the lane swap, lane l owns query l of the item
        const bool qmine = lane < item.nq;
        const int q_mine = item.q0 + (qmine ? lane : 0);
        const int qc_mine = p.q_col[q_mine];
        const float nq_mine = a.norm[qc_mine];
        // retire the item's own loads here: the stream loop then waits on operand loads only, so
        // the compiler can keep the second operand buffer's loads in flight (vmcnt(6), not 0)
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0), expcnt and lgkmcnt unconstrained

        // the item's list depth K' is a compile-time constant: slots past it hold KEY_EMPTY
        // between merges, so the merge networks lose their dead half
        auto body = [&](auto kp_c) {
        constexpr int kprime = decltype(kp_c)::value;
        int32_t L[LS];
#pragma unroll
        for (int i = 0; i < LS; ++i) L[i] = KEY_EMPTY;
        int32_t tau = qmine ? KEY_EMPTY : KEY_NONE;
        int cnt = 0;
        const int ib = ceil_log2(ncand > 2 ? ncand : 2);
        const uint32_t pmask = (1u << ib) - 1u, hmask = ~pmask;
        // the position mask in a vector register, so the key is one v_bfi_b32 with the scalar
        // position as its only scalar operand
        uint32_t pmask_v;
        asm volatile("v_mov_b32 %0, %1" : "=v"(pmask_v) : "s"(pmask));
        int run = 0;

        // the lane's K'-th kept key: the append threshold (L may hold up to LS >= K' keys; those
        // past K' are larger than the threshold and never change the result)
        auto kth = [&]() -> int32_t { return L[kprime - 1]; };
        int32_t tau_init = tau;
        // merge the oldest 32 buffered keys into L, move the rest (at most 32) down
        // Buffer rows of a lane past its count hold KEY_EMPTY or keys that failed the threshold
        // when written (larger than L[K'-1] then and ever after, so a merge never keeps them);
        // rows already merged are cleared.  Without the pre-pass the merge therefore reads the
        // rows unmasked; with it the threshold can sit below L[K'-1] and the rows are masked.
        auto compact_half = [&]() {
            int32_t B[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) {
                const int32_t v = (int32_t)sbuf[i * 64 + lane];
                B[i] = (!HIST || i < cnt) ? v : KEY_EMPTY;
            }
            oddeven_sort<32>(B);
#pragma unroll
            for (int i = 0; i < 32; ++i) {
                const int32_t v = B[31 - i];
                L[LS - 32 + i] = L[LS - 32 + i] < v ? L[LS - 32 + i] : v;
            }
            bitonic_clean<LS>(L);
#pragma unroll
            for (int i = kprime; i < LS; ++i) L[i] = KEY_EMPTY;
            if (__any(cnt > 32)) {
#pragma unroll
                for (int i = 0; i < 32; ++i) sbuf[i * 64 + lane] = sbuf[(32 + i) * 64 + lane];
                if constexpr (!HIST) {
#pragma unroll
                    for (int i = 32; i < 64; ++i) sbuf[i * 64 + lane] = (uint32_t)KEY_EMPTY;
                }
            } else if constexpr (!HIST) {
#pragma unroll
                for (int i = 0; i < 32; ++i) sbuf[i * 64 + lane] = (uint32_t)KEY_EMPTY;
            }
            cnt = cnt > 32 ? cnt - 32 : 0;
            const int32_t t = kth();
            tau = qmine ? (t < tau_init ? t : tau_init) : KEY_NONE;
        };

        // operand rows of the lane's candidate at stream position pos (bf16: hi then lo halves)
        constexpr int NA = PREC == SCREEN_BF16 ? 2 * KB : 2 * KB;
        float4 An0[NA], An1[NA];   // operands of the next two tiles
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
        auto issue = [&](int pos0, float4 (&An)[NA]) {
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
            // tiled layout: a load instruction reads piece p of 32 consecutive columns from
            // 512 contiguous bytes (two such runs when the tile straddles a 32-column group)
            if constexpr (PREC == SCREEN_BF16) {
                const float4* rh = reinterpret_cast<const float4*>(a.hi);
                const float4* rl = reinterpret_cast<const float4*>(a.lo);
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) {
                    const int64_t o = screen_piece(col, 2 * kb + h, 2 * KB);
                    An[2 * kb] = rh[o];
                    An[2 * kb + 1] = rl[o];
                }
            } else {
                const float4* r = reinterpret_cast<const float4*>(a.hi);
#pragma unroll
                for (int g = 0; g < 2 * KB; ++g) An[g] = r[screen_piece(col, 2 * g + h, 4 * KB)];
            }
        };
        // one pass over the candidate stream: MFMA distance tiles, then consume(v, cpos, valid)
        // for the 32 distances of the lane's query, then after_tile()
        auto stream = [&](auto&& consume, auto&& after_tile) {
            run_enter(0);
            issue(0, An0);
            issue(32, An1);
            // tiles alternate between the two operand buffers; every tile issues its MFMAs and
            // the loads two tiles ahead unconditionally (positions past the stream read a valid
            // row), so the count of loads in flight is the same on every path
            auto tile = [&](int base, float4 (&Ac)[NA]) {
                f32x16 acc0 = {}, acc1 = {};
                if constexpr (PREC == SCREEN_BF16) {
#pragma unroll
                    for (int kb = 0; kb < KB; ++kb) {
                        const bf16x8 ah = __builtin_bit_cast(bf16x8, Ac[2 * kb]);
                        const bf16x8 al = __builtin_bit_cast(bf16x8, Ac[2 * kb + 1]);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[0][kb], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[1][kb], acc1, 0, 0, 0);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[0][kb], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[1][kb], acc1, 0, 0, 0);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[0][kb], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[1][kb], acc1, 0, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int g = 0; g < 2 * KB; ++g) {
                        const float4 av = Ac[g];
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av.x, bq[0][g].x, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av.x, bq[1][g].x, acc1, 0, 0, 0);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av.y, bq[0][g].y, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av.y, bq[1][g].y, acc1, 0, 0, 0);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av.z, bq[0][g].z, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av.z, bq[1][g].z, acc1, 0, 0, 0);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av.w, bq[0][g].w, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av.w, bq[1][g].w, acc1, 0, 0, 0);
                    }
                }
                // the operand registers are free once the MFMAs have read them: the loads of the
                // tile after next are in flight while this tile and the next are consumed
                issue(base + 64, Ac);
                if (base >= ncand) return;
                // lane l < 32: query l of tile 0; lane l >= 32: query l-32 of tile 1
                uint32_t u[32];
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc0[r]),
                                                                    __float_as_uint(acc1[r]), false, false);
                    u[2 * r] = sw[0];
                    u[2 * r + 1] = sw[1];
                }
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
                after_tile();
            };
            for (int base = 0; base < ncand; base += 64) {
                tile(base, An0);
                tile(base + 32, An1);
            }
        };

        // ---- pass 1 (HIST): per-lane histogram of the screened distances in log buckets (4 per
        // octave, 16 octaves from |q|^2 * 2^-14) gives a starting threshold that already admits
        // at least K' candidates, so pass 2 appends about K' keys instead of K'(1 + ln(n/K')).
        // Row 64 of sbuf absorbs the writes of masked-out elements (no divergent branches).
        if constexpr (HIST) {
            if (ncand > kprime) {
                __syncthreads();
                uint4* z = reinterpret_cast<uint4*>(sbuf);
#pragma unroll
                for (int i = 0; i < 16; ++i) z[i * 64 + lane] = make_uint4(0u, 0u, 0u, 0u);
                __syncthreads();
                const int hbase = (int)(__float_as_uint(fmaxf(nq_mine, 1e-30f) * 0x1p-14f) & 0xffe00000u);
                stream(
                    [&](int bits, int, bool valid) {
                        // negative screened values give negative bits: clamped into bucket 0
                        int b = (bits - hbase) >> 21;
                        b = b < 0 ? 0 : (b > 63 ? 63 : b);
                        atomicAdd(&sbuf[(valid ? b : 64) * 64 + lane], 1u);
                    },
                    [&]() {});
                __syncthreads();
                int cum = 0, tb = 63;
                for (int b = 0; b < 63; ++b) {
                    cum += (int)sbuf[b * 64 + lane];
                    if (cum >= kprime) { tb = b; break; }
                }
                if (qmine && tb < 63) tau = hbase + ((tb + 1) << 21);
                tau_init = tau;
                __syncthreads();
            }
        }

        // ---- pass 2: keys (the screened distance's bits, sign kept, above the candidate's stream
        // position) below the threshold go to the lane's append buffer; full buffers merge into L
        if constexpr (!HIST) {
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 64; ++i) sbuf[i * 64 + lane] = (uint32_t)KEY_EMPTY;
        }
        stream(
            [&](int sbits, int cpos, bool valid) {
                // key: one bit-field insert (low ib bits from the wave-uniform position, the rest
                // from the value); slot cnt (< 64: at most 32 buffered when a tile starts) is
                // scratch until counted; the count takes the compare's carry in one add
                // (positions past the stream end -- wave-uniform -- write nothing: the buffer rows
                // past a lane's count must only hold keys that can never enter the list)
                if (!valid) return;
                int32_t key;
                asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(key) : "v"(pmask_v), "s"(cpos), "v"(sbits));
                sbuf[cnt * 64 + lane] = key;
                asm("v_cmp_lt_i32 vcc, %1, %2\n\tv_addc_co_u32 %0, vcc, 0, %0, vcc"
                    : "+v"(cnt) : "v"(key), "v"(tau) : "vcc");
            },
            [&]() {
                if (__any(cnt > 32)) compact_half();
            });
        while (__any(cnt > 0)) compact_half();
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
        if (qmine) {
            const int k = p.k;
            const int nl = ncand < kprime ? ncand : kprime;
            const int kk = ncand < k ? ncand : k;
            const double eps = a.eta * ((double)nq_mine + (double)nmax_w) + 1e-30;
            const double qf = ldexp(1.0, -(22 - ib));
            const double dkt = key_dist((int32_t)sbuf[(kk - 1) * 64 + lane], hmask);
            // every one of the first kk kept candidates has d^2 <= U
            const double U = dkt * (1.0 + qf) + eps;
            // every candidate outside L has a key at least min(L[K'-1], starting threshold)
            const int32_t kl = (int32_t)sbuf[(kprime - 1) * 64 + lane];
            const int32_t kexcl = kl < tau_init ? kl : tau_init;
            const bool ok = ncand <= kprime || ((double)key_dist(kexcl, hmask) - eps > U);
            if (ok) {
                // the re-rank needs every kept candidate whose lower bound reaches U
                // (the list is sorted and its screened distances nondecreasing, so the test
                // holds for a prefix of the slots past kk: count it from the registers)
                int r = kk;
#pragma unroll
                for (int j = 0; j < kprime; ++j)
                    r += (j >= kk && j < nl && (double)key_dist(L[j], hmask) - eps <= U) ? 1 : 0;
                for (int j = 0; j < r; ++j) {
                    const int32_t key = (int32_t)sbuf[j * 64 + lane];
                    const int pos = (int)(key & pmask);
                    const int lo = use_tbl ? (int)tbl[pos] : run_of(pos);
                    a.cand_col[(int64_t)j * a.n_q + q_mine] = s_rb[lo] + (pos - s_ro[lo]);
                    a.cand_d[(int64_t)j * a.n_q + q_mine] = key_dist(key, hmask);
                }
                a.cand_n[q_mine] = r | (kk << 16);
                a.cand_eps[q_mine] = eps;
            } else {
                a.cand_n[q_mine] = -1;
                const int f = atomicAdd(a.n_fallback, 1);
                a.fallback[f] = q_mine;
            }
        }
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
// overflows there goes to the exact kernel.
constexpr int WIN1 = 8;    // first pass: every certified query
constexpr int WIN2 = 16;   // second pass: the first pass's overflows

// one group of up to 64 queries (lane l = query q, valid when qv); overflows are appended to spill
template <int NV, int W>
__device__ __forceinline__ void rerank_group(const DevPlan& p, int64_t n_q, int64_t q, bool qv,
                                             const int* __restrict__ cand_n, const int* __restrict__ cand_col,
                                             const float* __restrict__ cand_d, const double* __restrict__ cand_eps,
                                             int* __restrict__ knn_col, double* __restrict__ dk2,
                                             int* __restrict__ spill, int* __restrict__ n_spill, float4* stage,
                                             int* skey, float* sdn, int* obuf) {
    const int l = threadIdx.x;
    const int rn = qv ? cand_n[q] : -1;
    const int r = rn < 0 ? 0 : (rn & 0xffff);
    int rmax = r;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) rmax = max(rmax, __shfl_xor(rmax, off, 64));
    if (rmax == 0) return;
    const int qc = p.q_col[qv ? q : 0];
    float4 qr[NV];
    {
        const float4* q4 = reinterpret_cast<const float4*>(p.feat + (int64_t)qc * p.DP);
#pragma unroll
        for (int i = 0; i < NV; ++i) qr[i] = q4[i];
    }
    // lanes past their own list fetch their query's row (a valid address) and ignore it
    auto slot_col = [&](int j) { return j < r ? cand_col[(int64_t)j * n_q + q] : qc; };
    // slot j's rows, ids and the next slot's screened distances, straight into LDS
    auto gather = [&](int j, int colj) {
        const int jn = j + 1 < r ? j + 1 : (r > 0 ? r - 1 : 0);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(p.col_key + colj),
                                         (__attribute__((address_space(3))) void*)skey, 4, 0, 0);
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void*)(cand_d + (int64_t)jn * n_q + (qv ? q : 0)),
            (__attribute__((address_space(3))) void*)sdn, 4, 0, 0);
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
    int wn = 0, out = 0;
    double last_t = 0.0;
    int last_c = -1;
    bool ok = rn >= 0;
    int colj = slot_col(0);
    int coln = slot_col(1);
    gather(0, colj);
    for (int j = 0; j < rmax; ++j) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // the gather of slot j has landed in LDS
        float4 b[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) b[i] = stage[l * NV + i];
        const int kc = skey[l];
        const float dn = sdn[l];
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");   // LDS read before the next gather lands
        const bool act = ok && j < r && out < kk;
        const int coln2 = slot_col(j + 2);
        if (j + 1 < rmax) gather(j + 1, coln);
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
        coln = coln2;
    }
    const bool done = qv && rn >= 0 && ok && out == kk;
    if (done) {
        for (int j = kk; j < k; ++j) {
            if (k <= 32) obuf[j * 64 + l] = last_c;
            else knn_col[(int64_t)j * n_q + q] = last_c;
        }
        dk2[q] = last_t;
    } else if (qv && rn >= 0) {
        const int f = atomicAdd(n_spill, 1);
        spill[f] = (int)q;
    }
    // the buffered lists leave slot by slot: 64 queries per store
    if (k <= 32) {
        for (int j = 0; j < k; ++j)
            if (done) knn_col[(int64_t)j * n_q + q] = obuf[j * 64 + l];
    }
}

template <int NV>
__global__ __launch_bounds__(64) void k_rerank(DevPlan p, int64_t n_q, const int* __restrict__ cand_n,
                                              const int* __restrict__ cand_col, const float* __restrict__ cand_d,
                                              const double* __restrict__ cand_eps, int* __restrict__ knn_col,
                                              double* __restrict__ dk2, int* __restrict__ spill,
                                              int* __restrict__ n_spill) {
    __shared__ float4 stage[64 * NV];   // [row][piece]: the slot's 64 candidate rows
    __shared__ int skey[64];            // their spike ids
    __shared__ float sdn[64];           // each lane's next screened distance (lower bounds)
    extern __shared__ int obuf[];       // [k][64] emitted columns when k <= 32 (else written directly)
    // workgroups are dispatched round-robin over the 8 XCDs: give each XCD a contiguous eighth
    // of the query groups, so the candidate rows its groups share stay in its L2
    const unsigned G = gridDim.x, x = blockIdx.x & 7, per = (G + 7) >> 3, lo8 = G & 7;
    const unsigned grp = (lo8 == 0 || x < lo8) ? x * per + (blockIdx.x >> 3)
                                               : lo8 * per + (x - lo8) * (per - 1) + (blockIdx.x >> 3);
    const int64_t q = (int64_t)grp * 64 + threadIdx.x;
    rerank_group<NV, WIN1>(p, n_q, q, q < n_q, cand_n, cand_col, cand_d, cand_eps, knn_col, dk2, spill, n_spill,
                           stage, skey, sdn, obuf);
}

// second pass: the queries of list[0 .. *n_list), 64 per group, groups strided over the grid
template <int NV>
__global__ __launch_bounds__(64) void k_rerank_list(DevPlan p, int64_t n_q, const int* __restrict__ list,
                                                   const int* __restrict__ n_list, const int* __restrict__ cand_n,
                                                   const int* __restrict__ cand_col, const float* __restrict__ cand_d,
                                                   const double* __restrict__ cand_eps, int* __restrict__ knn_col,
                                                   double* __restrict__ dk2, int* __restrict__ fallback,
                                                   int* __restrict__ n_fallback) {
    __shared__ float4 stage[64 * NV];
    __shared__ int skey[64];
    __shared__ float sdn[64];
    extern __shared__ int obuf[];
    const int64_t n = *n_list;
    for (int64_t g = blockIdx.x; g * 64 < n; g += gridDim.x) {
        const int64_t i = g * 64 + threadIdx.x;
        const bool qv = i < n;
        rerank_group<NV, WIN2>(p, n_q, qv ? (int64_t)list[i] : 0, qv, cand_n, cand_col, cand_d, cand_eps, knn_col,
                               dk2, fallback, n_fallback, stage, skey, sdn, obuf);
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

// Lane layout probe of the two MFMA forms and the lane swap the screen relies on.
// out[0..1023]   : 32x32x16 bf16 with A[m][k] = (k == 0) * (m + 1), B[k][n] = (k == 0) * 64 * (n + 1)
//                  -> acc[r] of lane l, stored at out[r * 64 + l]
// out[1024..2047]: 32x32x2 f32 with the same A, B
// out[2048..4095]: permlane32_swap(acc0, acc1) of the bf16 result with acc1 = acc0 + 0.5
//                  (first and second result at out[2048 + r*64 + l], out[3072 + r*64 + l])
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
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        out[r * 64 + l] = acc[r];
        out[1024 + r * 64 + l] = acc2[r];
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[r]), __float_as_uint(acc[r] + 0.5f),
                                                        false, false);
        out[2048 + r * 64 + l] = __uint_as_float(sw[0]);
        out[3072 + r * 64 + l] = __uint_as_float(sw[1]);
    }
}

template <int PREC, int KB>
hipError_t launch_kb(hipStream_t st, const ScreenArgs& a, int grid, int* seg_ctr, const int* seg_beg) {
#define IRC_SK(KPB) k_screen_knn<PREC, KB, KPB, false><<<grid, 64, 0, st>>>(a, seg_ctr, seg_beg)
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

double screen_eta(int prec, int D, int KP) {
    // unit roundoff taken as 2^-23 (twice FP32's) so the bound also covers truncating accumulation
    const double u = ldexp(1.0, -23);
    auto gamma = [&](double n) { return n * u / (1.0 - n * u); };
    const double gD = gamma(D);
    double eta;
    if (prec == SCREEN_FP32) {
        eta = 2.0 * (1.0 + gD) * gamma(KP + 2) + gD;
    } else {
        eta = (3.02 * ldexp(1.0, -16) + 1.02 * gamma(3 * KP + 2)) * 2.0 * (1.0 + gD) + gD;
    }
    return 1.25 * eta + 1e-12;
}

hipError_t launch_screen_pack(hipStream_t st, int prec, int64_t n_cols, int D, int FS, int KP, const float* F,
                              void* hi, void* lo, float* norm, float* feat, int DP) {
    if (feat && FS != D) return hipErrorInvalidValue;
    if (n_cols <= 0) return hipSuccess;
    const int64_t blocks = (n_cols + PACK_COLS - 1) / PACK_COLS;
    k_screen_pack<<<(unsigned)blocks, PACK_COLS * (KP / 8), 0, st>>>(prec, n_cols, D, FS, KP, F, hi, lo, norm,
                                                                     feat, DP);
    return hipGetLastError();
}

hipError_t launch_screen_knn_q(hipStream_t st, int prec, int KP, const ScreenArgs& a, int grid, int* seg_ctr,
                               const int* seg_beg) {
    if (a.n_items <= 0) return hipSuccess;
    k_zero_ints<<<1, 64, 0, st>>>(seg_ctr, 8);
    const int KB = KP / 16;
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
    const unsigned grid2 = grid < 1024 ? grid : 1024;
    const size_t obytes = a.p.k <= 32 ? sizeof(int) * 64 * (size_t)a.p.k : 0;
#define IRC_RR(NV)                                                                                         \
    case NV:                                                                                               \
        k_rerank<NV><<<grid, 64, obytes, st>>>(a.p, a.n_q, a.cand_n, a.cand_col, a.cand_d, a.cand_eps, knn_col, \
                                          dk2, spill, n_spill);                                            \
        k_rerank_list<NV><<<grid2, 64, obytes, st>>>(a.p, a.n_q, spill, n_spill, a.cand_n, a.cand_col, a.cand_d, \
                                                a.cand_eps, knn_col, dk2, a.fallback, a.n_fallback);       \
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

}  // namespace irc
