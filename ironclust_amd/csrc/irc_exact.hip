// irc_exact.hip -- exact (FP64) kernels of the density-peak hot path for gfx950,
// plus the device-side planning kernels.
//
// Canonical distance: d2 = fma-chain over dimensions of (double(q_d) - double(x_d))^2.
// Zero padding beyond D adds fma(0,0,acc) = acc exactly, so loops run over the
// padded dimension DP.
//
// Reference semantics restated (paths relative to the ironclust repository):
//   kNN over the linked drift chunks of a site   matlab/irc2.m:3653-3668 (knn_cpu_ 4244-4263)
//   rho = 1/d_k                                   matlab/irc2.m:3669
//   nearest denser neighbour                      matlab/irc2.m:3703-3725 (delta_cpu_ 3981-3994)
#include "irc_common.h"
#include "irc_kernels.h"

#pragma clang fp contract(off)

namespace {

// wave-wide argmin of (t, key) carrying col; ties of identical (t,key) go to the lower lane
__device__ __forceinline__ void wave_argmin(double& t, int& key, int& col, int& lane) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double ot = __shfl_xor(t, off, 64);
        int ok = __shfl_xor(key, off, 64);
        int oc = __shfl_xor(col, off, 64);
        int ol = __shfl_xor(lane, off, 64);
        bool take = irc_less(ot, ok, t, key) || (ot == t && ok == key && ol < lane);
        if (take) { t = ot; key = ok; col = oc; lane = ol; }
    }
}

// (t, key) < (t2, key of column c2); the column's key is read only on an exact tie
__device__ __forceinline__ bool less_tc(const DevPlan& p, double t, int key, double t2, int c2) {
    return t < t2 || (t == t2 && key < p.col_key[c2]);
}

}  // namespace

// ---------------------------------------------------------------------------
// Exact kNN: one wavefront per query.  The candidate runs of the linked chunks are consecutive
// rows of feat, streamed 64 rows at a time straight into LDS (global_load_lds: each load
// instruction moves 1 KB of consecutive bytes); lane j takes row j of the batch against the
// query row held in registers, and keeps its own sorted list of up to k entries in LDS; k
// rounds of a wave-wide argmin then merge the 64 lists.  The next batch is in flight while a
// batch is scored.
// Writes knn_col (slot-major: entry j of query q at j * p.nq + q) and dk2 per query.
template <int NV>
__global__ __launch_bounds__(64) void k_knn_exact(DevPlan p, const int* __restrict__ qlist, int nq_host,
                                                  const int* __restrict__ nq_dev,
                                                  int* __restrict__ knn_col, double* __restrict__ dk2,
                                                  int* __restrict__ n_cand) {
    extern __shared__ double smem[];
    float4* stage = reinterpret_cast<float4*>(smem);                  // [64][NV] rows of the batch
    double* lt = smem + 64 * NV * 2;                                   // [k][64] distances
    int* lc = reinterpret_cast<int*>(lt + p.k * 64);                   // [k][64] columns (key = col_key[col])
    const int lane = threadIdx.x;
    const int k = p.k;
    const int nq = nq_dev ? *nq_dev : nq_host;
    for (int qi = blockIdx.x; qi < nq; qi += gridDim.x) {
        const int q = qlist ? qlist[qi] : qi;
        const int qc = p.q_col[q];
        const int blk = p.q_blk[q];
        const int s = blk / p.C, c = blk - s * p.C;
        float4 qr[NV];
        {
            const float4* q4 = reinterpret_cast<const float4*>(p.feat + (int64_t)qc * p.DP);
#pragma unroll
            for (int i = 0; i < NV; ++i) qr[i] = q4[i];
        }
        // batches: (run, offset) pairs walked in order; a batch is up to 64 rows of one run
        const int l0 = p.link_ptr[c], l1 = p.link_ptr[c + 1];
        int li = l0, part = 0, off = 0;
        int beg = 0, len = 0;
        auto next_run = [&]() {
            // advance (li, part) to the next non-empty run at or after the current one
            while (li < l1) {
                const int b2 = s * p.C + p.link_idx[li];
                len = p.run_len[2 * b2 + part];
                beg = p.run_beg[2 * b2 + part];
                if (len > 0) return true;
                if (++part == 2) { part = 0; ++li; }
            }
            return false;
        };
        auto advance = [&]() {
            off += 64;
            if (off >= len) {
                off = 0;
                if (++part == 2) { part = 0; ++li; }
                return next_run();
            }
            return true;
        };
        auto load = [&](int rbeg, int nrow) {
            // rows rbeg .. rbeg + nrow - 1 (nrow <= 64) are contiguous; pieces past them re-read
            // the batch's first row (a valid address) and are ignored
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int piece = i * 64 + lane;
                const int row = piece / NV;
                const float* src = p.feat + (int64_t)(rbeg + (row < nrow ? row : 0)) * p.DP + 4 * (piece - row * NV);
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                                 (__attribute__((address_space(3))) void*)(stage + i * 64), 16, 0, 0);
            }
        };
        __syncthreads();   // the previous query's LDS use is over
        int n = 0, total = 0;
        bool more = next_run();
        int cbeg = beg + off, cn = len - off < 64 ? len - off : 64;
        if (more) load(cbeg, cn);
        while (more) {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");     // the batch has landed in LDS
            float4 b[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) b[i] = stage[lane * NV + i];
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");   // batch read before the next lands
            const int bbeg = cbeg, bn = cn;
            more = advance();
            if (more) {
                cbeg = beg + off;
                cn = len - off < 64 ? len - off : 64;
                load(cbeg, cn);
            }
            total += bn;
            if (lane < bn) {
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
                const int col = bbeg + lane;
                const int key = p.col_key[col];
                int pos = -1;
                if (n < k) pos = n++;
                else if (less_tc(p, t, key, lt[(k - 1) * 64 + lane], lc[(k - 1) * 64 + lane])) pos = k - 1;
                if (pos >= 0) {
                    while (pos > 0 && less_tc(p, t, key, lt[(pos - 1) * 64 + lane], lc[(pos - 1) * 64 + lane])) {
                        lt[pos * 64 + lane] = lt[(pos - 1) * 64 + lane];
                        lc[pos * 64 + lane] = lc[(pos - 1) * 64 + lane];
                        --pos;
                    }
                    lt[pos * 64 + lane] = t;
                    lc[pos * 64 + lane] = col;
                }
            }
        }
        const int kk = total < k ? total : k;
        int head = 0;
        double last_t = 0.0;
        int last_c = -1;
        for (int j = 0; j < kk; ++j) {
            double t = head < n ? lt[head * 64 + lane] : __builtin_inf();
            int col = head < n ? lc[head * 64 + lane] : -1;
            int key = head < n ? p.col_key[col] : 0x7fffffff;
            int wl = lane;
            wave_argmin(t, key, col, wl);
            if (wl == lane) ++head;
            if (lane == 0) {
                knn_col[(int64_t)j * p.nq + q] = col;
            }
            last_t = t;
            last_c = col;
        }
        if (lane == 0) {
            for (int j = kk; j < k; ++j) {
                knn_col[(int64_t)j * p.nq + q] = last_c;
            }
            dk2[q] = last_t;
            if (n_cand) n_cand[q] = total;
        }
    }
}

// ---------------------------------------------------------------------------
// Finish the kNN pass: rho (batch mode) or d_k (site mode) and the reported ids.
//   mode 0 (batch): rho_slot[slot] = float(1/sqrt(dk2)); dk_slot (opt) = float(sqrt(dk2));
//                   knn64[slot*k+j] = col_ret (opt; int64 spike ids, miKnn after
//                   int64(viSpk_out(miKnn_in)), matlab/irc2.m:3670)
//   mode 1 (site) : dk_slot[slot] = float(sqrt(dk2)); knn32[slot*k+j] = col_ret (int32 column
//                   indices, search_knn_drift_'s miKnn_in)
// A workgroup takes 64 queries: the slot-major lists are read with coalesced loads into an LDS
// tile, then written out query-major (k contiguous entries per reported slot).
constexpr int FIN_Q = 64;

__global__ __launch_bounds__(256) void k_knn_finish(DevPlan p, int nq, int mode, const int* __restrict__ knn_col,
                                                    const double* __restrict__ dk2, float* __restrict__ rho_slot,
                                                    float* __restrict__ dk_slot, int* __restrict__ knn32,
                                                    int64_t* __restrict__ knn64) {
    __shared__ int tile[IRC_KMAX * FIN_Q];
    const int64_t i0 = (int64_t)blockIdx.x * FIN_Q;
    const int nb = nq - i0 < FIN_Q ? (int)(nq - i0) : FIN_Q;
    const int k = p.k;
    const int t = threadIdx.x;
    if (t < nb) {
        const int slot = p.q_out[i0 + t];
        const double d = dk2[i0 + t];
        if (rho_slot) rho_slot[slot] = (float)(1.0 / sqrt(d));
        if (dk_slot) dk_slot[slot] = (float)sqrt(d);
    }
    if (!knn32 && !knn64) return;
    for (int e = t; e < k * FIN_Q; e += blockDim.x) {
        const int j = e / FIN_Q, c = e - j * FIN_Q;
        if (c < nb) tile[j * FIN_Q + c] = p.col_ret[knn_col[(int64_t)j * p.nq + i0 + c]];
    }
    __syncthreads();
    for (int e = t; e < k * nb; e += blockDim.x) {
        const int c = e / k, j = e - c * k;
        const int64_t slot = p.q_out[i0 + c];
        const int v = tile[j * FIN_Q + c];
        if (knn64) knn64[slot * k + j] = v;
        else knn32[slot * k + j] = v;
    }
}

// The batch finish (mode 0 of k_knn_finish) for the queries list[0 .. *n_list): one thread per query,
// grid-strided -- the queries the exact kernel completed after the re-rank finished the others.
__global__ void k_knn_finish_list(DevPlan p, const int* __restrict__ list, const int* __restrict__ n_list,
                                  const int* __restrict__ knn_col, const double* __restrict__ dk2,
                                  float* __restrict__ rho_slot, float* __restrict__ dk_slot, int64_t* __restrict__ knn64,
                                  float* __restrict__ rho_col) {
    const int n = *n_list;
    const int k = p.k;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int q = list[i];
        const int64_t slot = p.q_out[q];
        const double d = dk2[q];
        if (rho_slot) rho_slot[slot] = (float)(1.0 / sqrt(d));
        if (rho_col) rho_col[p.q_col[q]] = (float)(1.0 / sqrt(d));
        if (dk_slot) dk_slot[slot] = (float)sqrt(d);
        if (knn64)
            for (int j = 0; j < k; ++j) knn64[slot * k + j] = p.col_ret[knn_col[(int64_t)j * p.nq + q]];
    }
}

// Every listed neighbour's distance (the cuda_knn_full output, cuda_knn_full.cu:97-100): one thread
// per (query, slot), the canonical FP64 chain over the padded rows.
__global__ void k_knn_dists(DevPlan p, int nq, const int* __restrict__ knn_col, float* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nq * p.k) return;
    const int q = (int)(e / p.k), j = (int)(e - (int64_t)q * p.k);
    const float4* a = reinterpret_cast<const float4*>(p.feat + (int64_t)p.q_col[q] * p.DP);
    const float4* b = reinterpret_cast<const float4*>(p.feat + (int64_t)knn_col[(int64_t)j * p.nq + q] * p.DP);
    double t = 0.0;
    for (int i = 0; i < p.DP / 4; ++i) {
        const float4 x = a[i], y = b[i];
        double d;
        d = (double)x.x - (double)y.x; t = fma(d, d, t);
        d = (double)x.y - (double)y.y; t = fma(d, d, t);
        d = (double)x.z - (double)y.z; t = fma(d, d, t);
        d = (double)x.w - (double)y.w; t = fma(d, d, t);
    }
    out[(int64_t)p.q_out[q] * p.k + j] = (float)sqrt(t);
}

// rho per column from rho per slot (batch mode: slot = spike id - 1 = col_key - 1)
__global__ void k_rho_cols(int64_t n_cols, const int* __restrict__ col_key, const float* __restrict__ rho_slot,
                           float* __restrict__ rho_col) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_cols) rho_col[i] = rho_slot[col_key[i] - 1];
}

// ---------------------------------------------------------------------------
// Delta from the kNN list: the first kNN entry (in (d2,id) order) whose rho exceeds
// the query's is the nearest denser spike, because kNN(a) is a prefix of C(a) in key
// order.  Its canonical distance is taken from d2head (the exact d^2 of the first IRC_D2HEAD
// entries, written by the re-rank; NaN where unknown) or recomputed here (one row pair per query).
// Queries without such an entry are appended to `unres`.
template <int NV>
__global__ __launch_bounds__(64) void k_delta_from_knn(DevPlan p, int nq, const int* __restrict__ knn_col,
                                                       const float* __restrict__ rho_col,
                                                       const double* __restrict__ d2head,
                                                       float* __restrict__ delta_slot,
                                                       int64_t* __restrict__ nneigh_slot, int* __restrict__ unres,
                                                       int* __restrict__ n_unres) {
    // one wavefront per 64 consecutive queries; rows are gathered cooperatively into LDS
    // (consecutive lanes fetch consecutive 16-byte pieces, global_load_lds) and read back per
    // lane, so a load instruction touches a few rows instead of 64
    __shared__ float4 stage[64 * NV];
    const int l = threadIdx.x;
    const int q = blockIdx.x * 64 + l;
    const bool qv = q < nq;
    const int qc = p.q_col[qv ? q : nq - 1];
    const float ra = rho_col[qc];
    // the list is scanned eight entries at a time: their columns, then their rho, each as eight
    // independent loads (two memory round trips per step instead of two per entry)
    int found = -1, jf = -1;
    if (qv) {
        for (int j0 = 0; j0 < p.k && found < 0; j0 += 8) {
            int cols[8];
            float rh[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) cols[u] = j0 + u < p.k ? knn_col[(int64_t)(j0 + u) * p.nq + q] : qc;
#pragma unroll
            for (int u = 0; u < 8; ++u) rh[u] = rho_col[cols[u]];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (found < 0 && j0 + u < p.k && rh[u] > ra) {
                    found = cols[u];
                    jf = j0 + u;
                }
            }
        }
    }
    double t = __builtin_nan("");
    if (found >= 0 && d2head && jf < IRC_D2HEAD) t = d2head[(int64_t)jf * p.nq + q];
    const bool need = found >= 0 && __builtin_isnan(t);
    const uint64_t nmask = __ballot(need);
    // the rows of the lanes in nmask: load instruction i carries pieces of the rows of lanes
    // i*64/NV .. (i*64+63)/NV and is skipped when none of them is needed; the pieces of unneeded rows
    // it does carry come from row 0
    auto gather = [&](int col) {
        int crow[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) crow[i] = __shfl(col, (i * 64 + l) / NV, 64);
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            constexpr uint64_t one = 1;
            const int o_lo = i * 64 / NV, o_hi = (i * 64 + 63) / NV;
            const uint64_t span = (o_hi - o_lo == 63 ? ~(uint64_t)0 : ((one << (o_hi - o_lo + 1)) - 1)) << o_lo;
            if ((nmask & span) == 0) continue;
            const int piece = i * 64 + l;
            const int owner = piece / NV;
            const int part = piece - owner * NV;
            const int row = ((nmask >> owner) & 1ull) ? crow[i] : 0;
            const float* src = p.feat + (int64_t)row * p.DP + 4 * part;
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                             (__attribute__((address_space(3))) void*)(stage + i * 64), 16, 0, 0);
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // the rows have landed in LDS
    };
    if (nmask) {
        gather(qc);
        float4 a[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) a[i] = stage[l * NV + i];
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");   // read back before the next gather lands
        gather(found >= 0 ? found : qc);
        if (need) {
            double u = 0.0;
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const float4 x = a[i], y = stage[l * NV + i];
                double d;
                d = (double)x.x - (double)y.x; u = fma(d, d, u);
                d = (double)x.y - (double)y.y; u = fma(d, d, u);
                d = (double)x.z - (double)y.z; u = fma(d, d, u);
                d = (double)x.w - (double)y.w; u = fma(d, d, u);
            }
            t = u;
        }
    }
    if (found >= 0) {
        const int slot = p.q_out[q];
        const float dl = (float)(sqrt(t) * (double)ra);
        if (__builtin_isnan(dl) || __builtin_isinf(dl)) {
            delta_slot[slot] = (float)sqrt(IRC_SINGLE_INF_D);
            nneigh_slot[slot] = p.col_ret[qc];
        } else {
            delta_slot[slot] = dl;
            nneigh_slot[slot] = p.col_ret[found];
        }
    }
    if (qv && found < 0) {
        const int u = atomicAdd(n_unres, 1);
        unres[u] = q;
    }
}

// ---------------------------------------------------------------------------
// Exact nearest denser neighbour: one wavefront per query, lanes over candidates.  Only denser
// candidates (rho_b > rho_a, usually a few percent for the queries that reach this kernel) are
// scored: each 64-candidate chunk is filtered on rho and its denser columns are appended to an
// LDS queue (ballot + prefix count), and the FP64 chain runs once per 64 queued columns, not once
// per chunk that holds any.  The block's candidate runs (up to 64 at a time) are flattened into
// one position stream by a prefix sum held in LDS, so a step loads the rho of 512 consecutive
// positions across run boundaries: a few dependent load rounds per query instead of one per run.
// The (d2, id) minimum does not depend on the scoring order.  The query row is held in registers.
//   mode 0 (batch): delta = float(sqrt(d2)*rho_a) with the sentinel rule; nneigh int64
//   mode 1 (site) : dmin = float(sqrt(d2)) or +Inf; nneigh32 = col_ret (own when none)
// The query list is qlist[0 .. *nq_dev) when nq_dev is set, else qlist[0..nq) / identity.
template <int NV>
__global__ __launch_bounds__(64) void k_delta_exact(DevPlan p, const int* __restrict__ qlist, int nq,
                                                    const int* __restrict__ nq_dev, int mode,
                                                    const float* __restrict__ rho_col,
                                                    float* __restrict__ delta_slot,
                                                    int64_t* __restrict__ nneigh64,
                                                    int* __restrict__ nneigh32) {
    __shared__ int qbuf[128];   // queued denser columns (fewer than 64 carried + one chunk)
    __shared__ int s_rb[64];     // a batch of the block's runs: first column
    __shared__ int s_ro[65];     // and first stream position ([64 or fewer] = the batch's length)
    const int lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    const int n = nq_dev ? *nq_dev : nq;
    for (int qi = blockIdx.x; qi < n; qi += gridDim.x) {
        const int q = qlist ? qlist[qi] : qi;
        const int qc = p.q_col[q];
        const int blk = p.q_blk[q];
        const int s = blk / p.C, c = blk - s * p.C;
        const float ra = rho_col[qc];
        float4 qr[NV];
        {
            const float4* q4 = reinterpret_cast<const float4*>(p.feat + (int64_t)qc * p.DP);
#pragma unroll
            for (int i = 0; i < NV; ++i) qr[i] = q4[i];
        }
        double bt = __builtin_inf();
        int bk = 0x7fffffff, bc = -1;
        // lane l scores queued column l (l < cnt)
        auto score = [&](int cnt) {
            if (lane < cnt) {
                const int col = qbuf[lane];
                const float4* x4 = reinterpret_cast<const float4*>(p.feat + (int64_t)col * p.DP);
                float4 b[NV];
#pragma unroll
                for (int i = 0; i < NV; ++i) b[i] = x4[i];
                double t = 0.0;
#pragma unroll
                for (int i = 0; i < NV; ++i) {
                    double d;
                    d = (double)qr[i].x - (double)b[i].x; t = fma(d, d, t);
                    d = (double)qr[i].y - (double)b[i].y; t = fma(d, d, t);
                    d = (double)qr[i].z - (double)b[i].z; t = fma(d, d, t);
                    d = (double)qr[i].w - (double)b[i].w; t = fma(d, d, t);
                }
                const int key = p.col_key[col];
                if (bc < 0 || irc_less(t, key, bt, bk)) { bt = t; bk = key; bc = col; }
            }
        };
        int nbuf = 0;   // wave-uniform queue length, < 64 between chunks
        const int lp0 = p.link_ptr[c], ne = 2 * (p.link_ptr[c + 1] - lp0);   // runs: (link, part)
        for (int e0 = 0; e0 < ne; e0 += 64) {
            // this batch's runs, one lane each, and their stream offsets (inclusive prefix sum)
            const int e = e0 + lane;
            int len = 0, beg = 0;
            if (e < ne) {
                const int b2 = s * p.C + p.link_idx[lp0 + (e >> 1)];
                len = p.run_len[2 * b2 + (e & 1)];
                beg = p.run_beg[2 * b2 + (e & 1)];
            }
            int incl = len;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int t = __shfl_up(incl, d, 64);
                if (lane >= d) incl += t;
            }
            const int total = __shfl(incl, 63, 64);
            __syncthreads();   // the previous batch's table is no longer read
            s_rb[lane] = beg;
            s_ro[lane] = incl - len;
            if (lane == 63) s_ro[64] = total;
            __syncthreads();
            // eight chunks' rho loads in flight at once, then filtered chunk by chunk; each lane's
            // positions ascend, so its run index only moves forward (empty runs are stepped over)
            int r = 0;
            for (int j8 = 0; j8 < total; j8 += 512) {
                int cj[8];
                float rj[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int j = j8 + 64 * u + lane;
                    cj[u] = -1;
                    if (j < total) {
                        while (s_ro[r + 1] <= j) ++r;
                        cj[u] = s_rb[r] + (j - s_ro[r]);
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) rj[u] = cj[u] >= 0 ? rho_col[cj[u]] : 0.0f;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool den = cj[u] >= 0 && rj[u] > ra;
                    const uint64_t m = __ballot(den);
                    if (den) qbuf[nbuf + __popcll(m & below)] = cj[u];
                    nbuf += __popcll(m);
                    if (nbuf >= 64) {
                        __syncthreads();
                        score(64);
                        const int rest = nbuf - 64;
                        const int mv = lane < rest ? qbuf[64 + lane] : 0;
                        __syncthreads();
                        if (lane < rest) qbuf[lane] = mv;
                        nbuf = rest;
                    }
                }
            }
        }
        __syncthreads();
        score(nbuf);
        __syncthreads();   // the queue is reused by the next query
        int wl = lane;
        if (bc < 0) { bt = __builtin_inf(); bk = 0x7fffffff; }
        wave_argmin(bt, bk, bc, wl);
        if (lane == 0) {
            const int slot = p.q_out[q];
            if (mode == 0) {
                float dl = bc < 0 ? __builtin_inff() : (float)(sqrt(bt) * (double)ra);
                if (bc < 0 || __builtin_isnan(dl) || __builtin_isinf(dl)) {
                    delta_slot[slot] = (float)sqrt(IRC_SINGLE_INF_D);
                    nneigh64[slot] = p.col_ret[qc];
                } else {
                    delta_slot[slot] = dl;
                    nneigh64[slot] = p.col_ret[bc];
                }
            } else {
                if (bc < 0) {
                    delta_slot[slot] = __builtin_inff();
                    nneigh32[slot] = p.col_ret[qc];
                } else {
                    delta_slot[slot] = (float)sqrt(bt);
                    nneigh32[slot] = p.col_ret[bc];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Planning kernels (batch form).

// Pack D x n (column-major) features into [n][DP] rows, zero padded.
// grid-stride: a dispatch covers at most 2^32 work-items, fewer than n_cols * DP at 100M columns
__global__ void k_pack_fp32(int64_t n_cols, int D, int DP, const float* __restrict__ F, float* __restrict__ feat) {
    const int64_t total = n_cols * DP;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t col = e / DP;
        const int d = (int)(e - col * DP);
        feat[e] = d < D ? F[col * D + d] : 0.0f;
    }
}

__device__ __forceinline__ int64_t lower_bound_i32(const int* a, int64_t lo, int64_t hi, int64_t v) {
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((int64_t)a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// One thread per (site, chunk, part): the column run of that block part.
__global__ void k_block_runs(int S, int C, const int64_t* __restrict__ site_ptr, const int64_t* __restrict__ site_np,
                             const int64_t* __restrict__ lim, const int* __restrict__ col_spike,
                             int* __restrict__ run_beg, int* __restrict__ run_len) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)S * C * 2) return;
    const int part = (int)(e & 1);
    const int64_t b = e >> 1;
    const int s = (int)(b / C), c = (int)(b - (int64_t)s * C);
    const int64_t lo = part == 0 ? site_ptr[s] : site_ptr[s] + site_np[s];
    const int64_t hi = part == 0 ? site_ptr[s] + site_np[s] : site_ptr[s + 1];
    const int64_t a0 = lower_bound_i32(col_spike, lo, hi, lim[c]);
    const int64_t a1 = lower_bound_i32(col_spike, a0, hi, lim[c + 1]);
    run_beg[e] = (int)a0;
    run_len[e] = (int)(a1 - a0);
}

// One thread per query (primary column): column, block and output slot.
// Flags bit 0: spike id out of range, bit 1: id outside the drift limits,
// bit 2: ids not ascending within a site part.
__global__ void k_batch_queries(int64_t N, int S, int C, const int64_t* __restrict__ site_ptr,
                                const int64_t* __restrict__ prim_ptr, const int64_t* __restrict__ lim,
                                const int* __restrict__ col_spike, int* __restrict__ q_col,
                                int* __restrict__ q_blk, int* __restrict__ q_out, int* __restrict__ flags) {
    const int64_t qi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= N) return;
    int lo = 0, hi = S;  // largest s with prim_ptr[s] <= qi
    while (hi - lo > 1) { int mid = (lo + hi) >> 1; if (prim_ptr[mid] <= qi) lo = mid; else hi = mid; }
    const int s = lo;
    const int64_t col = site_ptr[s] + (qi - prim_ptr[s]);
    const int id = col_spike[col];
    if (id < 1 || id > N) { atomicOr(flags, 1); return; }
    int cl = 0, ch = C;  // largest c with lim[c] <= id
    if (id < lim[0] || id >= lim[C]) { atomicOr(flags, 2); return; }
    while (ch - cl > 1) { int mid = (cl + ch) >> 1; if (lim[mid] <= id) cl = mid; else ch = mid; }
    if (qi > prim_ptr[s] && col_spike[col - 1] >= id) atomicOr(flags, 4);
    q_col[qi] = (int)col;
    q_blk[qi] = s * C + cl;
    q_out[qi] = id - 1;
}

// Secondary-part ordering and range check.
__global__ void k_check_secondary(int64_t N, int S, const int64_t* __restrict__ site_ptr,
                                  const int64_t* __restrict__ site_np, const int* __restrict__ col_spike,
                                  int* __restrict__ flags) {
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= site_ptr[S]) return;
    int lo = 0, hi = S;
    while (hi - lo > 1) { int mid = (lo + hi) >> 1; if (site_ptr[mid] <= col) lo = mid; else hi = mid; }
    const int64_t b = site_ptr[lo] + site_np[lo];
    if (col < b) return;
    const int id = col_spike[col];
    if (id < 1 || id > N) atomicOr(flags, 1);
    if (col > b && col_spike[col - 1] >= id) atomicOr(flags, 4);
}

// ---------------------------------------------------------------------------
// Host launchers
namespace irc {

hipError_t launch_pack_fp32(hipStream_t st, int64_t n_cols, int D, int DP, const float* F, float* feat) {
    const int64_t total = n_cols * DP;
    if (total == 0) return hipSuccess;
    const int64_t blocks = (total + 255) / 256;
    k_pack_fp32<<<(unsigned)(blocks < (1 << 20) ? blocks : (1 << 20)), 256, 0, st>>>(n_cols, D, DP, F, feat);
    return hipGetLastError();
}

hipError_t launch_block_runs(hipStream_t st, int S, int C, const int64_t* site_ptr, const int64_t* site_np,
                             const int64_t* lim, const int* col_spike, int* run_beg, int* run_len) {
    const int64_t total = (int64_t)S * C * 2;
    k_block_runs<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(S, C, site_ptr, site_np, lim, col_spike, run_beg, run_len);
    return hipGetLastError();
}

hipError_t launch_batch_queries(hipStream_t st, int64_t N, int S, int C, const int64_t* site_ptr,
                                const int64_t* site_np, const int64_t* prim_ptr, const int64_t* lim,
                                const int* col_spike, int64_t n_cols, int* q_col, int* q_blk, int* q_out, int* flags) {
    k_batch_queries<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(N, S, C, site_ptr, prim_ptr, lim, col_spike,
                                                                 q_col, q_blk, q_out, flags);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (n_cols > 0)
        k_check_secondary<<<(unsigned)((n_cols + 255) / 256), 256, 0, st>>>(N, S, site_ptr, site_np, col_spike, flags);
    return hipGetLastError();
}

size_t knn_exact_lds_bytes(int k, int DP) { return (size_t)64 * DP * 4 + (size_t)k * 64 * (8 + 4); }

namespace {
hipError_t knn_exact_go(hipStream_t st, const DevPlan& p, const int* qlist, int nq, const int* nq_dev, int grid,
                        int* knn_col, double* dk2, int* n_cand) {
    const int g = grid < nq ? grid : nq;
    const size_t lds = knn_exact_lds_bytes(p.k, p.DP);
#define IRC_EX(NV)                                                                                           \
    case NV:                                                                                                 \
        k_knn_exact<NV><<<g, 64, lds, st>>>(p, qlist, nq, nq_dev, knn_col, dk2, n_cand);                     \
        break;
    switch (p.DP / 4) {
        IRC_EX(1) IRC_EX(2) IRC_EX(3) IRC_EX(4) IRC_EX(5) IRC_EX(6) IRC_EX(7) IRC_EX(8)
        IRC_EX(9) IRC_EX(10) IRC_EX(11) IRC_EX(12) IRC_EX(13) IRC_EX(14) IRC_EX(15) IRC_EX(16)
        default: return hipErrorInvalidValue;
    }
#undef IRC_EX
    return hipGetLastError();
}
}  // namespace

hipError_t launch_knn_exact(hipStream_t st, const DevPlan& p, const int* qlist, int nq, int grid,
                            int* knn_col, double* dk2, int* n_cand) {
    if (nq <= 0) return hipSuccess;
    return knn_exact_go(st, p, qlist, nq, nullptr, grid, knn_col, dk2, n_cand);
}

hipError_t launch_knn_exact_dev(hipStream_t st, const DevPlan& p, const int* qlist, const int* nq_dev, int nq_max,
                                int grid, int* knn_col, double* dk2) {
    if (nq_max <= 0) return hipSuccess;
    return knn_exact_go(st, p, qlist, nq_max, nq_dev, grid, knn_col, dk2, nullptr);
}

hipError_t launch_knn_finish(hipStream_t st, const DevPlan& p, int nq, int mode, const int* knn_col,
                             const double* dk2, float* rho_slot, float* dk_slot, int* knn32, int64_t* knn64) {
    if (nq <= 0) return hipSuccess;
    k_knn_finish<<<(unsigned)((nq + FIN_Q - 1) / FIN_Q), 256, 0, st>>>(p, nq, mode, knn_col, dk2, rho_slot, dk_slot,
                                                                       knn32, knn64);
    return hipGetLastError();
}

hipError_t launch_knn_finish_list(hipStream_t st, const DevPlan& p, const int* list, const int* n_list, int nq_max,
                                  const int* knn_col, const double* dk2, float* rho_slot, float* dk_slot,
                                  int64_t* knn64, float* rho_col) {
    if (nq_max <= 0) return hipSuccess;
    const int grid = (nq_max + 255) / 256 < 1024 ? (nq_max + 255) / 256 : 1024;
    k_knn_finish_list<<<grid, 256, 0, st>>>(p, list, n_list, knn_col, dk2, rho_slot, dk_slot, knn64, rho_col);
    return hipGetLastError();
}

hipError_t launch_knn_dists(hipStream_t st, const DevPlan& p, int nq, const int* knn_col, float* out) {
    const int64_t n = (int64_t)nq * p.k;
    if (n <= 0) return hipSuccess;
    k_knn_dists<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, nq, knn_col, out);
    return hipGetLastError();
}

hipError_t launch_rho_cols(hipStream_t st, int64_t n_cols, const int* col_key, const float* rho_slot, float* rho_col) {
    if (n_cols <= 0) return hipSuccess;
    k_rho_cols<<<(unsigned)((n_cols + 255) / 256), 256, 0, st>>>(n_cols, col_key, rho_slot, rho_col);
    return hipGetLastError();
}

hipError_t launch_delta_from_knn(hipStream_t st, const DevPlan& p, int nq, const int* knn_col,
                                 const float* rho_col, float* delta_slot, int64_t* nneigh_slot, int* unres,
                                 int* n_unres, const double* d2head) {
    if (nq <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((nq + 63) / 64);
#define IRC_DK(NV)                                                                                       \
    case NV:                                                                                             \
        k_delta_from_knn<NV><<<grid, 64, 0, st>>>(p, nq, knn_col, rho_col, d2head, delta_slot, nneigh_slot, \
                                                  unres, n_unres);                                       \
        break;
    switch (p.DP / 4) {
        IRC_DK(1) IRC_DK(2) IRC_DK(3) IRC_DK(4) IRC_DK(5) IRC_DK(6) IRC_DK(7) IRC_DK(8)
        IRC_DK(9) IRC_DK(10) IRC_DK(11) IRC_DK(12) IRC_DK(13) IRC_DK(14) IRC_DK(15) IRC_DK(16)
        default: return hipErrorInvalidValue;
    }
#undef IRC_DK
    return hipGetLastError();
}

hipError_t launch_delta_exact(hipStream_t st, const DevPlan& p, const int* qlist, int nq, const int* nq_dev,
                              int grid, int mode, const float* rho_col, float* delta_slot, int64_t* nneigh64,
                              int* nneigh32) {
    if (nq <= 0) return hipSuccess;
    const int g = grid < nq ? grid : nq;
#define IRC_DX(NV)                                                                                        \
    case NV:                                                                                              \
        k_delta_exact<NV><<<g, 64, 0, st>>>(p, qlist, nq, nq_dev, mode, rho_col, delta_slot, nneigh64, nneigh32); \
        break;
    switch (p.DP / 4) {
        IRC_DX(1) IRC_DX(2) IRC_DX(3) IRC_DX(4) IRC_DX(5) IRC_DX(6) IRC_DX(7) IRC_DX(8)
        IRC_DX(9) IRC_DX(10) IRC_DX(11) IRC_DX(12) IRC_DX(13) IRC_DX(14) IRC_DX(15) IRC_DX(16)
        default: return hipErrorInvalidValue;
    }
#undef IRC_DX
    return hipGetLastError();
}

}  // namespace irc
