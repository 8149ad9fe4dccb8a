// irc_api.hip -- C ABI of the gfx950 density-peak library (see include/irc_hip.h).
//
// Host-side planning restates the reference's data preparation:
//   prepare_page_site_ (matlab/irc2.m:3869-3910): site-local columns, drift per column
//   search_knn_drift_  (matlab/irc2.m:3744-3748, 3768-3772): candidate runs of linked chunks
// and replaces the per-(site, chunk) kernel launches by whole-problem launches.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "irc_common.h"
#include "irc_kernels.h"
#include "irc_assign.h"
#include "irc_drift.h"
#include "irc_plan.h"
#include "irc_screen.h"
#include "../../include/irc_hip.h"

// Scratch of irc_rho_delta_batch_noalloc: every buffer such a call needs is carved from this array.
// Each GPU holds its own copy for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (irc_ctx::scratch), never through its name.
#define IRC_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) char scratch[IRC_NOALLOC_SCRATCH_BYTES];

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct Bufs {
    DevBuf b_F, b_colspk, b_feat, b_key, b_ret, b_runb, b_runl, b_lptr, b_lidx, b_qcol, b_qblk, b_qout;
    DevBuf b_small, b_flags, b_knncol, b_dk2, b_rhoslot, b_rhocol, b_unres, b_out_delta, b_out_nn,
        b_out_knn, b_out_dk, b_out_nn32, b_out_knn32, b_ncand, b_d2head;
    DevBuf b_scr[6];   // screen: [0] operand rows, [1] FP8 site scales, [2] norms + run maxima, [3] items,
                       // [4] fallback/spill lists, [5] counters
    DevBuf b_cand_n, b_cand, b_cand_eps;   // certified candidate lists of the screen
    DevBuf b_plan;                                         // device-side work planning
    DevBuf b_drift;                                        // drift linkage scratch
    DevBuf b_assign;                                       // chain labelling scratch
    DevBuf b_siteseg;     // per-site segment tables (IRC_FLAG_SITE_LAUNCHES)
    DevBuf b_tune;        // the autotuner's copy of the base setting's outputs
    DevBuf b_dstat;       // device counters of the call: [0] fallbacks after the screen, [1] after the
                          // re-rank, [2] pairs, [3] queries the re-rank's second pass redid, [4] queries
                          // whose denser neighbour was outside the kNN list
    // every buffer, for release
    std::vector<DevBuf*> all() {
        std::vector<DevBuf*> v{&b_F, &b_colspk, &b_feat, &b_key, &b_ret, &b_runb, &b_runl, &b_lptr, &b_lidx,
                               &b_qcol, &b_qblk, &b_qout, &b_small, &b_flags, &b_knncol, &b_dk2, &b_rhoslot,
                               &b_rhocol, &b_unres, &b_out_delta, &b_out_nn, &b_out_knn, &b_out_dk, &b_out_nn32,
                               &b_out_knn32, &b_ncand, &b_cand_n, &b_cand, &b_cand_eps, &b_plan,
                               &b_drift, &b_assign, &b_siteseg, &b_tune, &b_dstat, &b_d2head};
        for (auto& b : b_scr) v.push_back(&b);
        return v;
    }
};

}  // namespace

struct irc_ctx {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    std::string err;
    irc_stats stats{};
    int flags = 0;       // irc_options.flags of the current call
    int waves_per_cu = 8;  // persistent grid of the screen
    int large_ncand = 8192;
    Bufs heap;           // device buffers, grown on demand and kept across calls
    Bufs arena;          // the same buffers carved from the static scratch array (no-allocation calls)
    Bufs* bb = &heap;    // the set in use
    char* scratch = nullptr;   // device address of the scratch array on this context's device
    size_t scratch_used = 0;   // bytes carved in the current no-allocation call
    bool arena_on = false;
    bool dstat_screen = false;  // [0], [1] hold this call's screen counters
    bool capturing = false;     // the stream is being captured into a graph: no events, no host copies
    hipGraphExec_t graph_exec = nullptr;  // the batch call's device phase, for replays
    uint64_t graph_key = 0;     // descriptor hash of the captured call
    uint64_t buf_gen = 0;       // bumped whenever a device buffer is (re)allocated
    uint64_t graph_gen = 0;     // buf_gen when the graph was captured: a replay needs it unchanged
    uint64_t desc_key = 0;      // descriptor whose runs/queries/links the planning buffers hold (0: none)
    std::vector<std::pair<uint64_t, irc_options>> tuned;   // autotuned options per workload shape
};

namespace {

irc_status fail(irc_ctx* c, irc_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return st;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(ctx, IRC_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                  \
    } while (0)

#define CHK(expr)                          \
    do {                                   \
        irc_status s_ = (expr);            \
        if (s_ != IRC_OK) return s_;       \
    } while (0)

irc_status ensure(irc_ctx* ctx, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (bytes <= b.cap) return IRC_OK;
    if (ctx->arena_on) {
        // a no-allocation call: carve from the scratch array (never freed; the set is reset per call)
        const size_t need = (bytes + 255) & ~(size_t)255;
        if (ctx->scratch_used + need > IRC_NOALLOC_SCRATCH_BYTES)
            return fail(ctx, IRC_E_UNSUPPORTED,
                        "problem needs more than the %d MiB scratch of irc_rho_delta_batch_noalloc",
                        (int)(IRC_NOALLOC_SCRATCH_BYTES >> 20));
        b.p = ctx->scratch + ctx->scratch_used;
        b.cap = need;
        ctx->scratch_used += need;
        return IRC_OK;
    }
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return fail(ctx, IRC_E_HIP, "hipFree failed: %s", hipGetErrorString(e));
    }
    size_t want = bytes + bytes / 8;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return fail(ctx, IRC_E_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    b.cap = want;
    // kernels captured in a graph hold the old address: a replay must not run on it
    ++ctx->buf_gen;
    return IRC_OK;
}

template <class T>
T* ptr(DevBuf& b) { return reinterpret_cast<T*>(b.p); }

int pad_dim(int D) { return (D + 3) / 4 * 4; }

irc_status check_opt(irc_ctx* ctx, const irc_options* opt) {
    if (opt && opt->force_fail) return fail(ctx, IRC_E_FORCED, "failure forced by irc_options.force_fail");
    ctx->flags = opt ? opt->flags : 0;
    ctx->waves_per_cu = opt && opt->waves_per_cu > 0 ? opt->waves_per_cu : 8;
    ctx->large_ncand = opt && opt->large_ncand > 0 ? opt->large_ncand : 8192;
    if (ctx->waves_per_cu > 32) return fail(ctx, IRC_E_ARG, "waves_per_cu=%d above 32", ctx->waves_per_cu);
    return IRC_OK;
}

int persistent_grid(const irc_ctx* ctx, int waves_per_cu) { return ctx->n_cu * waves_per_cu; }

// Link lists L(c) = {r : link[r + c*C] != 0}, ascending r.
void build_links(int C, const uint8_t* link, std::vector<int>& lptr, std::vector<int>& lidx) {
    lptr.assign((size_t)C + 1, 0);
    lidx.clear();
    for (int c = 0; c < C; ++c) {
        for (int r = 0; r < C; ++r)
            if (link[(size_t)r + (size_t)c * C]) lidx.push_back(r);
        lptr[c + 1] = (int)lidx.size();
    }
}

// Every chunk must be linked to itself (calc_drift_ always keeps a chunk among its own
// nearest chunks, matlab/irc2.m:6433-6436), so every query has at least itself as candidate.
bool links_have_diagonal(int C, const uint8_t* link) {
    for (int c = 0; c < C; ++c)
        if (!link[(size_t)c + (size_t)c * C]) return false;
    return true;
}

// Host-side plan of one site problem (per-site API and the legacy A/B forms).
struct HostPlan {
    int D = 0, DP = 0, C = 0, k = 0;
    int64_t n_cols = 0;
    std::vector<float> feat;             // [n_cols][DP], chunk-grouped column order
    std::vector<int> key, ret;           // per (permuted) column
    std::vector<int> runb, runl;         // [C*2]
    std::vector<int> lptr, lidx;
    std::vector<int> qcol, qblk, qout;
    std::vector<int64_t> inv;            // original column -> permuted column
};

// Upload a host plan into the context buffers and return the device view.
irc_status upload_plan(irc_ctx* ctx, const HostPlan& h, DevPlan& p) {
    const int64_t n = h.n_cols;
    ctx->desc_key = 0;   // the batch call's planning buffers are overwritten below
    CHK(ensure(ctx, ctx->bb->b_feat, sizeof(float) * (size_t)n * h.DP));
    CHK(ensure(ctx, ctx->bb->b_key, sizeof(int) * (size_t)n));
    CHK(ensure(ctx, ctx->bb->b_ret, sizeof(int) * (size_t)n));
    CHK(ensure(ctx, ctx->bb->b_runb, sizeof(int) * h.runb.size()));
    CHK(ensure(ctx, ctx->bb->b_runl, sizeof(int) * h.runl.size()));
    CHK(ensure(ctx, ctx->bb->b_lptr, sizeof(int) * h.lptr.size()));
    CHK(ensure(ctx, ctx->bb->b_lidx, sizeof(int) * (h.lidx.size() + 1)));
    CHK(ensure(ctx, ctx->bb->b_qcol, sizeof(int) * h.qcol.size()));
    CHK(ensure(ctx, ctx->bb->b_qblk, sizeof(int) * h.qblk.size()));
    CHK(ensure(ctx, ctx->bb->b_qout, sizeof(int) * h.qout.size()));
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemcpyAsync(ctx->bb->b_feat.p, h.feat.data(), sizeof(float) * h.feat.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_key.p, h.key.data(), sizeof(int) * h.key.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_ret.p, h.ret.data(), sizeof(int) * h.ret.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_runb.p, h.runb.data(), sizeof(int) * h.runb.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_runl.p, h.runl.data(), sizeof(int) * h.runl.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_lptr.p, h.lptr.data(), sizeof(int) * h.lptr.size(), hipMemcpyHostToDevice, st));
    if (!h.lidx.empty())
        HIPCHK(hipMemcpyAsync(ctx->bb->b_lidx.p, h.lidx.data(), sizeof(int) * h.lidx.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_qcol.p, h.qcol.data(), sizeof(int) * h.qcol.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_qblk.p, h.qblk.data(), sizeof(int) * h.qblk.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_qout.p, h.qout.data(), sizeof(int) * h.qout.size(), hipMemcpyHostToDevice, st));
    p.feat = ptr<float>(ctx->bb->b_feat);
    p.DP = h.DP;
    p.D = h.D;
    p.C = h.C;
    p.k = h.k;
    p.col_key = ptr<int>(ctx->bb->b_key);
    p.col_ret = ptr<int>(ctx->bb->b_ret);
    p.run_beg = ptr<int>(ctx->bb->b_runb);
    p.run_len = ptr<int>(ctx->bb->b_runl);
    p.link_ptr = ptr<int>(ctx->bb->b_lptr);
    p.link_idx = ptr<int>(ctx->bb->b_lidx);
    p.q_col = ptr<int>(ctx->bb->b_qcol);
    p.q_blk = ptr<int>(ctx->bb->b_qblk);
    p.q_out = ptr<int>(ctx->bb->b_qout);
    p.nq = (int)h.qcol.size();
    p.n_blocks = (int)(h.runb.size() / 2);
    return IRC_OK;
}

// Plan of a per-site problem: columns regrouped by drift chunk (stable), one run per chunk.
irc_status plan_site(irc_ctx* ctx, const irc_site_problem* pr, HostPlan& h) {
    if (!pr || !pr->F || !pr->drift_of_col || !pr->link) return fail(ctx, IRC_E_ARG, "null pointer in irc_site_problem");
    if (pr->D <= 0 || pr->k <= 0 || pr->n_cols <= 0 || pr->n_drift <= 0)
        return fail(ctx, IRC_E_ARG, "invalid sizes (D=%d k=%d n_cols=%lld n_drift=%d)", pr->D, pr->k,
                    (long long)pr->n_cols, pr->n_drift);
    if (pr->D > IRC_DMAX) return fail(ctx, IRC_E_UNSUPPORTED, "D=%d exceeds %d", pr->D, IRC_DMAX);
    if (pr->k > IRC_KMAX) return fail(ctx, IRC_E_UNSUPPORTED, "k=%d exceeds %d", pr->k, IRC_KMAX);
    if (pr->n_cols >= (int64_t)1 << 31) return fail(ctx, IRC_E_UNSUPPORTED, "n_cols too large");
    if (pr->in_begin < 0 || pr->n_in < 0 || pr->in_begin + pr->n_in > pr->n_cols)
        return fail(ctx, IRC_E_ARG, "query range [%lld, +%lld) outside %lld columns", (long long)pr->in_begin,
                    (long long)pr->n_in, (long long)pr->n_cols);
    const int C = pr->n_drift;
    const int64_t n = pr->n_cols;
    if (!links_have_diagonal(C, pr->link)) return fail(ctx, IRC_E_ARG, "link(c,c) must be set for every chunk");
    std::vector<int64_t> cnt((size_t)C + 1, 0);
    for (int64_t j = 0; j < n; ++j) {
        const int r = pr->drift_of_col[j];
        if (r < 1 || r > C) return fail(ctx, IRC_E_ARG, "drift_of_col[%lld]=%d outside 1..%d", (long long)j, r, C);
        cnt[r]++;
    }
    for (int c = 0; c < C; ++c) cnt[c + 1] += cnt[c];
    h.D = pr->D;
    h.DP = pad_dim(pr->D);
    h.C = C;
    h.k = pr->k;
    h.n_cols = n;
    h.feat.assign((size_t)n * h.DP, 0.0f);
    h.key.resize(n);
    h.ret.resize(n);
    h.inv.resize(n);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t j = 0; j < n; ++j) {
        const int r = pr->drift_of_col[j] - 1;
        const int64_t dst = pos[r]++;
        h.inv[j] = dst;
        memcpy(&h.feat[(size_t)dst * h.DP], pr->F + (size_t)j * pr->D, sizeof(float) * pr->D);
        int64_t key = pr->id_of_col ? pr->id_of_col[j] : j;
        if (key < INT32_MIN || key > INT32_MAX) return fail(ctx, IRC_E_ARG, "id_of_col out of int32 range");
        h.key[dst] = (int)key;
        h.ret[dst] = (int)(j + 1);
    }
    h.runb.assign((size_t)C * 2, 0);
    h.runl.assign((size_t)C * 2, 0);
    for (int c = 0; c < C; ++c) {
        h.runb[2 * c] = (int)cnt[c];
        h.runl[2 * c] = (int)(cnt[c + 1] - cnt[c]);
        h.runb[2 * c + 1] = (int)cnt[c + 1];
    }
    build_links(C, pr->link, h.lptr, h.lidx);
    h.qcol.resize(pr->n_in);
    h.qblk.resize(pr->n_in);
    h.qout.resize(pr->n_in);
    for (int64_t i = 0; i < pr->n_in; ++i) {
        const int64_t j = pr->in_begin + i;
        h.qcol[i] = (int)h.inv[j];
        h.qblk[i] = pr->drift_of_col[j] - 1;
        h.qout[i] = (int)i;
    }
    return IRC_OK;
}

// Plan of an A/B problem: chunk 0 = B (candidates), chunk 1 = A (queries), link 0 -> 1.  With IB,
// chunk 0 holds the listed columns IB[j] - 1 of B in list order (key and reported index: the column).
void plan_ab(int D, int64_t nB, const float* B, int64_t nA, const float* A, int k, HostPlan& h,
             const int32_t* IB = nullptr, int64_t nIB = 0) {
    const int64_t n0 = IB ? nIB : nB;
    h.D = D;
    h.DP = pad_dim(D);
    h.C = 2;
    h.k = k;
    h.n_cols = n0 + nA;
    h.feat.assign((size_t)h.n_cols * h.DP, 0.0f);
    h.key.resize(h.n_cols);
    h.ret.resize(h.n_cols);
    for (int64_t j = 0; j < n0; ++j) {
        const int64_t col = IB ? (int64_t)IB[j] - 1 : j;
        memcpy(&h.feat[(size_t)j * h.DP], B + (size_t)col * D, sizeof(float) * D);
        h.key[j] = (int)col;
        h.ret[j] = (int)(col + 1);
    }
    for (int64_t i = 0; i < nA; ++i) {
        memcpy(&h.feat[(size_t)(n0 + i) * h.DP], A + (size_t)i * D, sizeof(float) * D);
        h.key[n0 + i] = (int)(n0 + i);
        h.ret[n0 + i] = 0;
    }
    h.runb = {0, (int)n0, (int)n0, (int)(n0 + nA)};
    h.runl = {(int)n0, 0, (int)nA, 0};
    h.lptr = {0, 0, 1};
    h.lidx = {0};
    h.qcol.resize(nA);
    h.qblk.assign(nA, 1);
    h.qout.resize(nA);
    for (int64_t i = 0; i < nA; ++i) {
        h.qcol[i] = (int)(n0 + i);
        h.qout[i] = (int)i;
    }
}

double event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms;
}

// query groups smaller than this go to the exact kernel (irc_options.tile; 0 or < 0 = never)
int min_item(const irc_options* opt) {
    const int t = opt ? opt->tile : 0;
    return t > 0 ? t : 0;
}

int resolve_precision(const irc_options* opt) {
    int prec = opt ? opt->precision : IRC_PREC_AUTO;
    if (prec == IRC_PREC_AUTO) prec = IRC_PREC_BF16_SCREEN;
    return prec;
}

// Query range of one (site, chunk) block: queries q0 .. q0+nq-1 share the candidate
// stream of the block's linked chunks (ncand columns in nruns runs).
struct QRange {
    int blk, q0, nq, ncand, nruns;
};

void block_ranges_cands(int C, const std::vector<int>& runl, const std::vector<int>& lptr,
                        const std::vector<int>& lidx, std::vector<QRange>& rg) {
    for (auto& r : rg) {
        const int s = r.blk / C, c = r.blk % C;
        int n = 0, nr = 0;
        for (int li = lptr[c]; li < lptr[c + 1]; ++li) {
            const int b2 = s * C + lidx[li];
            for (int part = 0; part < 2; ++part) {
                const int len = runl[2 * b2 + part];
                if (len > 0) { n += len; ++nr; }
            }
        }
        r.ncand = n;
        r.nruns = nr;
    }
}

// Device-side planning inputs of the batch path (the block runs are already on the device).
struct DevWork {
    int S, C;
    const int64_t* site_ptr;   // device [S+1]
    const int64_t* prim;       // device [S+1] prefix of primary counts
    int64_t max_ncand;         // largest candidate count of a block
};

// kNN pass: MFMA screen + certification + FP64 re-rank for every block that fits the
// screen, the exact FP64 kernel for the rest and for uncertified queries.  The work list comes
// from `ranges` (host planning, per-site entry points) or is built on the device (`dw`, batch).
// feat_out: when set, the [n_cols][DP] FP32 rows (p.feat) are written here from F (stride D) --
// by the screen's packing kernel in the same pass over F, or by the FP32 packer on the exact path
// fin: batch outputs (rho, d_k, int64 ids) the screened path finishes in the re-rank and, for the
// exact kernel's queries, in a list pass; *fin_done tells the caller whether that happened
struct FinishOut {
    float* rho;
    float* dk;
    int64_t* knn;
    double* d2head;   // exact d^2 of the first IRC_D2HEAD entries of every list, for the delta pass
    float* rho_col;   // rho of the queries' own columns
};
irc_status run_knn(irc_ctx* ctx, const DevPlan& p, int nq, int prec, int k_over, int min_item, const float* F,
                   int FS, int64_t n_cols, const std::vector<QRange>* ranges, const DevWork* dw, int* knn_col,
                   double* dk2, float* feat_out = nullptr, const FinishOut* fin = nullptr, bool* fin_done = nullptr) {
    if (fin_done) *fin_done = false;
    hipStream_t st = ctx->stream;
    int64_t pairs = 0;
    if (ranges)
        for (const auto& r : *ranges) pairs += (int64_t)r.nq * r.ncand;
    ctx->stats.n_pairs = pairs;
    ctx->stats.n_screen_fallback = 0;
    ctx->stats.ms_rerank = 0.0;
    ctx->stats.n_rerank_overflow = 0;
    ctx->stats.n_rerank_second = 0;
    ctx->dstat_screen = false;
    CHK(ensure(ctx, ctx->bb->b_dstat, sizeof(int64_t) * 8));
    HIPCHK(irc::launch_zero_i64(st, ptr<int64_t>(ctx->bb->b_dstat), 8));
    // the screen's operand rows hold D + 2 values in at most 4 MFMA K-blocks of 16 (D <= 62);
    // wider features take the exact path
    const int sp = prec == IRC_PREC_FP32 ? irc::SCREEN_FP32 : (prec == IRC_PREC_FP8_SCREEN ? irc::SCREEN_FP8 : irc::SCREEN_BF16);
    const bool screen = (prec == IRC_PREC_FP32 || prec == IRC_PREC_BF16_SCREEN || prec == IRC_PREC_FP8_SCREEN) &&
                        irc::screen_fits(sp, p.D);
    if (!screen) {
        if (feat_out) HIPCHK(irc::launch_pack_fp32(st, n_cols, p.D, p.DP, F, feat_out));
        ctx->stats.precision_used = IRC_PREC_EXACT;
        if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev[4], st));
        HIPCHK(irc::launch_knn_exact(st, p, nullptr, nq, persistent_grid(ctx, 8), knn_col, dk2, nullptr));
        if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev[5], st));
        ctx->stats.n_screen_fallback = nq;
        return IRC_OK;
    }
    ctx->stats.precision_used = prec;
    ctx->dstat_screen = true;
    // over-fetch: the looser bounds of the BF16 and FP8 forms need deeper lists to certify the k-th
    int kprime = k_over > 0 ? k_over : (sp == irc::SCREEN_FP8 ? 64 : (sp == irc::SCREEN_BF16 ? 48 : 40));
    if (kprime < p.k) kprime = p.k;
    kprime = kprime <= 32 ? 32 : (kprime <= 40 ? 40 : (kprime <= 48 ? 48 : 64));
    const int large_ncand = ctx->large_ncand;
    // [0, nq) exact-kernel list, [nq] its count, [nq + 1, 2 nq + 1) the first re-rank pass's overflows
    CHK(ensure(ctx, ctx->bb->b_scr[4], sizeof(int) * (2 * (size_t)nq + 1)));
    // [0, 8) unused, [8, 17) the XCD segment offsets, [20] the re-rank's spill count, [32, 57) the
    // screen launch's counters and split offsets
    CHK(ensure(ctx, ctx->bb->b_scr[5], sizeof(int) * 64));
    int* fb = ptr<int>(ctx->bb->b_scr[4]);
    int* nfb = fb + nq;
    int* scr5 = ptr<int>(ctx->bb->b_scr[5]);
    int* d_seg_beg = scr5 + 8;
    int* seg_ctr = scr5 + 32;
    int n_items_slots = 0;   // room of each of the two item arrays in b_scr[3]
    int n_items_cap = 0;
    bool any_large = false;
    irc::PlanBuf pb{};
    if (ranges) {
        // items (<= 64 queries each) in site-major order; queries outside the screen's limits go exact
        std::vector<irc::ScreenItem> items;
        std::vector<int> exact_q;
        std::vector<double> cost;
        for (const auto& r : *ranges) {
            if (r.nq <= 0) continue;
            if (r.ncand > IRC_SCREEN_MAXCAND || r.nruns > IRC_SCREEN_MAXRUNS) {
                for (int i = 0; i < r.nq; ++i) exact_q.push_back(r.q0 + i);
                continue;
            }
            for (int i = 0; i < r.nq; i += 64) {
                const int n = r.nq - i < 64 ? r.nq - i : 64;
                if (n < min_item) {
                    // a wavefront streams the whole candidate set whatever its query count: a few
                    // leftover queries cost less on the exact kernel
                    for (int j = 0; j < n; ++j) exact_q.push_back(r.q0 + i + j);
                    continue;
                }
                items.push_back({r.blk, r.q0 + i, n, r.ncand});
                cost.push_back((double)r.ncand + 64.0);
                any_large = any_large || r.ncand > large_ncand;
            }
        }
        // eight contiguous segments of about equal cost, one per XCD
        int seg_beg[9];
        {
            double tot = 0.0;
            for (double c : cost) tot += c;
            double acc = 0.0;
            int sgi = 1;
            seg_beg[0] = 0;
            for (size_t i = 0; i < items.size() && sgi < 8; ++i) {
                acc += cost[i];
                while (sgi < 8 && acc >= tot * sgi / 8.0) seg_beg[sgi++] = (int)i + 1;
            }
            while (sgi < 8) seg_beg[sgi++] = (int)items.size();
            seg_beg[8] = (int)items.size();
        }
        n_items_cap = (int)items.size();
        n_items_slots = (int)items.size() + 1;
        CHK(ensure(ctx, ctx->bb->b_scr[3], sizeof(irc::ScreenItem) * 2 * (size_t)n_items_slots));
        const int n_exact = (int)exact_q.size();
        if (!items.empty())
            HIPCHK(hipMemcpyAsync(ctx->bb->b_scr[3].p, items.data(), sizeof(irc::ScreenItem) * items.size(),
                                  hipMemcpyHostToDevice, st));
        if (n_exact) HIPCHK(hipMemcpyAsync(fb, exact_q.data(), sizeof(int) * n_exact, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(nfb, &n_exact, sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_seg_beg, seg_beg, sizeof(int) * 9, hipMemcpyHostToDevice, st));
    } else {
        const int64_t nb = (int64_t)dw->S * dw->C;
        n_items_cap = nq / 64 + (int)nb + 1;
        n_items_slots = n_items_cap;
        CHK(ensure(ctx, ctx->bb->b_scr[3], sizeof(irc::ScreenItem) * 2 * (size_t)n_items_slots));
        CHK(ensure(ctx, ctx->bb->b_plan, irc::plan_buf_bytes(nb)));
        pb = irc::plan_buf_carve(ctx->bb->b_plan.p, nb);
        HIPCHK(irc::launch_plan(st, dw->S, dw->C, p, dw->site_ptr, dw->prim, min_item, IRC_SCREEN_MAXCAND,
                                IRC_SCREEN_MAXRUNS, pb, ptr<irc::ScreenItem>(ctx->bb->b_scr[3]), fb, d_seg_beg, nfb));
        any_large = dw->max_ncand > large_ncand;
    }
    const int KP = irc::screen_kp(p.D);
    const int KB = KP / 16;
    const size_t pcols = (size_t)irc::screen_cols_padded(n_cols);   // whole 32-column groups
    CHK(ensure(ctx, ctx->bb->b_scr[0], (size_t)16 * irc::screen_np(sp, KB) * pcols));
    // FP8: the scale of every site ([1]: per segment; the per-site and A/B forms are one segment)
    const int n_seg = dw ? dw->S : 1;
    CHK(ensure(ctx, ctx->bb->b_scr[1], sizeof(int64_t) * 2 + sizeof(float) * (size_t)n_seg));
    const int64_t* seg_ptr = dw ? dw->site_ptr : ptr<int64_t>(ctx->bb->b_scr[1]);
    float* seg_scale = reinterpret_cast<float*>(ptr<int64_t>(ctx->bb->b_scr[1]) + 2);
    CHK(ensure(ctx, ctx->bb->b_scr[2], sizeof(float) * (size_t)n_cols + sizeof(float) * 2 * (size_t)p.n_blocks));
    // blocks past large_ncand keep 64 slots when the list is 64 deep in registers
    const int slots = (kprime > 32 && any_large) ? 64 : kprime;
    CHK(ensure(ctx, ctx->bb->b_cand_n, sizeof(int) * (size_t)nq));
    CHK(ensure(ctx, ctx->bb->b_cand, sizeof(int2) * (size_t)slots * nq));
    CHK(ensure(ctx, ctx->bb->b_cand_eps, sizeof(double) * (size_t)nq));
    if (sp == irc::SCREEN_FP8) {
        if (!dw) {
            const int64_t one[2] = {0, n_cols};
            HIPCHK(hipMemcpyAsync(ctx->bb->b_scr[1].p, one, sizeof(one), hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));   // `one` leaves scope
        }
        HIPCHK(irc::launch_site_scale(st, n_seg, seg_ptr, F, FS, p.D, seg_scale));
    }
    HIPCHK(irc::launch_screen_pack(st, sp, n_cols, p.D, FS, KB, F, ctx->bb->b_scr[0].p, ptr<float>(ctx->bb->b_scr[2]),
                                   feat_out, p.DP, seg_ptr, n_seg, seg_scale));
    float* run_nmax = ptr<float>(ctx->bb->b_scr[2]) + n_cols;
    HIPCHK(irc::launch_run_nmax(st, 2 * (int64_t)p.n_blocks, p.run_beg, p.run_len, ptr<float>(ctx->bb->b_scr[2]),
                                run_nmax));
    irc::ScreenArgs a;
    a.p = p;
    a.run_nmax = run_nmax;
    a.rows = ctx->bb->b_scr[0].p;
    a.norm = ptr<float>(ctx->bb->b_scr[2]);
    a.site_scale = seg_scale;
    a.items = ptr<irc::ScreenItem>(ctx->bb->b_scr[3]);
    a.items_kind = ptr<irc::ScreenItem>(ctx->bb->b_scr[3]) + n_items_slots;
    a.n_items = n_items_cap;
    a.kprime = kprime;
    a.large_ncand = large_ncand;
    irc::screen_eta(sp, p.D, KP, &a.eta, &a.eta2, &a.eta3);
    a.n_q = nq;
    a.cand_n = ptr<int>(ctx->bb->b_cand_n);
    a.cand = ptr<int2>(ctx->bb->b_cand);
    a.cand_eps = ptr<double>(ctx->bb->b_cand_eps);
    a.fallback = fb;
    a.n_fallback = nfb;
    a.fin_rho = fin ? fin->rho : nullptr;
    a.fin_dk = fin ? fin->dk : nullptr;
    a.fin_knn = fin ? fin->knn : nullptr;
    a.fin_d2head = fin ? fin->d2head : nullptr;
    a.fin_rho_col = fin ? fin->rho_col : nullptr;
    HIPCHK(irc::launch_fill_int(st, a.cand_n, nq, -1));
    if (dw && (ctx->flags & IRC_FLAG_SITE_LAUNCHES)) {
        // one launch per site over that site's items (blocks are site-major, so a site's items are
        // contiguous): the item offsets come back to the host once
        const int S = dw->S, C = dw->C;
        const int64_t nb = (int64_t)S * C;
        std::vector<int> ioff((size_t)nb);
        int64_t tot = 0;
        HIPCHK(hipMemcpyAsync(ioff.data(), pb.item_off, sizeof(int) * (size_t)nb, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&tot, pb.totals, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::vector<int> segs((size_t)S * 9);
        for (int s = 0; s < S; ++s) {
            const int lo = ioff[(size_t)s * C], hi = s + 1 < S ? ioff[(size_t)(s + 1) * C] : (int)tot;
            for (int j = 0; j <= 8; ++j) segs[(size_t)s * 9 + j] = lo + (int)((int64_t)(hi - lo) * j / 8);
        }
        CHK(ensure(ctx, ctx->bb->b_siteseg, sizeof(int) * segs.size()));
        HIPCHK(hipMemcpyAsync(ctx->bb->b_siteseg.p, segs.data(), sizeof(int) * segs.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipEventRecord(ctx->ev[4], st));
        for (int s = 0; s < S; ++s) {
            const int n = segs[(size_t)s * 9 + 8] - segs[(size_t)s * 9];
            if (n <= 0) continue;
            const int g = std::min(n, persistent_grid(ctx, ctx->waves_per_cu));
            HIPCHK(irc::launch_screen_knn_q(st, sp, KB, a, g, seg_ctr, ptr<int>(ctx->bb->b_siteseg) + (size_t)s * 9));
        }
        HIPCHK(hipEventRecord(ctx->ev[5], st));
    } else {
        if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev[4], st));
        HIPCHK(irc::launch_screen_knn_q(st, sp, KB, a, persistent_grid(ctx, ctx->waves_per_cu), seg_ctr, d_seg_beg));
        if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev[5], st));
    }
    int64_t* dstat = ptr<int64_t>(ctx->bb->b_dstat);
    HIPCHK(irc::launch_stat_copy(st, nfb, nullptr, dstat + 0));
    HIPCHK(irc::launch_screen_rerank(st, a, knn_col, dk2, nfb + 1, scr5 + 20));
    HIPCHK(irc::launch_stat_copy(st, scr5 + 20, nullptr, dstat + 3));
    if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev[6], st));
    // exact kNN for queries outside the screen's limits and for uncertified queries
    HIPCHK(irc::launch_knn_exact_dev(st, p, fb, nfb, nq, persistent_grid(ctx, 8), knn_col, dk2));
    if (irc::screen_rerank_finishes(a)) {
        HIPCHK(irc::launch_knn_finish_list(st, p, fb, nfb, nq, knn_col, dk2, a.fin_rho, a.fin_dk, a.fin_knn,
                                           a.fin_rho_col));
        if (fin_done) *fin_done = true;
    }
    HIPCHK(irc::launch_stat_copy(st, nfb, nullptr, dstat + 1));
    if (!ranges) HIPCHK(irc::launch_stat_copy(st, nullptr, pb.n_pairs, dstat + 2));
    return IRC_OK;
}

// After the call's stream has been synchronized: device counters and kernel times into stats.
irc_status collect_stats(irc_ctx* ctx, bool events) {
    int64_t d[8] = {0};
    if (ctx->bb->b_dstat.p) HIPCHK(hipMemcpy(d, ctx->bb->b_dstat.p, sizeof(d), hipMemcpyDeviceToHost));
    ctx->stats.n_delta_full = d[4];   // batch calls only (the kNN pass zeroes the counters)
    if (ctx->dstat_screen) {
        ctx->stats.n_screen_fallback = d[1];
        ctx->stats.n_rerank_overflow = d[1] - d[0];
        ctx->stats.n_rerank_second = d[3];
        if (d[2] > 0) ctx->stats.n_pairs = d[2];
        if (events) {
            ctx->stats.ms_screen = event_ms(ctx->ev[4], ctx->ev[5]);
            ctx->stats.ms_rerank = event_ms(ctx->ev[5], ctx->ev[6]);
        }
    } else if (events) {
        ctx->stats.ms_screen = event_ms(ctx->ev[4], ctx->ev[5]);
    }
    return IRC_OK;
}

// Workload shape the autotuner's choice is kept under: precision, D, k, sites, chunks, N / 65536.
uint64_t tune_signature(const irc_batch* b, int prec) {
    const int64_t v[] = {prec, b->D, b->k, b->n_sites, b->n_drift, b->n_spikes >> 16};
    uint64_t h = 1469598103934665603ull;
    const unsigned char* c = reinterpret_cast<const unsigned char*>(v);
    for (size_t i = 0; i < sizeof(v); ++i) h = (h ^ c[i]) * 1099511628211ull;
    return h;
}

// MATLAB round(): half away from zero
double round_half_away(double x) { return x < 0 ? -std::floor(-x + 0.5) : std::floor(x + 0.5); }

// subsample_vr_ (matlab/irc2.m:1557-1572) then quantile_vr_ (6478-6486): the sorted subsample at
// max(min(round(n q), n), 1) for q = (0:nq)/nq
std::vector<float> drift_quantiles(const float* v, int64_t n, int nq) {
    const int64_t n_max = 100000;
    std::vector<float> s;
    if (n <= n_max) {
        s.assign(v, v + n);
    } else {
        const int64_t skip = n / n_max;
        if (skip > 1) {
            for (int64_t i = 0; i < n; i += skip) s.push_back(v[i]);
        } else {
            s.assign(v, v + n);
        }
        if ((int64_t)s.size() > n_max) {
            const int64_t m = (int64_t)s.size(), n_remove = m - n_max;
            std::vector<char> keep((size_t)m, 1);
            for (int64_t j = 0; j < n_remove; ++j) {
                // linspace(1, m, n_remove)
                const double x = n_remove == 1 ? (double)m : 1.0 + (double)(m - 1) * (double)j / (double)(n_remove - 1);
                int64_t r = (int64_t)round_half_away(x);
                r = r < 1 ? 1 : (r > m ? m : r);
                keep[(size_t)(r - 1)] = 0;
            }
            std::vector<float> t;
            t.reserve((size_t)n_max);
            for (int64_t i = 0; i < m; ++i)
                if (keep[(size_t)i]) t.push_back(s[(size_t)i]);
            s.swap(t);
        }
    }
    std::stable_sort(s.begin(), s.end());
    const int64_t m = (int64_t)s.size();
    std::vector<float> q((size_t)nq + 1);
    for (int i = 0; i <= nq; ++i) {
        int64_t idx = (int64_t)round_half_away((double)m * ((double)i / (double)nq));
        idx = idx > m ? m : idx;
        idx = idx < 1 ? 1 : idx;
        q[(size_t)i] = s[(size_t)(idx - 1)];
    }
    return q;
}

}  // namespace

// ===========================================================================
extern "C" {

void irc_options_default(irc_options* opt) {
    if (!opt) return;
    memset(opt, 0, sizeof(*opt));
    opt->precision = IRC_PREC_AUTO;
    opt->k_over = 0;
    opt->mem = IRC_MEM_HOST;
}

const char* irc_version(void) { return "irc_hip 1.0 gfx950 (density-peak hot path)"; }

irc_status irc_ctx_create(irc_ctx** out, int device) {
    if (!out) return IRC_E_ARG;
    *out = nullptr;
    irc_ctx* ctx = new (std::nothrow) irc_ctx();
    if (!ctx) return IRC_E_NOMEM;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
        delete ctx;
        return IRC_E_HIP;
    }
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return IRC_E_HIP;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        ctx->n_cu = ncu;
    for (auto& ev : ctx->ev) {
        if (hipEventCreate(&ev) != hipSuccess) {
            irc_ctx_destroy(ctx);
            return IRC_E_HIP;
        }
    }
    // this device's copy of the no-allocation scratch (the device is current)
    void* sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) != hipSuccess || !sp) {
        (void)hipGetLastError();
        irc_ctx_destroy(ctx);
        return IRC_E_HIP;
    }
    ctx->scratch = reinterpret_cast<char*>(sp);
    *out = ctx;
    return IRC_OK;
}

void irc_ctx_destroy(irc_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    // the heap set only: the arena set points into the static scratch array
    for (DevBuf* b : ctx->heap.all())
        if (b->p) (void)hipFree(b->p);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    for (auto& ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* irc_last_error(const irc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

irc_status irc_selftest_mfma(irc_ctx* ctx, float* out) {
    if (!ctx || !out) return IRC_E_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    CHK(ensure(ctx, ctx->bb->b_ncand, sizeof(float) * 6144));
    hipStream_t st = ctx->stream;
    HIPCHK(irc::launch_mfma_layout_probe(st, ptr<float>(ctx->bb->b_ncand)));
    HIPCHK(hipMemcpyAsync(out, ctx->bb->b_ncand.p, sizeof(float) * 6144, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return IRC_OK;
}

irc_status irc_selftest_mfma_accum(irc_ctx* ctx, int prec, int K, int n_tiles, const float* A, const float* B,
                                   float* out) {
    if (!ctx) return IRC_E_ARG;
    if (!A || !B || !out) return fail(ctx, IRC_E_ARG, "null pointer");
    if (prec != 1 && prec != 2) return fail(ctx, IRC_E_ARG, "prec must be 1 (FP32 MFMA) or 2 (BF16 MFMA)");
    if (K < 2 || K > 64 || K % (prec == 1 ? 2 : 16) != 0)
        return fail(ctx, IRC_E_ARG, "K=%d: 2..64, even (prec 1) or a multiple of 16 (prec 2)", K);
    if (n_tiles < 1 || n_tiles > 65536) return fail(ctx, IRC_E_ARG, "n_tiles=%d outside 1..65536", n_tiles);
    HIPCHK(hipSetDevice(ctx->device));
    const size_t na = (size_t)n_tiles * 32 * K, no = (size_t)n_tiles * 1024;
    CHK(ensure(ctx, ctx->bb->b_scr[0], sizeof(float) * (2 * na + no)));
    float* dA = ptr<float>(ctx->bb->b_scr[0]);
    float* dB = dA + na;
    float* dO = dB + na;
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemcpyAsync(dA, A, sizeof(float) * na, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dB, B, sizeof(float) * na, hipMemcpyHostToDevice, st));
    HIPCHK(irc::launch_mfma_accum_probe(st, prec, K, n_tiles, dA, dB, dO));
    HIPCHK(hipMemcpyAsync(out, dO, sizeof(float) * no, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return IRC_OK;
}

irc_status irc_get_stats(const irc_ctx* ctx, irc_stats* st) {
    if (!ctx || !st) return IRC_E_ARG;
    *st = ctx->stats;
    return IRC_OK;
}

// ---------------------------------------------------------------------------
irc_status irc_knn_drift(irc_ctx* ctx, const irc_site_problem* prob, const irc_options* opt, float* dk_out,
                         int32_t* knn_out) {
    if (!ctx) return IRC_E_ARG;
    CHK(check_opt(ctx, opt));
    if (!dk_out || !knn_out) return fail(ctx, IRC_E_ARG, "null output");
    HIPCHK(hipSetDevice(ctx->device));
    HostPlan h;
    CHK(plan_site(ctx, prob, h));
    const int nq = (int)prob->n_in;
    if (nq == 0) return IRC_OK;
    DevPlan p;
    CHK(upload_plan(ctx, h, p));
    const int k = p.k;
    CHK(ensure(ctx, ctx->bb->b_knncol, sizeof(int) * (size_t)nq * k));
    CHK(ensure(ctx, ctx->bb->b_dk2, sizeof(double) * (size_t)nq));
    CHK(ensure(ctx, ctx->bb->b_out_dk, sizeof(float) * (size_t)nq));
    CHK(ensure(ctx, ctx->bb->b_out_knn32, sizeof(int) * (size_t)nq * k));
    hipStream_t st = ctx->stream;
    ctx->stats = irc_stats{};
    ctx->stats.n_queries = nq;
    // query ranges: maximal runs of queries in the same drift chunk
    std::vector<QRange> ranges;
    for (int i = 0; i < nq;) {
        int j = i;
        while (j < nq && h.qblk[j] == h.qblk[i]) ++j;
        ranges.push_back({h.qblk[i], i, j - i, 0, 0});
        i = j;
    }
    block_ranges_cands(h.C, h.runl, h.lptr, h.lidx, ranges);
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    CHK(run_knn(ctx, p, nq, resolve_precision(opt), opt ? opt->k_over : 0, min_item(opt), p.feat, p.DP, h.n_cols,
                &ranges, nullptr, ptr<int>(ctx->bb->b_knncol), ptr<double>(ctx->bb->b_dk2)));
    HIPCHK(irc::launch_knn_finish(st, p, nq, 1, ptr<int>(ctx->bb->b_knncol), ptr<double>(ctx->bb->b_dk2), nullptr,
                                  ptr<float>(ctx->bb->b_out_dk), ptr<int>(ctx->bb->b_out_knn32)));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    HIPCHK(hipMemcpyAsync(dk_out, ctx->bb->b_out_dk.p, sizeof(float) * nq, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(knn_out, ctx->bb->b_out_knn32.p, sizeof(int) * (size_t)nq * k, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->stats.ms_knn = ctx->stats.ms_total = event_ms(ctx->ev[0], ctx->ev[1]);
    CHK(collect_stats(ctx, true));
    return IRC_OK;
}

irc_status irc_delta_drift(irc_ctx* ctx, const irc_site_problem* prob, const irc_options* opt,
                           const float* rho_cols, float* dmin_out, int32_t* nneigh_out) {
    if (!ctx) return IRC_E_ARG;
    CHK(check_opt(ctx, opt));
    if (!rho_cols || !dmin_out || !nneigh_out) return fail(ctx, IRC_E_ARG, "null argument");
    HIPCHK(hipSetDevice(ctx->device));
    HostPlan h;
    CHK(plan_site(ctx, prob, h));
    const int nq = (int)prob->n_in;
    if (nq == 0) return IRC_OK;
    DevPlan p;
    CHK(upload_plan(ctx, h, p));
    std::vector<float> rho_perm((size_t)h.n_cols);
    for (int64_t j = 0; j < h.n_cols; ++j) rho_perm[h.inv[j]] = rho_cols[j];
    CHK(ensure(ctx, ctx->bb->b_rhocol, sizeof(float) * (size_t)h.n_cols));
    CHK(ensure(ctx, ctx->bb->b_out_delta, sizeof(float) * (size_t)nq));
    CHK(ensure(ctx, ctx->bb->b_out_nn32, sizeof(int) * (size_t)nq));
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemcpyAsync(ctx->bb->b_rhocol.p, rho_perm.data(), sizeof(float) * h.n_cols, hipMemcpyHostToDevice, st));
    ctx->stats = irc_stats{};
    ctx->stats.n_queries = nq;
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    HIPCHK(irc::launch_delta_exact(st, p, nullptr, nq, nullptr, persistent_grid(ctx, 16), 1, ptr<float>(ctx->bb->b_rhocol),
                                   ptr<float>(ctx->bb->b_out_delta), nullptr, ptr<int>(ctx->bb->b_out_nn32)));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    HIPCHK(hipMemcpyAsync(dmin_out, ctx->bb->b_out_delta.p, sizeof(float) * nq, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(nneigh_out, ctx->bb->b_out_nn32.p, sizeof(int) * nq, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->stats.ms_delta = ctx->stats.ms_total = event_ms(ctx->ev[0], ctx->ev[1]);
    ctx->stats.n_delta_full = nq;
    return IRC_OK;
}

// ---------------------------------------------------------------------------
irc_status irc_rho_delta_batch(irc_ctx* ctx, const irc_batch* b, const irc_options* opt, irc_outputs* out) {
    if (!ctx) return IRC_E_ARG;
    if (opt && (opt->flags & IRC_FLAG_AUTOTUNE) && !opt->force_fail && b) {
        // the context's tuned setting for this workload shape (tuned now when there is none)
        irc_options o = *opt;
        o.flags &= ~IRC_FLAG_AUTOTUNE;
        const uint64_t sig = tune_signature(b, resolve_precision(&o));
        const irc_options* hit = nullptr;
        for (const auto& e : ctx->tuned)
            if (e.first == sig) hit = &e.second;
        irc_options t;
        if (!hit) {
            CHK(irc_autotune(ctx, b, &o, out, &t, nullptr));
            hit = &t;
        }
        o.k_over = hit->k_over;
        o.waves_per_cu = hit->waves_per_cu;
        o.tile = hit->tile;
        o.large_ncand = hit->large_ncand;
        return irc_rho_delta_batch(ctx, b, &o, out);
    }
    CHK(check_opt(ctx, opt));
    if (!b || !out || !b->F || !b->col_spike || !b->site_ptr || !b->site_n_primary || !b->lim_drift || !b->link)
        return fail(ctx, IRC_E_ARG, "null pointer in irc_batch");
    if (!out->rho || !out->delta || !out->nneigh) return fail(ctx, IRC_E_ARG, "null output (rho, delta, nneigh)");
    const int S = b->n_sites, C = b->n_drift, D = b->D, k = b->k;
    const int64_t N = b->n_spikes;
    if (S <= 0 || C <= 0 || D <= 0 || k <= 0 || N <= 0) return fail(ctx, IRC_E_ARG, "invalid sizes");
    if (D > IRC_DMAX) return fail(ctx, IRC_E_UNSUPPORTED, "D=%d exceeds %d", D, IRC_DMAX);
    if (k > IRC_KMAX) return fail(ctx, IRC_E_UNSUPPORTED, "k=%d exceeds %d", k, IRC_KMAX);
    if (N >= ((int64_t)1 << 31) - 1) return fail(ctx, IRC_E_UNSUPPORTED, "n_spikes must be < 2^31");
    const bool dev = opt && opt->mem == IRC_MEM_DEVICE;
    // host-side checks of the small descriptors
    if (b->site_ptr[0] != 0) return fail(ctx, IRC_E_ARG, "site_ptr[0] must be 0");
    std::vector<int64_t> prim((size_t)S + 1, 0);
    for (int s = 0; s < S; ++s) {
        if (b->site_ptr[s + 1] < b->site_ptr[s]) return fail(ctx, IRC_E_ARG, "site_ptr not monotone at %d", s);
        if (b->site_n_primary[s] < 0 || b->site_n_primary[s] > b->site_ptr[s + 1] - b->site_ptr[s])
            return fail(ctx, IRC_E_ARG, "site_n_primary[%d] invalid", s);
        prim[s + 1] = prim[s] + b->site_n_primary[s];
    }
    const int64_t n_cols = b->site_ptr[S];
    if (b->n_cols != 0 && b->n_cols != n_cols) return fail(ctx, IRC_E_ARG, "n_cols does not match site_ptr[S]");
    if (prim[S] != N) return fail(ctx, IRC_E_ARG, "sum of site_n_primary (%lld) != n_spikes (%lld)", (long long)prim[S], (long long)N);
    // kernels launch up to 6 work-items per column; one dispatch covers at most 2^32
    if (n_cols > 500000000) return fail(ctx, IRC_E_UNSUPPORTED, "more than 500M columns");
    for (int c = 0; c < C; ++c)
        if (b->lim_drift[c + 1] < b->lim_drift[c]) return fail(ctx, IRC_E_ARG, "lim_drift not monotone");
    if (!links_have_diagonal(C, b->link)) return fail(ctx, IRC_E_ARG, "link(c,c) must be set for every chunk");

    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int DP = pad_dim(D);
    const bool graph = opt && (opt->flags & IRC_FLAG_GRAPH);
    if (graph && (opt->flags & IRC_FLAG_SITE_LAUNCHES))
        return fail(ctx, IRC_E_ARG, "IRC_FLAG_SITE_LAUNCHES cannot be captured in a graph (IRC_FLAG_GRAPH)");
    // descriptor of the call: device pointers, sizes, options and the host descriptors' contents
    uint64_t h = 1469598103934665603ull;
    if (graph) {
        auto mix = [&](const void* v, size_t n) {
            const unsigned char* c = reinterpret_cast<const unsigned char*>(v);
            for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
        };
        const void* ptrs[] = {b->F, b->col_spike, out->rho, out->delta, out->nneigh, out->knn, out->dk};
        const int64_t sizes[] = {N, S, C, D, k, n_cols, dev ? 1 : 0};
        mix(dev ? ptrs : ptrs + 7, dev ? sizeof(ptrs) : 0);
        const int64_t want[] = {out->knn != nullptr, out->dk != nullptr};
        mix(want, sizeof(want));
        mix(sizes, sizeof(sizes));
        mix(opt, sizeof(*opt));
        mix(b->site_ptr, sizeof(int64_t) * (S + 1));
        mix(b->site_n_primary, sizeof(int64_t) * S);
        mix(b->lim_drift, sizeof(int64_t) * (C + 1));
        mix(b->link, (size_t)C * C);
        if (h == 0) h = 1;
    }
    // The replay needs no set-up when the captured graph still matches: same descriptor, no context
    // buffer reallocated since the capture (buf_gen), and the shared planning buffers still hold this
    // descriptor's runs and queries (no other call has written them since, desc_key).
    const float* dF = b->F;
    const int* dspk = b->col_spike;
    if (!dev) {
        // both copies allocated before their addresses are taken
        CHK(ensure(ctx, ctx->bb->b_F, sizeof(float) * (size_t)n_cols * D));
        CHK(ensure(ctx, ctx->bb->b_colspk, sizeof(int) * (size_t)n_cols));
        dF = ptr<float>(ctx->bb->b_F);
        dspk = ptr<int>(ctx->bb->b_colspk);
    }
    float *o_rho, *o_delta, *o_dk = nullptr;
    int64_t* o_nn;
    int64_t* o_knn = nullptr;
    if (dev) {
        o_rho = out->rho;
        o_delta = out->delta;
        o_nn = out->nneigh;
        o_knn = out->knn;
        o_dk = out->dk;
    } else {
        CHK(ensure(ctx, ctx->bb->b_rhoslot, sizeof(float) * (size_t)N));
        CHK(ensure(ctx, ctx->bb->b_out_delta, sizeof(float) * (size_t)N));
        CHK(ensure(ctx, ctx->bb->b_out_nn, sizeof(int64_t) * (size_t)N));
        o_rho = ptr<float>(ctx->bb->b_rhoslot);
        o_delta = ptr<float>(ctx->bb->b_out_delta);
        o_nn = ptr<int64_t>(ctx->bb->b_out_nn);
        if (out->knn) {
            CHK(ensure(ctx, ctx->bb->b_out_knn, sizeof(int64_t) * (size_t)N * k));
            o_knn = ptr<int64_t>(ctx->bb->b_out_knn);
        }
        if (out->dk) {
            CHK(ensure(ctx, ctx->bb->b_out_dk, sizeof(float) * (size_t)N));
            o_dk = ptr<float>(ctx->bb->b_out_dk);
        }
    }
    // decided after this call's own allocations above, so a reallocation in them is seen too
    const bool fast = graph && ctx->graph_exec && ctx->graph_key == h && ctx->graph_gen == ctx->buf_gen &&
                      ctx->desc_key == h;
    ctx->desc_key = 0;
    ctx->stats = irc_stats{};
    ctx->stats.n_queries = N;
    bool replayed = false;
    if (fast) {
        // features in, then the captured device phase; no host round trip before the final one
        HIPCHK(hipEventRecord(ctx->ev[0], st));
        if (!dev)
            HIPCHK(hipMemcpyAsync(ctx->bb->b_F.p, b->F, sizeof(float) * (size_t)n_cols * D, hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(ctx->ev[1], st));
        HIPCHK(hipGraphLaunch(ctx->graph_exec, st));
        replayed = true;
    } else {
        std::vector<int> lptr, lidx;
        build_links(C, b->link, lptr, lidx);
        // small descriptors in one buffer: site_ptr[S+1], site_np[S], prim[S+1], lim[C+1]
        const size_t n_small = (size_t)(S + 1) + S + (S + 1) + (C + 1);
        CHK(ensure(ctx, ctx->bb->b_small, sizeof(int64_t) * n_small));
        CHK(ensure(ctx, ctx->bb->b_lptr, sizeof(int) * lptr.size()));
        CHK(ensure(ctx, ctx->bb->b_lidx, sizeof(int) * (lidx.size() + 1)));
        CHK(ensure(ctx, ctx->bb->b_feat, sizeof(float) * (size_t)n_cols * DP));
        CHK(ensure(ctx, ctx->bb->b_runb, sizeof(int) * (size_t)S * C * 2));
        CHK(ensure(ctx, ctx->bb->b_runl, sizeof(int) * (size_t)S * C * 2));
        CHK(ensure(ctx, ctx->bb->b_qcol, sizeof(int) * (size_t)N));
        CHK(ensure(ctx, ctx->bb->b_qblk, sizeof(int) * (size_t)N));
        CHK(ensure(ctx, ctx->bb->b_qout, sizeof(int) * (size_t)N));
        CHK(ensure(ctx, ctx->bb->b_flags, sizeof(int) * 4));
        CHK(ensure(ctx, ctx->bb->b_knncol, sizeof(int) * (size_t)N * k));
        CHK(ensure(ctx, ctx->bb->b_dk2, sizeof(double) * (size_t)N));
        CHK(ensure(ctx, ctx->bb->b_rhocol, sizeof(float) * (size_t)n_cols));
        CHK(ensure(ctx, ctx->bb->b_unres, sizeof(int) * ((size_t)N + 1)));
        CHK(ensure(ctx, ctx->bb->b_d2head, sizeof(double) * IRC_D2HEAD * (size_t)N));
        int64_t* d_site_ptr = ptr<int64_t>(ctx->bb->b_small);
        int64_t* d_site_np = d_site_ptr + (S + 1);
        int64_t* d_prim = d_site_np + S;
        int64_t* d_lim = d_prim + (S + 1);
        {
            std::vector<int64_t> smallv(n_small);
            memcpy(&smallv[0], b->site_ptr, sizeof(int64_t) * (S + 1));
            memcpy(&smallv[S + 1], b->site_n_primary, sizeof(int64_t) * S);
            memcpy(&smallv[2 * S + 1], prim.data(), sizeof(int64_t) * (S + 1));
            memcpy(&smallv[3 * S + 2], b->lim_drift, sizeof(int64_t) * (C + 1));
            HIPCHK(hipMemcpyAsync(d_site_ptr, smallv.data(), sizeof(int64_t) * n_small, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(ctx->bb->b_lptr.p, lptr.data(), sizeof(int) * lptr.size(), hipMemcpyHostToDevice, st));
            if (!lidx.empty())
                HIPCHK(hipMemcpyAsync(ctx->bb->b_lidx.p, lidx.data(), sizeof(int) * lidx.size(), hipMemcpyHostToDevice, st));
            // the host vectors are released at the end of this scope
            HIPCHK(hipStreamSynchronize(st));
        }
        if (!dev) {
            HIPCHK(hipMemcpyAsync(ctx->bb->b_F.p, b->F, sizeof(float) * (size_t)n_cols * D, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(ctx->bb->b_colspk.p, b->col_spike, sizeof(int) * (size_t)n_cols, hipMemcpyHostToDevice, st));
        }
        DevPlan p;
        p.feat = ptr<float>(ctx->bb->b_feat);
        p.DP = DP;
        p.D = D;
        p.C = C;
        p.k = k;
        p.col_key = dspk;
        p.col_ret = dspk;
        p.run_beg = ptr<int>(ctx->bb->b_runb);
        p.run_len = ptr<int>(ctx->bb->b_runl);
        p.link_ptr = ptr<int>(ctx->bb->b_lptr);
        p.link_idx = ptr<int>(ctx->bb->b_lidx);
        p.q_col = ptr<int>(ctx->bb->b_qcol);
        p.q_blk = ptr<int>(ctx->bb->b_qblk);
        p.q_out = ptr<int>(ctx->bb->b_qout);
        p.nq = (int)N;
        p.n_blocks = S * C;
        int* d_flags = ptr<int>(ctx->bb->b_flags);
        int* d_nunres = ptr<int>(ctx->bb->b_unres) + N;
        const int nq = (int)N;

        HIPCHK(irc::launch_fill_int(st, d_flags, 3, 0));
        HIPCHK(hipEventRecord(ctx->ev[0], st));
        HIPCHK(irc::launch_block_runs(st, S, C, d_site_ptr, d_site_np, d_lim, dspk, ptr<int>(ctx->bb->b_runb),
                                      ptr<int>(ctx->bb->b_runl)));
        HIPCHK(irc::launch_batch_queries(st, N, S, C, d_site_ptr, d_site_np, d_prim, d_lim, dspk, n_cols,
                                         ptr<int>(ctx->bb->b_qcol), ptr<int>(ctx->bb->b_qblk), ptr<int>(ctx->bb->b_qout), d_flags));
        HIPCHK(irc::launch_max_ncand(st, S, C, p.link_ptr, p.link_idx, p.run_len, d_flags + 2));
        int hflags[3] = {0, 0, 0};
        HIPCHK(hipMemcpyAsync(hflags, d_flags, sizeof(hflags), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        const int flags = hflags[0];
        if (flags & 1) return fail(ctx, IRC_E_ARG, "col_spike holds ids outside 1..n_spikes");
        if (flags & 2) return fail(ctx, IRC_E_ARG, "spike ids outside lim_drift");
        if (flags & 4) return fail(ctx, IRC_E_ARG, "spike ids not ascending within a site part");
        // the work list is built on the device from the block runs (no host round trip)
        DevWork dw;
        dw.S = S;
        dw.C = C;
        dw.site_ptr = d_site_ptr;
        dw.prim = d_prim;
        dw.max_ncand = hflags[2];
        // the device phase: kNN (plan, pack, screen, re-rank, exact), rho, delta.  With
        // IRC_FLAG_GRAPH it is captured once into a HIP graph and replayed while the call's
        // descriptor is unchanged and no context buffer has been reallocated since the capture.
        auto device_phase = [&]() -> irc_status {
            HIPCHK(irc::launch_fill_int(st, d_nunres, 1, 0));
            if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev[1], st));
            const FinishOut fo{o_rho, o_dk, o_knn, ptr<double>(ctx->bb->b_d2head), ptr<float>(ctx->bb->b_rhocol)};
            bool fin_done = false;
            CHK(run_knn(ctx, p, nq, resolve_precision(opt), opt ? opt->k_over : 0, min_item(opt), dF, D, n_cols,
                        nullptr, &dw, ptr<int>(ctx->bb->b_knncol), ptr<double>(ctx->bb->b_dk2), ptr<float>(ctx->bb->b_feat),
                        &fo, &fin_done));
            if (!fin_done)
                HIPCHK(irc::launch_knn_finish(st, p, nq, 0, ptr<int>(ctx->bb->b_knncol), ptr<double>(ctx->bb->b_dk2),
                                              o_rho, o_dk, nullptr, o_knn));
            if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev[2], st));
            // the finish wrote rho for every query's own column: only secondary columns are left
            if (!fin_done || n_cols > N)
                HIPCHK(irc::launch_rho_cols(st, n_cols, dspk, o_rho, ptr<float>(ctx->bb->b_rhocol)));
            HIPCHK(irc::launch_delta_from_knn(st, p, nq, ptr<int>(ctx->bb->b_knncol), ptr<float>(ctx->bb->b_rhocol),
                                              o_delta, o_nn, ptr<int>(ctx->bb->b_unres), d_nunres,
                                              fin_done ? ptr<double>(ctx->bb->b_d2head) : nullptr));
            HIPCHK(irc::launch_delta_exact(st, p, ptr<int>(ctx->bb->b_unres), nq, d_nunres, persistent_grid(ctx, 16), 0,
                                           ptr<float>(ctx->bb->b_rhocol), o_delta, o_nn, nullptr));
            HIPCHK(irc::launch_stat_copy(st, d_nunres, nullptr, ptr<int64_t>(ctx->bb->b_dstat) + 4));
            return IRC_OK;
        };
        if (graph && ctx->graph_exec && ctx->graph_key == h && ctx->graph_gen == ctx->buf_gen) {
            // the set-up above rewrote the shared planning buffers; the graph itself still matches
            HIPCHK(hipEventRecord(ctx->ev[1], st));
            HIPCHK(hipGraphLaunch(ctx->graph_exec, st));
            replayed = true;
        } else if (graph) {
            CHK(device_phase());
            // record the same work (all buffers now exist) for the next call with this descriptor
            if (ctx->graph_exec) {
                (void)hipGraphExecDestroy(ctx->graph_exec);
                ctx->graph_exec = nullptr;
            }
            const uint64_t gen0 = ctx->buf_gen;
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            ctx->capturing = true;
            const irc_status cs = device_phase();
            ctx->capturing = false;
            const hipError_t ce = hipStreamEndCapture(st, &g);
            if (cs != IRC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return cs;
            }
            if (ce != hipSuccess) return fail(ctx, IRC_E_HIP, "graph capture failed: %s", hipGetErrorString(ce));
            if (ctx->buf_gen != gen0) {
                // a buffer grew while capturing (it cannot: the eager run sized them all): drop the graph
                (void)hipGraphDestroy(g);
                return fail(ctx, IRC_E_HIP, "a context buffer was reallocated during graph capture");
            }
            const hipError_t ie = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                ctx->graph_exec = nullptr;
                return fail(ctx, IRC_E_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
            }
            ctx->graph_key = h;
            ctx->graph_gen = ctx->buf_gen;
        } else {
            CHK(device_phase());
        }
    }
    HIPCHK(hipEventRecord(ctx->ev[3], st));

    if (!dev) {
        HIPCHK(hipMemcpyAsync(out->rho, o_rho, sizeof(float) * N, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(out->delta, o_delta, sizeof(float) * N, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(out->nneigh, o_nn, sizeof(int64_t) * N, hipMemcpyDeviceToHost, st));
        if (out->knn) HIPCHK(hipMemcpyAsync(out->knn, o_knn, sizeof(int64_t) * (size_t)N * k, hipMemcpyDeviceToHost, st));
        if (out->dk) HIPCHK(hipMemcpyAsync(out->dk, o_dk, sizeof(float) * N, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    CHK(collect_stats(ctx, !replayed));
    if (!replayed) {
        ctx->stats.ms_knn = event_ms(ctx->ev[1], ctx->ev[2]);
        ctx->stats.ms_delta = event_ms(ctx->ev[2], ctx->ev[3]);
    }
    ctx->stats.ms_graph = replayed ? event_ms(ctx->ev[1], ctx->ev[3]) : 0.0;
    ctx->stats.ms_total = event_ms(ctx->ev[0], ctx->ev[3]);
    // the shared planning buffers (runs, queries, links, small descriptors) now hold this descriptor
    ctx->desc_key = graph ? h : 0;
    return IRC_OK;
}

// ---------------------------------------------------------------------------
irc_status irc_rho_delta_batch_noalloc(irc_ctx* ctx, const irc_batch* b, const irc_options* opt, irc_outputs* out) {
    if (!ctx) return IRC_E_ARG;
    if (!ctx->scratch) return fail(ctx, IRC_E_HIP, "the scratch array is not mapped on this device");
    if (opt && (opt->flags & (IRC_FLAG_GRAPH | IRC_FLAG_AUTOTUNE | IRC_FLAG_SITE_LAUNCHES)))
        return fail(ctx, IRC_E_ARG, "irc_rho_delta_batch_noalloc takes no graph, autotune or per-site launch flag");
    // the same call with every device buffer carved from the scratch array: a fresh set per call
    ctx->arena = Bufs{};
    ctx->scratch_used = 0;
    ctx->arena_on = true;
    ctx->bb = &ctx->arena;
    ctx->desc_key = 0;   // the heap set's planning buffers are not the ones this call writes
    const irc_status st = irc_rho_delta_batch(ctx, b, opt, out);
    ctx->bb = &ctx->heap;
    ctx->arena_on = false;
    return st;
}

irc_status irc_autotune(irc_ctx* ctx, const irc_batch* b, const irc_options* base, irc_outputs* out,
                        irc_options* best, irc_tune_report* rep) {
    if (!ctx) return IRC_E_ARG;
    if (!b || !base || !out || !best) return fail(ctx, IRC_E_ARG, "null argument in irc_autotune");
    if (base->force_fail) return fail(ctx, IRC_E_FORCED, "failure forced by irc_options.force_fail");
    irc_options o0 = *base;
    o0.flags &= ~(IRC_FLAG_AUTOTUNE | IRC_FLAG_GRAPH | IRC_FLAG_SITE_LAUNCHES);
    const int prec = resolve_precision(&o0);
    const bool dev = o0.mem == IRC_MEM_DEVICE;
    // candidates: the base setting, then one parameter changed at a time
    std::vector<irc_options> cand{o0};
    if (prec == IRC_PREC_BF16_SCREEN || prec == IRC_PREC_FP32) {
        const int kd = o0.k_over > 0 ? o0.k_over : (prec == IRC_PREC_FP32 ? 40 : 48);
        for (int kv : {32, 40, 48, 64})
            if (kv != kd && kv >= b->k) { irc_options o = o0; o.k_over = kv; cand.push_back(o); }
        const int wd = o0.waves_per_cu > 0 ? o0.waves_per_cu : 8;
        for (int wv : {4, 16})
            if (wv != wd) { irc_options o = o0; o.waves_per_cu = wv; cand.push_back(o); }
        for (int tv : {8, 16, 24})
            if (tv != o0.tile) { irc_options o = o0; o.tile = tv; cand.push_back(o); }
        const int ld = o0.large_ncand > 0 ? o0.large_ncand : 8192;
        for (int lv : {4096, 16384})
            if (lv != ld) { irc_options o = o0; o.large_ncand = lv; cand.push_back(o); }
    }
    if (cand.size() > 15) cand.resize(15);   // one slot of the report is kept for the combined setting
    const int64_t N = b->n_spikes;
    const int k = b->k;
    // the outputs the trials are compared against: base's, kept in the context (device) or on the host
    const size_t sz[5] = {sizeof(float) * (size_t)N, sizeof(float) * (size_t)N, sizeof(int64_t) * (size_t)N,
                          out->knn ? sizeof(int64_t) * (size_t)N * k : 0, out->dk ? sizeof(float) * (size_t)N : 0};
    void* optr[5] = {out->rho, out->delta, out->nneigh, out->knn, out->dk};
    size_t off[6] = {0};
    for (int i = 0; i < 5; ++i) off[i + 1] = off[i] + ((sz[i] + 255) & ~(size_t)255);
    std::vector<char> host_ref;
    auto run = [&](const irc_options& o, double& ms) -> irc_status {
        ms = 1e30;
        for (int r = 0; r < 3; ++r) {   // the first run also warms the buffers and code objects up
            const irc_status st = irc_rho_delta_batch(ctx, b, &o, out);
            if (st != IRC_OK) return st;
            if (r > 0) ms = std::min(ms, ctx->stats.ms_total);
        }
        return IRC_OK;
    };
    double ms0 = 0.0;
    CHK(run(o0, ms0));
    hipStream_t st = ctx->stream;
    if (dev) {
        CHK(ensure(ctx, ctx->bb->b_tune, off[5] + 256));
        for (int i = 0; i < 5; ++i)
            if (sz[i]) HIPCHK(hipMemcpyAsync((char*)ctx->bb->b_tune.p + off[i], optr[i], sz[i], hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    } else {
        host_ref.resize(off[5]);
        for (int i = 0; i < 5; ++i)
            if (sz[i]) memcpy(host_ref.data() + off[i], optr[i], sz[i]);
    }
    auto same = [&](bool& eq) -> irc_status {
        eq = true;
        if (dev) {
            unsigned long long* cnt = reinterpret_cast<unsigned long long*>((char*)ctx->bb->b_tune.p + off[5]);
            HIPCHK(irc::launch_zero_i64(st, reinterpret_cast<int64_t*>(cnt), 1));
            for (int i = 0; i < 5; ++i)
                if (sz[i])
                    HIPCHK(irc::launch_count_diff(st, reinterpret_cast<const uint32_t*>((char*)ctx->bb->b_tune.p + off[i]),
                                                  reinterpret_cast<const uint32_t*>(optr[i]), (int64_t)(sz[i] / 4), cnt));
            unsigned long long nd = 0;
            HIPCHK(hipMemcpyAsync(&nd, cnt, sizeof(nd), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            eq = nd == 0;
        } else {
            for (int i = 0; i < 5 && eq; ++i)
                if (sz[i]) eq = memcmp(host_ref.data() + off[i], optr[i], sz[i]) == 0;
        }
        return IRC_OK;
    };
    std::vector<double> ms(cand.size(), 1e30);
    std::vector<int> ident(cand.size(), 0);
    ms[0] = ms0;
    ident[0] = 1;
    for (size_t i = 1; i < cand.size(); ++i) {
        if (run(cand[i], ms[i]) != IRC_OK) continue;   // an unsupported setting is not a candidate
        bool eq = false;
        CHK(same(eq));
        ident[i] = eq;
    }
    // the combined setting: every parameter at its best single-change value
    irc_options comb = o0;
    auto better = [&](size_t i) { return ident[i] && ms[i] < ms0; };
    double bk = ms0, bw = ms0, bt = ms0, bl = ms0;
    for (size_t i = 1; i < cand.size(); ++i) {
        if (!better(i)) continue;
        if (cand[i].k_over != o0.k_over && ms[i] < bk) { bk = ms[i]; comb.k_over = cand[i].k_over; }
        if (cand[i].waves_per_cu != o0.waves_per_cu && ms[i] < bw) { bw = ms[i]; comb.waves_per_cu = cand[i].waves_per_cu; }
        if (cand[i].tile != o0.tile && ms[i] < bt) { bt = ms[i]; comb.tile = cand[i].tile; }
        if (cand[i].large_ncand != o0.large_ncand && ms[i] < bl) { bl = ms[i]; comb.large_ncand = cand[i].large_ncand; }
    }
    if (memcmp(&comb, &o0, sizeof(comb)) != 0) {
        cand.push_back(comb);
        ms.push_back(1e30);
        ident.push_back(0);
        const size_t i = cand.size() - 1;
        if (run(comb, ms[i]) == IRC_OK) {
            bool eq = false;
            CHK(same(eq));
            ident[i] = eq;
        }
    }
    size_t bi = 0;
    for (size_t i = 1; i < cand.size(); ++i)
        if (ident[i] && ms[i] < ms[bi]) bi = i;
    // leave base's outputs in place
    if (dev) {
        for (int i = 0; i < 5; ++i)
            if (sz[i]) HIPCHK(hipMemcpyAsync(optr[i], (char*)ctx->bb->b_tune.p + off[i], sz[i], hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    } else {
        for (int i = 0; i < 5; ++i)
            if (sz[i]) memcpy(optr[i], host_ref.data() + off[i], sz[i]);
    }
    *best = cand[bi];
    best->flags = base->flags & ~IRC_FLAG_AUTOTUNE;
    best->mem = base->mem;
    const uint64_t sig = tune_signature(b, prec);
    bool stored = false;
    for (auto& e : ctx->tuned)
        if (e.first == sig) { e.second = cand[bi]; stored = true; }
    if (!stored) ctx->tuned.emplace_back(sig, cand[bi]);
    if (rep) {
        memset(rep, 0, sizeof(*rep));
        rep->n_trials = (int32_t)std::min<size_t>(cand.size(), 16);
        rep->best_trial = (int32_t)bi;
        rep->ms_base = ms0;
        rep->ms_best = ms[bi];
        for (int i = 0; i < rep->n_trials; ++i) {
            rep->k_over[i] = cand[i].k_over;
            rep->waves_per_cu[i] = cand[i].waves_per_cu;
            rep->tile[i] = cand[i].tile;
            rep->large_ncand[i] = cand[i].large_ncand;
            rep->ms[i] = ms[i];
            rep->identical[i] = ident[i];
        }
    }
    return IRC_OK;
}

}  // extern "C"

namespace {
// The A/B kNN forms: exact kNN of the columns of A among the columns of B (or the listed columns
// IB of B).  K_all (k x nA): every listed distance; K_out (nA): the k-th; I_out (k x nA): indices.
irc_status knn_ab_impl(irc_ctx* ctx, int32_t D, int64_t nB, const float* B, const int32_t* IB, int64_t nIB,
                       int64_t nA, const float* A, int32_t k, float* K_all, float* K_out, int32_t* I_out) {
    if (!ctx) return IRC_E_ARG;
    if (D <= 0 || k <= 0 || nB <= 0 || nA < 0 || !B || (nA > 0 && (!A || (!K_out && !K_all))))
        return fail(ctx, IRC_E_ARG, "invalid arguments");
    if (IB && nIB <= 0) return fail(ctx, IRC_E_ARG, "empty candidate list IB");
    if (D > IRC_DMAX || k > IRC_KMAX) return fail(ctx, IRC_E_UNSUPPORTED, "D or k too large");
    const int64_t n0 = IB ? nIB : nB;
    if (nA + n0 >= ((int64_t)1 << 31) || nB >= ((int64_t)1 << 31))
        return fail(ctx, IRC_E_UNSUPPORTED, "too many columns");
    for (int64_t j = 0; IB && j < nIB; ++j)
        if (IB[j] < 1 || IB[j] > nB) return fail(ctx, IRC_E_ARG, "IB[%lld]=%d outside 1..%lld", (long long)j, IB[j],
                                                 (long long)nB);
    if (nA == 0) return IRC_OK;
    HIPCHK(hipSetDevice(ctx->device));
    HostPlan h;
    plan_ab(D, nB, B, nA, A, k, h, IB, nIB);
    DevPlan p;
    CHK(upload_plan(ctx, h, p));
    const int nq = (int)nA;
    CHK(ensure(ctx, ctx->bb->b_knncol, sizeof(int) * (size_t)nq * k));
    CHK(ensure(ctx, ctx->bb->b_dk2, sizeof(double) * (size_t)nq));
    CHK(ensure(ctx, ctx->bb->b_out_dk, sizeof(float) * (size_t)nq * (K_all ? k : 1)));
    CHK(ensure(ctx, ctx->bb->b_out_knn32, sizeof(int) * (size_t)nq * k));
    hipStream_t st = ctx->stream;
    ctx->stats = irc_stats{};
    ctx->stats.n_queries = nq;
    const std::vector<QRange> no_ranges;
    CHK(run_knn(ctx, p, nq, IRC_PREC_EXACT, 0, 0, p.feat, p.DP, h.n_cols, &no_ranges, nullptr,
                ptr<int>(ctx->bb->b_knncol), ptr<double>(ctx->bb->b_dk2)));
    HIPCHK(irc::launch_knn_finish(st, p, nq, 1, ptr<int>(ctx->bb->b_knncol), ptr<double>(ctx->bb->b_dk2), nullptr,
                                  K_all ? nullptr : ptr<float>(ctx->bb->b_out_dk), ptr<int>(ctx->bb->b_out_knn32)));
    if (K_all) {
        HIPCHK(irc::launch_knn_dists(st, p, nq, ptr<int>(ctx->bb->b_knncol), ptr<float>(ctx->bb->b_out_dk)));
        HIPCHK(hipMemcpyAsync(K_all, ctx->bb->b_out_dk.p, sizeof(float) * (size_t)nq * k, hipMemcpyDeviceToHost, st));
    } else {
        HIPCHK(hipMemcpyAsync(K_out, ctx->bb->b_out_dk.p, sizeof(float) * nq, hipMemcpyDeviceToHost, st));
    }
    if (I_out) HIPCHK(hipMemcpyAsync(I_out, ctx->bb->b_out_knn32.p, sizeof(int) * (size_t)nq * k, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return IRC_OK;
}
}  // namespace

extern "C" {

irc_status irc_knn_ab(irc_ctx* ctx, int32_t D, int64_t nB, const float* B, int64_t nA, const float* A, int32_t k,
                      float* K_out, int32_t* I_out) {
    if (ctx && nA > 0 && !K_out) return fail(ctx, IRC_E_ARG, "null K_out");
    return knn_ab_impl(ctx, D, nB, B, nullptr, 0, nA, A, k, nullptr, K_out, I_out);
}

irc_status irc_knn_ab_full(irc_ctx* ctx, int32_t D, int64_t nB, const float* B, int64_t nA, const float* A,
                           int32_t k, float* K_out, int32_t* I_out) {
    if (ctx && nA > 0 && (!K_out || !I_out)) return fail(ctx, IRC_E_ARG, "null output");
    return knn_ab_impl(ctx, D, nB, B, nullptr, 0, nA, A, k, K_out, nullptr, I_out);
}

irc_status irc_knn_ab_ib(irc_ctx* ctx, int32_t D, int64_t nB, const float* B, int64_t nIB, const int32_t* IB,
                         int64_t nA, const float* A, int32_t k, float* K_out, int32_t* I_out) {
    if (ctx && (!IB || (nA > 0 && !K_out))) return fail(ctx, IRC_E_ARG, "null IB or K_out");
    return knn_ab_impl(ctx, D, nB, B, IB, nIB, nA, A, k, nullptr, K_out, I_out);
}

irc_status irc_delta_knn_ab(irc_ctx* ctx, int32_t D, int64_t nB, const float* B, const float* RB, int64_t nA,
                            const float* A, const float* RA, float* D_out, int32_t* N_out) {
    if (!ctx) return IRC_E_ARG;
    if (D <= 0 || nB <= 0 || nA < 0 || !B || !RB || (nA > 0 && (!A || !RA || !D_out || !N_out)))
        return fail(ctx, IRC_E_ARG, "invalid arguments");
    if (D > IRC_DMAX) return fail(ctx, IRC_E_UNSUPPORTED, "D too large");
    if (nA + nB >= ((int64_t)1 << 31)) return fail(ctx, IRC_E_UNSUPPORTED, "too many columns");
    if (nA == 0) return IRC_OK;
    HIPCHK(hipSetDevice(ctx->device));
    HostPlan h;
    plan_ab(D, nB, B, nA, A, 1, h);
    DevPlan p;
    CHK(upload_plan(ctx, h, p));
    const int nq = (int)nA;
    std::vector<float> rho((size_t)(nB + nA));
    memcpy(rho.data(), RB, sizeof(float) * nB);
    memcpy(rho.data() + nB, RA, sizeof(float) * nA);
    CHK(ensure(ctx, ctx->bb->b_rhocol, sizeof(float) * rho.size()));
    CHK(ensure(ctx, ctx->bb->b_out_delta, sizeof(float) * (size_t)nq));
    CHK(ensure(ctx, ctx->bb->b_out_nn32, sizeof(int) * (size_t)nq));
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemcpyAsync(ctx->bb->b_rhocol.p, rho.data(), sizeof(float) * rho.size(), hipMemcpyHostToDevice, st));
    HIPCHK(irc::launch_delta_exact(st, p, nullptr, nq, nullptr, persistent_grid(ctx, 16), 1, ptr<float>(ctx->bb->b_rhocol),
                                   ptr<float>(ctx->bb->b_out_delta), nullptr, ptr<int>(ctx->bb->b_out_nn32)));
    HIPCHK(hipMemcpyAsync(D_out, ctx->bb->b_out_delta.p, sizeof(float) * nq, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(N_out, ctx->bb->b_out_nn32.p, sizeof(int) * nq, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return IRC_OK;
}

irc_status irc_rho_drift_dc(irc_ctx* ctx, int32_t D, int64_t n1, int64_t n12, const int32_t* miSite12,
                            const float* fet1, const float* fet2, float dc2, float* rho_out) {
    if (!ctx) return IRC_E_ARG;
    if (D <= 0 || n1 < 0 || n12 < n1 || !miSite12 || !fet1 || !fet2 || (n1 > 0 && !rho_out))
        return fail(ctx, IRC_E_ARG, "invalid arguments");
    if (n12 >= ((int64_t)1 << 31)) return fail(ctx, IRC_E_UNSUPPORTED, "too many spikes");
    if (n1 == 0) return IRC_OK;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t fb = sizeof(float) * (size_t)n12 * D;
    CHK(ensure(ctx, ctx->bb->b_F, 2 * fb));
    CHK(ensure(ctx, ctx->bb->b_key, sizeof(int) * 2 * (size_t)n12));
    CHK(ensure(ctx, ctx->bb->b_out_delta, sizeof(float) * (size_t)n1));
    float* d_f1 = ptr<float>(ctx->bb->b_F);
    float* d_f2 = d_f1 + (size_t)n12 * D;
    HIPCHK(hipMemcpyAsync(d_f1, fet1, fb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_f2, fet2, fb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_key.p, miSite12, sizeof(int) * 2 * n12, hipMemcpyHostToDevice, st));
    HIPCHK(irc::launch_rho_dc(st, D, (int)n1, (int)n12, ptr<int>(ctx->bb->b_key), d_f1, d_f2, dc2,
                              ptr<float>(ctx->bb->b_out_delta), persistent_grid(ctx, 16)));
    HIPCHK(hipMemcpyAsync(rho_out, ctx->bb->b_out_delta.p, sizeof(float) * n1, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return IRC_OK;
}

irc_status irc_delta_drift_dc(irc_ctx* ctx, int32_t D, int64_t n1, int64_t n12, const int32_t* miSite12,
                              const float* fet1, const float* fet2, const float* rho12, float dc2, float* delta_out,
                              uint32_t* nneigh_out) {
    if (!ctx) return IRC_E_ARG;
    if (D <= 0 || n1 < 0 || n12 < n1 || !miSite12 || !fet1 || !fet2 || !rho12 ||
        (n1 > 0 && (!delta_out || !nneigh_out)))
        return fail(ctx, IRC_E_ARG, "invalid arguments");
    if (n12 >= ((int64_t)1 << 31)) return fail(ctx, IRC_E_UNSUPPORTED, "too many spikes");
    if (n1 == 0) return IRC_OK;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t fb = sizeof(float) * (size_t)n12 * D;
    CHK(ensure(ctx, ctx->bb->b_F, 2 * fb));
    CHK(ensure(ctx, ctx->bb->b_key, sizeof(int) * 2 * (size_t)n12));
    CHK(ensure(ctx, ctx->bb->b_rhocol, sizeof(float) * (size_t)n12));
    CHK(ensure(ctx, ctx->bb->b_out_delta, sizeof(float) * (size_t)n1));
    CHK(ensure(ctx, ctx->bb->b_out_nn32, sizeof(int) * (size_t)n1));
    float* d_f1 = ptr<float>(ctx->bb->b_F);
    float* d_f2 = d_f1 + (size_t)n12 * D;
    HIPCHK(hipMemcpyAsync(d_f1, fet1, fb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_f2, fet2, fb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_key.p, miSite12, sizeof(int) * 2 * n12, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->bb->b_rhocol.p, rho12, sizeof(float) * n12, hipMemcpyHostToDevice, st));
    HIPCHK(irc::launch_delta_dc(st, D, (int)n1, (int)n12, ptr<int>(ctx->bb->b_key), d_f1, d_f2, ptr<float>(ctx->bb->b_rhocol),
                                dc2, ptr<float>(ctx->bb->b_out_delta), ptr<unsigned>(ctx->bb->b_out_nn32),
                                persistent_grid(ctx, 16)));
    HIPCHK(hipMemcpyAsync(delta_out, ctx->bb->b_out_delta.p, sizeof(float) * n1, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(nneigh_out, ctx->bb->b_out_nn32.p, sizeof(uint32_t) * n1, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return IRC_OK;
}

irc_status irc_calc_drift(irc_ctx* ctx, const irc_drift_problem* pr, int32_t* n_drift_out, int64_t* lim_out,
                          uint8_t* link_out, int32_t* counts_out) {
    if (!ctx) return IRC_E_ARG;
    if (!pr || !pr->amp || !pr->pos || !n_drift_out || !lim_out || !link_out)
        return fail(ctx, IRC_E_ARG, "null pointer in irc_calc_drift");
    const int64_t N = pr->n_spikes;
    const int C = pr->n_drift, na = pr->n_amp, np = pr->n_pos;
    if (N <= 0 || C <= 0 || na <= 0 || np <= 0 || pr->n_batch <= 0 || pr->n_time_max < 4)
        return fail(ctx, IRC_E_ARG, "invalid sizes in irc_calc_drift");
    // one chunk's histogram and the bin edges live in LDS (64 KB)
    if (((int64_t)na * np + na + np + 2) * 4 > 65536)
        return fail(ctx, IRC_E_UNSUPPORTED, "n_amp * n_pos too large (%d x %d)", na, np);
    if (pr->n_batch >= C) {
        // one chunk linked to itself (matlab/irc2.m:6388-6391)
        *n_drift_out = 1;
        lim_out[0] = 1;
        lim_out[1] = N + 1;
        link_out[0] = 1;
        return IRC_OK;
    }
    if (C > 16384) return fail(ctx, IRC_E_UNSUPPORTED, "n_drift above 16384");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    std::vector<int64_t> lim((size_t)C + 1);
    lim[0] = 1;
    for (int c = 1; c <= C; ++c) lim[(size_t)c] = (int64_t)std::ceil((double)c / C * (double)N) + 1;
    const std::vector<float> qa = drift_quantiles(pr->amp, N, na);
    const std::vector<float> qp = drift_quantiles(pr->pos, N, np);
    const int nb = na * np;
    // one buffer: amp, pos, edges, lim, counts, d2, link
    const size_t o_amp = 0, o_pos = o_amp + sizeof(float) * (size_t)N;
    const size_t o_ea = (o_pos + sizeof(float) * (size_t)N + 255) & ~(size_t)255;
    const size_t o_ep = o_ea + sizeof(float) * (size_t)(na + 1);
    const size_t o_lim = (o_ep + sizeof(float) * (size_t)(np + 1) + 255) & ~(size_t)255;
    const size_t o_cnt = (o_lim + sizeof(int64_t) * (size_t)(C + 1) + 255) & ~(size_t)255;
    const size_t o_d2 = (o_cnt + sizeof(int) * (size_t)nb * C + 255) & ~(size_t)255;
    const size_t o_lnk = (o_d2 + sizeof(int64_t) * (size_t)C * C + 255) & ~(size_t)255;
    const size_t total = o_lnk + (size_t)C * C;
    CHK(ensure(ctx, ctx->bb->b_drift, total));
    char* base = reinterpret_cast<char*>(ctx->bb->b_drift.p);
    HIPCHK(hipMemcpyAsync(base + o_amp, pr->amp, sizeof(float) * N, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(base + o_pos, pr->pos, sizeof(float) * N, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(base + o_ea, qa.data(), sizeof(float) * qa.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(base + o_ep, qp.data(), sizeof(float) * qp.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(base + o_lim, lim.data(), sizeof(int64_t) * lim.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    HIPCHK(irc::launch_drift(st, N, C, reinterpret_cast<int64_t*>(base + o_lim), reinterpret_cast<float*>(base + o_amp),
                             reinterpret_cast<float*>(base + o_pos), reinterpret_cast<float*>(base + o_ea), na,
                             reinterpret_cast<float*>(base + o_ep), np, pr->n_batch, pr->n_time_max,
                             reinterpret_cast<int*>(base + o_cnt), reinterpret_cast<int64_t*>(base + o_d2),
                             reinterpret_cast<uint8_t*>(base + o_lnk)));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    HIPCHK(hipMemcpyAsync(link_out, base + o_lnk, (size_t)C * C, hipMemcpyDeviceToHost, st));
    if (counts_out)
        HIPCHK(hipMemcpyAsync(counts_out, base + o_cnt, sizeof(int) * (size_t)nb * C, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    memcpy(lim_out, lim.data(), sizeof(int64_t) * lim.size());
    *n_drift_out = C;
    ctx->stats = irc_stats{};
    ctx->stats.ms_total = event_ms(ctx->ev[0], ctx->ev[1]);
    return IRC_OK;
}

irc_status irc_assign_chain(irc_ctx* ctx, int64_t n, const int64_t* nneigh, int32_t* labels,
                            const irc_options* opt) {
    if (!ctx) return IRC_E_ARG;
    if (n < 0 || (n > 0 && (!nneigh || !labels))) return fail(ctx, IRC_E_ARG, "null pointer in irc_assign_chain");
    if (n >= ((int64_t)1 << 31)) return fail(ctx, IRC_E_UNSUPPORTED, "n must be < 2^31");
    if (n == 0) return IRC_OK;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const bool dev = opt && opt->mem == IRC_MEM_DEVICE;
    // scratch: jump targets, and the host form's copies of nneigh and the labels
    const size_t o_nxt = 0, o_lab = (sizeof(int32_t) * (size_t)n + 255) & ~(size_t)255;
    const size_t o_nn = (o_lab + sizeof(int32_t) * (size_t)n + 255) & ~(size_t)255;
    CHK(ensure(ctx, ctx->bb->b_assign, dev ? o_lab : o_nn + sizeof(int64_t) * (size_t)n));
    char* base = reinterpret_cast<char*>(ctx->bb->b_assign.p);
    int32_t* d_nxt = reinterpret_cast<int32_t*>(base + o_nxt);
    int32_t* d_lab = dev ? labels : reinterpret_cast<int32_t*>(base + o_lab);
    const int64_t* d_nn = dev ? nneigh : reinterpret_cast<const int64_t*>(base + o_nn);
    if (!dev) {
        HIPCHK(hipMemcpyAsync(base + o_nn, nneigh, sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_lab, labels, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    HIPCHK(irc::launch_assign_chain(st, n, d_nn, d_lab, d_nxt));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    if (!dev) HIPCHK(hipMemcpyAsync(labels, d_lab, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->stats = irc_stats{};
    ctx->stats.ms_total = event_ms(ctx->ev[0], ctx->ev[1]);
    return IRC_OK;
}

irc_status irc_isolation_counts(irc_ctx* ctx, int64_t n, int32_t k, const int64_t* knn, const int32_t* clu,
                                int32_t n_clu, uint64_t* counts, const irc_options* opt) {
    if (!ctx) return IRC_E_ARG;
    if (n < 0 || k < 1 || n_clu < 1 || !counts || (n > 0 && (!knn || !clu)))
        return fail(ctx, IRC_E_ARG, "invalid arguments in irc_isolation_counts");
    if (n_clu > 4096) return fail(ctx, IRC_E_UNSUPPORTED, "n_clu=%d exceeds 4096", n_clu);
    if (n >= ((int64_t)1 << 31)) return fail(ctx, IRC_E_UNSUPPORTED, "n must be < 2^31");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const bool dev = opt && opt->mem == IRC_MEM_DEVICE;
    const size_t nc = (size_t)n_clu * (size_t)(n_clu + 1);
    const size_t o_cnt = 0, o_clu = (sizeof(uint64_t) * nc + 255) & ~(size_t)255;
    const size_t o_knn = (o_clu + sizeof(int32_t) * (size_t)n + 255) & ~(size_t)255;
    CHK(ensure(ctx, ctx->bb->b_assign, dev ? o_clu : o_knn + sizeof(int64_t) * (size_t)n * k));
    char* base = reinterpret_cast<char*>(ctx->bb->b_assign.p);
    unsigned long long* d_cnt = dev ? reinterpret_cast<unsigned long long*>(counts)
                                    : reinterpret_cast<unsigned long long*>(base + o_cnt);
    const int32_t* d_clu = dev ? clu : reinterpret_cast<const int32_t*>(base + o_clu);
    const int64_t* d_knn = dev ? knn : reinterpret_cast<const int64_t*>(base + o_knn);
    if (!dev) {
        HIPCHK(hipMemcpyAsync(base + o_clu, clu, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(base + o_knn, knn, sizeof(int64_t) * (size_t)n * k, hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemsetAsync(d_cnt, 0, sizeof(uint64_t) * nc, st));
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    HIPCHK(irc::launch_isolation_counts(st, n, k, d_knn, d_clu, n_clu, d_cnt));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    if (!dev) HIPCHK(hipMemcpyAsync(counts, d_cnt, sizeof(uint64_t) * nc, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->stats = irc_stats{};
    ctx->stats.ms_total = event_ms(ctx->ev[0], ctx->ev[1]);
    return IRC_OK;
}

}  // extern "C"
