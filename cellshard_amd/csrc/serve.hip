// Serving loop: resident CSC shards, feature-subset grouped queries admitted
// into output slots under a device-memory budget, batched into launches
// that run while new queries keep arriving (continuous batching).
//
// No reference counterpart: the reference runtime is one call per query
// (mask_groups.cuh:159-224).  Masking follows its rules: a row takes part
// iff it is kept and its label is < G (mask_groups.cu:65-67); a stored entry
// counts iff its value is nonzero (mask_groups.cu:155).
//
// Device side: one workgroup per (query, feature) item, taken from a
// per-launch counter (the next item is taken when the current one starts,
// so the counter's round trip overlaps the item).  The item's column of the
// shard's CSC is streamed once in chunks of 8 entries per thread through two
// register buffers (coalesced row index / value loads, the next chunk in
// flight while the current one is consumed); each kept entry lands in an LDS
// accumulator indexed by the row's label; the G accumulators are then
// stored contiguously at out[j * G] (feature-major output).  A query's work
// is a few tens of MB, so batching many queries per launch is what keeps
// HBM busy.
#include "common.hh"

#include <chrono>
#include <cstring>
#include <deque>
#include <vector>

namespace csamd {
namespace {

__host__ __device__ __forceinline__ unsigned int serve_hash(unsigned int x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__host__ __device__ __forceinline__ unsigned int seeded_label(unsigned int seed, unsigned int r, unsigned int G) {
    return (unsigned int) (((unsigned long long) serve_hash(seed ^ (r * 0x9E3779B9u)) * G) >> 32);
}

__host__ __device__ __forceinline__ bool seeded_keep(unsigned int seed, unsigned int r, unsigned int thr24) {
    return (serve_hash(seed ^ (r * 0x85EBCA6Bu)) >> 8) < thr24;
}

struct query_dev {
    const unsigned int *colptr;
    const unsigned int *rowidx;
    const void *val;
    const unsigned int *features;
    const unsigned int *labels;
    const unsigned char *keep;
    void *out;
    unsigned int rows;
    unsigned int value_type;
    unsigned int k;
    unsigned int G;
    unsigned int op;
    unsigned int label_seed;
    unsigned int mask_seed;
    unsigned int keep_thr24;
};

constexpr unsigned int SERVE_THREADS = 256u;
constexpr unsigned int SERVE_MAX_BATCH = 256u;
constexpr unsigned int SERVE_UNROLL = 8u;

// One item's column [b, e): chunks of SERVE_THREADS x SERVE_UNROLL entries
// (thread t takes entries t, t + SERVE_THREADS, ...) through two register
// buffers in turn, the next chunk's loads in flight while the current one
// is consumed.  Loads are unconditional (indices clamped to the column's
// last entry; entries past the end are masked when consumed), so the
// compiler waits only for the chunk it consumes.
template<int VT> struct serve_chunk {
    unsigned int r[SERVE_UNROLL];
    typename value_traits<VT>::storage v[SERVE_UNROLL];
};

template<int VT>
__device__ __forceinline__ void serve_load(const CSAMD_GLOBAL unsigned int *rowidx, const CSAMD_GLOBAL typename value_traits<VT>::storage *val,
                                           unsigned int i0, unsigned int last, serve_chunk<VT> &c) {
#pragma unroll
    for (unsigned int u = 0; u < SERVE_UNROLL; ++u) {
        const unsigned int i = min(i0 + u * SERVE_THREADS, last);
        c.r[u] = rowidx[i];
        c.v[u] = val[i];
    }
}

template<int VT>
__device__ __forceinline__ void serve_consume(const query_dev &Q, unsigned int i0, unsigned int e, const serve_chunk<VT> &c, double *acc_d,
                                              unsigned int *acc_u) {
#pragma unroll
    for (unsigned int u = 0; u < SERVE_UNROLL; ++u) {
        const unsigned int r = c.r[u];
        const float v = value_traits<VT>::decode(c.v[u]);
        if (i0 + u * SERVE_THREADS >= e || r >= Q.rows || v == 0.0f) continue;
        const bool kept = Q.keep != nullptr ? gmem(Q.keep)[r] != 0u : seeded_keep(Q.mask_seed, r, Q.keep_thr24);
        if (!kept) continue;
        const unsigned int g = Q.labels != nullptr ? gmem(Q.labels)[r] : seeded_label(Q.label_seed, r, Q.G);
        if (g >= Q.G) continue;
        if (Q.op == CSAMD_SERVE_NNZ) atomicAdd(acc_u + g, 1u);
        else {
            const double d = (double) v;
            atomicAdd(acc_d + g, Q.op == CSAMD_SERVE_SUMSQ ? d * d : d);
        }
    }
}

template<int VT>
__device__ __forceinline__ void serve_column(const query_dev &Q, unsigned int b, unsigned int e, double *acc_d, unsigned int *acc_u) {
    using S = typename value_traits<VT>::storage;
    if (b >= e) return;  // uniform
    const CSAMD_GLOBAL S *__restrict__ val = gmem((const S *) Q.val);
    const CSAMD_GLOBAL unsigned int *__restrict__ rowidx = gmem(Q.rowidx);
    constexpr unsigned int step = SERVE_THREADS * SERVE_UNROLL;
    const unsigned int last = e - 1u;
    unsigned int i0 = b + threadIdx.x;
    serve_chunk<VT> A, B;
    serve_load<VT>(rowidx, val, i0, last, A);
    // chunks start at b, b + step, ... (uniform count)
    for (unsigned int c0 = b; c0 < e; c0 += 2u * step) {
        serve_load<VT>(rowidx, val, i0 + step, last, B);
        serve_consume<VT>(Q, i0, e, A, acc_d, acc_u);
        if (c0 + step >= e) break;
        serve_load<VT>(rowidx, val, i0 + 2u * step, last, A);
        serve_consume<VT>(Q, i0 + step, e, B, acc_d, acc_u);
        i0 += 2u * step;
    }
}

__global__ void __launch_bounds__(SERVE_THREADS) serve_batch_kernel(const query_dev *__restrict__ qs,
                                                                    const unsigned int *__restrict__ prefix, unsigned int nq,
                                                                    unsigned int *__restrict__ counter) {
    extern __shared__ double acc_d[];  // [gmax] (u32 view for NNZ)
    unsigned int *acc_u = (unsigned int *) acc_d;
    __shared__ unsigned int pre[SERVE_MAX_BATCH + 1u];
    __shared__ unsigned int item_s;
    for (unsigned int i = threadIdx.x; i <= nq; i += blockDim.x) pre[i] = prefix[i];
    __syncthreads();
    const unsigned int total = pre[nq];
    if (threadIdx.x == 0u) item_s = atomicAdd(counter, 1u);
    __syncthreads();
    unsigned int item = item_s;
    while (item < total) {  // uniform
        // the next item, taken now so that the counter's round trip
        // overlaps this item's work
        unsigned int nxt = 0u;
        if (threadIdx.x == 0u) nxt = atomicAdd(counter, 1u);
        unsigned int lo = 0u, hi = nq - 1u;  // last q with pre[q] <= item
        while (lo < hi) {
            const unsigned int mid = (lo + hi + 1u) >> 1;
            if (pre[mid] <= item) lo = mid;
            else hi = mid - 1u;
        }
        const query_dev Q = qs[lo];
        const unsigned int j = item - pre[lo];
        const unsigned int c = gmem(Q.features)[j];
        const unsigned int b = gmem(Q.colptr)[c], e = gmem(Q.colptr)[c + 1u];
        for (unsigned int g = threadIdx.x; g < Q.G; g += blockDim.x) acc_d[g] = 0.0;
        __syncthreads();
        switch (Q.value_type) {
            case CSAMD_VALUE_F16: serve_column<CSAMD_VALUE_F16>(Q, b, e, acc_d, acc_u); break;
            case CSAMD_VALUE_BF16: serve_column<CSAMD_VALUE_BF16>(Q, b, e, acc_d, acc_u); break;
            case CSAMD_VALUE_F32: serve_column<CSAMD_VALUE_F32>(Q, b, e, acc_d, acc_u); break;
            default: serve_column<CSAMD_VALUE_FP8_E4M3>(Q, b, e, acc_d, acc_u); break;
        }
        __syncthreads();
        const unsigned long long o = (unsigned long long) j * Q.G;
        if (Q.op == CSAMD_SERVE_NNZ)
            for (unsigned int g = threadIdx.x; g < Q.G; g += blockDim.x) gmem((unsigned int *) Q.out)[o + g] = acc_u[g];
        else
            for (unsigned int g = threadIdx.x; g < Q.G; g += blockDim.x) gmem((double *) Q.out)[o + g] = acc_d[g];
        if (threadIdx.x == 0u) item_s = nxt;  // read by all after this item's first barrier
        __syncthreads();                        // accumulators are reused; item_s published
        item = item_s;
    }
}

__global__ void serve_materialize_kernel(unsigned int rows, unsigned int G, unsigned int label_seed, unsigned int mask_seed,
                                         unsigned int keep_thr24, unsigned int *__restrict__ labels,
                                         unsigned char *__restrict__ keep) {
    for (unsigned int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        if (labels != nullptr) labels[r] = seeded_label(label_seed, r, G);
        if (keep != nullptr) keep[r] = seeded_keep(mask_seed, r, keep_thr24) ? 1u : 0u;
    }
}

// ---------------------------------------------------------------------------
// Host runtime.
// ---------------------------------------------------------------------------
struct ticket_rec {
    csamd_serve_query q;
    int status;  // 0 queued, 1 running, 2 done, 3 released
    int slot;
    double t_submit;  // microseconds since the server's epoch
    double t_done;
};

struct batch_rec {
    bool active;
    hipStream_t stream;
    hipEvent_t done;
    std::vector<unsigned long long> tickets;
    query_dev *d_desc;
    unsigned int *d_prefix;  // [SERVE_MAX_BATCH + 1], then the counter
    query_dev *h_desc;       // pinned
    unsigned int *h_prefix;  // pinned
};

struct serve_state {
    csamd_serve_config cfg;
    std::vector<csamd_compressed_view> shards;
    std::vector<unsigned int> shard_types;
    unsigned long long resident;
    char *d_slots;
    std::vector<int> free_slots;
    std::deque<unsigned long long> waiting;
    std::vector<ticket_rec> tickets;
    batch_rec batches[4];
    unsigned int cu;
    std::chrono::steady_clock::time_point epoch;
};

double now_us(const serve_state *S) {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - S->epoch).count();
}

serve_state *state_of(csamd_serve_workspace *ws) {
    return ws != nullptr ? (serve_state *) ws->state : nullptr;
}

int retire(serve_state *S, unsigned int *completed) {
    for (unsigned int bi = 0; bi < S->cfg.max_batches_in_flight; ++bi) {
        batch_rec &B = S->batches[bi];
        if (!B.active) continue;
        const hipError_t st = hipEventQuery(B.done);
        if (st == hipErrorNotReady) continue;
        CSAMD_HIP(st, "serve batch");
        const double t = now_us(S);
        for (unsigned long long tk : B.tickets) {
            S->tickets[tk].status = 2;
            S->tickets[tk].t_done = t;
        }
        if (completed != nullptr) *completed += (unsigned int) B.tickets.size();
        B.tickets.clear();
        B.active = false;
    }
    return 1;
}

int launch_ready(serve_state *S) {
    for (unsigned int bi = 0; bi < S->cfg.max_batches_in_flight && !S->waiting.empty(); ++bi) {
        batch_rec &B = S->batches[bi];
        if (B.active) continue;
        unsigned int nq = 0u, items = 0u, gmax = 1u;
        while (!S->waiting.empty() && nq < S->cfg.max_batch_queries && !S->free_slots.empty()) {
            const unsigned long long tk = S->waiting.front();
            S->waiting.pop_front();
            ticket_rec &T = S->tickets[tk];
            T.slot = S->free_slots.back();
            S->free_slots.pop_back();
            T.status = 1;
            const csamd_compressed_view &sh = S->shards[T.q.shard];
            query_dev &d = B.h_desc[nq];
            d.colptr = sh.majorPtr;
            d.rowidx = sh.minorIdx;
            d.val = sh.val;
            d.features = T.q.features;
            d.labels = T.q.row_labels;
            d.keep = T.q.row_keep;
            d.out = S->d_slots + (size_t) T.slot * S->cfg.slot_bytes;
            d.rows = sh.rows;
            d.value_type = S->shard_types[T.q.shard];
            d.k = T.q.feature_count;
            d.G = T.q.group_count;
            d.op = T.q.op;
            d.label_seed = T.q.label_seed;
            d.mask_seed = T.q.mask_seed;
            d.keep_thr24 = T.q.keep_thr24;
            B.h_prefix[nq] = items;
            items += T.q.feature_count;
            if (T.q.group_count > gmax) gmax = T.q.group_count;
            B.tickets.push_back(tk);
            ++nq;
        }
        if (nq == 0u) break;  // no free slot: admitted queries wait
        B.h_prefix[nq] = items;
        B.h_prefix[nq + 1u] = 0u;  // the work counter
        CSAMD_HIP(hipMemcpyAsync(B.d_desc, B.h_desc, nq * sizeof(query_dev), hipMemcpyHostToDevice, B.stream), "serve desc H2D");
        CSAMD_HIP(hipMemcpyAsync(B.d_prefix, B.h_prefix, (nq + 2u) * sizeof(unsigned int), hipMemcpyHostToDevice, B.stream),
                  "serve prefix H2D");
        const unsigned int grid = items < 4u * S->cu ? items : 4u * S->cu;
        serve_batch_kernel<<<grid, SERVE_THREADS, (size_t) gmax * sizeof(double), B.stream>>>(B.d_desc, B.d_prefix, nq,
                                                                                             B.d_prefix + nq + 1u);
        CSAMD_HIP(hipGetLastError(), "serve_batch_kernel");
        CSAMD_HIP(hipEventRecord(B.done, B.stream), "serve event");
        B.active = true;
    }
    return 1;
}

int submit_at(serve_state *S, const csamd_serve_query *q, double t_submit, unsigned long long *ticket) {
    CSAMD_REQUIRE(q != nullptr && ticket != nullptr, "serve submit: null argument");
    CSAMD_REQUIRE(q->shard < S->shards.size(), "serve submit: unknown shard %u", q->shard);
    CSAMD_REQUIRE(q->op <= CSAMD_SERVE_NNZ, "serve submit: bad op %u", q->op);
    CSAMD_REQUIRE(q->group_count >= 1u && q->group_count <= CSAMD_SERVE_MAX_GROUPS, "serve submit: group_count %u out of range",
                  q->group_count);
    CSAMD_REQUIRE(q->feature_count >= 1u && q->features != nullptr, "serve submit: empty feature subset");
    CSAMD_REQUIRE((unsigned long long) q->feature_count * q->group_count * 8ull <= S->cfg.slot_bytes,
                  "serve submit: output %u x %u exceeds the slot size", q->feature_count, q->group_count);
    CSAMD_REQUIRE(q->feature_count <= S->shards[q->shard].cols, "serve submit: more features than columns");
    ticket_rec T;
    T.q = *q;
    T.status = 0;
    T.slot = -1;
    T.t_submit = t_submit;
    T.t_done = 0.0;
    *ticket = S->tickets.size();
    S->tickets.push_back(T);
    S->waiting.push_back(*ticket);
    return 1;
}

int release_ticket(serve_state *S, unsigned long long ticket) {
    CSAMD_REQUIRE(ticket < S->tickets.size(), "serve release: unknown ticket");
    ticket_rec &T = S->tickets[ticket];
    CSAMD_REQUIRE(T.status == 2, "serve release: query not finished");
    S->free_slots.push_back(T.slot);
    T.slot = -1;
    T.status = 3;
    return 1;
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_serve_init(csamd_serve_workspace *ws) {
    if (ws == nullptr) return;
    ws->device = -1;
    ws->state = nullptr;
}

void csamd_serve_clear(csamd_serve_workspace *ws) {
    serve_state *S = state_of(ws);
    if (S == nullptr) return;
    (void) hipSetDevice(S->cfg.device);
    for (batch_rec &B : S->batches) {
        if (B.stream != nullptr) (void) hipStreamSynchronize(B.stream);
        if (B.done != nullptr) (void) hipEventDestroy(B.done);
        if (B.stream != nullptr) (void) hipStreamDestroy(B.stream);
        if (B.d_desc != nullptr) (void) hipFree(B.d_desc);
        if (B.h_desc != nullptr) (void) hipHostFree(B.h_desc);
    }
    if (S->d_slots != nullptr) (void) hipFree(S->d_slots);
    delete S;
    csamd_serve_init(ws);
}

int csamd_serve_setup(csamd_serve_workspace *ws, const csamd_serve_config *cfg) {
    CSAMD_REQUIRE(ws != nullptr && cfg != nullptr, "serve setup: null argument");
    CSAMD_REQUIRE(cfg->max_slots >= 1u && cfg->max_batch_queries >= 1u && cfg->max_batch_queries <= SERVE_MAX_BATCH,
                  "serve setup: slots / batch size out of range");
    CSAMD_REQUIRE(cfg->max_batches_in_flight >= 1u && cfg->max_batches_in_flight <= 4u, "serve setup: 1..4 batches in flight");
    CSAMD_REQUIRE(cfg->slot_bytes >= 8u, "serve setup: slot_bytes too small");
    const unsigned long long slots_bytes = (unsigned long long) cfg->max_slots * cfg->slot_bytes;
    CSAMD_REQUIRE(slots_bytes <= cfg->budget_bytes, "serve setup: %u output slots need %llu bytes, budget %llu", cfg->max_slots,
                  slots_bytes, cfg->budget_bytes);
    csamd_serve_clear(ws);
    CSAMD_HIP(hipSetDevice(cfg->device), "hipSetDevice serve");
    serve_state *S = new serve_state();
    S->cfg = *cfg;
    S->resident = slots_bytes;
    S->d_slots = nullptr;
    S->cu = (unsigned int) device_cu_count();
    S->epoch = std::chrono::steady_clock::now();
    for (batch_rec &B : S->batches) {
        B.active = false;
        B.stream = nullptr;
        B.done = nullptr;
        B.d_desc = nullptr;
        B.h_desc = nullptr;
    }
    ws->state = S;
    ws->device = cfg->device;
    CSAMD_HIP(hipMalloc(&S->d_slots, slots_bytes), "hipMalloc serve slots");
    for (unsigned int i = 0; i < cfg->max_slots; ++i) S->free_slots.push_back((int) (cfg->max_slots - 1u - i));
    const size_t desc_bytes = align_up(SERVE_MAX_BATCH * sizeof(query_dev), 256u);
    const size_t pre_bytes = (SERVE_MAX_BATCH + 2u) * sizeof(unsigned int);
    for (unsigned int bi = 0; bi < cfg->max_batches_in_flight; ++bi) {
        batch_rec &B = S->batches[bi];
        CSAMD_HIP(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking), "serve stream");
        CSAMD_HIP(hipEventCreateWithFlags(&B.done, hipEventDisableTiming), "serve event");
        void *d = nullptr, *h = nullptr;
        CSAMD_HIP(hipMalloc(&d, desc_bytes + pre_bytes), "hipMalloc serve batch");
        CSAMD_HIP(hipHostMalloc(&h, desc_bytes + pre_bytes, hipHostMallocDefault), "hipHostMalloc serve batch");
        B.d_desc = (query_dev *) d;
        B.d_prefix = (unsigned int *) ((char *) d + desc_bytes);
        B.h_desc = (query_dev *) h;
        B.h_prefix = (unsigned int *) ((char *) h + desc_bytes);
    }
    return 1;
}

int csamd_serve_add_shard(csamd_serve_workspace *ws, const csamd_compressed_view *csc, unsigned int value_type,
                          unsigned int *shard_id) {
    serve_state *S = state_of(ws);
    CSAMD_REQUIRE(S != nullptr && csc != nullptr && shard_id != nullptr, "serve add_shard: bad arguments");
    CSAMD_REQUIRE(csc->axis == CSAMD_COMPRESSED_BY_COL, "serve add_shard: shard must be column-major (CSC)");
    CSAMD_REQUIRE(value_type_decodable(value_type), "serve add_shard: unsupported value type %u", value_type);
    CSAMD_REQUIRE(csc->cols == 0u || csc->majorPtr != nullptr, "serve add_shard: null column pointers");
    const unsigned long long bytes = ((unsigned long long) csc->cols + 1ull) * 4ull +
                                     (unsigned long long) csc->nnz * (4ull + (unsigned long long) value_type_bytes(value_type));
    CSAMD_REQUIRE(S->resident + bytes <= S->cfg.budget_bytes, "serve add_shard: %llu bytes over the budget (%llu of %llu used)",
                  bytes, S->resident, S->cfg.budget_bytes);
    S->resident += bytes;
    *shard_id = (unsigned int) S->shards.size();
    S->shards.push_back(*csc);
    S->shard_types.push_back(value_type);
    return 1;
}

int csamd_serve_submit(csamd_serve_workspace *ws, const csamd_serve_query *q, unsigned long long *ticket) {
    serve_state *S = state_of(ws);
    CSAMD_REQUIRE(S != nullptr, "serve submit: not set up");
    return submit_at(S, q, now_us(S), ticket);
}

int csamd_serve_pump(csamd_serve_workspace *ws, unsigned int *completed) {
    serve_state *S = state_of(ws);
    CSAMD_REQUIRE(S != nullptr, "serve pump: not set up");
    if (completed != nullptr) *completed = 0u;
    if (!retire(S, completed)) return 0;
    return launch_ready(S);
}

int csamd_serve_drain(csamd_serve_workspace *ws) {
    serve_state *S = state_of(ws);
    CSAMD_REQUIRE(S != nullptr, "serve drain: not set up");
    while (true) {
        if (!retire(S, nullptr) || !launch_ready(S)) return 0;
        bool active = false;
        for (unsigned int bi = 0; bi < S->cfg.max_batches_in_flight; ++bi) active |= S->batches[bi].active;
        if (!active && S->waiting.empty()) return 1;
        CSAMD_REQUIRE(active, "serve drain: queries wait for output slots held by unreleased results");
    }
}

int csamd_serve_get_result(csamd_serve_workspace *ws, unsigned long long ticket, csamd_serve_result *out) {
    serve_state *S = state_of(ws);
    CSAMD_REQUIRE(S != nullptr && out != nullptr, "serve result: bad arguments");
    std::memset(out, 0, sizeof(*out));
    if (ticket >= S->tickets.size()) {
        out->status = -1;
        return 1;
    }
    const ticket_rec &T = S->tickets[ticket];
    out->status = T.status;
    out->op = T.q.op;
    out->group_count = T.q.group_count;
    out->feature_count = T.q.feature_count;
    out->out = T.status == 2 ? (const void *) (S->d_slots + (size_t) T.slot * S->cfg.slot_bytes) : nullptr;
    out->latency_us = T.status >= 2 ? T.t_done - T.t_submit : 0.0;
    return 1;
}

int csamd_serve_release(csamd_serve_workspace *ws, unsigned long long ticket) {
    serve_state *S = state_of(ws);
    CSAMD_REQUIRE(S != nullptr, "serve release: not set up");
    return release_ticket(S, ticket);
}

int csamd_serve_replay(csamd_serve_workspace *ws, const csamd_serve_query *queries, const double *arrival_us, unsigned int n,
                       unsigned long long *tickets, double *latency_us, double *wall_us) {
    serve_state *S = state_of(ws);
    CSAMD_REQUIRE(S != nullptr && (n == 0u || (queries != nullptr && arrival_us != nullptr)), "serve replay: bad arguments");
    CSAMD_REQUIRE(S->waiting.empty(), "serve replay: queries already queued");
    for (unsigned int bi = 0; bi < S->cfg.max_batches_in_flight; ++bi)
        CSAMD_REQUIRE(!S->batches[bi].active, "serve replay: launches already running");
    const bool hold = tickets != nullptr;
    const double t0 = now_us(S);
    unsigned int next = 0u, done = 0u;
    const size_t first = S->tickets.size();
    while (done < n) {
        const double t = now_us(S) - t0;
        while (next < n && arrival_us[next] <= t) {
            unsigned long long tk = 0ull;
            if (!submit_at(S, queries + next, t0 + arrival_us[next], &tk)) return 0;
            if (hold) tickets[next] = tk;
            ++next;
        }
        unsigned int c = 0u;
        if (!retire(S, &c)) return 0;
        if (!hold && c != 0u)
            for (size_t i = first; i < S->tickets.size(); ++i)
                if (S->tickets[i].status == 2 && !release_ticket(S, i)) return 0;
        if (!launch_ready(S)) return 0;
        done += c;
        if (c == 0u) {
            bool active = false;
            for (unsigned int bi = 0; bi < S->cfg.max_batches_in_flight; ++bi) active |= S->batches[bi].active;
            CSAMD_REQUIRE(active || next < n || S->waiting.empty(),
                          "serve replay: queries wait for output slots held by unreleased results");
        }
    }
    double last = t0;
    for (unsigned int i = 0; i < n; ++i) {
        const ticket_rec &T = S->tickets[first + i];
        if (latency_us != nullptr) latency_us[i] = T.t_done - T.t_submit;
        if (T.t_done > last) last = T.t_done;
    }
    if (wall_us != nullptr) *wall_us = last - t0;
    return 1;
}

/* Seeded row labels / row mask of a serving query as device arrays (for the
 * per-query row-scan path and for tests).  Either output may be NULL. */
int csamd_serve_materialize(unsigned int rows, unsigned int group_count, unsigned int label_seed, unsigned int mask_seed,
                            unsigned int keep_thr24, unsigned int *d_labels, unsigned char *d_keep, csamd_stream_t stream) {
    CSAMD_REQUIRE(group_count >= 1u, "serve materialize: group_count 0");
    if (rows == 0u) return 1;
    serve_materialize_kernel<<<grid_for(rows, 256u, 4096u), 256, 0, (hipStream_t) stream>>>(rows, group_count, label_seed,
                                                                                          mask_seed, keep_thr24, d_labels, d_keep);
    return hip_ok(hipGetLastError(), "serve_materialize_kernel");
}

} // extern "C"
