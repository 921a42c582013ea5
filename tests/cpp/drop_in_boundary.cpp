// A caller of the drop-in C++ layer: includes only <cellshard_amd/cellshard.hh>
// and drives the hot path through the reference's namespaces and names.
//
// Part 1 follows the call sequence of the reference's
// tests/mask_groups_runtime_test.cu:55-198 (workspace lifecycle, Blocked-ELL /
// compressed / Sliced-ELL reductions with and without groups, masked
// reoptimize) on the same 4 x 5 matrices, with its two assertions that
// contradict the reference kernel corrected: keep[3] is 1 (row 3 holds
// nothing of group 0, the filter's fraction group, so its fraction is 0 <=
// 0.8), and the Blocked-ELL group_counts[8] is 0 (that matrix stores row 2's
// value 2 at column 3, which belongs to no group; the compressed and
// Sliced-ELL matrices store it at column 4, group 2).
// Part 2 calls the host-struct converters (convert/) on the same matrix and
// checks them against known answers.
// Part 3 exercises the sharded<T> metadata (the only part that runs with
// --host-only, without a device) and part 4 the sharded_device<T> residency.
//
// Exit status 0 when every check passes; each failed check is printed.
#include <cellshard_amd/cellshard.hh>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace rt = cellshard::runtime;
namespace dv = cellshard::device;
namespace sp = cellshard::sparse;
namespace cv = cellshard::convert;

static int g_failed = 0;
static int g_checked = 0;

#define CHECK(cond)                                                                 \
    do {                                                                            \
        ++g_checked;                                                                \
        if (!(cond)) {                                                              \
            ++g_failed;                                                             \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);  \
        }                                                                           \
    } while (0)

static bool near(float a, float b) { return std::fabs(a - b) <= 1e-3f; }

template<typename T>
static T *to_device(const T *h, std::size_t n) {
    T *d = nullptr;
    if (hipMalloc((void **) &d, n * sizeof(T)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

template<typename T>
static std::vector<T> to_host(const T *d, std::size_t n) {
    std::vector<T> h(n);
    if (n != 0 && hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) h.assign(n, T());
    return h;
}

static std::vector<__half> halves(const float *v, std::size_t n) {
    std::vector<__half> h(n);
    for (std::size_t i = 0; i < n; ++i) h[i] = __float2half(v[i]);
    return h;
}

struct reduce_host {
    std::vector<float> totals, groups, pct;
    std::vector<unsigned int> detected;
    std::vector<unsigned char> keep;
};

static reduce_host fetch(rt::sparse_group_reduce_workspace *ws, const rt::sparse_group_reduce_result &r) {
    reduce_host h;
    CHECK(hipStreamSynchronize(ws->stream) == hipSuccess);
    h.totals = to_host(r.row_totals, r.rows);
    h.detected = to_host(r.detected_features, r.rows);
    h.keep = to_host(r.row_keep, r.rows);
    if (r.group_count != 0u) {
        h.groups = to_host(r.group_counts, (std::size_t) r.rows * r.group_count);
        h.pct = to_host(r.group_percentages, (std::size_t) r.rows * r.group_count);
    }
    return h;
}

// ---------------------------------------------------------------------------
static void part1_mask_groups() {
    const unsigned int rows = 4u, cols = 5u;
    const std::uint32_t gm[5] = {1u << 0, 1u << 1, (1u << 0) | (1u << 1), 0u, 1u << 2};
    const unsigned char rk_h[4] = {1u, 0u, 1u, 1u};
    const unsigned char fk_h[5] = {1u, 1u, 0u, 1u, 1u};
    unsigned char *rk = to_device(rk_h, 4), *fk = to_device(fk_h, 5);
    CHECK(rk != nullptr && fk != nullptr);

    rt::sparse_group_reduce_workspace ws;
    rt::init(&ws);
    CHECK(rt::setup(&ws, 0) == 1);
    CHECK(rt::upload_feature_group_masks(&ws, cols, gm) == 1);
    rt::group_mask_config_view groups{3u, nullptr, ws.feature_group_masks};
    rt::row_feature_mask_view masks{rk, fk, nullptr, nullptr};
    rt::sparse_group_filter_params filter{1.0f, 1u, 0.80f, 0u};
    rt::sparse_group_reduce_result res{};

    // Blocked-ELL, block 2, two block slots per row block
    const unsigned int bci_h[8] = {0u, 1u, 0u, 1u, 0u, 1u, 1u, 2u};
    const float bv[16] = {5, 3, 7, 0, 9, 0, 0, 0, 1, 4, 6, 2, 0, 8, 0, 10};
    std::vector<__half> bvh = halves(bv, 16);
    dv::blocked_ell_view bell{};
    bell.rows = rows;
    bell.cols = cols;
    bell.nnz = 10u;
    bell.block_size = 2u;
    bell.ell_cols = 4u;
    bell.blockColIdx = to_device(bci_h, 8);
    bell.val = to_device(bvh.data(), 16);
    CHECK(rt::compute_sparse_group_reduce(&bell, &ws, &groups, &masks, &filter, &res) == 1);
    reduce_host h = fetch(&ws, res);
    CHECK(res.rows == rows && res.group_count == 3u);
    CHECK(near(h.totals[0], 8) && near(h.totals[1], 0) && near(h.totals[2], 7) && near(h.totals[3], 18));
    CHECK(h.detected[0] == 2u && h.detected[1] == 0u && h.detected[2] == 3u && h.detected[3] == 2u);
    CHECK(h.keep[0] == 1u && h.keep[1] == 0u && h.keep[2] == 1u && h.keep[3] == 1u);
    CHECK(near(h.groups[0], 5) && near(h.groups[1], 3) && near(h.groups[2], 0));
    CHECK(near(h.groups[6], 1) && near(h.groups[7], 4) && near(h.groups[8], 0));
    CHECK(near(h.groups[9], 0) && near(h.groups[10], 8) && near(h.groups[11], 0));
    CHECK(near(h.pct[0], 62.5f) && near(h.pct[4], 0.0f));

    // no groups, no masks: every stored nonzero of a row counts
    rt::group_mask_config_view none{0u, nullptr, nullptr};
    CHECK(rt::compute_sparse_group_reduce(&bell, &ws, &none, nullptr, &filter, &res) == 1);
    h = fetch(&ws, res);
    CHECK(res.group_count == 0u && near(h.totals[0], 15) && h.detected[0] == 3u && near(h.totals[3], 18));

    // compressed fallback: row 2 = {0: 1, 1: 4, 2: 6 (feature masked), 4: 2}
    const unsigned int cp_h[5] = {0u, 3u, 4u, 8u, 10u};
    const unsigned int ci_h[10] = {0u, 1u, 2u, 0u, 0u, 1u, 2u, 4u, 1u, 4u};
    const float cvf[10] = {5, 3, 7, 9, 1, 4, 6, 2, 8, 10};
    std::vector<__half> cvh = halves(cvf, 10);
    dv::compressed_view csr{rows, cols, 10u, sp::compressed_by_row, to_device(cp_h, 5), to_device(ci_h, 10), to_device(cvh.data(), 10)};
    CHECK(rt::compute_sparse_group_reduce_compressed_fallback(&csr, &ws, &groups, &masks, &filter, &res) == 1);
    h = fetch(&ws, res);
    CHECK(near(h.totals[2], 7) && near(h.groups[6], 1) && near(h.groups[7], 4) && near(h.groups[8], 2));
    CHECK(near(h.totals[3], 18) && near(h.groups[10], 8) && near(h.groups[11], 10) && h.keep[3] == 1u);

    // Sliced-ELL, one slice of 4 rows, width 3
    const unsigned int so_h[2] = {0u, rows}, sw_h[1] = {3u}, ss_h[1] = {0u};
    const unsigned int INV = sp::sliced_ell_invalid_col;
    const unsigned int sc_h[12] = {0u, 1u, 2u, 0u, INV, INV, 0u, 1u, 4u, 1u, 4u, INV};
    const float sv[12] = {5, 3, 7, 9, 0, 0, 1, 4, 2, 8, 10, 0};
    std::vector<__half> svh = halves(sv, 12);
    dv::sliced_ell_view sell{rows, cols, 9u, 1u, rows, to_device(so_h, 2), to_device(sw_h, 1), to_device(ss_h, 1), to_device(sc_h, 12),
                             to_device(svh.data(), 12)};
    CHECK(rt::compute_sparse_group_reduce(&sell, &ws, &groups, &masks, &filter, &res) == 1);
    h = fetch(&ws, res);
    CHECK(near(h.totals[0], 8) && near(h.totals[3], 18) && near(h.totals[2], 7) && h.detected[2] == 3u);
    CHECK(near(h.groups[8], 2) && h.keep[1] == 0u);

    // masked reoptimize: kept rows {0, 2, 3}, kept features {0, 1, 3, 4}
    rt::masked_sparse_reoptimize_workspace rws;
    rt::init(&rws);
    CHECK(rt::setup(&rws, 0) == 1);
    const unsigned int rr_h[4] = {0u, 0xffffffffu, 1u, 2u};
    const unsigned int fr_h[5] = {0u, 1u, 0xffffffffu, 2u, 3u};
    rt::row_feature_mask_view rmasks{rk, fk, to_device(rr_h, 4), to_device(fr_h, 5)};
    rt::masked_sparse_reoptimize_config rcfg{3u, 4u, 0u, 2u, rt::sparse_mask_layout_compressed};
    rt::masked_sparse_reoptimize_result rres{};
    CHECK(rt::manual_reoptimize_masked_sparse(&bell, &rmasks, &rcfg, &rws, &rres) == 1);
    CHECK(hipStreamSynchronize(rws.stream) == hipSuccess);
    CHECK(rres.metadata.kept_rows == 3u);
    CHECK(rres.metadata.kept_features == 4u);
    CHECK(rres.metadata.live_nnz == 7u);
    CHECK(rres.metadata.output_layout == rt::sparse_mask_layout_compressed);

    // the reference's no-op predictor
    rt::auto_reoptimize_prediction_input pin{};
    rt::auto_reoptimize_prediction pout{};
    CHECK(rt::predict_masked_sparse_reoptimization(&pin, &pout) == 1);

    rt::clear(&rws);
    rt::clear(&ws);
    std::printf("part 1 (mask_groups call sequence): done\n");
}

// ---------------------------------------------------------------------------
static void fill_compressed(sp::compressed *m, unsigned int rows, unsigned int cols, unsigned int axis, const unsigned int *ptr,
                            const unsigned int *idx, const float *v, unsigned int nnz) {
    sp::init(m, rows, cols, nnz, axis);
    CHECK(sp::allocate(m) == 1);
    const unsigned int major = axis == sp::compressed_by_col ? cols : rows;
    std::memcpy(m->majorPtr, ptr, (major + 1u) * 4u);
    std::memcpy(m->minorIdx, idx, nnz * 4u);
    for (unsigned int i = 0; i < nnz; ++i) m->val[i] = __float2half(v[i]);
}

static void part2_host_converters() {
    const unsigned int rows = 4u, cols = 5u;
    const unsigned int cp[5] = {0u, 3u, 4u, 8u, 10u};
    const unsigned int ci[10] = {0u, 1u, 2u, 0u, 0u, 1u, 2u, 4u, 1u, 4u};
    const float cvv[10] = {5, 3, 7, 9, 1, 4, 6, 2, 8, 10};
    sp::compressed csr;
    fill_compressed(&csr, rows, cols, sp::compressed_by_row, cp, ci, cvv, 10u);
    const unsigned int INV = sp::blocked_ell_invalid_col;

    // Blocked-ELL block 2: row block 0 uses blocks {0, 1}, row block 1 {0, 1, 2}
    sp::blocked_ell bell;
    sp::init(&bell);
    CHECK(cv::blocked_ell_from_compressed(&csr, 2u, &bell) == 1);
    CHECK(bell.rows == rows && bell.cols == cols && bell.block_size == 2u && bell.ell_cols == 6u && bell.nnz == 10u);
    const unsigned int want_bci[6] = {0u, 1u, INV, 0u, 1u, 2u};
    CHECK(std::memcmp(bell.blockColIdx, want_bci, sizeof(want_bci)) == 0);
    // row 2 = [1 4 | 6 0 | 2 0]: columns 0, 1, 2, 4 at slot * 2 + column % 2
    const float want_r2[6] = {1, 4, 6, 0, 2, 0};
    bool r2 = true;
    for (unsigned int e = 0; e < 6u; ++e) r2 = r2 && __half2float(bell.val[2u * 6u + e]) == want_r2[e];
    CHECK(r2);
    CHECK(sp::row_block_count(&bell) == 2u && sp::ell_width_blocks(&bell) == 3u);

    // tuner: every candidate has the same width in blocks here, so the
    // smallest padded bytes (B = 8) wins
    const unsigned int cands[3] = {8u, 16u, 32u};
    cv::blocked_ell_tune_result tr{};
    CHECK(cv::choose_blocked_ell_block_size(&csr, cands, 3u, &tr) == 1);
    CHECK(tr.block_size == 8u);
    sp::blocked_ell bell_auto;
    sp::init(&bell_auto);
    cv::blocked_ell_tune_result picked{};
    CHECK(cv::blocked_ell_from_compressed_auto(&csr, cands, 3u, &bell_auto, &picked) == 1);
    CHECK(picked.block_size == 8u && bell_auto.block_size == 8u && bell_auto.ell_cols == 8u && bell_auto.nnz == 10u);

    // Sliced-ELL, slices of 2 rows: widths max(3, 1) = 3 and max(4, 2) = 4
    sp::sliced_ell sell;
    sp::init(&sell);
    CHECK(cv::sliced_ell_from_compressed(&csr, 2u, &sell) == 1);
    CHECK(sell.slice_count == 2u && sell.slice_widths[0] == 3u && sell.slice_widths[1] == 4u && sell.nnz == 10u);
    CHECK(sp::total_slots(&sell) == 14u && sp::uniform_slice_rows(&sell) == 2u);
    const unsigned int want_cols[14] = {0u, 1u, 2u, 0u, INV, INV, 0u, 1u, 2u, 4u, 1u, 4u, INV, INV};
    CHECK(std::memcmp(sell.col_idx, want_cols, sizeof(want_cols)) == 0);
    CHECK(sp::slice_slot_base(&sell, 1u) == 6u && sp::find_slice(&sell, 3u) == 1u && sp::row_nnz(&sell, 3u) == 2u);

    // the same matrix stored by column (CSC) converts to the same Sliced-ELL
    const unsigned int kp[6] = {0u, 3u, 6u, 8u, 8u, 10u};
    const unsigned int ki[10] = {0u, 1u, 2u, 0u, 2u, 3u, 0u, 2u, 2u, 3u};
    const float kv[10] = {5, 9, 1, 3, 4, 8, 7, 6, 2, 10};
    sp::compressed csc;
    fill_compressed(&csc, rows, cols, sp::compressed_by_col, kp, ki, kv, 10u);
    sp::sliced_ell sell2;
    sp::init(&sell2);
    CHECK(cv::sliced_ell_from_compressed(&csc, 2u, &sell2) == 1);
    CHECK(sell2.slice_count == 2u && std::memcmp(sell2.col_idx, want_cols, sizeof(want_cols)) == 0);
    bool same_vals = true;
    for (unsigned int k = 0; k < 14u; ++k) same_vals = same_vals && __half2float(sell.val[k]) == __half2float(sell2.val[k]);
    CHECK(same_vals);

    // Blocked-ELL -> Sliced-ELL keeps the stored nonzeros in slot order
    sp::sliced_ell sell3;
    sp::init(&sell3);
    CHECK(cv::sliced_ell_from_blocked_ell(&bell, 2u, &sell3) == 1);
    CHECK(sell3.slice_count == 2u && std::memcmp(sell3.col_idx, want_cols, sizeof(want_cols)) == 0);

    // slice-height tuner: {1, 2, 4} scores 1 + 0.25, 1.4 + 0.125, 1.6 + 0.0625
    const unsigned int sc[3] = {1u, 2u, 4u};
    cv::sliced_ell_tune_result st{};
    CHECK(cv::choose_sliced_ell_slice_rows_from_compressed(&csr, sc, 3u, &st) == 1);
    CHECK(st.slice_rows == 1u && st.padded_slots == 10u);

    sp::clear(&sell3);
    sp::clear(&sell2);
    sp::clear(&csc);
    sp::clear(&sell);
    sp::clear(&bell_auto);
    sp::clear(&bell);
    sp::clear(&csr);
    std::printf("part 2 (host converters): done\n");
}

// ---------------------------------------------------------------------------
// three CSR parts of 5 columns: 3 rows (5 entries), 0 rows, 2 rows (3 entries)
static sp::compressed *make_part(unsigned int which) {
    static const unsigned int p0[4] = {0u, 2u, 2u, 5u}, i0[5] = {0u, 4u, 1u, 2u, 3u};
    static const unsigned int p1[1] = {0u};
    static const unsigned int p2[3] = {0u, 1u, 3u}, i2[3] = {2u, 0u, 4u};
    static const float v0[5] = {1, 2, 3, 4, 5}, v2[3] = {6, 7, 8};
    sp::compressed *m = new sp::compressed;
    if (which == 0u) fill_compressed(m, 3u, 5u, sp::compressed_by_row, p0, i0, v0, 5u);
    else if (which == 1u) fill_compressed(m, 0u, 5u, sp::compressed_by_row, p1, nullptr, nullptr, 0u);
    else fill_compressed(m, 2u, 5u, sp::compressed_by_row, p2, i2, v2, 3u);
    return m;
}

static void part3_sharded_metadata(cellshard::sharded<sp::compressed> *m) {
    cellshard::init(m);
    for (unsigned int k = 0; k < 3u; ++k) CHECK(cellshard::append_partition(m, make_part(k)) == 1);
    CHECK(m->rows == 5u && m->cols == 5u && m->nnz == 8u && m->num_partitions == 3u && m->num_shards == 3u);
    CHECK(m->partition_offsets[1] == 3u && m->partition_offsets[2] == 3u && m->partition_offsets[3] == 5u);
    CHECK(cellshard::find_partition(m, 3ul) == 2u && cellshard::find_partition(m, 5ul) == 3u);
    CHECK(cellshard::partition_aux(m->parts[0]) == sp::compressed_by_row && cellshard::partition_nnz(m->parts[2]) == 3u);
    CHECK(cellshard::partition_bytes(m, 0ul) == sp::bytes(m->parts[0]));
    // two shards: rows {0..2} (partition 0) and {3, 4} (the empty partition 1 and partition 2)
    CHECK(cellshard::set_equal_shards(m, 2ul) == 1);
    CHECK(m->num_shards == 2u && m->shard_offsets[1] == 3u && cellshard::first_partition_in_shard(m, 1ul) == 1u);
    CHECK(cellshard::partition_count_in_shard(m, 1ul) == 2u && cellshard::rows_in_shard(m, 1ul) == 2u && cellshard::nnz_in_shard(m, 1ul) == 3u);
    CHECK(cellshard::shard_bytes(m, 0ul) == sp::bytes(m->parts[0]) && cellshard::shard_loaded(m, 1ul) == 1);
    unsigned long *offs = nullptr, count = 0ul;
    CHECK(cellshard::build_shard_offsets_by_rows(m, 2ul, &offs, &count) == 1);
    CHECK(count == 2u && offs != nullptr && offs[0] == 0u && offs[1] == 3u && offs[2] == 5u);
    std::free(offs);
    const unsigned long bad[3] = {0ul, 4ul, 5ul};
    CHECK(cellshard::reshard(m, 2ul, bad) == 0);  // row 4 is inside partition 2
    std::printf("part 3 (sharded metadata): done\n");
}

static void part4_sharded_device(cellshard::sharded<sp::compressed> *m) {
    dv::sharded_device<sp::compressed> st;
    dv::init(&st);
    CHECK(dv::reserve(&st, m->num_partitions) == 1);
    CHECK(dv::upload_shard(&st, m, 1ul, 0) == hipSuccess);
    // partitions 1 and 2 share one packed allocation, owned by the first
    CHECK(st.parts[1].view != nullptr && st.parts[2].view != nullptr && st.parts[0].view == nullptr);
    CHECK(st.parts[1].allocation != nullptr && st.parts[2].allocation == nullptr && st.parts[2].group_begin == 1u);
    CHECK(st.parts[1].allocation_bytes <= dv::device_shard_bytes(m, 1ul));
    dv::compressed_view dview{};
    CHECK(hipMemcpy(&dview, st.parts[2].view, sizeof(dview), hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(dview.rows == 2u && dview.nnz == 3u);
    std::vector<unsigned int> idx = to_host(dview.minorIdx, 3);
    CHECK(idx[0] == 2u && idx[1] == 0u && idx[2] == 4u);
    std::vector<__half> val = to_host(dview.val, 3);
    CHECK(__half2float(val[2]) == 8.0f);
    // swap: shard 0 in, shard 1 out (its allocation goes to the cache)
    CHECK(dv::swap_shard(&st, m, (const csamd_cspack_file *) nullptr, 1ul, 0ul, 0, 0, 0) == hipSuccess);
    CHECK(st.parts[0].view != nullptr && st.parts[1].view == nullptr && st.parts[2].view == nullptr && st.cache_count == 1u);
    CHECK(hipMemcpy(&dview, st.parts[0].view, sizeof(dview), hipMemcpyDeviceToHost) == hipSuccess);
    std::vector<unsigned int> p0 = to_host(dview.majorPtr, 4);
    CHECK(dview.rows == 3u && p0[3] == 5u);
    // releasing one partition of a packed shard releases the group
    CHECK(dv::upload_shard(&st, m, 1ul, 0) == hipSuccess && st.cache_count == 0u);
    CHECK(dv::release_partition(&st, 2ul) == hipSuccess && st.parts[1].view == nullptr && st.cache_count == 1u);
    // stage with drop: the host parts of shard 0 go away, the device copy stays
    CHECK(dv::stage_shard(&st, m, (const csamd_cspack_file *) nullptr, 0ul, 0, 1) == hipSuccess);
    CHECK(m->parts[0] == nullptr && st.parts[0].view != nullptr);
    // a shard with parts that are not materialized and no pack to read fails cleanly
    CHECK(dv::release_shard(&st, m, 0ul) == hipSuccess);
    CHECK(dv::stage_shard(&st, m, (const csamd_cspack_file *) nullptr, 0ul, 0, 0) != hipSuccess);
    dv::clear(&st);
    CHECK(st.parts == nullptr && st.cache_count == 0u);
    std::printf("part 4 (sharded_device residency): done\n");
}

int main(int argc, char **argv) {
    const bool host_only = argc > 1 && std::strcmp(argv[1], "--host-only") == 0;
    cellshard::sharded<sp::compressed> m;
    if (host_only) {
        part3_sharded_metadata(&m);
        cellshard::clear(&m);
        std::printf("%d checks, %d failed\n", g_checked, g_failed);
        return g_failed == 0 ? 0 : 1;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        std::fprintf(stderr, "no HIP device\n");
        return 2;
    }
    part1_mask_groups();
    part2_host_converters();
    part3_sharded_metadata(&m);
    part4_sharded_device(&m);
    cellshard::clear(&m);
    std::printf("%d checks, %d failed\n", g_checked, g_failed);
    return g_failed == 0 ? 0 : 1;
}
