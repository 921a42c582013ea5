// Host-struct converters: the reference's convert:: entry points take host
// sparse::compressed / blocked_ell structs and (re)allocate the host output
// with one packed malloc (formats/*.cuh allocate()).  Here the input is
// staged to HBM, converted by the HIP kernels, and copied back into an output
// laid out exactly as the reference's allocate() lays it out.
// Reference: src/convert/blocked_ell_from_compressed.cuh:157-190,258-347,449-461;
// src/convert/sliced_ell_from_compressed.cuh:86-164,256-284;
// src/convert/sliced_ell_from_blocked_ell.cuh:233-364; formats/*.cuh clear/allocate.
#include "common.hh"

#include <cstdlib>

namespace csamd {
namespace {

constexpr unsigned int HOST_PINNED = 1u;  // storage from hipHostMalloc

struct dev_arrays {
    void *p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    ~dev_arrays() {
        for (void *q : p)
            if (q != nullptr) (void) hipFree(q);
    }
    void *alloc(int slot, size_t bytes) {
        if (bytes == 0u) bytes = 4u;
        if (!hip_ok(hipMalloc(&p[slot], bytes), "hipMalloc host converter")) return nullptr;
        return p[slot];
    }
};

struct scoped_convert_ws {
    csamd_convert_workspace ws;
    bool ok = false;
    scoped_convert_ws() {
        csamd_convert_init(&ws);
        int dev = 0;
        (void) hipGetDevice(&dev);
        ok = csamd_convert_setup(&ws, dev, nullptr) == 1;
    }
    ~scoped_convert_ws() { csamd_convert_clear(&ws); }
};

int upload_csr(const csamd_host_compressed *src, dev_arrays &d, csamd_compressed_view *v) {
    const unsigned int major = src->axis == CSAMD_COMPRESSED_BY_COL ? src->cols : src->rows;
    void *pp = d.alloc(0, (size_t) (major + 1u) * 4u), *pi = d.alloc(1, (size_t) src->nnz * 4u), *pv = d.alloc(2, (size_t) src->nnz * 2u);
    if (pp == nullptr || pi == nullptr || pv == nullptr) return 0;
    CSAMD_HIP(hipMemcpy(pp, src->majorPtr, (size_t) (major + 1u) * 4u, hipMemcpyHostToDevice), "H2D ptr");
    if (src->nnz != 0u) {
        CSAMD_HIP(hipMemcpy(pi, src->minorIdx, (size_t) src->nnz * 4u, hipMemcpyHostToDevice), "H2D idx");
        CSAMD_HIP(hipMemcpy(pv, src->val, (size_t) src->nnz * 2u, hipMemcpyHostToDevice), "H2D val");
    }
    std::memset(v, 0, sizeof(*v));
    v->rows = src->rows;
    v->cols = src->cols;
    v->nnz = src->nnz;
    v->axis = src->axis;
    v->majorPtr = (unsigned int *) pp;
    v->minorIdx = (unsigned int *) pi;
    v->val = (csamd_half *) pv;
    return 1;
}

// Transpose a device CSC view into a CSR view (ascending columns per row).
int csc_to_csr(const csamd_compressed_view *csc, dev_arrays &d, csamd_compressed_view *csr, hipStream_t s) {
    const size_t scratch = csamd_compressed_transpose_scratch_bytes(csc->cols, csc->rows, csc->nnz);
    void *pp = d.alloc(3, (size_t) (csc->rows + 1u) * 4u), *pi = d.alloc(4, (size_t) csc->nnz * 4u), *pv = d.alloc(5, (size_t) csc->nnz * 2u);
    void *tmp = nullptr;
    if (pp == nullptr || pi == nullptr || pv == nullptr) return 0;
    CSAMD_HIP(hipMalloc(&tmp, scratch), "hipMalloc transpose scratch");
    const int ok = csamd_build_compressed_transpose_raw_typed(csc->cols, csc->rows, csc->nnz, csc->majorPtr, csc->minorIdx, csc->val,
                                                             CSAMD_VALUE_F16, (unsigned int *) pp, (unsigned int *) pi, pv, tmp, scratch,
                                                             (csamd_stream_t) s);
    const hipError_t e = hipStreamSynchronize(s);
    (void) hipFree(tmp);
    if (!ok) return 0;
    CSAMD_HIP(e, "transpose sync");
    *csr = *csc;
    csr->axis = CSAMD_COMPRESSED_BY_ROW;
    csr->majorPtr = (unsigned int *) pp;
    csr->minorIdx = (unsigned int *) pi;
    csr->val = (csamd_half *) pv;
    return 1;
}

int prepare_csr(const csamd_host_compressed *src, dev_arrays &d, csamd_compressed_view *csr, hipStream_t s) {
    csamd_compressed_view v;
    if (!upload_csr(src, d, &v)) return 0;
    if (src->axis == CSAMD_COMPRESSED_BY_ROW) {
        *csr = v;
        return 1;
    }
    return csc_to_csr(&v, d, csr, s);
}

int host_bell_allocate(csamd_host_blocked_ell *m) {
    const size_t rb = m->block_size != 0u ? (m->rows + m->block_size - 1u) / m->block_size : 0u;
    const size_t w = m->block_size != 0u && m->ell_cols != 0u ? m->ell_cols / m->block_size : 0u;
    const size_t idx_bytes = rb * w * 4u;
    const size_t val_off = align_up(idx_bytes, 2u);
    const size_t total = val_off + (size_t) m->rows * m->ell_cols * 2u;
    m->storage = nullptr;
    m->blockColIdx = nullptr;
    m->val = nullptr;
    if (total == 0u) return 1;
    m->storage = std::malloc(total);
    CSAMD_REQUIRE(m->storage != nullptr, "blocked_ell allocate: malloc(%zu) failed", total);
    m->blockColIdx = idx_bytes != 0u ? (unsigned int *) m->storage : nullptr;
    m->val = m->rows != 0u && m->ell_cols != 0u ? (csamd_half *) ((char *) m->storage + val_off) : nullptr;
    return 1;
}

int host_sell_allocate(csamd_host_sliced_ell *m, unsigned int S, size_t total_slots) {
    const size_t off_bytes = (size_t) (S + 1u) * 4u;
    const size_t w_off = align_up(off_bytes, 4u);
    const size_t col_off = align_up(w_off + (size_t) S * 4u, 4u);
    const size_t val_off = align_up(col_off + total_slots * 4u, 2u);
    const size_t total = val_off + total_slots * 2u;
    m->storage = nullptr;
    m->slice_row_offsets = m->slice_widths = m->col_idx = nullptr;
    m->val = nullptr;
    m->slice_count = S;
    if (total == 0u) return 1;
    m->storage = std::malloc(total);
    CSAMD_REQUIRE(m->storage != nullptr, "sliced_ell allocate: malloc(%zu) failed", total);
    char *b = (char *) m->storage;
    m->slice_row_offsets = (unsigned int *) b;
    m->slice_widths = (unsigned int *) (b + w_off);
    m->col_idx = (unsigned int *) (b + col_off);
    m->val = (csamd_half *) (b + val_off);
    return 1;
}

// Device SELL (layout + emit) into a host struct.
int sell_to_host(const csamd_sliced_ell_view &v, size_t total, const csamd_host_sliced_ell *meta, csamd_host_sliced_ell *dst) {
    csamd_host_sliced_ell_clear(dst);
    dst->rows = meta->rows;
    dst->cols = meta->cols;
    dst->nnz = meta->nnz;
    if (!host_sell_allocate(dst, v.slice_count, total)) return 0;
    if (dst->storage == nullptr) return 1;
    CSAMD_HIP(hipMemcpy(dst->slice_row_offsets, v.slice_row_offsets, (size_t) (v.slice_count + 1u) * 4u, hipMemcpyDeviceToHost), "D2H");
    if (v.slice_count != 0u)
        CSAMD_HIP(hipMemcpy(dst->slice_widths, v.slice_widths, (size_t) v.slice_count * 4u, hipMemcpyDeviceToHost), "D2H");
    if (total != 0u) {
        CSAMD_HIP(hipMemcpy(dst->col_idx, v.col_idx, total * 4u, hipMemcpyDeviceToHost), "D2H");
        CSAMD_HIP(hipMemcpy(dst->val, v.val, total * 2u, hipMemcpyDeviceToHost), "D2H");
    }
    return 1;
}

int sell_build(const csamd_compressed_view *csr, unsigned int slice_rows, csamd_convert_workspace *ws, csamd_host_sliced_ell *dst,
               unsigned int nnz_meta) {
    const unsigned int rows = csr->rows;
    const unsigned int S = rows == 0u ? 0u : (slice_rows > rows ? rows : slice_rows);
    const unsigned int count = rows == 0u ? 0u : (rows + S - 1u) / S;
    dev_arrays d;
    unsigned int *off = (unsigned int *) d.alloc(0, (size_t) (count + 1u) * 4u), *w = (unsigned int *) d.alloc(1, (size_t) count * 4u + 4u),
                 *so = (unsigned int *) d.alloc(2, (size_t) count * 4u + 4u);
    if (off == nullptr || w == nullptr || so == nullptr) return 0;
    unsigned int got = 0u;
    unsigned long long total = 0ull;
    if (!csamd_sliced_ell_layout(csr, slice_rows, off, w, so, ws, &got, &total)) return 0;
    void *col = d.alloc(3, (size_t) total * 4u), *val = d.alloc(4, (size_t) total * 2u);
    if (col == nullptr || val == nullptr) return 0;
    csamd_sliced_ell_view v;
    std::memset(&v, 0, sizeof(v));
    v.rows = rows;
    v.cols = csr->cols;
    v.slice_count = got;
    v.slice_rows = S;
    v.slice_row_offsets = off;
    v.slice_widths = w;
    v.slice_slot_offsets = so;
    v.col_idx = (unsigned int *) col;
    v.val = (csamd_half *) val;
    if (!csamd_sliced_ell_from_compressed(csr, CSAMD_VALUE_F16, &v, ws)) return 0;
    CSAMD_HIP(hipStreamSynchronize((hipStream_t) ws->stream), "sell sync");
    csamd_host_sliced_ell meta;
    std::memset(&meta, 0, sizeof(meta));
    meta.rows = rows;
    meta.cols = csr->cols;
    meta.nnz = nnz_meta;
    return sell_to_host(v, (size_t) total, &meta, dst);
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_host_compressed_clear(csamd_host_compressed *m) {
    if (m == nullptr) return;
    if ((m->flags & HOST_PINNED) != 0u && m->storage != nullptr) (void) hipHostFree(m->storage);
    else if (m->storage != nullptr) std::free(m->storage);
    else {
        std::free(m->majorPtr);
        std::free(m->minorIdx);
        std::free(m->val);
    }
    std::memset(m, 0, sizeof(*m));
}

void csamd_host_blocked_ell_clear(csamd_host_blocked_ell *m) {
    if (m == nullptr) return;
    if ((m->flags & HOST_PINNED) != 0u && m->storage != nullptr) (void) hipHostFree(m->storage);
    else if (m->storage != nullptr) std::free(m->storage);
    else {
        std::free(m->blockColIdx);
        std::free(m->val);
    }
    std::memset(m, 0, sizeof(*m));
}

void csamd_host_sliced_ell_clear(csamd_host_sliced_ell *m) {
    if (m == nullptr) return;
    if ((m->flags & HOST_PINNED) != 0u && m->storage != nullptr) (void) hipHostFree(m->storage);
    else if (m->storage != nullptr) std::free(m->storage);
    else {
        std::free(m->slice_row_offsets);
        std::free(m->slice_widths);
        std::free(m->col_idx);
        std::free(m->val);
    }
    std::memset(m, 0, sizeof(*m));
}

int csamd_host_choose_blocked_ell_block_size(const csamd_host_compressed *src, const unsigned int *candidates,
                                             unsigned int candidate_count, csamd_blocked_ell_tune_result *out) {
    CSAMD_REQUIRE(src != nullptr && out != nullptr && candidates != nullptr && candidate_count != 0u,
                  "choose_blocked_ell_block_size: bad arguments");
    CSAMD_REQUIRE(src->axis == CSAMD_COMPRESSED_BY_ROW, "choose_blocked_ell_block_size: source axis must be row");
    scoped_convert_ws w;
    if (!w.ok) return 0;
    dev_arrays d;
    csamd_compressed_view v;
    if (!upload_csr(src, d, &v)) return 0;
    return csamd_choose_blocked_ell_block_size(&v, candidates, candidate_count, 2u, &w.ws, out);
}

int csamd_host_blocked_ell_from_compressed(const csamd_host_compressed *src, unsigned int block_size, csamd_host_blocked_ell *dst) {
    CSAMD_REQUIRE(src != nullptr && dst != nullptr && block_size != 0u, "blocked_ell_from_compressed: bad arguments");
    CSAMD_REQUIRE(src->axis == CSAMD_COMPRESSED_BY_ROW, "blocked_ell_from_compressed: source axis must be row");
    scoped_convert_ws w;
    if (!w.ok) return 0;
    dev_arrays d;
    csamd_compressed_view v;
    if (!upload_csr(src, d, &v)) return 0;
    unsigned int width = 0u;
    unsigned long long total = 0ull;
    double fill = 0.0;
    if (!csamd_blocked_ell_shape(&v, block_size, &w.ws, &width, &total, &fill)) return 0;
    const size_t rb = (src->rows + block_size - 1u) / block_size;
    void *bci = d.alloc(3, rb * width * 4u), *val = d.alloc(4, (size_t) src->rows * width * block_size * 2u);
    if (bci == nullptr || val == nullptr) return 0;
    csamd_blocked_ell_view bv;
    std::memset(&bv, 0, sizeof(bv));
    bv.block_size = block_size;
    bv.ell_cols = width * block_size;
    bv.blockColIdx = (unsigned int *) bci;
    bv.val = (csamd_half *) val;
    if (!csamd_blocked_ell_from_compressed(&v, CSAMD_VALUE_F16, &bv, &w.ws)) return 0;
    csamd_host_blocked_ell_clear(dst);
    dst->rows = src->rows;
    dst->cols = src->cols;
    dst->nnz = bv.nnz;
    dst->block_size = block_size;
    dst->ell_cols = width * block_size;
    if (!host_bell_allocate(dst)) return 0;
    if (dst->blockColIdx != nullptr) CSAMD_HIP(hipMemcpy(dst->blockColIdx, bci, rb * width * 4u, hipMemcpyDeviceToHost), "D2H bci");
    if (dst->val != nullptr) CSAMD_HIP(hipMemcpy(dst->val, val, (size_t) src->rows * dst->ell_cols * 2u, hipMemcpyDeviceToHost), "D2H val");
    return 1;
}

int csamd_host_blocked_ell_from_compressed_auto(const csamd_host_compressed *src, const unsigned int *candidates,
                                                unsigned int candidate_count, csamd_host_blocked_ell *dst,
                                                csamd_blocked_ell_tune_result *picked) {
    csamd_blocked_ell_tune_result t;
    std::memset(&t, 0, sizeof(t));
    if (!csamd_host_choose_blocked_ell_block_size(src, candidates, candidate_count, &t)) return 0;
    if (!csamd_host_blocked_ell_from_compressed(src, t.block_size, dst)) return 0;
    if (picked != nullptr) *picked = t;
    return 1;
}

int csamd_host_choose_sliced_ell_slice_rows(const csamd_host_compressed *src, const unsigned int *candidates,
                                            unsigned int candidate_count, csamd_sliced_ell_tune_result *out) {
    CSAMD_REQUIRE(src != nullptr && out != nullptr, "choose_sliced_ell_slice_rows: null argument");
    scoped_convert_ws w;
    if (!w.ok) return 0;
    dev_arrays d;
    csamd_compressed_view v;
    if (!prepare_csr(src, d, &v, (hipStream_t) w.ws.stream)) return 0;
    return csamd_choose_sliced_ell_slice_rows(&v, candidates, candidate_count, &w.ws, out);
}

int csamd_host_sliced_ell_from_compressed(const csamd_host_compressed *src, unsigned int slice_rows, csamd_host_sliced_ell *dst) {
    CSAMD_REQUIRE(src != nullptr && dst != nullptr && slice_rows != 0u, "sliced_ell_from_compressed: bad arguments");
    scoped_convert_ws w;
    if (!w.ok) return 0;
    dev_arrays d;
    csamd_compressed_view v;
    if (!prepare_csr(src, d, &v, (hipStream_t) w.ws.stream)) return 0;
    return sell_build(&v, slice_rows, &w.ws, dst, src->nnz);
}

int csamd_host_sliced_ell_from_compressed_auto(const csamd_host_compressed *src, const unsigned int *candidates,
                                               unsigned int candidate_count, csamd_host_sliced_ell *dst,
                                               csamd_sliced_ell_tune_result *picked) {
    csamd_sliced_ell_tune_result t;
    std::memset(&t, 0, sizeof(t));
    if (!csamd_host_choose_sliced_ell_slice_rows(src, candidates, candidate_count, &t)) return 0;
    if (!csamd_host_sliced_ell_from_compressed(src, t.slice_rows == 0u ? 1u : t.slice_rows, dst)) return 0;
    if (picked != nullptr) *picked = t;
    return 1;
}

int csamd_host_sliced_ell_from_blocked_ell(const csamd_host_blocked_ell *src, unsigned int slice_rows, csamd_host_sliced_ell *dst) {
    CSAMD_REQUIRE(src != nullptr && dst != nullptr && slice_rows != 0u && src->block_size != 0u,
                  "sliced_ell_from_blocked_ell: bad arguments");
    scoped_convert_ws w;
    if (!w.ok) return 0;
    dev_arrays d;
    const size_t rb = (src->rows + src->block_size - 1u) / src->block_size;
    const size_t wb = src->ell_cols / src->block_size;
    void *bci = d.alloc(0, rb * wb * 4u), *val = d.alloc(1, (size_t) src->rows * src->ell_cols * 2u);
    if (bci == nullptr || val == nullptr) return 0;
    if (rb * wb != 0u) CSAMD_HIP(hipMemcpy(bci, src->blockColIdx, rb * wb * 4u, hipMemcpyHostToDevice), "H2D bci");
    if ((size_t) src->rows * src->ell_cols != 0u)
        CSAMD_HIP(hipMemcpy(val, src->val, (size_t) src->rows * src->ell_cols * 2u, hipMemcpyHostToDevice), "H2D val");
    csamd_blocked_ell_view bv;
    std::memset(&bv, 0, sizeof(bv));
    bv.rows = src->rows;
    bv.cols = src->cols;
    bv.nnz = src->nnz;
    bv.block_size = src->block_size;
    bv.ell_cols = src->ell_cols;
    bv.blockColIdx = (unsigned int *) bci;
    bv.val = (csamd_half *) val;
    const unsigned int rows = src->rows;
    const unsigned int S = rows == 0u ? 0u : (slice_rows > rows ? rows : slice_rows);
    const unsigned int count = rows == 0u ? 0u : (rows + S - 1u) / S;
    unsigned int *off = (unsigned int *) d.alloc(2, (size_t) (count + 1u) * 4u), *wd = (unsigned int *) d.alloc(3, (size_t) count * 4u + 4u),
                 *so = (unsigned int *) d.alloc(4, (size_t) count * 4u + 4u);
    if (off == nullptr || wd == nullptr || so == nullptr) return 0;
    unsigned int got = 0u;
    unsigned long long total = 0ull;
    if (!csamd_sliced_ell_layout_from_blocked_ell(&bv, CSAMD_VALUE_F16, slice_rows, off, wd, so, &w.ws, &got, &total)) return 0;
    void *col = d.alloc(5, (size_t) total * 4u);
    void *sval = nullptr;
    CSAMD_HIP(hipMalloc(&sval, (size_t) total * 2u + 4u), "hipMalloc sell val");
    csamd_sliced_ell_view v;
    std::memset(&v, 0, sizeof(v));
    v.rows = rows;
    v.cols = src->cols;
    v.slice_count = got;
    v.slice_rows = S;
    v.slice_row_offsets = off;
    v.slice_widths = wd;
    v.slice_slot_offsets = so;
    v.col_idx = (unsigned int *) col;
    v.val = (csamd_half *) sval;
    int ok = col != nullptr && csamd_sliced_ell_from_blocked_ell(&bv, CSAMD_VALUE_F16, &v, &w.ws) &&
             hip_ok(hipStreamSynchronize((hipStream_t) w.ws.stream), "sync");
    if (ok) {
        csamd_host_sliced_ell meta;
        std::memset(&meta, 0, sizeof(meta));
        meta.rows = rows;
        meta.cols = src->cols;
        meta.nnz = src->nnz;
        ok = sell_to_host(v, (size_t) total, &meta, dst);
    }
    (void) hipFree(sval);
    return ok;
}

} // extern "C"
