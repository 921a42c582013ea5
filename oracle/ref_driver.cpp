// Thin C driver over the reference's own CPU Blocked-ELL converter, compiled
// from the reference sources where they lie (see build_ref.sh).  Used only by
// tests to pin the oracle (oracle.c) and as an optional CPU baseline.
// Reference functions called: src/convert/blocked_ell_from_compressed.cuh
// choose_blocked_ell_block_size (:157-190) and blocked_ell_from_compressed
// (:258-347).
#include "src/convert/blocked_ell_from_compressed.cuh"

#include <cstdint>
#include <cstring>

namespace {
void make_csr(unsigned rows, unsigned cols, unsigned nnz, const unsigned *ptr, const unsigned *idx, const uint16_t *val,
              cellshard::sparse::compressed *m) {
    cellshard::sparse::init(m, rows, cols, nnz, cellshard::sparse::compressed_by_row);
    m->majorPtr = const_cast<unsigned *>(ptr);
    m->minorIdx = const_cast<unsigned *>(idx);
    m->val = reinterpret_cast<__half *>(const_cast<uint16_t *>(val));
}
} // namespace

extern "C" int ref_choose_blocked_ell_block_size(unsigned rows, unsigned cols, unsigned nnz, const unsigned *ptr,
                                                 const unsigned *idx, const uint16_t *val, const unsigned *cands,
                                                 unsigned n, unsigned *out_bs, double *out_fill,
                                                 unsigned long long *out_padded) {
    cellshard::sparse::compressed m;
    make_csr(rows, cols, nnz, ptr, idx, val, &m);
    cellshard::convert::blocked_ell_tune_result r{};
    if (!cellshard::convert::choose_blocked_ell_block_size(&m, cands, n, &r)) return 0;
    *out_bs = r.block_size;
    *out_fill = r.fill_ratio;
    *out_padded = (unsigned long long) r.padded_bytes;
    return 1;
}

// Converts and copies into the caller's buffers when they are large enough;
// always reports ell_cols and nnz.
extern "C" int ref_blocked_ell_from_compressed(unsigned rows, unsigned cols, unsigned nnz, const unsigned *ptr,
                                               const unsigned *idx, const uint16_t *val, unsigned block_size,
                                               unsigned *out_ell_cols, unsigned *out_nnz, unsigned *bci,
                                               unsigned long long bci_cap, uint16_t *dval,
                                               unsigned long long val_cap) {
    cellshard::sparse::compressed m;
    make_csr(rows, cols, nnz, ptr, idx, val, &m);
    cellshard::sparse::blocked_ell d;
    cellshard::sparse::init(&d);
    if (!cellshard::convert::blocked_ell_from_compressed(&m, block_size, &d)) return 0;
    *out_ell_cols = d.ell_cols;
    *out_nnz = d.nnz;
    const unsigned long long nb = (unsigned long long) cellshard::sparse::row_block_count(&d) *
                                  cellshard::sparse::ell_width_blocks(&d);
    const unsigned long long nv = (unsigned long long) d.rows * d.ell_cols;
    if (bci != nullptr && bci_cap >= nb && nb != 0) std::memcpy(bci, d.blockColIdx, nb * sizeof(unsigned));
    if (dval != nullptr && val_cap >= nv && nv != 0) std::memcpy(dval, d.val, nv * sizeof(uint16_t));
    cellshard::sparse::clear(&d);
    return 1;
}
