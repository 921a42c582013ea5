// Thin C driver over the reference's raw payload codec, compiled from the
// reference sources where they lie (see build_ref.sh).  Used only by tests to
// pin the CSPACK restatement (oracle.py) and as the CPU baseline of partition
// loading.  Reference functions called: src/io/pack/packfile.cu
// store_blocked_ell_raw (:329-369), load_blocked_ell_raw (:371-425),
// store_sliced_ell_raw (:574-621), load_sliced_ell_raw (:623-697) and
// load(FILE *, sparse::blocked_ell *) (:886-902).
#include "src/io/pack/packfile.cu"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sys/types.h>

extern "C" int ref_store_blocked_ell_raw(const char *path, unsigned rows, unsigned cols, unsigned nnz, unsigned block_size,
                                         unsigned ell_cols, const unsigned *bci, const void *val, size_t value_size) {
    return cellshard::store_blocked_ell_raw(path, rows, cols, nnz, block_size, ell_cols, bci, val, value_size);
}

extern "C" int ref_store_sliced_ell_raw(const char *path, unsigned rows, unsigned cols, unsigned nnz, unsigned slice_count,
                                        const unsigned *offsets, const unsigned *widths, const unsigned *col_idx, const void *val,
                                        size_t value_size) {
    return cellshard::store_sliced_ell_raw(path, rows, cols, nnz, slice_count, offsets, widths, col_idx, val, value_size);
}

// hdr: format, rows, cols, nnz, block_size, ell_cols.  Arrays are copied when
// the caller's buffers are large enough.
extern "C" int ref_load_blocked_ell_raw(const char *path, size_t value_size, unsigned *hdr, unsigned *bci,
                                        unsigned long long bci_cap, void *val, unsigned long long val_cap) {
    cellshard::blocked_ell_load_result r;
    std::memset(&r, 0, sizeof(r));
    if (!cellshard::load_blocked_ell_raw(path, value_size, &r)) return 0;
    hdr[0] = (unsigned) r.h.format;
    hdr[1] = r.h.rows;
    hdr[2] = r.h.cols;
    hdr[3] = r.h.nnz;
    hdr[4] = r.block_size;
    hdr[5] = r.ell_cols;
    const unsigned long long nb = r.block_size ? ((unsigned long long) (r.h.rows + r.block_size - 1u) / r.block_size) *
                                                     (r.ell_cols / r.block_size)
                                               : 0ull;
    const unsigned long long nv = (unsigned long long) r.h.rows * r.ell_cols;
    if (bci != nullptr && bci_cap >= nb && nb) std::memcpy(bci, r.blockColIdx, nb * 4u);
    if (val != nullptr && val_cap >= nv && nv) std::memcpy(val, r.val, nv * value_size);
    std::free(r.storage);
    return 1;
}

// hdr: format, rows, cols, nnz, slice_count, total_slots.
extern "C" int ref_load_sliced_ell_raw(const char *path, size_t value_size, unsigned *hdr, unsigned *offsets, unsigned *widths,
                                       unsigned *col_idx, void *val, unsigned long long slots_cap) {
    cellshard::sliced_ell_load_result r;
    std::memset(&r, 0, sizeof(r));
    if (!cellshard::load_sliced_ell_raw(path, value_size, &r)) return 0;
    unsigned long long slots = 0u;
    for (unsigned s = 0; s < r.slice_count; ++s)
        slots += (unsigned long long) (r.slice_row_offsets[s + 1u] - r.slice_row_offsets[s]) * r.slice_widths[s];
    hdr[0] = (unsigned) r.h.format;
    hdr[1] = r.h.rows;
    hdr[2] = r.h.cols;
    hdr[3] = r.h.nnz;
    hdr[4] = r.slice_count;
    hdr[5] = (unsigned) slots;
    if (offsets != nullptr) std::memcpy(offsets, r.slice_row_offsets, ((size_t) r.slice_count + 1u) * 4u);
    if (widths != nullptr && r.slice_count) std::memcpy(widths, r.slice_widths, (size_t) r.slice_count * 4u);
    if (col_idx != nullptr && slots_cap >= slots && slots) std::memcpy(col_idx, r.col_idx, slots * 4u);
    if (val != nullptr && slots_cap >= slots && slots) std::memcpy(val, r.val, slots * value_size);
    std::free(r.storage);
    return 1;
}

// One Blocked-ELL execution partition read the way the reference loads it
// (fread of the fixed fields and maps into fresh host arrays, then the
// reference's own load(FILE *, blocked_ell *) per segment), for the CPU
// baseline.  Returns the bytes read, or -1.
extern "C" long long ref_load_bell_partition(const char *path, unsigned long long offset) {
    std::FILE *fp = std::fopen(path, "rb");
    if (fp == nullptr) return -1;
    std::setvbuf(fp, 0, _IOFBF, (std::size_t) 8u << 20u);
    long long bytes = -1;
    unsigned head[4];
    if (fseeko(fp, (off_t) offset, SEEK_SET) == 0 && std::fread(head, 4, 4, fp) == 4) {
        const unsigned rows = head[0], cols = head[1], nseg = head[3];
        const size_t counts[5] = {(size_t) nseg + 1u, rows, rows, cols, cols};
        unsigned *arrays[5] = {};
        bool ok = true;
        long long total = 16;
        for (int i = 0; i < 5 && ok; ++i) {
            arrays[i] = (unsigned *) std::calloc(counts[i] ? counts[i] : 1u, 4u);
            ok = arrays[i] != nullptr && std::fread(arrays[i], 4, counts[i], fp) == counts[i];
            total += (long long) counts[i] * 4;
        }
        for (unsigned s = 0; ok && s < nseg; ++s) {
            cellshard::sparse::blocked_ell m;
            cellshard::sparse::init(&m);
            ok = cellshard::load(fp, &m) != 0;
            if (ok) {
                const unsigned long long nb = m.block_size ? ((unsigned long long) (m.rows + m.block_size - 1u) / m.block_size) *
                                                                 (m.ell_cols / m.block_size)
                                                           : 0ull;
                total += 13 + 8 + (long long) nb * 4 + (long long) m.rows * m.ell_cols * 2;
            }
            cellshard::sparse::clear(&m);
        }
        for (int i = 0; i < 5; ++i) std::free(arrays[i]);
        if (ok) bytes = total;
    }
    std::fclose(fp);
    return bytes;
}
