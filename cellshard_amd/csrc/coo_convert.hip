// COO -> compressed (row-major) and COO -> Blocked-ELL on the GPU.
//
// Reference: src/convert/blocked_ell_from_coo_cuda.cuh:514-966
// (evaluate_block_size, choose_blocked_ell_block_size_from_coo_cuda,
// blocked_ell_from_coo_cuda, blocked_ell_from_coo_cuda_auto): the COO is
// uploaded, columns optionally remapped through feature_to_global, entries
// sorted by (row block, column block), and a Blocked-ELL with ascending
// distinct column blocks per row block is scattered; nnz counts the stored
// nonzero values.
//
// Here the entries are sorted once by (row, column) with the stable LSD
// radix sort (so repeated (row, column) entries keep their input order), the
// sorted list is a CSR, and the CSR -> Blocked-ELL shape / tuner / emit
// kernels build the output: the same layout, and a repeated entry resolves
// to its last occurrence in input order (the reference's parallel scatter
// leaves that choice open).  Entries whose row or (remapped) column is out
// of range are rejected.
#include "common.hh"

#include <vector>

namespace csamd {
namespace {

// keys[i] = row << 32 | column (remapped), pos[i] = i; flags[0] |= 1 for an
// entry outside the matrix.
__global__ void coo_keys_kernel(unsigned int nnz, unsigned int rows, unsigned int cols, const unsigned int *__restrict__ row_idx,
                                const unsigned int *__restrict__ col_idx, const unsigned int *__restrict__ feature_map,
                                unsigned int map_len, unsigned long long *__restrict__ keys, unsigned int *__restrict__ pos,
                                unsigned int *__restrict__ flags) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
        const unsigned int r = row_idx[i];
        unsigned int c = col_idx[i];
        if (feature_map != nullptr) c = c < map_len ? feature_map[c] : CSAMD_INVALID_COL;
        if (r >= rows || c >= cols) {
            flags[0] = 1u;
            keys[i] = 0ull;
        } else {
            keys[i] = ((unsigned long long) r << 32) | c;
        }
        pos[i] = i;
    }
}

// Sorted entries -> CSR: ptr[r] = first entry of row r (rows with no entry
// take the next row's start), idx = column, val gathered from the input.
template<typename S>
__global__ void coo_to_csr_kernel(unsigned int nnz, unsigned int rows, const unsigned long long *__restrict__ keys,
                                  const unsigned int *__restrict__ pos, const S *__restrict__ val_in, unsigned int *__restrict__ ptr,
                                  unsigned int *__restrict__ idx, S *__restrict__ val) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
        const unsigned long long k = keys[i];
        const unsigned int r = (unsigned int) (k >> 32);
        const unsigned int prev = i == 0u ? 0u : (unsigned int) (keys[i - 1u] >> 32) + 1u;  // rows before r not yet started
        for (unsigned int rr = i == 0u ? 0u : prev; rr <= r; ++rr) ptr[rr] = i;
        if (i == nnz - 1u)
            for (unsigned int rr = r + 1u; rr <= rows; ++rr) ptr[rr] = nnz;
        idx[i] = (unsigned int) k;
        val[i] = val_in[pos[i]];
    }
}

__global__ void fill_u32_kernel(unsigned int *__restrict__ p, unsigned int n, unsigned int v) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

unsigned int bits_for(unsigned int n) {
    unsigned int b = 0u;
    while (b < 32u && (n >> b) != 0u) ++b;
    return b;
}

struct coo_scratch {
    size_t off_keys_in, off_keys_out, off_pos_in, off_pos_out, off_flags, off_sort, total;
};

coo_scratch coo_layout(unsigned int nnz) {
    coo_scratch L;
    size_t p = 0;
    L.off_flags = p;
    p = align_up(p + 256u, 256u);
    L.off_keys_in = p;
    p = align_up(p + (size_t) nnz * 8u, 256u);
    L.off_keys_out = p;
    p = align_up(p + (size_t) nnz * 8u, 256u);
    L.off_pos_in = p;
    p = align_up(p + (size_t) nnz * 4u, 256u);
    L.off_pos_out = p;
    p = align_up(p + (size_t) nnz * 4u, 256u);
    L.off_sort = p;
    p = align_up(p + radix_sort_pairs_u64_scratch_bytes(nnz), 256u);
    L.total = p;
    return L;
}

struct dev_mem {
    std::vector<void *> ptrs;
    ~dev_mem() {
        for (void *q : ptrs)
            if (q != nullptr) (void) hipFree(q);
    }
    void *alloc(size_t bytes, const char *what) {
        void *q = nullptr;
        if (!hip_ok(hipMalloc(&q, bytes == 0u ? 16u : bytes), what)) return nullptr;
        ptrs.push_back(q);
        return q;
    }
};

int coo_to_csr_device(unsigned int rows, unsigned int cols, unsigned int nnz, const unsigned int *d_row, const unsigned int *d_col,
                      const void *d_val, unsigned int value_type, const unsigned int *d_map, unsigned int map_len, unsigned int *d_ptr,
                      unsigned int *d_idx, void *d_out_val, void *scratch, size_t scratch_bytes, hipStream_t s) {
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "compressed_from_coo: unsupported value type %u", value_type);
    CSAMD_REQUIRE(d_ptr != nullptr, "compressed_from_coo: null output");
    if (nnz == 0u) {
        fill_u32_kernel<<<grid_for(rows + 1ull, 256u, 4096u), 256, 0, s>>>(d_ptr, rows + 1u, 0u);
        return hip_ok(hipGetLastError(), "fill_u32_kernel");
    }
    CSAMD_REQUIRE(d_row != nullptr && d_col != nullptr && d_val != nullptr && d_idx != nullptr && d_out_val != nullptr,
                  "compressed_from_coo: null input");
    const coo_scratch L = coo_layout(nnz);
    CSAMD_REQUIRE(scratch != nullptr && scratch_bytes >= L.total, "compressed_from_coo: scratch %zu < required %zu", scratch_bytes,
                  L.total);
    char *b = (char *) scratch;
    unsigned int *flags = (unsigned int *) (b + L.off_flags);
    unsigned long long *kin = (unsigned long long *) (b + L.off_keys_in), *kout = (unsigned long long *) (b + L.off_keys_out);
    unsigned int *pin = (unsigned int *) (b + L.off_pos_in), *pout = (unsigned int *) (b + L.off_pos_out);
    if (!zero_bytes(flags, 16u, s)) return 0;
    const unsigned int grid = grid_for(nnz, 256u, 8192u);
    coo_keys_kernel<<<grid, 256, 0, s>>>(nnz, rows, cols, d_row, d_col, d_map, map_len, kin, pin, flags);
    CSAMD_HIP(hipGetLastError(), "coo_keys_kernel");
    unsigned int hflag = 0u;
    CSAMD_HIP(hipMemcpyAsync(&hflag, flags, 4u, hipMemcpyDeviceToHost, s), "D2H coo flags");
    CSAMD_HIP(hipStreamSynchronize(s), "coo sync");
    CSAMD_REQUIRE(hflag == 0u, "compressed_from_coo: an entry lies outside the %u x %u matrix", rows, cols);
    const unsigned int key_bits = 32u + bits_for(rows == 0u ? 0u : rows - 1u);
    if (!radix_sort_pairs_u64(kin, kout, pin, pout, nnz, key_bits, b + L.off_sort, L.total - L.off_sort, s)) return 0;
    switch (vb) {
        case 1: coo_to_csr_kernel<uint8_t><<<grid, 256, 0, s>>>(nnz, rows, kout, pout, (const uint8_t *) d_val, d_ptr, d_idx, (uint8_t *) d_out_val); break;
        case 2: coo_to_csr_kernel<uint16_t><<<grid, 256, 0, s>>>(nnz, rows, kout, pout, (const uint16_t *) d_val, d_ptr, d_idx, (uint16_t *) d_out_val); break;
        default: coo_to_csr_kernel<uint32_t><<<grid, 256, 0, s>>>(nnz, rows, kout, pout, (const uint32_t *) d_val, d_ptr, d_idx, (uint32_t *) d_out_val); break;
    }
    return hip_ok(hipGetLastError(), "coo_to_csr_kernel");
}

struct scoped_ws {
    csamd_convert_workspace ws;
    bool ok = false;
    explicit scoped_ws(int device) {
        csamd_convert_init(&ws);
        ok = csamd_convert_setup(&ws, device, nullptr) == 1;
    }
    ~scoped_ws() { csamd_convert_clear(&ws); }
};

// Host COO (fp16 values) -> device CSR view, arrays owned by `mem`.
int stage_coo(const csamd_host_coo *src, unsigned int cols, const unsigned int *feature_to_global, dev_mem &mem,
              csamd_compressed_view *csr, hipStream_t s) {
    CSAMD_REQUIRE(src != nullptr, "blocked_ell_from_coo: null source");
    CSAMD_REQUIRE(src->nnz == 0u || (src->rowIdx != nullptr && src->colIdx != nullptr && src->val != nullptr),
                  "blocked_ell_from_coo: null source arrays");
    const unsigned int nnz = src->nnz, rows = src->rows;
    unsigned int *d_row = (unsigned int *) mem.alloc((size_t) nnz * 4u, "hipMalloc coo rows");
    unsigned int *d_col = (unsigned int *) mem.alloc((size_t) nnz * 4u, "hipMalloc coo cols");
    void *d_val = mem.alloc((size_t) nnz * 2u, "hipMalloc coo vals");
    unsigned int *d_ptr = (unsigned int *) mem.alloc((size_t) (rows + 1u) * 4u, "hipMalloc csr ptr");
    unsigned int *d_idx = (unsigned int *) mem.alloc((size_t) nnz * 4u, "hipMalloc csr idx");
    void *d_out = mem.alloc((size_t) nnz * 2u, "hipMalloc csr vals");
    unsigned int *d_map = nullptr;
    const coo_scratch L = coo_layout(nnz);
    void *scratch = mem.alloc(L.total, "hipMalloc coo scratch");
    if (d_row == nullptr || d_col == nullptr || d_val == nullptr || d_ptr == nullptr || d_idx == nullptr || d_out == nullptr ||
        scratch == nullptr)
        return 0;
    if (feature_to_global != nullptr && src->cols != 0u) {
        d_map = (unsigned int *) mem.alloc((size_t) src->cols * 4u, "hipMalloc feature map");
        if (d_map == nullptr) return 0;
        CSAMD_HIP(hipMemcpyAsync(d_map, feature_to_global, (size_t) src->cols * 4u, hipMemcpyHostToDevice, s), "H2D feature map");
    }
    if (nnz != 0u) {
        CSAMD_HIP(hipMemcpyAsync(d_row, src->rowIdx, (size_t) nnz * 4u, hipMemcpyHostToDevice, s), "H2D coo rows");
        CSAMD_HIP(hipMemcpyAsync(d_col, src->colIdx, (size_t) nnz * 4u, hipMemcpyHostToDevice, s), "H2D coo cols");
        CSAMD_HIP(hipMemcpyAsync(d_val, src->val, (size_t) nnz * 2u, hipMemcpyHostToDevice, s), "H2D coo vals");
    }
    if (!coo_to_csr_device(rows, cols, nnz, d_row, d_col, d_val, CSAMD_VALUE_F16, d_map, d_map != nullptr ? src->cols : 0u, d_ptr, d_idx,
                           d_out, scratch, L.total, s))
        return 0;
    std::memset(csr, 0, sizeof(*csr));
    csr->rows = rows;
    csr->cols = cols;
    csr->nnz = nnz;
    csr->axis = CSAMD_COMPRESSED_BY_ROW;
    csr->majorPtr = d_ptr;
    csr->minorIdx = d_idx;
    csr->val = (csamd_half *) d_out;
    return 1;
}

int bell_host_allocate(csamd_host_blocked_ell *m) {
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

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

size_t csamd_compressed_from_coo_scratch_bytes(unsigned int nnz) { return coo_layout(nnz).total; }

int csamd_compressed_from_coo_raw(unsigned int rows, unsigned int cols, unsigned int nnz, const unsigned int *d_row_idx,
                                  const unsigned int *d_col_idx, const void *d_val, unsigned int value_type,
                                  const unsigned int *d_feature_to_global, unsigned int feature_count, unsigned int *d_out_ptr,
                                  unsigned int *d_out_idx, void *d_out_val, void *d_scratch, size_t scratch_bytes,
                                  csamd_stream_t stream) {
    return coo_to_csr_device(rows, cols, nnz, d_row_idx, d_col_idx, d_val, value_type, d_feature_to_global, feature_count, d_out_ptr,
                             d_out_idx, d_out_val, d_scratch, scratch_bytes, (hipStream_t) stream);
}

int csamd_choose_blocked_ell_block_size_from_coo(const csamd_host_coo *src, unsigned int cols, const unsigned int *feature_to_global,
                                                 const unsigned int *candidates, unsigned int candidate_count, int device,
                                                 csamd_blocked_ell_tune_result *out) {
    CSAMD_REQUIRE(src != nullptr && out != nullptr && cols != 0u && candidates != nullptr && candidate_count != 0u,
                  "choose_blocked_ell_block_size_from_coo: bad arguments");
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice coo");
    scoped_ws w(device);
    CSAMD_REQUIRE(w.ok, "choose_blocked_ell_block_size_from_coo: workspace setup failed");
    dev_mem mem;
    csamd_compressed_view csr;
    if (!stage_coo(src, cols, feature_to_global, mem, &csr, (hipStream_t) w.ws.stream)) return 0;
    return csamd_choose_blocked_ell_block_size(&csr, candidates, candidate_count, 2u, &w.ws, out);
}

int csamd_blocked_ell_from_coo(const csamd_host_coo *src, unsigned int cols, const unsigned int *feature_to_global,
                               unsigned int block_size, csamd_host_blocked_ell *dst, int device) {
    CSAMD_REQUIRE(src != nullptr && dst != nullptr && cols != 0u && block_size != 0u, "blocked_ell_from_coo: bad arguments");
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice coo");
    scoped_ws w(device);
    CSAMD_REQUIRE(w.ok, "blocked_ell_from_coo: workspace setup failed");
    dev_mem mem;
    csamd_compressed_view csr;
    if (!stage_coo(src, cols, feature_to_global, mem, &csr, (hipStream_t) w.ws.stream)) return 0;
    unsigned int width = 0u;
    unsigned long long total = 0ull;
    double fill = 0.0;
    if (!csamd_blocked_ell_shape(&csr, block_size, &w.ws, &width, &total, &fill)) return 0;
    const size_t rb = (src->rows + block_size - 1u) / block_size;
    csamd_blocked_ell_view bv;
    std::memset(&bv, 0, sizeof(bv));
    bv.rows = src->rows;
    bv.cols = cols;
    bv.block_size = block_size;
    bv.ell_cols = width * block_size;
    bv.blockColIdx = (unsigned int *) mem.alloc(rb * width * 4u, "hipMalloc bell idx");
    bv.val = (csamd_half *) mem.alloc((size_t) src->rows * width * block_size * 2u, "hipMalloc bell vals");
    if (bv.blockColIdx == nullptr || bv.val == nullptr) return 0;
    if (!csamd_blocked_ell_from_compressed(&csr, CSAMD_VALUE_F16, &bv, &w.ws)) return 0;
    CSAMD_HIP(hipStreamSynchronize((hipStream_t) w.ws.stream), "bell sync");
    csamd_host_blocked_ell_clear(dst);
    dst->rows = src->rows;
    dst->cols = cols;
    dst->nnz = bv.nnz;
    dst->block_size = block_size;
    dst->ell_cols = width * block_size;
    if (!bell_host_allocate(dst)) return 0;
    if (dst->blockColIdx != nullptr) CSAMD_HIP(hipMemcpy(dst->blockColIdx, bv.blockColIdx, rb * width * 4u, hipMemcpyDeviceToHost), "D2H bci");
    if (dst->val != nullptr) CSAMD_HIP(hipMemcpy(dst->val, bv.val, (size_t) src->rows * dst->ell_cols * 2u, hipMemcpyDeviceToHost), "D2H val");
    return 1;
}

int csamd_blocked_ell_from_coo_auto(const csamd_host_coo *src, unsigned int cols, const unsigned int *feature_to_global,
                                    const unsigned int *candidates, unsigned int candidate_count, csamd_host_blocked_ell *dst,
                                    int device, csamd_blocked_ell_tune_result *picked) {
    csamd_blocked_ell_tune_result t;
    std::memset(&t, 0, sizeof(t));
    if (!csamd_choose_blocked_ell_block_size_from_coo(src, cols, feature_to_global, candidates, candidate_count, device, &t)) return 0;
    if (!csamd_blocked_ell_from_coo(src, cols, feature_to_global, t.block_size, dst, device)) return 0;
    if (picked != nullptr) *picked = t;
    return 1;
}

} // extern "C"
