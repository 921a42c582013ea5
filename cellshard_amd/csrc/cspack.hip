// CSPACK01 shard packs of bucketed execution partitions: the writer, the
// container header checks, and staging of a partition straight into device
// memory.
//
// Staging: the partition's byte range is read by parallel pread() workers into
// pinned host memory; each piece is queued as a host-to-device copy into a
// device staging buffer as soon as it lands, so the copies run under the
// reads.  The blob's arrays sit at arbitrary byte offsets (the raw segment
// header is 13 bytes, values may be 1 byte wide), so one unpack kernel then
// moves every array to a 256-byte aligned place in the partition's single
// device allocation, 16 bytes per thread with funnel shifts.  Headers are
// parsed and checked on the host, from the pinned copy, before any device
// work is queued for the partition.
//
// References: docs/SPEC_CSPACK_V1.md:87-155 (container, blob layouts);
// src/io/csh5/execution_builders_part.hh:2116-2132, :2146-2197 (Blocked-ELL
// blob); src/io/csh5/execution_runtime_part.hh:168-258 (Sliced-ELL blob),
// :569-641 (partition loads), :643-796 (materialize); src/io/pack/packfile.cu
// :13-45 (raw header), :351-369 and :390-425 (raw Blocked-ELL), :597-621 and
// :643-697 (raw Sliced-ELL); include/CellShard/io/common/raw_format.hh:10-31.
#include "common.hh"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/stat.h>
#include <sys/uio.h>
#include <unistd.h>

namespace csamd {
namespace {

const unsigned char kMagic[8] = {'C', 'S', 'P', 'A', 'C', 'K', '0', '1'};
constexpr unsigned char kFormatBlockedEll = 6u;  // raw_format.hh:10-24
constexpr unsigned char kFormatSlicedEll = 7u;
constexpr uint64_t kContainerHeader = 24u;       // magic, shard_id, partition_count
constexpr uint64_t kRawHeader = 13u;             // format u8, rows, cols, nnz (packfile.cu:13-15)
constexpr size_t kAlign = 256u;                  // device sub-array alignment
constexpr size_t kPad = 64u;                     // readable bytes past every staged source
constexpr unsigned int kUnpackThreads = 256u;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// One array moved from its place in a staged blob to its aligned place in
// the partition's allocation.  `first` is the move's first 16-byte group in
// the launch; moves are ordered by it and none is empty.
struct move_desc {
    unsigned long long src;
    unsigned long long dst;
    unsigned long long bytes;
    unsigned long long first;
};

// ---------------------------------------------------------------------------
// Unpack kernel.  Thread g writes 16 bytes of one move: it reads the five
// aligned dwords covering source bytes [s, s + 16) and shifts them into
// place.  Sources have kPad readable bytes past their end; bytes past a
// move's end are written as zero (every destination is padded to kAlign).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kUnpackThreads) cs_unpack_kernel(const move_desc *__restrict__ moves, unsigned int count,
                                                                   unsigned long long groups) {
    const unsigned long long stride = (unsigned long long) gridDim.x * blockDim.x;
    for (unsigned long long g = (unsigned long long) blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
        unsigned int lo = 0u, hi = count - 1u;
        while (lo < hi) {
            const unsigned int mid = (lo + hi + 1u) >> 1;
            if (moves[mid].first <= g) lo = mid;
            else hi = mid - 1u;
        }
        const move_desc m = moves[lo];
        const unsigned long long k = (g - m.first) << 4;
        const unsigned long long s = m.src + k;
        const unsigned int *w = (const unsigned int *) (s & ~3ull);
        const unsigned int sh = (unsigned int) (s & 3ull) * 8u;
        unsigned int a[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) a[i] = w[i];
        unsigned int o[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = (unsigned int) (((((unsigned long long) a[i + 1]) << 32) | a[i]) >> sh);
        const unsigned long long left = m.bytes - k;
        if (left < 16ull) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long long keep = (long long) left - 4ll * i;
                if (keep <= 0) o[i] = 0u;
                else if (keep < 4) o[i] &= (1u << (8u * (unsigned int) keep)) - 1u;
            }
        }
        *(u32x4 *) (m.dst + k) = u32x4{o[0], o[1], o[2], o[3]};
    }
}

// ---------------------------------------------------------------------------
// Host helpers.
// ---------------------------------------------------------------------------
inline uint32_t rd32(const unsigned char *p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

inline uint64_t rd64(const unsigned char *p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}

// Reads exactly n bytes at off; 0 on success, else an errno value (EIO for
// an early end of file).
int pread_all(int fd, unsigned char *dst, uint64_t n, uint64_t off) {
    while (n != 0u) {
        const size_t ask = (size_t) std::min<uint64_t>(n, 1ull << 30);
        const ssize_t r = ::pread(fd, dst, ask, (off_t) off);
        if (r < 0) {
            if (errno == EINTR) continue;
            return errno;
        }
        if (r == 0) return EIO;
        dst += r;
        off += (uint64_t) r;
        n -= (uint64_t) r;
    }
    return 0;
}

int write_all(int fd, const void *p, size_t n) {
    const unsigned char *q = (const unsigned char *) p;
    while (n != 0u) {
        const ssize_t r = ::write(fd, q, std::min<size_t>(n, 1u << 30));
        if (r < 0) {
            if (errno == EINTR) continue;
            return errno;
        }
        q += r;
        n -= (size_t) r;
    }
    return 0;
}

// Bounds-checked cursor over one blob in host memory.
struct blob_reader {
    const unsigned char *b;
    uint64_t len;
    uint64_t pos;
    const char *fail;  // first failure

    int need(uint64_t n, const char *what) {
        if (fail == nullptr && n > len - pos) fail = what;
        return fail == nullptr;
    }
    uint32_t u32(const char *what) {
        if (!need(4u, what)) return 0u;
        const uint32_t v = rd32(b + pos);
        pos += 4u;
        return v;
    }
    unsigned char u8(const char *what) {
        if (!need(1u, what)) return 0u;
        return b[pos++];
    }
    // n elements of e bytes; returns their offset in the blob.
    uint64_t array(uint64_t n, uint64_t e, const char *what) {
        uint64_t bytes = 0u;
        if (fail == nullptr && __builtin_mul_overflow(n, e, &bytes)) fail = what;
        if (!need(bytes, what)) return 0u;
        const uint64_t at = pos;
        pos += bytes;
        return at;
    }
};

struct bell_segment {
    uint32_t rows, cols, nnz, block_size, ell_cols;
    uint64_t bci, bci_bytes, val, val_bytes;
};

struct bell_blob {
    uint32_t rows, cols, nnz, nseg;
    uint64_t sro, e2cr, c2er, e2cc, c2ec;
    std::vector<bell_segment> seg;
};

struct sell_segment {
    uint32_t rows, cols, nnz, slice_count, slice_rows;
    uint64_t offsets, widths, col, col_bytes, val, val_bytes;
    std::vector<uint32_t> slot_offsets;
};

struct sell_blob {
    uint32_t rows, cols, nnz, nseg, canonical_slices;
    uint64_t sro, e2cr, c2er, cso, csw;
    std::vector<sell_segment> seg;
};

// Segment row offsets start at 0, never decrease and end at `rows`; segment
// k holds rows [off[k], off[k+1]).
int check_segment_offsets(const unsigned char *sro, uint32_t nseg, uint32_t rows, unsigned long long part) {
    CSAMD_REQUIRE(rd32(sro) == 0u, "cspack partition %llu: segment row offsets do not start at 0", part);
    for (uint32_t k = 0; k < nseg; ++k)
        CSAMD_REQUIRE(rd32(sro + 4u * (k + 1u)) >= rd32(sro + 4u * k), "cspack partition %llu: segment row offsets decrease", part);
    CSAMD_REQUIRE(rd32(sro + 4ull * nseg) == rows, "cspack partition %llu: segment row offsets end at %u, not at rows %u", part,
                  rd32(sro + 4ull * nseg), rows);
    return 1;
}

// load_execution_partition_blob (execution_builders_part.hh:2146-2197) over
// bytes in memory, with the raw Blocked-ELL payload of packfile.cu:390-425
// per segment.  Beyond the reference's reads: the format tag, each
// segment's rows against the segment row offsets, its cols against the
// partition's, and ell_cols a multiple of a non-zero block size.
int parse_bell(const unsigned char *b, uint64_t len, unsigned int vb, unsigned long long part, bell_blob &o) {
    blob_reader r{b, len, 0u, nullptr};
    o.rows = r.u32("header");
    o.cols = r.u32("header");
    o.nnz = r.u32("header");
    o.nseg = r.u32("header");
    o.sro = r.array((uint64_t) o.nseg + 1u, 4u, "segment row offsets");
    o.e2cr = r.array(o.rows, 4u, "exec_to_canonical_rows");
    o.c2er = r.array(o.rows, 4u, "canonical_to_exec_rows");
    o.e2cc = r.array(o.cols, 4u, "exec_to_canonical_cols");
    o.c2ec = r.array(o.cols, 4u, "canonical_to_exec_cols");
    CSAMD_REQUIRE(r.fail == nullptr, "cspack partition %llu: blob truncated in %s", part, r.fail);
    if (!check_segment_offsets(b + o.sro, o.nseg, o.rows, part)) return 0;
    CSAMD_REQUIRE((uint64_t) o.nseg <= (len - r.pos) / (kRawHeader + 8u), "cspack partition %llu: blob truncated in segment headers",
                  part);
    o.seg.assign(o.nseg, bell_segment{});
    for (uint32_t k = 0; k < o.nseg; ++k) {
        bell_segment &s = o.seg[k];
        const unsigned char fmt = r.u8("segment header");
        s.rows = r.u32("segment header");
        s.cols = r.u32("segment header");
        s.nnz = r.u32("segment header");
        s.block_size = r.u32("segment header");
        s.ell_cols = r.u32("segment header");
        CSAMD_REQUIRE(r.fail == nullptr, "cspack partition %llu segment %u: blob truncated in %s", part, k, r.fail);
        CSAMD_REQUIRE(fmt == kFormatBlockedEll, "cspack partition %llu segment %u: format %u is not blocked_ell (%u)", part, k,
                      (unsigned int) fmt, (unsigned int) kFormatBlockedEll);
        CSAMD_REQUIRE(s.rows == rd32(b + o.sro + 4u * (k + 1u)) - rd32(b + o.sro + 4u * k),
                      "cspack partition %llu segment %u: %u rows, segment row offsets say %u", part, k, s.rows,
                      rd32(b + o.sro + 4u * (k + 1u)) - rd32(b + o.sro + 4u * k));
        CSAMD_REQUIRE(s.cols == o.cols, "cspack partition %llu segment %u: %u cols, partition has %u", part, k, s.cols, o.cols);
        CSAMD_REQUIRE(s.block_size != 0u || (s.rows == 0u && s.ell_cols == 0u), "cspack partition %llu segment %u: block_size 0", part, k);
        CSAMD_REQUIRE(s.block_size == 0u || s.ell_cols % s.block_size == 0u,
                      "cspack partition %llu segment %u: ell_cols %u is not a multiple of block_size %u", part, k, s.ell_cols,
                      s.block_size);
        const uint64_t rb = s.block_size != 0u ? ((uint64_t) s.rows + s.block_size - 1u) / s.block_size : 0u;
        const uint64_t w = s.block_size != 0u ? (uint64_t) s.ell_cols / s.block_size : 0u;
        s.bci = r.array(rb * w, 4u, "block column indices");
        s.bci_bytes = rb * w * 4u;
        s.val = r.array((uint64_t) s.rows * s.ell_cols, vb, "values");
        s.val_bytes = (uint64_t) s.rows * s.ell_cols * vb;
        CSAMD_REQUIRE(r.fail == nullptr, "cspack partition %llu segment %u: blob truncated in %s", part, k, r.fail);
    }
    return 1;
}

// load_sliced_execution_partition_blob (execution_runtime_part.hh:197-258)
// with the raw Sliced-ELL payload of packfile.cu:643-697 per segment (which
// requires slice_row_offsets[slice_count] == rows).  Beyond the reference:
// the format tag, slice row offsets from 0 and non-decreasing, a segment
// with rows has slices, rows against the segment row offsets, cols against
// the partition's, and fewer than 2^32 slots.  Derives each segment's slot
// offsets and its uniform slice height (0 when the heights differ).
int parse_sell(const unsigned char *b, uint64_t len, unsigned int vb, unsigned long long part, sell_blob &o) {
    blob_reader r{b, len, 0u, nullptr};
    o.rows = r.u32("header");
    o.cols = r.u32("header");
    o.nnz = r.u32("header");
    o.nseg = r.u32("header");
    o.canonical_slices = r.u32("header");
    o.sro = r.array((uint64_t) o.nseg + 1u, 4u, "segment row offsets");
    o.e2cr = r.array(o.rows, 4u, "exec_to_canonical_rows");
    o.c2er = r.array(o.rows, 4u, "canonical_to_exec_rows");
    o.cso = r.array((uint64_t) o.canonical_slices + 1u, 4u, "canonical slice row offsets");
    o.csw = r.array(o.canonical_slices, 4u, "canonical slice widths");
    CSAMD_REQUIRE(r.fail == nullptr, "cspack partition %llu: blob truncated in %s", part, r.fail);
    if (!check_segment_offsets(b + o.sro, o.nseg, o.rows, part)) return 0;
    CSAMD_REQUIRE((uint64_t) o.nseg <= (len - r.pos) / (kRawHeader + 8u), "cspack partition %llu: blob truncated in segment headers",
                  part);
    o.seg.assign(o.nseg, sell_segment{});
    for (uint32_t k = 0; k < o.nseg; ++k) {
        sell_segment &s = o.seg[k];
        const unsigned char fmt = r.u8("segment header");
        s.rows = r.u32("segment header");
        s.cols = r.u32("segment header");
        s.nnz = r.u32("segment header");
        s.slice_count = r.u32("segment header");
        s.offsets = r.array((uint64_t) s.slice_count + 1u, 4u, "slice row offsets");
        s.widths = r.array(s.slice_count, 4u, "slice widths");
        CSAMD_REQUIRE(r.fail == nullptr, "cspack partition %llu segment %u: blob truncated in %s", part, k, r.fail);
        CSAMD_REQUIRE(fmt == kFormatSlicedEll, "cspack partition %llu segment %u: format %u is not sliced_ell (%u)", part, k,
                      (unsigned int) fmt, (unsigned int) kFormatSlicedEll);
        CSAMD_REQUIRE(s.rows == rd32(b + o.sro + 4u * (k + 1u)) - rd32(b + o.sro + 4u * k),
                      "cspack partition %llu segment %u: %u rows, segment row offsets say %u", part, k, s.rows,
                      rd32(b + o.sro + 4u * (k + 1u)) - rd32(b + o.sro + 4u * k));
        CSAMD_REQUIRE(s.cols == o.cols, "cspack partition %llu segment %u: %u cols, partition has %u", part, k, s.cols, o.cols);
        CSAMD_REQUIRE(s.slice_count != 0u || s.rows == 0u, "cspack partition %llu segment %u: %u rows in no slice", part, k, s.rows);
        uint64_t slots = 0u;
        uint32_t height = 0u;
        bool uniform = true;
        s.slot_offsets.assign(s.slice_count, 0u);
        if (s.slice_count != 0u) {
            const unsigned char *off = b + s.offsets, *wid = b + s.widths;
            CSAMD_REQUIRE(rd32(off) == 0u, "cspack partition %llu segment %u: slice row offsets do not start at 0", part, k);
            CSAMD_REQUIRE(rd32(off + 4ull * s.slice_count) == s.rows,
                          "cspack partition %llu segment %u: slice row offsets end at %u, not at rows %u", part, k,
                          rd32(off + 4ull * s.slice_count), s.rows);
            height = rd32(off + 4u) - rd32(off);
            for (uint32_t q = 0; q < s.slice_count; ++q) {
                const uint32_t a = rd32(off + 4ull * q), e = rd32(off + 4ull * (q + 1u));
                CSAMD_REQUIRE(e >= a, "cspack partition %llu segment %u: slice row offsets decrease", part, k);
                CSAMD_REQUIRE(slots <= 0xffffffffull, "cspack partition %llu segment %u: 2^32 slots or more", part, k);
                s.slot_offsets[q] = (uint32_t) slots;
                slots += (uint64_t) (e - a) * rd32(wid + 4ull * q);
                if (e - a != height && !(q + 1u == s.slice_count && e - a <= height)) uniform = false;
            }
            CSAMD_REQUIRE(slots <= 0xffffffffull, "cspack partition %llu segment %u: 2^32 slots or more", part, k);
        }
        s.slice_rows = uniform ? height : 0u;
        s.col = r.array(slots, 4u, "column indices");
        s.col_bytes = slots * 4u;
        s.val = r.array(slots, vb, "values");
        s.val_bytes = slots * vb;
        CSAMD_REQUIRE(r.fail == nullptr, "cspack partition %llu segment %u: blob truncated in %s", part, k, r.fail);
    }
    return 1;
}

// ---------------------------------------------------------------------------
// Workspace private state: two staging slots (pinned blob, pinned side
// table, device blob, device side table, completion event) so that the read
// of partition p + 1 overlaps the copies and unpack of p, and the writer's
// pinned buffer.
// ---------------------------------------------------------------------------
struct stage_slot {
    unsigned char *pinned = nullptr;
    size_t pinned_cap = 0u;
    unsigned char *pinned_x = nullptr;
    size_t pinned_x_cap = 0u;
    unsigned char *dev = nullptr;
    size_t dev_cap = 0u;
    unsigned char *dev_x = nullptr;
    size_t dev_x_cap = 0u;
    hipEvent_t done = nullptr;
    bool busy = false;
};

struct cs_private {
    stage_slot slot[2];
    unsigned char *wpinned = nullptr;
    size_t wpinned_cap = 0u;
};

struct cs_file_impl {
    int fd;
};

size_t grown(size_t need) { return align_up(need + need / 4u, 1u << 20); }

int grow_pinned(unsigned char **p, size_t *cap, size_t need) {
    if (*cap >= need) return 1;
    if (*p != nullptr) (void) hipHostFree(*p);
    *p = nullptr;
    *cap = 0u;
    const size_t n = grown(need);
    CSAMD_HIP(hipHostMalloc((void **) p, n, hipHostMallocDefault), "hipHostMalloc cspack staging");
    *cap = n;
    return 1;
}

int grow_device(unsigned char **p, size_t *cap, size_t need) {
    if (*cap >= need) return 1;
    if (*p != nullptr) (void) hipFree(*p);
    *p = nullptr;
    *cap = 0u;
    const size_t n = grown(need);
    CSAMD_HIP(hipMalloc((void **) p, n), "hipMalloc cspack staging");
    *cap = n;
    return 1;
}

int wait_slot(stage_slot &s) {
    if (s.busy) {
        s.busy = false;
        CSAMD_HIP(hipEventSynchronize(s.done), "cspack staging wait");
    }
    return 1;
}

void free_private(void *p) {
    cs_private *q = (cs_private *) p;
    if (q == nullptr) return;
    for (stage_slot &s : q->slot) {
        if (s.busy && s.done != nullptr) (void) hipEventSynchronize(s.done);
        if (s.pinned != nullptr) (void) hipHostFree(s.pinned);
        if (s.pinned_x != nullptr) (void) hipHostFree(s.pinned_x);
        if (s.dev != nullptr) (void) hipFree(s.dev);
        if (s.dev_x != nullptr) (void) hipFree(s.dev_x);
        if (s.done != nullptr) (void) hipEventDestroy(s.done);
    }
    if (q->wpinned != nullptr) (void) hipHostFree(q->wpinned);
    delete q;
}

int check_ws(csamd_cspack_workspace *ws, const char *what) {
    CSAMD_REQUIRE(ws != nullptr && ws->device >= 0, "%s: workspace not set up", what);
    CSAMD_HIP(hipSetDevice(ws->device), "hipSetDevice cspack");
    if (ws->private_workspace == nullptr) {
        cs_private *q = new cs_private();
        for (stage_slot &s : q->slot) {
            if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
                free_private(q);
                set_error("%s: hipEventCreate failed", what);
                return 0;
            }
        }
        ws->private_workspace = q;
    }
    return 1;
}

// Reads bytes [file_off, file_off + bytes) into `pinned` with up to `threads`
// workers taking `chunk`-sized pieces in turn; each piece is queued as a copy
// to dev + its offset on `s` as soon as it is read.
int read_and_copy(int fd, uint64_t file_off, uint64_t bytes, unsigned char *pinned, unsigned char *dev, size_t chunk,
                  unsigned int threads, int device, hipStream_t s, unsigned long long part) {
    if (bytes == 0u) return 1;
    const uint64_t pieces = (bytes + chunk - 1u) / chunk;
    std::atomic<uint64_t> next{0u};
    std::atomic<int> failed{0};
    std::mutex mu;
    std::string why;
    auto fail = [&](const std::string &m) {
        std::lock_guard<std::mutex> g(mu);
        if (why.empty()) why = m;
        failed.store(1);
    };
    auto work = [&](bool own_thread) {
        if (own_thread && hipSetDevice(device) != hipSuccess) {
            fail("hipSetDevice in a cspack reader");
            return;
        }
        for (;;) {
            const uint64_t i = next.fetch_add(1u);
            if (i >= pieces || failed.load() != 0) break;
            const uint64_t o = i * chunk, n = std::min<uint64_t>(chunk, bytes - o);
            const int e = pread_all(fd, pinned + o, n, file_off + o);
            if (e != 0) {
                fail(std::string("read failed: ") + std::strerror(e));
                break;
            }
            const hipError_t h = hipMemcpyAsync(dev + o, pinned + o, n, hipMemcpyHostToDevice, s);
            if (h != hipSuccess) {
                fail(std::string("hipMemcpyAsync: ") + hipGetErrorString(h));
                break;
            }
        }
    };
    const unsigned int T = (unsigned int) std::min<uint64_t>(std::max(1u, threads), pieces);
    if (T <= 1u) {
        work(false);
    } else {
        std::vector<std::thread> pool;
        pool.reserve(T - 1u);
        for (unsigned int t = 1; t < T; ++t) pool.emplace_back(work, true);
        work(false);
        for (std::thread &t : pool) t.join();
    }
    CSAMD_REQUIRE(failed.load() == 0, "cspack partition %llu: %s", part, why.c_str());
    return 1;
}

// Device layout of one partition: each array at a kAlign boundary, padded
// to a multiple of kAlign (at least 16 bytes).
struct dev_layout {
    size_t total = 0u;
    size_t place(uint64_t bytes) {
        const size_t at = total;
        total += align_up((size_t) std::max<uint64_t>(bytes, 16u), kAlign);
        return at;
    }
};

// Side table of one staging step, built on the host: side arrays (derived
// on the host), then the moves.  Sources inside the blob are given as blob
// offsets and sources in the side table as side offsets until the device
// addresses are known.
struct side_table {
    std::vector<unsigned char> bytes;
    struct mv {
        bool side;
        uint64_t src, dst_off, n;
    };
    std::vector<mv> moves;

    uint64_t add(const void *p, size_t n) {
        const uint64_t at = align_up(bytes.size(), 16u);
        bytes.resize(at + n);
        if (n != 0u) std::memcpy(bytes.data() + at, p, n);
        return at;
    }
    void move_blob(uint64_t blob_off, uint64_t dst_off, uint64_t n) {
        if (n != 0u) moves.push_back(mv{false, blob_off, dst_off, n});
    }
    void move_side(uint64_t side_off, uint64_t dst_off, uint64_t n) {
        if (n != 0u) moves.push_back(mv{true, side_off, dst_off, n});
    }
};

// Copies the side table to the slot's device side buffer and queues the
// unpack of every move into `base`.  Records the slot's completion event.
int launch_unpack(stage_slot &sl, side_table &t, unsigned char *base, hipStream_t s, unsigned long long part) {
    const size_t moves_at = align_up(t.bytes.size(), 16u);
    const size_t xbytes = moves_at + t.moves.size() * sizeof(move_desc);
    if (!grow_pinned(&sl.pinned_x, &sl.pinned_x_cap, xbytes + kPad)) return 0;
    if (!grow_device(&sl.dev_x, &sl.dev_x_cap, xbytes + kPad)) return 0;
    if (!t.bytes.empty()) std::memcpy(sl.pinned_x, t.bytes.data(), t.bytes.size());
    move_desc *md = (move_desc *) (sl.pinned_x + moves_at);
    unsigned long long groups = 0u;
    for (size_t i = 0; i < t.moves.size(); ++i) {
        const side_table::mv &m = t.moves[i];
        md[i].src = (unsigned long long) (uintptr_t) ((m.side ? sl.dev_x : sl.dev) + m.src);
        md[i].dst = (unsigned long long) (uintptr_t) (base + m.dst_off);
        md[i].bytes = m.n;
        md[i].first = groups;
        groups += (m.n + 15u) / 16u;
    }
    CSAMD_HIP(hipMemcpyAsync(sl.dev_x, sl.pinned_x, xbytes, hipMemcpyHostToDevice, s), "H2D cspack side table");
    if (groups != 0u) {
        const unsigned int grid = grid_for(groups, kUnpackThreads, (unsigned int) std::max(1, device_cu_count()) * 16u);
        cs_unpack_kernel<<<grid, kUnpackThreads, 0, s>>>((const move_desc *) (sl.dev_x + moves_at), (unsigned int) t.moves.size(),
                                                         groups);
        CSAMD_HIP(hipGetLastError(), "cs_unpack_kernel");
    }
    CSAMD_HIP(hipEventRecord(sl.done, s), "cspack staging event");
    sl.busy = true;
    (void) part;
    return 1;
}

// Reads partition `p` of `f` into slot `sl` (pinned + device staging).
// Returns the blob length read (the partition's byte range) in *len.
int stage_read(const csamd_cspack_file *f, unsigned long long p, csamd_cspack_workspace *ws, stage_slot &sl, uint64_t *len) {
    const uint64_t n = f->partition_bytes[p];
    if (!wait_slot(sl)) return 0;
    if (!grow_pinned(&sl.pinned, &sl.pinned_cap, n + kPad)) return 0;
    if (!grow_device(&sl.dev, &sl.dev_cap, n + kPad)) return 0;
    const size_t chunk = ws->chunk_bytes != 0u ? ws->chunk_bytes : (size_t) 8u << 20;
    const unsigned int threads = ws->reader_threads != 0u ? std::min(ws->reader_threads, 64u) : 8u;
    const cs_file_impl *fi = (const cs_file_impl *) f->impl;
    if (!read_and_copy(fi->fd, f->partition_offsets[p], n, sl.pinned, sl.dev, chunk, threads, ws->device, (hipStream_t) ws->stream, p))
        return 0;
    *len = n;
    return 1;
}

int check_stage_args(const csamd_cspack_file *f, unsigned long long p, unsigned int value_type, csamd_cspack_workspace *ws,
                     const void *out, const char *what) {
    CSAMD_REQUIRE(f != nullptr && f->impl != nullptr && out != nullptr, "%s: null argument", what);
    CSAMD_REQUIRE(p < f->partition_count, "%s: partition %llu of %llu", what, p, f->partition_count);
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "%s: unsupported value type %u", what, value_type);
    return check_ws(ws, what);
}

// Blocked-ELL: read, parse, allocate, unpack (no wait at the end).
int stage_bell_one(const csamd_cspack_file *f, unsigned long long p, unsigned int value_type, csamd_cspack_workspace *ws,
                   stage_slot &sl, csamd_bucketed_blocked_ell *out) {
    const unsigned int vb = (unsigned int) value_type_bytes(value_type);
    hipStream_t s = (hipStream_t) ws->stream;
    uint64_t len = 0u;
    csamd_bucketed_blocked_ell_clear(out);
    if (!stage_read(f, p, ws, sl, &len)) return 0;
    bell_blob B;
    if (!parse_bell(sl.pinned, len, vb, p, B)) return 0;
    dev_layout L;
    const size_t o_e2cr = L.place(4ull * B.rows), o_c2er = L.place(4ull * B.rows);
    const size_t o_e2cc = L.place(4ull * B.cols), o_c2ec = L.place(4ull * B.cols);
    std::vector<size_t> o_bci(B.nseg), o_val(B.nseg);
    for (uint32_t k = 0; k < B.nseg; ++k) {
        o_bci[k] = L.place(B.seg[k].bci_bytes);
        o_val[k] = L.place(B.seg[k].val_bytes);
    }
    void *st = exec_storage_new();
    out->storage = st;
    unsigned char *base = (unsigned char *) exec_storage_dev(st, L.total);
    out->segments = (csamd_blocked_ell_view *) exec_storage_host(st, (size_t) B.nseg * sizeof(csamd_blocked_ell_view));
    out->segment_row_offsets = (unsigned int *) exec_storage_host(st, ((size_t) B.nseg + 1u) * 4u);
    CSAMD_REQUIRE(base != nullptr && out->segments != nullptr && out->segment_row_offsets != nullptr,
                  "cspack partition %llu: allocation of %zu device bytes failed", p, L.total);
    out->rows = B.rows;
    out->cols = B.cols;
    out->nnz = B.nnz;
    out->segment_count = B.nseg;
    std::memcpy(out->segment_row_offsets, sl.pinned + B.sro, ((size_t) B.nseg + 1u) * 4u);
    out->exec_to_canonical_rows = (unsigned int *) (base + o_e2cr);
    out->canonical_to_exec_rows = (unsigned int *) (base + o_c2er);
    out->exec_to_canonical_cols = (unsigned int *) (base + o_e2cc);
    out->canonical_to_exec_cols = (unsigned int *) (base + o_c2ec);
    side_table t;
    t.move_blob(B.e2cr, o_e2cr, 4ull * B.rows);
    t.move_blob(B.c2er, o_c2er, 4ull * B.rows);
    t.move_blob(B.e2cc, o_e2cc, 4ull * B.cols);
    t.move_blob(B.c2ec, o_c2ec, 4ull * B.cols);
    unsigned long long bytes = 0u;
    for (uint32_t k = 0; k < B.nseg; ++k) {
        const bell_segment &g = B.seg[k];
        csamd_blocked_ell_view &v = out->segments[k];
        std::memset(&v, 0, sizeof(v));
        v.rows = g.rows;
        v.cols = g.cols;
        v.nnz = g.nnz;
        v.block_size = g.block_size;
        v.ell_cols = g.ell_cols;
        v.blockColIdx = (unsigned int *) (base + o_bci[k]);
        v.val = (csamd_half *) (base + o_val[k]);
        t.move_blob(g.bci, o_bci[k], g.bci_bytes);
        t.move_blob(g.val, o_val[k], g.val_bytes);
        bytes += exec_bell_packed_bytes(g.rows, g.ell_cols, g.block_size);
    }
    out->bytes = bytes;
    return launch_unpack(sl, t, base, s, p);
}

int stage_sell_one(const csamd_cspack_file *f, unsigned long long p, unsigned int value_type, csamd_cspack_workspace *ws,
                   stage_slot &sl, csamd_bucketed_sliced_ell *out) {
    const unsigned int vb = (unsigned int) value_type_bytes(value_type);
    hipStream_t s = (hipStream_t) ws->stream;
    uint64_t len = 0u;
    csamd_bucketed_sliced_ell_clear(out);
    if (!stage_read(f, p, ws, sl, &len)) return 0;
    sell_blob B;
    if (!parse_sell(sl.pinned, len, vb, p, B)) return 0;
    dev_layout L;
    const size_t o_e2cr = L.place(4ull * B.rows), o_c2er = L.place(4ull * B.rows);
    const size_t o_cso = L.place(4ull * ((uint64_t) B.canonical_slices + 1u)), o_csw = L.place(4ull * B.canonical_slices);
    std::vector<size_t> o_off(B.nseg), o_wid(B.nseg), o_slot(B.nseg), o_col(B.nseg), o_val(B.nseg);
    for (uint32_t k = 0; k < B.nseg; ++k) {
        const sell_segment &g = B.seg[k];
        o_off[k] = L.place(4ull * ((uint64_t) g.slice_count + 1u));
        o_wid[k] = L.place(4ull * g.slice_count);
        o_slot[k] = L.place(4ull * g.slice_count);
        o_col[k] = L.place(g.col_bytes);
        o_val[k] = L.place(g.val_bytes);
    }
    void *st = exec_storage_new();
    out->storage = st;
    unsigned char *base = (unsigned char *) exec_storage_dev(st, L.total);
    out->segments = (csamd_sliced_ell_view *) exec_storage_host(st, (size_t) B.nseg * sizeof(csamd_sliced_ell_view));
    out->segment_row_offsets = (unsigned int *) exec_storage_host(st, ((size_t) B.nseg + 1u) * 4u);
    CSAMD_REQUIRE(base != nullptr && out->segments != nullptr && out->segment_row_offsets != nullptr,
                  "cspack partition %llu: allocation of %zu device bytes failed", p, L.total);
    out->rows = B.rows;
    out->cols = B.cols;
    out->nnz = B.nnz;
    out->segment_count = B.nseg;
    out->canonical_slice_count = B.canonical_slices;
    std::memcpy(out->segment_row_offsets, sl.pinned + B.sro, ((size_t) B.nseg + 1u) * 4u);
    out->exec_to_canonical_rows = (unsigned int *) (base + o_e2cr);
    out->canonical_to_exec_rows = (unsigned int *) (base + o_c2er);
    out->canonical_slice_row_offsets = (unsigned int *) (base + o_cso);
    out->canonical_slice_widths = (unsigned int *) (base + o_csw);
    side_table t;
    t.move_blob(B.e2cr, o_e2cr, 4ull * B.rows);
    t.move_blob(B.c2er, o_c2er, 4ull * B.rows);
    t.move_blob(B.cso, o_cso, 4ull * ((uint64_t) B.canonical_slices + 1u));
    t.move_blob(B.csw, o_csw, 4ull * B.canonical_slices);
    unsigned long long bytes = 0u;
    for (uint32_t k = 0; k < B.nseg; ++k) {
        const sell_segment &g = B.seg[k];
        csamd_sliced_ell_view &v = out->segments[k];
        std::memset(&v, 0, sizeof(v));
        v.rows = g.rows;
        v.cols = g.cols;
        v.nnz = g.nnz;
        v.slice_count = g.slice_count;
        v.slice_rows = g.slice_rows;
        v.slice_row_offsets = (unsigned int *) (base + o_off[k]);
        v.slice_widths = (unsigned int *) (base + o_wid[k]);
        v.slice_slot_offsets = (unsigned int *) (base + o_slot[k]);
        v.col_idx = (unsigned int *) (base + o_col[k]);
        v.val = (csamd_half *) (base + o_val[k]);
        t.move_blob(g.offsets, o_off[k], 4ull * ((uint64_t) g.slice_count + 1u));
        t.move_blob(g.widths, o_wid[k], 4ull * g.slice_count);
        t.move_side(t.add(g.slot_offsets.data(), 4u * g.slot_offsets.size()), o_slot[k], 4ull * g.slice_count);
        t.move_blob(g.col, o_col[k], g.col_bytes);
        t.move_blob(g.val, o_val[k], g.val_bytes);
        bytes += exec_sell_packed_bytes(g.slice_count, g.col_bytes / 4u);
    }
    out->bytes = bytes;
    return launch_unpack(sl, t, base, s, p);
}

// ---------------------------------------------------------------------------
// Writer.  Each blob is a list of pieces: host bytes (headers, host arrays)
// or device arrays copied into the writer's pinned buffer at aligned places;
// the pieces go out with writev in blob order.
// ---------------------------------------------------------------------------
struct blob_plan {
    std::vector<unsigned char> hdr;  // header bytes, reserved up front
    struct piece {
        int kind;  // 0: hdr offset, 1: host pointer, 2: device pointer (pinned offset after planning)
        const void *p;
        uint64_t off, n;
    };
    std::vector<piece> pieces;
    uint64_t bytes = 0u;
    uint64_t pinned = 0u;

    void put_u8(unsigned char v) {
        pieces.push_back(piece{0, nullptr, hdr.size(), 1u});
        hdr.push_back(v);
        bytes += 1u;
    }
    void put_u32(uint32_t v) {
        pieces.push_back(piece{0, nullptr, hdr.size(), 4u});
        const unsigned char *q = (const unsigned char *) &v;
        hdr.insert(hdr.end(), q, q + 4);
        bytes += 4u;
    }
    void put_host(const void *p, uint64_t n) {
        if (n == 0u) return;
        pieces.push_back(piece{1, p, 0u, n});
        bytes += n;
    }
    void put_device(const void *p, uint64_t n) {
        if (n == 0u) return;
        pieces.push_back(piece{2, p, pinned, n});
        pinned += align_up(n, kAlign);
        bytes += n;
    }
};

// write_execution_partition_blob (execution_builders_part.hh:2116-2132) and
// store_blocked_ell_raw (packfile.cu:351-369) per segment.
int plan_bell(const csamd_bucketed_blocked_ell &m, unsigned int vb, unsigned int idx, blob_plan &b) {
    CSAMD_REQUIRE(m.segment_count == 0u || (m.segments != nullptr && m.segment_row_offsets != nullptr),
                  "write cspack: partition %u has no segments", idx);
    CSAMD_REQUIRE(m.segment_row_offsets != nullptr, "write cspack: partition %u has no segment row offsets", idx);
    CSAMD_REQUIRE(m.rows == 0u || (m.exec_to_canonical_rows != nullptr && m.canonical_to_exec_rows != nullptr),
                  "write cspack: partition %u has no row maps", idx);
    CSAMD_REQUIRE(m.cols == 0u || (m.exec_to_canonical_cols != nullptr && m.canonical_to_exec_cols != nullptr),
                  "write cspack: partition %u has no column maps", idx);
    b.hdr.reserve(16u + 21u * (size_t) m.segment_count);
    b.put_u32(m.rows);
    b.put_u32(m.cols);
    b.put_u32(m.nnz);
    b.put_u32(m.segment_count);
    b.put_host(m.segment_row_offsets, 4ull * ((uint64_t) m.segment_count + 1u));
    b.put_device(m.exec_to_canonical_rows, 4ull * m.rows);
    b.put_device(m.canonical_to_exec_rows, 4ull * m.rows);
    b.put_device(m.exec_to_canonical_cols, 4ull * m.cols);
    b.put_device(m.canonical_to_exec_cols, 4ull * m.cols);
    for (unsigned int k = 0; k < m.segment_count; ++k) {
        const csamd_blocked_ell_view &v = m.segments[k];
        const uint64_t rb = v.block_size != 0u ? ((uint64_t) v.rows + v.block_size - 1u) / v.block_size : 0u;
        const uint64_t w = v.block_size != 0u ? (uint64_t) v.ell_cols / v.block_size : 0u;
        b.put_u8(kFormatBlockedEll);
        b.put_u32(v.rows);
        b.put_u32(v.cols);
        b.put_u32(v.nnz);
        b.put_u32(v.block_size);
        b.put_u32(v.ell_cols);
        b.put_device(v.blockColIdx, rb * w * 4u);
        b.put_device(v.val, (uint64_t) v.rows * v.ell_cols * vb);
    }
    return 1;
}

// write_sliced_execution_partition_blob (execution_runtime_part.hh:168-195)
// and store_sliced_ell_raw (packfile.cu:597-621) per segment.  The slice
// tables are read back first to size the slot arrays.
int plan_sell(const csamd_bucketed_sliced_ell &m, unsigned int vb, unsigned int idx, hipStream_t s, blob_plan &b,
              std::vector<std::vector<uint32_t>> &tables) {
    CSAMD_REQUIRE(m.segment_row_offsets != nullptr && (m.segment_count == 0u || m.segments != nullptr),
                  "write cspack: partition %u has no segments", idx);
    CSAMD_REQUIRE(m.rows == 0u || (m.exec_to_canonical_rows != nullptr && m.canonical_to_exec_rows != nullptr),
                  "write cspack: partition %u has no row maps", idx);
    CSAMD_REQUIRE(m.canonical_slice_row_offsets != nullptr && (m.canonical_slice_count == 0u || m.canonical_slice_widths != nullptr),
                  "write cspack: partition %u has no canonical slice tables", idx);
    tables.assign(m.segment_count, {});
    for (unsigned int k = 0; k < m.segment_count; ++k) {
        const csamd_sliced_ell_view &v = m.segments[k];
        std::vector<uint32_t> &t = tables[k];
        t.assign(2u * (size_t) v.slice_count + 1u, 0u);
        CSAMD_HIP(hipMemcpyAsync(t.data(), v.slice_row_offsets, 4ull * ((size_t) v.slice_count + 1u), hipMemcpyDeviceToHost, s),
                  "D2H slice row offsets");
        if (v.slice_count != 0u)
            CSAMD_HIP(hipMemcpyAsync(t.data() + v.slice_count + 1u, v.slice_widths, 4ull * v.slice_count, hipMemcpyDeviceToHost, s),
                      "D2H slice widths");
    }
    CSAMD_HIP(hipStreamSynchronize(s), "write cspack slice tables");
    b.hdr.reserve(20u + 17u * (size_t) m.segment_count);
    b.put_u32(m.rows);
    b.put_u32(m.cols);
    b.put_u32(m.nnz);
    b.put_u32(m.segment_count);
    b.put_u32(m.canonical_slice_count);
    b.put_host(m.segment_row_offsets, 4ull * ((uint64_t) m.segment_count + 1u));
    b.put_device(m.exec_to_canonical_rows, 4ull * m.rows);
    b.put_device(m.canonical_to_exec_rows, 4ull * m.rows);
    b.put_device(m.canonical_slice_row_offsets, 4ull * ((uint64_t) m.canonical_slice_count + 1u));
    b.put_device(m.canonical_slice_widths, 4ull * m.canonical_slice_count);
    for (unsigned int k = 0; k < m.segment_count; ++k) {
        const csamd_sliced_ell_view &v = m.segments[k];
        const std::vector<uint32_t> &t = tables[k];
        uint64_t slots = 0u;
        for (unsigned int q = 0; q < v.slice_count; ++q) slots += (uint64_t) (t[q + 1u] - t[q]) * t[v.slice_count + 1u + q];
        b.put_u8(kFormatSlicedEll);
        b.put_u32(v.rows);
        b.put_u32(v.cols);
        b.put_u32(v.nnz);
        b.put_u32(v.slice_count);
        b.put_host(t.data(), 4ull * ((uint64_t) v.slice_count + 1u));
        b.put_host(t.data() + v.slice_count + 1u, 4ull * v.slice_count);
        b.put_device(v.col_idx, slots * 4u);
        b.put_device(v.val, slots * vb);
    }
    return 1;
}

// Copies a planned blob's device pieces to pinned memory and writes all
// pieces at the file's current position.
int emit_blob(int fd, blob_plan &b, cs_private *q, hipStream_t s, unsigned int idx) {
    if (!grow_pinned(&q->wpinned, &q->wpinned_cap, b.pinned + kPad)) return 0;
    for (const blob_plan::piece &pc : b.pieces)
        if (pc.kind == 2)
            CSAMD_HIP(hipMemcpyAsync(q->wpinned + pc.off, pc.p, pc.n, hipMemcpyDeviceToHost, s), "D2H cspack array");
    CSAMD_HIP(hipStreamSynchronize(s), "write cspack copies");
    std::vector<struct iovec> iov;
    iov.reserve(b.pieces.size());
    for (const blob_plan::piece &pc : b.pieces) {
        const void *p = pc.kind == 0 ? (const void *) (b.hdr.data() + pc.off)
                                     : pc.kind == 1 ? pc.p : (const void *) (q->wpinned + pc.off);
        iov.push_back(iovec{const_cast<void *>(p), (size_t) pc.n});
    }
    size_t at = 0u;
    while (at < iov.size()) {
        const int n = (int) std::min<size_t>(iov.size() - at, 1024u);
        const ssize_t r = ::writev(fd, iov.data() + at, n);
        if (r < 0) {
            if (errno == EINTR) continue;
            set_error("write cspack partition %u: %s", idx, std::strerror(errno));
            return 0;
        }
        size_t done = (size_t) r;
        while (at < iov.size() && done >= iov[at].iov_len) {
            done -= iov[at].iov_len;
            ++at;
        }
        if (done != 0u) {
            iov[at].iov_base = (char *) iov[at].iov_base + done;
            iov[at].iov_len -= done;
        }
    }
    return 1;
}

// materialize_*_cspack (execution_runtime_part.hh:643-796): the container
// header with the offsets of blobs laid end to end, then the blobs, written
// to `path`.tmp and renamed into place.
template<typename Part, typename Plan>
int write_pack(const char *path, unsigned long long shard_id, const Part *parts, unsigned int count, unsigned int value_type,
               csamd_cspack_workspace *ws, Plan plan) {
    CSAMD_REQUIRE(path != nullptr && (count == 0u || parts != nullptr), "write cspack: null argument");
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "write cspack: unsupported value type %u", value_type);
    if (!check_ws(ws, "write cspack")) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    std::vector<blob_plan> plans(count);
    for (unsigned int i = 0; i < count; ++i)
        if (!plan(parts[i], (unsigned int) vb, i, s, plans[i])) return 0;
    std::vector<uint64_t> head(3u + count);
    std::memcpy(&head[0], kMagic, 8u);
    head[1] = shard_id;
    head[2] = count;
    uint64_t off = kContainerHeader + 8ull * count;
    for (unsigned int i = 0; i < count; ++i) {
        head[3u + i] = off;
        off += plans[i].bytes;
    }
    const std::string tmp = std::string(path) + ".tmp";
    const int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    CSAMD_REQUIRE(fd >= 0, "write cspack: cannot create %s: %s", tmp.c_str(), std::strerror(errno));
    int ok = 1;
    const int e = write_all(fd, head.data(), head.size() * 8u);
    if (e != 0) {
        set_error("write cspack %s: %s", tmp.c_str(), std::strerror(e));
        ok = 0;
    }
    for (unsigned int i = 0; ok && i < count; ++i) ok = emit_blob(fd, plans[i], (cs_private *) ws->private_workspace, s, i);
    if (::close(fd) != 0 && ok) {
        set_error("write cspack %s: close: %s", tmp.c_str(), std::strerror(errno));
        ok = 0;
    }
    if (ok && ::rename(tmp.c_str(), path) != 0) {
        set_error("write cspack: rename to %s: %s", path, std::strerror(errno));
        ok = 0;
    }
    if (!ok) (void) ::unlink(tmp.c_str());
    return ok;
}

typedef unsigned int rc_u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Canonical CSR of a bucketed Blocked-ELL execution partition
// (bucketed_partition_to_canonical_coo, execution_builders_part.hh:909-968):
// one wave per exec row; the stored values that are not +-0 at exec columns
// < cols, in slot order, land in the canonical row's CSR range with their
// columns mapped to canonical ones.
// ---------------------------------------------------------------------------
struct rc_segment {
    const unsigned int *bci;
    const unsigned char *val;
    unsigned int row0, rows, block_size, ell_cols, width;
    unsigned int vec;  // 16-byte chunks of a value row never straddle a block (see rc_vec_ok)
};

// Values in a 16-byte chunk, and whether a segment can be read in chunks:
// a block holds whole chunks and every value row starts on a chunk (with a
// 16-byte aligned value array).
template<int VB> constexpr unsigned int rc_per_chunk() { return 16u / (unsigned int) VB; }

inline bool rc_vec_ok(unsigned int vb, unsigned int block_size, unsigned int ell_cols, const void *val) {
    const unsigned int per = 16u / vb;
    return block_size % per == 0u && ell_cols % per == 0u && ((uintptr_t) val & 15u) == 0u;
}

template<int VB>
__device__ __forceinline__ unsigned int rc_elem(rc_u32x4 q, unsigned int e) {
    if constexpr (VB == 4) return q[e];
    else if constexpr (VB == 2) return (q[e >> 1] >> ((e & 1u) * 16u)) & 0xffffu;
    else return (q[e >> 2] >> ((e & 3u) * 8u)) & 0xffu;
}

// Chunk c of exec row `local` of segment g: bit e of the result is set when
// value e of the chunk is an entry; *col0 is the chunk's first exec column.
template<int VB>
__device__ __forceinline__ unsigned int rc_chunk(const rc_segment &g, unsigned int local, unsigned int c, unsigned int cols,
                                                 unsigned int zero_mask, rc_u32x4 *q, unsigned long long *col0) {
    constexpr unsigned int PER = rc_per_chunk<VB>();
    const unsigned int t0 = c * PER, slot = t0 / g.block_size, cib0 = t0 - slot * g.block_size;
    const unsigned int bc = g.bci[(size_t) (local / g.block_size) * g.width + slot];
    *q = ((const rc_u32x4 *) (g.val + (size_t) local * g.ell_cols * VB))[c];
    *col0 = (unsigned long long) bc * g.block_size + cib0;
    unsigned int m = 0u;
    if (bc != CSAMD_INVALID_COL) {
#pragma unroll
        for (unsigned int e = 0; e < PER; ++e)
            if ((rc_elem<VB>(*q, e) & zero_mask) != 0u && *col0 + e < cols) m |= 1u << e;
    }
    return m;
}

struct rc_args {
    const rc_segment *seg;
    unsigned int nseg, rows, cols, zero_mask;
    const unsigned int *e2c_rows;
    const unsigned int *e2c_cols;  // null: identity
    unsigned int *counts;          // [rows], by canonical row
    unsigned int *hits;            // [rows], zeroed
    unsigned int *flags;           // bit 0: row map not a permutation, bit 1: mapped column >= cols
};

template<int VB>
__device__ __forceinline__ unsigned int rc_bits(const unsigned char *v, size_t i) {
    if constexpr (VB == 1) return v[i];
    else if constexpr (VB == 2) return ((const unsigned short *) v)[i];
    else return ((const unsigned int *) v)[i];
}

__device__ __forceinline__ unsigned int rc_segment_of(const rc_segment *seg, unsigned int nseg, unsigned int x) {
    unsigned int lo = 0u, hi = nseg - 1u;
    while (lo < hi) {
        const unsigned int mid = (lo + hi + 1u) >> 1;
        if (seg[mid].row0 <= x) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

// Exec row x, slot position t of its value row: whether it holds an entry
// and its exec column.
template<int VB>
__device__ __forceinline__ bool rc_entry(const rc_segment &g, unsigned int local, unsigned int t, unsigned int cols,
                                         unsigned int zero_mask, unsigned long long *exec_col, unsigned int *bits) {
    const unsigned int slot = t / g.block_size, cib = t - slot * g.block_size;
    const unsigned int bc = g.bci[(size_t) (local / g.block_size) * g.width + slot];
    *bits = rc_bits<VB>(g.val, (size_t) local * g.ell_cols + t);
    *exec_col = (unsigned long long) bc * g.block_size + cib;
    return bc != CSAMD_INVALID_COL && (*bits & zero_mask) != 0u && *exec_col < cols;
}

template<int VB>
__global__ void rc_count_kernel(rc_args a) {
    const unsigned int nw = (gridDim.x * blockDim.x) >> 6;
    for (unsigned int x = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; x < a.rows; x += nw) {
        const rc_segment g = a.seg[rc_segment_of(a.seg, a.nseg, x)];
        const unsigned int local = x - g.row0;
        unsigned int n = 0u;
        if (g.vec) {
            const unsigned int chunks = g.ell_cols / rc_per_chunk<VB>();
            for (unsigned int c = lane_id(); c < chunks; c += 64u) {
                rc_u32x4 q;
                unsigned long long c0;
                n += (unsigned int) __popc(rc_chunk<VB>(g, local, c, a.cols, a.zero_mask, &q, &c0));
            }
        } else {
            for (unsigned int t = lane_id(); t < g.ell_cols; t += 64u) {
                unsigned long long ec;
                unsigned int bits;
                n += rc_entry<VB>(g, local, t, a.cols, a.zero_mask, &ec, &bits) ? 1u : 0u;
            }
        }
        n = wave_sum_u32(n);
        if (lane_id() == 0u) {
            const unsigned int c = a.e2c_rows[x];
            if (c >= a.rows || atomicAdd(&a.hits[c], 1u) != 0u) atomicOr(a.flags, 1u);
            else a.counts[c] = n;
        }
    }
}

template<int VB>
__global__ void rc_emit_kernel(rc_args a, const unsigned int *__restrict__ ptr, unsigned int *__restrict__ out_idx,
                               unsigned char *__restrict__ out_val, unsigned int limit) {
    const unsigned int nw = (gridDim.x * blockDim.x) >> 6;
    for (unsigned int x = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; x < a.rows; x += nw) {
        const rc_segment g = a.seg[rc_segment_of(a.seg, a.nseg, x)];
        const unsigned int local = x - g.row0;
        const unsigned int c = a.e2c_rows[x];  // a permutation (checked by the count pass)
        if (c >= a.rows) continue;             // uniform: one row per wave
        const unsigned int base = ptr[c], end = min(ptr[c + 1u], limit);
        unsigned int run = 0u;
        if (g.vec) {
            constexpr unsigned int PER = rc_per_chunk<VB>();
            const unsigned int chunks = g.ell_cols / PER;
            for (unsigned int k0 = 0u; k0 < chunks; k0 += 64u) {
                const unsigned int k = k0 + lane_id();
                rc_u32x4 q = rc_u32x4{0u, 0u, 0u, 0u};
                unsigned long long c0 = 0u;
                const unsigned int m = k < chunks ? rc_chunk<VB>(g, local, k, a.cols, a.zero_mask, &q, &c0) : 0u;
                const unsigned int cnt = (unsigned int) __popc(m);
                const unsigned int incl = wave_inclusive_scan_u32(cnt);
                unsigned int pos = base + run + incl - cnt;
#pragma unroll
                for (unsigned int e = 0; e < PER; ++e) {
                    if ((m >> e) & 1u) {
                        const unsigned int col = a.e2c_cols != nullptr ? a.e2c_cols[c0 + e] : (unsigned int) (c0 + e);
                        if (col >= a.cols) atomicOr(a.flags, 2u);
                        if (pos < end) {
                            const unsigned int bits = rc_elem<VB>(q, e);
                            out_idx[pos] = col;
                            if constexpr (VB == 1) out_val[pos] = (unsigned char) bits;
                            else if constexpr (VB == 2) ((unsigned short *) out_val)[pos] = (unsigned short) bits;
                            else ((unsigned int *) out_val)[pos] = bits;
                        }
                        ++pos;
                    }
                }
                run += __shfl(incl, 63);
            }
            continue;
        }
        for (unsigned int t0 = 0u; t0 < g.ell_cols; t0 += 64u) {
            const unsigned int t = t0 + lane_id();
            unsigned long long ec = 0u;
            unsigned int bits = 0u;
            const bool ok = t < g.ell_cols && rc_entry<VB>(g, local, t, a.cols, a.zero_mask, &ec, &bits);
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const unsigned int pos = base + run + (unsigned int) __popcll(m & lanemask_lt());
                const unsigned int col = a.e2c_cols != nullptr ? a.e2c_cols[ec] : (unsigned int) ec;
                if (col >= a.cols) atomicOr(a.flags, 2u);
                if (pos < end) {
                    out_idx[pos] = col;
                    if constexpr (VB == 1) out_val[pos] = (unsigned char) bits;
                    else if constexpr (VB == 2) ((unsigned short *) out_val)[pos] = (unsigned short) bits;
                    else ((unsigned int *) out_val)[pos] = bits;
                }
            }
            run += (unsigned int) __popcll(m);
        }
    }
}

// ---------------------------------------------------------------------------
// Canonical Sliced-ELL of a bucketed Sliced-ELL execution partition
// (rebuild_canonical_sliced_partition, execution_runtime_part.hh:275-318):
// one wave per exec row; its valid slots (column != invalid), in slot order,
// fill the front of its canonical row, the rest of the row is invalid / 0.
// ---------------------------------------------------------------------------
struct rs_segment {
    const unsigned int *offsets, *widths, *slots, *col;
    const unsigned char *val;
    unsigned int row0, rows, slice_count, pad;
};

struct rs_args {
    const rs_segment *seg;
    unsigned int nseg, rows, cslices, pad;
    const unsigned int *e2c_rows;
    const unsigned int *c_off, *c_wid, *c_slot;  // canonical slices (device)
    unsigned int *hits;                          // [rows], zeroed
    unsigned int *flags;                         // bit 0: row map, bit 1: a row's valid slots exceed its canonical width
};

__device__ __forceinline__ unsigned int rs_find(const unsigned int *off, unsigned int n, unsigned int row) {
    unsigned int lo = 0u, hi = n - 1u;  // last slice whose start <= row
    while (lo < hi) {
        const unsigned int mid = (lo + hi + 1u) >> 1;
        if (off[mid] <= row) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

template<int VB>
__global__ void rs_rebuild_kernel(rs_args a, unsigned int *__restrict__ out_col, unsigned char *__restrict__ out_val) {
    const unsigned int nw = (gridDim.x * blockDim.x) >> 6;
    for (unsigned int x = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; x < a.rows; x += nw) {
        const unsigned int c = a.e2c_rows[x];
        unsigned int bad = 0u;
        if (lane_id() == 0u) bad = (c >= a.rows || atomicAdd(&a.hits[c], 1u) != 0u) ? 1u : 0u;
        bad = __shfl(bad, 0);
        if (bad) {
            if (lane_id() == 0u) atomicOr(a.flags, 1u);
            continue;
        }
        // the canonical slices cover [0, rows) (checked on the host)
        const unsigned int ds = rs_find(a.c_off, a.cslices, c);
        const unsigned int dw = a.c_wid[ds];
        const size_t dbase = (size_t) a.c_slot[ds] + (size_t) (c - a.c_off[ds]) * dw;
        unsigned int lo = 0u, hi = a.nseg - 1u;
        while (lo < hi) {
            const unsigned int mid = (lo + hi + 1u) >> 1;
            if (a.seg[mid].row0 <= x) lo = mid;
            else hi = mid - 1u;
        }
        const rs_segment g = a.seg[lo];
        const unsigned int local = x - g.row0;
        const unsigned int ss = rs_find(g.offsets, g.slice_count, local);
        const unsigned int sw = g.widths[ss];
        const size_t sbase = (size_t) g.slots[ss] + (size_t) (local - g.offsets[ss]) * sw;
        unsigned int run = 0u;
        for (unsigned int t0 = 0u; t0 < sw; t0 += 64u) {
            const unsigned int t = t0 + lane_id();
            const unsigned int col = t < sw ? g.col[sbase + t] : CSAMD_INVALID_COL;
            const bool ok = col != CSAMD_INVALID_COL;
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const unsigned int pos = run + (unsigned int) __popcll(m & lanemask_lt());
                if (pos < dw) {
                    out_col[dbase + pos] = col;
                    if constexpr (VB == 1) out_val[dbase + pos] = g.val[sbase + t];
                    else if constexpr (VB == 2) ((unsigned short *) out_val)[dbase + pos] = ((const unsigned short *) g.val)[sbase + t];
                    else ((unsigned int *) out_val)[dbase + pos] = ((const unsigned int *) g.val)[sbase + t];
                } else {
                    atomicOr(a.flags, 2u);
                }
            }
            run += (unsigned int) __popcll(m);
        }
        for (unsigned int j = run + lane_id(); j < dw; j += 64u) {
            out_col[dbase + j] = CSAMD_INVALID_COL;
            if constexpr (VB == 1) out_val[dbase + j] = 0u;
            else if constexpr (VB == 2) ((unsigned short *) out_val)[dbase + j] = 0u;
            else ((unsigned int *) out_val)[dbase + j] = 0u;
        }
    }
}

// Emit pass of the canonical CSR (row pointers in place), then the mapped
// column check.
int canonical_emit(const rc_args &a, int vb, unsigned int grid, const unsigned int *d_ptr, unsigned int *d_idx, void *d_val,
                   unsigned int rows, unsigned int nnz, unsigned int *flags, unsigned int cols, hipStream_t s) {
    if (rows != 0u && nnz != 0u) {
        switch (vb) {
            case 1: rc_emit_kernel<1><<<grid, 256, 0, s>>>(a, d_ptr, d_idx, (unsigned char *) d_val, nnz); break;
            case 2: rc_emit_kernel<2><<<grid, 256, 0, s>>>(a, d_ptr, d_idx, (unsigned char *) d_val, nnz); break;
            default: rc_emit_kernel<4><<<grid, 256, 0, s>>>(a, d_ptr, d_idx, (unsigned char *) d_val, nnz); break;
        }
        CSAMD_HIP(hipGetLastError(), "rc_emit_kernel");
    }
    unsigned int h = 0u;
    CSAMD_HIP(hipMemcpyAsync(&h, flags, 4u, hipMemcpyDeviceToHost, s), "D2H flags");
    CSAMD_HIP(hipStreamSynchronize(s), "canonical emit sync");
    CSAMD_REQUIRE((h & 2u) == 0u, "canonical_compressed_from_bucketed_blocked_ell: a mapped column is >= cols %u", cols);
    return 1;
}

} // namespace
} // namespace csamd

using namespace csamd;

extern "C" {

void csamd_cspack_init(csamd_cspack_workspace *ws) {
    if (ws == nullptr) return;
    std::memset(ws, 0, sizeof(*ws));
    ws->device = -1;
}

void csamd_cspack_clear(csamd_cspack_workspace *ws) {
    if (ws == nullptr) return;
    if (ws->device >= 0) (void) hipSetDevice(ws->device);
    free_private(ws->private_workspace);
    if (ws->owns_stream && ws->stream != nullptr) (void) hipStreamDestroy((hipStream_t) ws->stream);
    const unsigned int threads = ws->reader_threads, chunk = ws->chunk_bytes;
    csamd_cspack_init(ws);
    ws->reader_threads = threads;
    ws->chunk_bytes = chunk;
}

int csamd_cspack_setup(csamd_cspack_workspace *ws, int device, csamd_stream_t stream) {
    CSAMD_REQUIRE(ws != nullptr, "cspack setup: null workspace");
    csamd_cspack_clear(ws);
    CSAMD_HIP(hipSetDevice(device), "hipSetDevice cspack setup");
    ws->device = device;
    if (stream == nullptr) {
        hipStream_t s = nullptr;
        CSAMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags cspack");
        ws->stream = (csamd_stream_t) s;
        ws->owns_stream = 1;
    } else {
        ws->stream = stream;
        ws->owns_stream = 0;
    }
    return 1;
}

int csamd_cspack_write_blocked_ell(const char *path, unsigned long long shard_id, const csamd_bucketed_blocked_ell *parts,
                                   unsigned int part_count, unsigned int value_type, csamd_cspack_workspace *ws) {
    return write_pack(path, shard_id, parts, part_count, value_type, ws,
                      [](const csamd_bucketed_blocked_ell &m, unsigned int vb, unsigned int i, hipStream_t, blob_plan &b) {
                          return plan_bell(m, vb, i, b);
                      });
}

int csamd_cspack_write_sliced_ell(const char *path, unsigned long long shard_id, const csamd_bucketed_sliced_ell *parts,
                                  unsigned int part_count, unsigned int value_type, csamd_cspack_workspace *ws) {
    std::vector<std::vector<std::vector<uint32_t>>> tables(part_count);
    return write_pack(path, shard_id, parts, part_count, value_type, ws,
                      [&](const csamd_bucketed_sliced_ell &m, unsigned int vb, unsigned int i, hipStream_t s, blob_plan &b) {
                          return plan_sell(m, vb, i, s, b, tables[i]);
                      });
}

int csamd_cspack_open(const char *path, unsigned long long shard_id, csamd_cspack_file *out) {
    CSAMD_REQUIRE(path != nullptr && out != nullptr, "cspack open: null argument");
    std::memset(out, 0, sizeof(*out));
    const int fd = ::open(path, O_RDONLY | O_CLOEXEC);
    CSAMD_REQUIRE(fd >= 0, "cspack open: cannot open %s: %s", path, std::strerror(errno));
    struct stat sb;
    unsigned char head[kContainerHeader];
    uint64_t size = 0u, count = 0u, file_shard = 0u;
    std::vector<uint64_t> offs;
    const char *why = nullptr;
    if (::fstat(fd, &sb) != 0) why = "stat failed";
    if (why == nullptr) {
        size = (uint64_t) sb.st_size;
        if (size < kContainerHeader || pread_all(fd, head, kContainerHeader, 0u) != 0) why = "file shorter than the CSPACK header";
    }
    if (why == nullptr && std::memcmp(head, kMagic, 8u) != 0) why = "bad magic (not CSPACK01)";
    if (why == nullptr) {
        file_shard = rd64(head + 8);
        count = rd64(head + 16);
        if (shard_id != CSAMD_CSPACK_ANY_SHARD && file_shard != shard_id) why = "shard id differs";
        else if (count > (size - kContainerHeader) / 8u) why = "partition offset table runs past the end of the file";
    }
    if (why == nullptr) {
        offs.resize(count);
        if (count != 0u && pread_all(fd, (unsigned char *) offs.data(), 8ull * count, kContainerHeader) != 0)
            why = "cannot read the partition offset table";
    }
    const uint64_t table_end = kContainerHeader + 8ull * count;
    for (uint64_t i = 0; why == nullptr && i < count; ++i)
        if (offs[i] < table_end || offs[i] > size) why = "a partition offset lies outside the file's payload region";
    if (why != nullptr) {
        ::close(fd);
        if (std::strcmp(why, "shard id differs") == 0)
            set_error("cspack open %s: shard id %llu, expected %llu", path, (unsigned long long) file_shard, shard_id);
        else
            set_error("cspack open %s: %s", path, why);
        return 0;
    }
    std::vector<uint64_t> sorted(offs);
    std::sort(sorted.begin(), sorted.end());
    out->partition_offsets = (unsigned long long *) std::calloc(count != 0u ? count : 1u, 8u);
    out->partition_bytes = (unsigned long long *) std::calloc(count != 0u ? count : 1u, 8u);
    if (out->partition_offsets == nullptr || out->partition_bytes == nullptr) {
        std::free(out->partition_offsets);
        std::free(out->partition_bytes);
        std::memset(out, 0, sizeof(*out));
        ::close(fd);
        set_error("cspack open: out of host memory");
        return 0;
    }
    for (uint64_t i = 0; i < count; ++i) {
        const auto nx = std::upper_bound(sorted.begin(), sorted.end(), offs[i]);
        out->partition_offsets[i] = offs[i];
        out->partition_bytes[i] = (nx == sorted.end() ? size : *nx) - offs[i];
    }
    out->shard_id = file_shard;
    out->partition_count = count;
    out->file_bytes = size;
    out->impl = new cs_file_impl{fd};
    return 1;
}

void csamd_cspack_close(csamd_cspack_file *f) {
    if (f == nullptr) return;
    cs_file_impl *fi = (cs_file_impl *) f->impl;
    if (fi != nullptr) {
        ::close(fi->fd);
        delete fi;
    }
    std::free(f->partition_offsets);
    std::free(f->partition_bytes);
    std::memset(f, 0, sizeof(*f));
}

int csamd_cspack_stage_blocked_ell(const csamd_cspack_file *f, unsigned long long local_partition, unsigned int value_type,
                                   csamd_cspack_workspace *ws, csamd_bucketed_blocked_ell *out) {
    if (!check_stage_args(f, local_partition, value_type, ws, out, "cspack stage blocked_ell")) return 0;
    cs_private *q = (cs_private *) ws->private_workspace;
    int ok = stage_bell_one(f, local_partition, value_type, ws, q->slot[0], out);
    ok = hip_ok(hipStreamSynchronize((hipStream_t) ws->stream), "cspack stage sync") && ok;
    q->slot[0].busy = false;
    if (!ok) csamd_bucketed_blocked_ell_clear(out);
    return ok;
}

int csamd_cspack_stage_sliced_ell(const csamd_cspack_file *f, unsigned long long local_partition, unsigned int value_type,
                                  csamd_cspack_workspace *ws, csamd_bucketed_sliced_ell *out) {
    if (!check_stage_args(f, local_partition, value_type, ws, out, "cspack stage sliced_ell")) return 0;
    cs_private *q = (cs_private *) ws->private_workspace;
    int ok = stage_sell_one(f, local_partition, value_type, ws, q->slot[0], out);
    ok = hip_ok(hipStreamSynchronize((hipStream_t) ws->stream), "cspack stage sync") && ok;
    q->slot[0].busy = false;
    if (!ok) csamd_bucketed_sliced_ell_clear(out);
    return ok;
}

int csamd_cspack_stage_shard_blocked_ell(const csamd_cspack_file *f, unsigned int value_type, csamd_cspack_workspace *ws,
                                         csamd_bucketed_blocked_ell *out) {
    CSAMD_REQUIRE(f != nullptr && f->impl != nullptr && (out != nullptr || f->partition_count == 0u),
                  "cspack stage shard: null argument");
    const unsigned long long n = f->partition_count;
    for (unsigned long long p = 0; p < n; ++p) std::memset(&out[p], 0, sizeof(out[p]));
    if (n == 0u) return 1;
    if (!check_stage_args(f, 0u, value_type, ws, out, "cspack stage shard")) return 0;
    cs_private *q = (cs_private *) ws->private_workspace;
    int ok = 1;
    for (unsigned long long p = 0; ok && p < n; ++p) ok = stage_bell_one(f, p, value_type, ws, q->slot[p & 1u], &out[p]);
    ok = hip_ok(hipStreamSynchronize((hipStream_t) ws->stream), "cspack stage shard sync") && ok;
    q->slot[0].busy = q->slot[1].busy = false;
    if (!ok)
        for (unsigned long long p = 0; p < n; ++p) csamd_bucketed_blocked_ell_clear(&out[p]);
    return ok;
}

int csamd_canonical_compressed_from_bucketed_blocked_ell(const csamd_bucketed_blocked_ell *part, unsigned int value_type,
                                                         unsigned int *d_out_ptr, int ptr_ready, unsigned int *d_out_idx,
                                                         void *d_out_val, unsigned long long capacity, unsigned long long *nnz_out,
                                                         csamd_convert_workspace *ws) {
    CSAMD_REQUIRE(part != nullptr && d_out_ptr != nullptr && nnz_out != nullptr && ws != nullptr,
                  "canonical_compressed_from_bucketed_blocked_ell: null argument");
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "canonical_compressed_from_bucketed_blocked_ell: unsupported value type %u", value_type);
    const unsigned int rows = part->rows, nseg = part->segment_count;
    CSAMD_REQUIRE(rows == 0u || (part->exec_to_canonical_rows != nullptr && nseg != 0u && part->segments != nullptr &&
                                 part->segment_row_offsets != nullptr),
                  "canonical_compressed_from_bucketed_blocked_ell: partition without segments or row map");
    CSAMD_REQUIRE(rows == 0u || (part->segment_row_offsets[0] == 0u && part->segment_row_offsets[nseg] == rows),
                  "canonical_compressed_from_bucketed_blocked_ell: segment row offsets do not cover rows [0, %u)", rows);
    std::vector<rc_segment> segs(nseg != 0u ? nseg : 1u);
    for (unsigned int k = 0; k < nseg; ++k) {
        const csamd_blocked_ell_view &v = part->segments[k];
        const unsigned int r0 = part->segment_row_offsets[k], r1 = part->segment_row_offsets[k + 1u];
        CSAMD_REQUIRE(r1 >= r0 && r1 - r0 == v.rows && r1 <= rows, "canonical_compressed_from_bucketed_blocked_ell: segment %u rows", k);
        CSAMD_REQUIRE(v.rows == 0u || (v.block_size != 0u && v.ell_cols % v.block_size == 0u &&
                                       (v.ell_cols == 0u || (v.blockColIdx != nullptr && v.val != nullptr))),
                      "canonical_compressed_from_bucketed_blocked_ell: segment %u shape", k);
        segs[k] = rc_segment{v.blockColIdx, (const unsigned char *) v.val, r0, v.rows, v.block_size != 0u ? v.block_size : 1u,
                             v.ell_cols, v.block_size != 0u ? v.ell_cols / v.block_size : 0u,
                             v.block_size != 0u && rc_vec_ok((unsigned int) vb, v.block_size, v.ell_cols, v.val) ? 1u : 0u};
    }
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    const size_t seg_bytes = align_up((size_t) segs.size() * sizeof(rc_segment), 256u);
    const size_t row_bytes = align_up((size_t) rows * 4u + 4u, 256u);
    const size_t scan_bytes = exclusive_scan_scratch_bytes(rows);
    char *scratch = (char *) convert_scratch(ws, seg_bytes + 2u * row_bytes + 256u + scan_bytes);
    if (scratch == nullptr) return 0;
    rc_segment *d_seg = (rc_segment *) scratch;
    unsigned int *counts = (unsigned int *) (scratch + seg_bytes);
    unsigned int *hits = (unsigned int *) (scratch + seg_bytes + row_bytes);
    unsigned int *flags = (unsigned int *) (scratch + seg_bytes + 2u * row_bytes);
    void *scan_scratch = scratch + seg_bytes + 2u * row_bytes + 256u;
    CSAMD_HIP(hipMemcpyAsync(d_seg, segs.data(), segs.size() * sizeof(rc_segment), hipMemcpyHostToDevice, s), "H2D segments");
    if (!zero_bytes(hits, row_bytes + 256u, s)) return 0;  // hits and flags
    const unsigned int zero_mask = vb == 1 ? 0x7fu : vb == 2 ? 0x7fffu : 0x7fffffffu;
    const rc_args a{d_seg, nseg, rows, part->cols, zero_mask, part->exec_to_canonical_rows, part->exec_to_canonical_cols, counts, hits,
                    flags};
    const unsigned int grid = grid_for((unsigned long long) rows * 64u, 256u, (unsigned int) device_cu_count() * 16u);
    unsigned int h[3] = {0u, 0u, 0u};
    if (ptr_ready) {
        // the row pointers of a previous sizing call on this partition
        CSAMD_REQUIRE(d_out_idx != nullptr, "canonical_compressed_from_bucketed_blocked_ell: ptr_ready without outputs");
        CSAMD_HIP(hipMemcpyAsync(h + 2, d_out_ptr + rows, 4u, hipMemcpyDeviceToHost, s), "D2H nnz");
        CSAMD_HIP(hipStreamSynchronize(s), "canonical nnz sync");
        *nnz_out = h[2];
        CSAMD_REQUIRE(capacity >= h[2] && (h[2] == 0u || d_out_val != nullptr),
                      "canonical_compressed_from_bucketed_blocked_ell: capacity %llu < nnz %u", capacity, h[2]);
        return canonical_emit(a, vb, grid, d_out_ptr, d_out_idx, d_out_val, rows, h[2], flags, part->cols, s);
    }
    if (rows != 0u) {
        switch (vb) {
            case 1: rc_count_kernel<1><<<grid, 256, 0, s>>>(a); break;
            case 2: rc_count_kernel<2><<<grid, 256, 0, s>>>(a); break;
            default: rc_count_kernel<4><<<grid, 256, 0, s>>>(a); break;
        }
        CSAMD_HIP(hipGetLastError(), "rc_count_kernel");
    }
    if (!exclusive_scan_u32(counts, rows, d_out_ptr, flags + 1, scan_scratch, scan_bytes, s, true)) return 0;
    CSAMD_HIP(hipMemcpyAsync(h, flags, 8u, hipMemcpyDeviceToHost, s), "D2H flags");
    CSAMD_HIP(hipMemcpyAsync(h + 2, d_out_ptr + rows, 4u, hipMemcpyDeviceToHost, s), "D2H nnz");
    CSAMD_HIP(hipStreamSynchronize(s), "canonical count sync");
    CSAMD_REQUIRE((h[0] & 1u) == 0u, "canonical_compressed_from_bucketed_blocked_ell: exec_to_canonical_rows is not a permutation");
    CSAMD_REQUIRE(h[1] == 0u, "canonical_compressed_from_bucketed_blocked_ell: 2^32 entries or more");
    *nnz_out = h[2];
    if (d_out_idx == nullptr) return 1;
    CSAMD_REQUIRE(capacity >= h[2] && (h[2] == 0u || d_out_val != nullptr),
                  "canonical_compressed_from_bucketed_blocked_ell: capacity %llu < nnz %u", capacity, h[2]);
    return canonical_emit(a, vb, grid, d_out_ptr, d_out_idx, d_out_val, rows, h[2], flags, part->cols, s);
}

int csamd_canonical_sliced_ell_from_bucketed_sliced_ell(const csamd_bucketed_sliced_ell *part, unsigned int value_type,
                                                        unsigned int *d_slot_offsets, unsigned int *d_col_idx, void *d_val,
                                                        unsigned long long capacity, unsigned long long *total_slots_out,
                                                        csamd_convert_workspace *ws) {
    CSAMD_REQUIRE(part != nullptr && d_slot_offsets != nullptr && total_slots_out != nullptr && ws != nullptr,
                  "canonical_sliced_ell_from_bucketed_sliced_ell: null argument");
    const int vb = value_type_bytes(value_type);
    CSAMD_REQUIRE(vb == 1 || vb == 2 || vb == 4, "canonical_sliced_ell_from_bucketed_sliced_ell: unsupported value type %u", value_type);
    const unsigned int rows = part->rows, nseg = part->segment_count, ncs = part->canonical_slice_count;
    CSAMD_REQUIRE(part->canonical_slice_row_offsets != nullptr && (ncs == 0u || part->canonical_slice_widths != nullptr),
                  "canonical_sliced_ell_from_bucketed_sliced_ell: no canonical slice tables");
    CSAMD_REQUIRE(rows == 0u || (part->exec_to_canonical_rows != nullptr && nseg != 0u && part->segments != nullptr &&
                                 part->segment_row_offsets != nullptr && part->segment_row_offsets[0] == 0u &&
                                 part->segment_row_offsets[nseg] == rows && ncs != 0u),
                  "canonical_sliced_ell_from_bucketed_sliced_ell: partition without segments, row map or slices");
    if (!ensure_device(ws->device >= 0 ? ws->device : 0)) return 0;
    hipStream_t s = (hipStream_t) ws->stream;
    std::vector<unsigned int> off(ncs + 1u), wid(ncs != 0u ? ncs : 1u), slot(ncs != 0u ? ncs : 1u);
    CSAMD_HIP(hipMemcpyAsync(off.data(), part->canonical_slice_row_offsets, 4ull * (ncs + 1u), hipMemcpyDeviceToHost, s),
              "D2H canonical slice offsets");
    if (ncs != 0u)
        CSAMD_HIP(hipMemcpyAsync(wid.data(), part->canonical_slice_widths, 4ull * ncs, hipMemcpyDeviceToHost, s), "D2H canonical widths");
    std::vector<rs_segment> segs(nseg != 0u ? nseg : 1u);
    for (unsigned int k = 0; k < nseg; ++k) {
        const csamd_sliced_ell_view &v = part->segments[k];
        const unsigned int r0 = part->segment_row_offsets[k], r1 = part->segment_row_offsets[k + 1u];
        CSAMD_REQUIRE(r1 >= r0 && r1 - r0 == v.rows, "canonical_sliced_ell_from_bucketed_sliced_ell: segment %u rows", k);
        CSAMD_REQUIRE(v.rows == 0u || (v.slice_count != 0u && v.slice_row_offsets != nullptr && v.slice_widths != nullptr &&
                                       v.slice_slot_offsets != nullptr),
                      "canonical_sliced_ell_from_bucketed_sliced_ell: segment %u has rows but no slices", k);
        segs[k] = rs_segment{v.slice_row_offsets, v.slice_widths, v.slice_slot_offsets, v.col_idx, (const unsigned char *) v.val, r0,
                             v.rows, v.slice_count, 0u};
    }
    CSAMD_HIP(hipStreamSynchronize(s), "canonical slices sync");
    CSAMD_REQUIRE(off[0] == 0u && off[ncs] == rows, "canonical_sliced_ell_from_bucketed_sliced_ell: canonical slices cover [%u, %u), "
                  "not [0, %u)", off[0], off[ncs], rows);
    unsigned long long total = 0u;
    for (unsigned int q = 0; q < ncs; ++q) {
        CSAMD_REQUIRE(off[q + 1u] >= off[q], "canonical_sliced_ell_from_bucketed_sliced_ell: canonical slice offsets decrease");
        slot[q] = (unsigned int) total;
        total += (unsigned long long) (off[q + 1u] - off[q]) * wid[q];
        CSAMD_REQUIRE(total <= 0xffffffffull, "canonical_sliced_ell_from_bucketed_sliced_ell: 2^32 slots or more");
    }
    *total_slots_out = total;
    if (ncs != 0u)
        CSAMD_HIP(hipMemcpyAsync(d_slot_offsets, slot.data(), 4ull * ncs, hipMemcpyHostToDevice, s), "H2D canonical slot offsets");
    if (d_col_idx == nullptr) {
        CSAMD_HIP(hipStreamSynchronize(s), "canonical slot offsets sync");
        return 1;
    }
    CSAMD_REQUIRE(capacity >= total && (total == 0u || d_val != nullptr),
                  "canonical_sliced_ell_from_bucketed_sliced_ell: capacity %llu < %llu slots", capacity, total);
    const size_t seg_bytes = align_up(segs.size() * sizeof(rs_segment), 256u);
    const size_t row_bytes = align_up((size_t) rows * 4u + 4u, 256u);
    char *scratch = (char *) convert_scratch(ws, seg_bytes + row_bytes + 256u);
    if (scratch == nullptr) return 0;
    rs_segment *d_seg = (rs_segment *) scratch;
    unsigned int *hits = (unsigned int *) (scratch + seg_bytes);
    unsigned int *flags = (unsigned int *) (scratch + seg_bytes + row_bytes);
    CSAMD_HIP(hipMemcpyAsync(d_seg, segs.data(), segs.size() * sizeof(rs_segment), hipMemcpyHostToDevice, s), "H2D sliced segments");
    if (!zero_bytes(hits, row_bytes + 256u, s)) return 0;  // hits and flags
    if (rows != 0u) {
        const rs_args a{d_seg, nseg, rows, ncs, 0u, part->exec_to_canonical_rows, part->canonical_slice_row_offsets,
                        part->canonical_slice_widths, d_slot_offsets, hits, flags};
        const unsigned int grid = grid_for((unsigned long long) rows * 64u, 256u, (unsigned int) device_cu_count() * 16u);
        switch (vb) {
            case 1: rs_rebuild_kernel<1><<<grid, 256, 0, s>>>(a, d_col_idx, (unsigned char *) d_val); break;
            case 2: rs_rebuild_kernel<2><<<grid, 256, 0, s>>>(a, d_col_idx, (unsigned char *) d_val); break;
            default: rs_rebuild_kernel<4><<<grid, 256, 0, s>>>(a, d_col_idx, (unsigned char *) d_val); break;
        }
        CSAMD_HIP(hipGetLastError(), "rs_rebuild_kernel");
    }
    unsigned int h = 0u;
    CSAMD_HIP(hipMemcpyAsync(&h, flags, 4u, hipMemcpyDeviceToHost, s), "D2H flags");
    CSAMD_HIP(hipStreamSynchronize(s), "canonical sliced sync");
    CSAMD_REQUIRE((h & 1u) == 0u, "canonical_sliced_ell_from_bucketed_sliced_ell: exec_to_canonical_rows is not a permutation");
    CSAMD_REQUIRE((h & 2u) == 0u, "canonical_sliced_ell_from_bucketed_sliced_ell: a row's valid slots exceed its canonical width");
    return 1;
}

} // extern "C"
