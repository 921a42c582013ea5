/*
 * cellshard_amd C ABI.
 *
 * MI355X (gfx950) implementation of CellShard's sparse execution hot path:
 *   - masked grouped reductions (runtime/mask_groups.cuh call signatures),
 *   - the row-label grouped {sum, sumsq, nnz} reduction with mean/variance,
 *   - CSR -> Blocked-ELL / Sliced-ELL conversion, Blocked-ELL -> Sliced-ELL,
 *     CSR -> CSC transpose,
 *   - major-axis nnz bucketing (stable nnz order, equal-count and log2 cuts).
 *
 * Every entry point is plain C: POD structs, device pointers, sizes, and a
 * HIP stream handle.  Return convention follows the reference: 1 = ok, 0 =
 * failure; the failure text is available from csamd_last_error() (thread
 * local).  Nothing here throws.
 *
 * All struct layouts below are field-for-field the reference layouts
 * (alignas(16), LP64), with cudaStream_t -> hipStream_t and __half -> the same
 * 2-byte storage, so a caller's existing structs can be passed through.
 * Each declaration names the reference interface it replaces (file:line,
 * paths relative to the CellShard source tree).
 */
#ifndef CELLSHARD_AMD_H
#define CELLSHARD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSAMD_ALIGN16 __attribute__((aligned(16)))

/* Same handle type as hipStream_t (hip_runtime_api.h: typedef struct ihipStream_t *hipStream_t). */
typedef struct ihipStream_t *csamd_stream_t;

/* IEEE binary16 storage, bit-identical to __half.  A C++/HIP includer may
 * define CSAMD_HALF_T (e.g. to __half) before including this header; the ABI
 * is unchanged (the member is a pointer). */
#ifdef CSAMD_HALF_T
typedef CSAMD_HALF_T csamd_half;
#else
typedef uint16_t csamd_half;
#endif

/* Value type codes: include/CellShard/core/real.cuh:44-51. */
enum {
    CSAMD_VALUE_F16 = 3,
    CSAMD_VALUE_F32 = 4,
    CSAMD_VALUE_F64 = 5,
    CSAMD_VALUE_BF16 = 6,
    CSAMD_VALUE_FP8_E4M3 = 7,
    CSAMD_VALUE_FP8_E5M2 = 8
};

enum {
    CSAMD_COMPRESSED_BY_ROW = 0, /* formats/compressed.cuh:11-14 */
    CSAMD_COMPRESSED_BY_COL = 1,
    CSAMD_INVALID_COL = 0xffffffffu /* formats/blocked_ell.cuh:11-14, sliced_ell.cuh:12-15 */
};

#define CSAMD_MAX_MASK_GROUPS 32u /* runtime/mask_groups.cuh:14 */
#define CSAMD_MAX_LABEL_GROUPS 4096u

/* ---------------------------------------------------------------------------
 * Device views: runtime/device/sharded_device.cuh:67-98.
 * The value pointer is declared as csamd_half*; the *_typed entry points
 * reinterpret it according to an explicit value type code.
 * ------------------------------------------------------------------------- */
typedef struct CSAMD_ALIGN16 csamd_compressed_view {
    unsigned int rows;
    unsigned int cols;
    unsigned int nnz;
    unsigned int axis;
    unsigned int *majorPtr;
    unsigned int *minorIdx;
    csamd_half *val;
} csamd_compressed_view;

typedef struct CSAMD_ALIGN16 csamd_blocked_ell_view {
    unsigned int rows;
    unsigned int cols;
    unsigned int nnz;
    unsigned int block_size;
    unsigned int ell_cols;
    unsigned int *blockColIdx;
    csamd_half *val;
} csamd_blocked_ell_view;

typedef struct CSAMD_ALIGN16 csamd_sliced_ell_view {
    unsigned int rows;
    unsigned int cols;
    unsigned int nnz;
    unsigned int slice_count;
    unsigned int slice_rows;
    unsigned int *slice_row_offsets;
    unsigned int *slice_widths;
    unsigned int *slice_slot_offsets;
    unsigned int *col_idx;
    csamd_half *val;
} csamd_sliced_ell_view;

/* ---------------------------------------------------------------------------
 * mask_groups types: runtime/mask_groups.cuh:16-154.
 * ------------------------------------------------------------------------- */
typedef enum csamd_sparse_mask_layout {
    CSAMD_LAYOUT_UNKNOWN = 0,
    CSAMD_LAYOUT_BLOCKED_ELL = 1,
    CSAMD_LAYOUT_SLICED_ELL = 2,
    CSAMD_LAYOUT_COMPRESSED = 3
} csamd_sparse_mask_layout;

typedef struct CSAMD_ALIGN16 csamd_group_mask_config_view {
    unsigned int group_count;
    const char *const *group_names;
    const uint32_t *feature_group_masks;
} csamd_group_mask_config_view;

typedef struct CSAMD_ALIGN16 csamd_row_feature_mask_view {
    const unsigned char *row_keep;
    const unsigned char *feature_keep;
    const unsigned int *row_remap;
    const unsigned int *feature_remap;
} csamd_row_feature_mask_view;

typedef struct CSAMD_ALIGN16 csamd_sparse_group_filter_params {
    float min_total;
    unsigned int min_detected_features;
    float max_group_fraction;
    unsigned int fraction_group_index;
} csamd_sparse_group_filter_params;

typedef struct CSAMD_ALIGN16 csamd_sparse_group_reduce_result {
    unsigned int rows;
    unsigned int group_count;
    float *row_totals;
    unsigned int *detected_features;
    float *max_values;
    unsigned char *row_keep;
    float *group_counts;
    float *group_percentages;
} csamd_sparse_group_reduce_result;

typedef struct CSAMD_ALIGN16 csamd_sparse_group_reduce_workspace {
    int device;
    csamd_stream_t stream;
    int owns_stream;

    unsigned int rows_capacity;
    unsigned int cols_capacity;
    unsigned int group_capacity;
    unsigned int values_capacity;

    void *row_block;
    void *feature_block;
    void *scratch_block;
    size_t scratch_bytes;

    float *row_totals;
    unsigned int *detected_features;
    float *max_values;
    unsigned char *row_keep;
    float *group_counts;
    float *group_percentages;
    uint32_t *feature_group_masks;
} csamd_sparse_group_reduce_workspace;

/* One-GPU fleet: the reference fleet (mask_groups.cuh:78-86) holds a
 * distributed::local_context over every visible GPU; on this build the shard
 * set of one GPU is one slot, so the context collapses to the slot table. */
typedef struct CSAMD_ALIGN16 csamd_sparse_group_reduce_fleet_workspace {
    unsigned int slot_count;
    int *device_ids;
    unsigned int *slots;
    csamd_sparse_group_reduce_workspace *devices;
    csamd_sparse_group_reduce_result *results;
} csamd_sparse_group_reduce_fleet_workspace;

typedef struct CSAMD_ALIGN16 csamd_sparse_group_reduce_fleet_result {
    unsigned int slot_count;
    unsigned int leader_index;
    csamd_sparse_group_reduce_result *slot_results;
} csamd_sparse_group_reduce_fleet_result;

typedef struct CSAMD_ALIGN16 csamd_sparse_group_reduce_fleet_config {
    const int *device_ids;
    unsigned int device_count;
    unsigned int enable_peer_access;
    unsigned int stream_flags;
    const void *ranked_nccl; /* must be NULL: no collectives on one GPU */
} csamd_sparse_group_reduce_fleet_config;

typedef struct CSAMD_ALIGN16 csamd_masked_sparse_reoptimize_config {
    unsigned int output_rows;
    unsigned int output_cols;
    unsigned int requested_bucket_count;
    unsigned int output_block_size;
    csamd_sparse_mask_layout requested_layout;
} csamd_masked_sparse_reoptimize_config;

typedef struct CSAMD_ALIGN16 csamd_masked_sparse_reoptimize_metadata {
    unsigned int kept_rows;
    unsigned int kept_features;
    unsigned int live_nnz;
    unsigned long long bytes_before;
    unsigned long long bytes_after;
    csamd_sparse_mask_layout output_layout;
} csamd_masked_sparse_reoptimize_metadata;

typedef struct CSAMD_ALIGN16 csamd_masked_sparse_reoptimize_result {
    csamd_compressed_view compressed;
    csamd_blocked_ell_view blocked_ell;
    csamd_masked_sparse_reoptimize_metadata metadata;
} csamd_masked_sparse_reoptimize_result;

typedef struct CSAMD_ALIGN16 csamd_masked_sparse_reoptimize_workspace {
    int device;
    csamd_stream_t stream;
    int owns_stream;
    void *private_workspace;
} csamd_masked_sparse_reoptimize_workspace;

typedef struct CSAMD_ALIGN16 csamd_auto_reoptimize_prediction_input {
    float density_change;
    float row_drop_ratio;
    float feature_drop_ratio;
    float layout_padding_change;
    float estimated_rebuild_cost_ms;
    unsigned int expected_reuse_count;
} csamd_auto_reoptimize_prediction_input;

typedef struct CSAMD_ALIGN16 csamd_auto_reoptimize_prediction {
    unsigned int should_reoptimize;
    float estimated_net_gain_ms;
} csamd_auto_reoptimize_prediction;

/* ---------------------------------------------------------------------------
 * Errors and version.
 * ------------------------------------------------------------------------- */
/* Last failure message of the calling thread ("" when none).  Replaces the
 * reference's stderr-only reporting (mask_groups.cu:44-48). */
const char *csamd_last_error(void);
int csamd_abi_version(void);
/* Number of visible HIP devices (0 on a host without a GPU; never fails). */
int csamd_device_count(void);
/* hipSetDevice with the 1/0 convention. */
int csamd_set_device(int device);

/* ---------------------------------------------------------------------------
 * mask_groups lifecycle and reductions (reference ABI).
 * ------------------------------------------------------------------------- */
/* runtime/mask_groups.cuh:198-202 (init), src/runtime/mask_groups.cu:534-538 */
void csamd_group_reduce_init(csamd_sparse_group_reduce_workspace *workspace);
/* mask_groups.cuh:202 (clear), mask_groups.cu:555-563 */
void csamd_group_reduce_clear(csamd_sparse_group_reduce_workspace *workspace);
/* mask_groups.cuh:206 (setup), mask_groups.cu:604-618 */
int csamd_group_reduce_setup(csamd_sparse_group_reduce_workspace *workspace, int device, csamd_stream_t stream);
/* mask_groups.cuh:211-215 (reserve), mask_groups.cu:723-788 */
int csamd_group_reduce_reserve(csamd_sparse_group_reduce_workspace *workspace,
                               unsigned int rows, unsigned int cols, unsigned int values, unsigned int group_count);
/* mask_groups.cuh:217-219, mask_groups.cu:790-809 */
int csamd_upload_feature_group_masks(csamd_sparse_group_reduce_workspace *workspace,
                                     unsigned int cols, const uint32_t *host_masks);
/* mask_groups.cuh:221-226 (BELL overload), mask_groups.cu:811-845 */
int csamd_compute_sparse_group_reduce_blocked_ell(const csamd_blocked_ell_view *src,
                                                  csamd_sparse_group_reduce_workspace *workspace,
                                                  const csamd_group_mask_config_view *groups,
                                                  const csamd_row_feature_mask_view *masks,
                                                  const csamd_sparse_group_filter_params *filter,
                                                  csamd_sparse_group_reduce_result *out);
/* mask_groups.cuh:228-233 (SELL overload), mask_groups.cu:847-881 */
int csamd_compute_sparse_group_reduce_sliced_ell(const csamd_sliced_ell_view *src,
                                                 csamd_sparse_group_reduce_workspace *workspace,
                                                 const csamd_group_mask_config_view *groups,
                                                 const csamd_row_feature_mask_view *masks,
                                                 const csamd_sparse_group_filter_params *filter,
                                                 csamd_sparse_group_reduce_result *out);
/* mask_groups.cuh:235-240 (compressed fallback), mask_groups.cu:883-918 */
int csamd_compute_sparse_group_reduce_compressed(const csamd_compressed_view *src,
                                                 csamd_sparse_group_reduce_workspace *workspace,
                                                 const csamd_group_mask_config_view *groups,
                                                 const csamd_row_feature_mask_view *masks,
                                                 const csamd_sparse_group_filter_params *filter,
                                                 csamd_sparse_group_reduce_result *out);
/* Same three reductions with an explicit value type for src->val
 * (f16, bf16, f32, fp8-e4m3).  Extension: the reference reads __half only. */
int csamd_compute_sparse_group_reduce_compressed_typed(const csamd_compressed_view *src, unsigned int value_type,
                                                       csamd_sparse_group_reduce_workspace *workspace,
                                                       const csamd_group_mask_config_view *groups,
                                                       const csamd_row_feature_mask_view *masks,
                                                       const csamd_sparse_group_filter_params *filter,
                                                       csamd_sparse_group_reduce_result *out);
/* No counterpart in the reference.  compute_sparse_group_reduce over
 * compressed rows for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * streaming kernels as csamd_compute_sparse_group_reduce_compressed_typed,
 * with their scratch in a fixed 512 KiB device array per GPU instead of a
 * workspace, so that no call allocates or frees anything.
 *   - out: caller-owned device arrays for out->rows >= src->rows rows and
 *     out->group_count == groups->group_count groups (row-major
 *     group_counts / group_percentages); written, never rebound.
 *   - groups->feature_group_masks must be a device array when groups are
 *     asked for (there is no workspace to hold uploaded masks).
 *   - refused with an error: more than 122,303 columns (the fixed scratch),
 *     or a source the streaming path does not take (index / value arrays not
 *     aligned to 4-entry units).
 *   - enqueued on `stream` with the calling thread's current device; calls
 *     that overlap on one GPU share its scratch array and race. */
int csamd_compute_sparse_group_reduce_compressed_noalloc(const csamd_compressed_view *src, unsigned int value_type,
                                                         const csamd_group_mask_config_view *groups,
                                                         const csamd_row_feature_mask_view *masks,
                                                         const csamd_sparse_group_filter_params *filter,
                                                         const csamd_sparse_group_reduce_result *out, csamd_stream_t stream);
int csamd_compute_sparse_group_reduce_blocked_ell_typed(const csamd_blocked_ell_view *src, unsigned int value_type,
                                                        csamd_sparse_group_reduce_workspace *workspace,
                                                        const csamd_group_mask_config_view *groups,
                                                        const csamd_row_feature_mask_view *masks,
                                                        const csamd_sparse_group_filter_params *filter,
                                                        csamd_sparse_group_reduce_result *out);
int csamd_compute_sparse_group_reduce_sliced_ell_typed(const csamd_sliced_ell_view *src, unsigned int value_type,
                                                       csamd_sparse_group_reduce_workspace *workspace,
                                                       const csamd_group_mask_config_view *groups,
                                                       const csamd_row_feature_mask_view *masks,
                                                       const csamd_sparse_group_filter_params *filter,
                                                       csamd_sparse_group_reduce_result *out);

/* Fleet, collapsed to the shard set of one GPU.
 * mask_groups.cuh:199,203,207-208 and mask_groups.cu:540-547,565-588,620-705 */
void csamd_fleet_init(csamd_sparse_group_reduce_fleet_workspace *fleet);
void csamd_fleet_clear(csamd_sparse_group_reduce_fleet_workspace *fleet);
int csamd_setup_fleet(csamd_sparse_group_reduce_fleet_workspace *fleet,
                      const csamd_sparse_group_reduce_fleet_config *config);
/* mask_groups.cuh:242-254, mask_groups.cu:920-958 */
int csamd_compute_sparse_group_reduce_fleet_blocked_ell(const csamd_blocked_ell_view *src_by_slot,
                                                        csamd_sparse_group_reduce_fleet_workspace *fleet,
                                                        const csamd_group_mask_config_view *groups,
                                                        const csamd_row_feature_mask_view *masks,
                                                        const csamd_sparse_group_filter_params *filter,
                                                        csamd_sparse_group_reduce_fleet_result *out);
int csamd_compute_sparse_group_reduce_fleet_sliced_ell(const csamd_sliced_ell_view *src_by_slot,
                                                       csamd_sparse_group_reduce_fleet_workspace *fleet,
                                                       const csamd_group_mask_config_view *groups,
                                                       const csamd_row_feature_mask_view *masks,
                                                       const csamd_sparse_group_filter_params *filter,
                                                       csamd_sparse_group_reduce_fleet_result *out);

/* Masked Blocked-ELL -> CSR compaction.
 * mask_groups.cuh:195-196,200,204,209,256-260 and mask_groups.cu:526-532,548-552,590-602,706-721,960-1034 */
int csamd_predict_masked_sparse_reoptimization(const csamd_auto_reoptimize_prediction_input *input,
                                               csamd_auto_reoptimize_prediction *out);
void csamd_reoptimize_init(csamd_masked_sparse_reoptimize_workspace *workspace);
void csamd_reoptimize_clear(csamd_masked_sparse_reoptimize_workspace *workspace);
int csamd_reoptimize_setup(csamd_masked_sparse_reoptimize_workspace *workspace, int device, csamd_stream_t stream);
int csamd_manual_reoptimize_masked_sparse(const csamd_blocked_ell_view *src,
                                          const csamd_row_feature_mask_view *masks,
                                          const csamd_masked_sparse_reoptimize_config *config,
                                          csamd_masked_sparse_reoptimize_workspace *workspace,
                                          csamd_masked_sparse_reoptimize_result *out);

/* ---------------------------------------------------------------------------
 * Row-label grouped reduction (project extension; no reference counterpart:
 * the reference reduces per row with <= 32 feature-group bitmasks).
 *
 * For every kept row r (row_keep[r] != 0 and row_labels[r] < group_count) and
 * every kept feature c (feature_keep[c] != 0), with f = rank of c among the
 * kept features in ascending column order:
 *   sum[g][f]   += v(r,c)          g = row_labels[r]
 *   sumsq[g][f] += v(r,c)^2
 *   nnz[g][f]   += (v(r,c) != 0)
 * Stored zeros and invalid/out-of-range columns contribute nothing.  Partials
 * are combined in FP64; mean = sum/n_g and var = sumsq/n_g - mean^2 (ddof 0),
 * n_g = number of kept rows with label g.  With feature_scale, v(r,c) is the
 * decoded value times scale[c] in FP64; the scale is per feature, so it is
 * applied once per (g, f) after the sums (s*sum, s^2*sumsq, nnz 0 if s == 0).
 * ------------------------------------------------------------------------- */
typedef struct CSAMD_ALIGN16 csamd_label_reduce_query {
    unsigned int group_count;          /* G, 1..CSAMD_MAX_LABEL_GROUPS */
    unsigned int value_type;           /* CSAMD_VALUE_* of src->val */
    const unsigned int *row_labels;    /* device [rows]; label >= G drops the row */
    const unsigned char *row_keep;     /* device [rows] or NULL (all rows) */
    const unsigned char *feature_keep; /* device [cols] or NULL (all features) */
    const float *feature_scale;        /* device [cols] or NULL; multiplies decoded values in FP64 (FP8 absmax scale) */
    unsigned int compute_moments;      /* 1: also fill mean/var */
    unsigned int reserved;
} csamd_label_reduce_query;

typedef struct CSAMD_ALIGN16 csamd_label_reduce_result {
    unsigned int group_count;
    unsigned int kept_features;        /* F: width of every [G x F] output */
    unsigned int kept_rows;
    unsigned int reserved;
    double *sum;                       /* device [G x F] */
    double *sumsq;                     /* device [G x F] */
    unsigned int *nnz;                 /* device [G x F] */
    unsigned int *group_rows;          /* device [G] */
    double *mean;                      /* device [G x F] (compute_moments) */
    double *var;                       /* device [G x F] (compute_moments) */
    unsigned int *kept_feature_ids;    /* device [F]: column of output feature f */
} csamd_label_reduce_result;

typedef struct CSAMD_ALIGN16 csamd_label_reduce_workspace {
    int device;
    csamd_stream_t stream;
    int owns_stream;
    void *private_workspace;
} csamd_label_reduce_workspace;

void csamd_label_reduce_init(csamd_label_reduce_workspace *workspace);
void csamd_label_reduce_clear(csamd_label_reduce_workspace *workspace);
int csamd_label_reduce_setup(csamd_label_reduce_workspace *workspace, int device, csamd_stream_t stream);
int csamd_label_reduce_compressed(const csamd_compressed_view *src, const csamd_label_reduce_query *query,
                                  csamd_label_reduce_workspace *workspace, csamd_label_reduce_result *out);
int csamd_label_reduce_sliced_ell(const csamd_sliced_ell_view *src, const csamd_label_reduce_query *query,
                                  csamd_label_reduce_workspace *workspace, csamd_label_reduce_result *out);
int csamd_label_reduce_blocked_ell(const csamd_blocked_ell_view *src, const csamd_label_reduce_query *query,
                                   csamd_label_reduce_workspace *workspace, csamd_label_reduce_result *out);
/* Batched partitions (project extension; the reference launches once per
 * partition): partitions with the same column count, each compressed
 * (row axis) or sliced-ELL, reduced by one launch.  The query's row arrays
 * (row_labels, row_keep) cover the partitions' rows in partition order;
 * results are those of the row-wise concatenation of the partitions. */
typedef struct CSAMD_ALIGN16 csamd_label_reduce_partition {
    unsigned int layout;               /* CSAMD_LAYOUT_COMPRESSED or CSAMD_LAYOUT_SLICED_ELL */
    unsigned int reserved[3];
    csamd_compressed_view compressed;  /* layout == CSAMD_LAYOUT_COMPRESSED */
    csamd_sliced_ell_view sliced_ell;  /* layout == CSAMD_LAYOUT_SLICED_ELL */
} csamd_label_reduce_partition;

int csamd_label_reduce_batched(const csamd_label_reduce_partition *partitions, unsigned int partition_count,
                               const csamd_label_reduce_query *query, csamd_label_reduce_workspace *workspace,
                               csamd_label_reduce_result *out);

/* Captured calls (project extension; the reference has no graphs): the
 * batched call's launches (descriptor table, label sort, work ranges,
 * reduce, finish) captured once into a HIP graph with buffers of its own,
 * then replayed with one hipGraphLaunch per call.  The views' arrays and the
 * query's arrays are bound at capture: replays read their current contents.
 * The graph holds no memset nodes and no host copies; results alias the
 * graph's buffers until the next launch or destroy. */
typedef struct CSAMD_ALIGN16 csamd_label_reduce_graph {
    int device;
    void *state;
} csamd_label_reduce_graph;

int csamd_label_reduce_graph_capture(const csamd_label_reduce_partition *partitions, unsigned int partition_count,
                                     const csamd_label_reduce_query *query, int device, csamd_label_reduce_graph *graph);
/* Staging -> conversion -> reduce in one captured graph (project extension:
 * the batched-partition pipeline).  Per partition: the layout the reduce
 * streams after conversion, and where its CSR comes from.
 *   - staged_from (optional): a host CSR in pinned memory whose arrays the
 *     graph copies into `compressed`'s device arrays on every launch.
 *   - layout CSAMD_LAYOUT_SLICED_ELL: the graph converts `compressed` into
 *     `sliced_ell` (arrays preallocated by the caller for slice_rows and the
 *     padded slot count the layout tuner reports: slice_row_offsets
 *     [slices + 1], slice_widths / slice_slot_offsets [slices], col_idx / val
 *     [slots]); all such partitions in one batched conversion; the reduce
 *     streams the Sliced-ELL.  A layout that needs more slots than
 *     slots_capacity is not written past it and fails the graph's result.
 *   - layout CSAMD_LAYOUT_BLOCKED_ELL: the graph converts `compressed` into
 *     `blocked_ell` (block_size in {8, 16, 32} and ell_cols = ell_width x
 *     block_size from the tuner; blockColIdx [row blocks x ell_width], val
 *     [rows x ell_cols]) in one batched conversion, as blocked_ell_from_compressed
 *     does, and writes the stored nonzero count to *d_nnz (device); the
 *     reduce streams the CSR (the batched reduce reads row-major layouts).
 *   - layout CSAMD_LAYOUT_COMPRESSED: reduced as staged. */
typedef struct CSAMD_ALIGN16 csamd_pipeline_partition {
    unsigned int layout;
    unsigned int slice_rows;
    unsigned int block_size;
    unsigned int ell_width;
    const struct csamd_host_compressed *staged_from;
    unsigned int *d_nnz;
    unsigned long long slots_capacity;  /* entries sliced_ell.col_idx / val hold */
    csamd_compressed_view compressed;
    csamd_sliced_ell_view sliced_ell;
    csamd_blocked_ell_view blocked_ell;
} csamd_pipeline_partition;

int csamd_label_reduce_pipeline_capture(const csamd_pipeline_partition *partitions, unsigned int partition_count,
                                        const csamd_label_reduce_query *query, int device, csamd_label_reduce_graph *graph);
int csamd_label_reduce_graph_launch(csamd_label_reduce_graph *graph, csamd_stream_t stream);
/* Waits for the last launch, then fills out (kept rows / features included). */
int csamd_label_reduce_graph_result(csamd_label_reduce_graph *graph, csamd_label_reduce_result *out);
void csamd_label_reduce_graph_destroy(csamd_label_reduce_graph *graph);

/* Kernel timing (project extension): with enable != 0, later calls record
 * HIP events around the main reduce kernel (first feature tile) on the
 * workspace stream; off by default, since each event record costs a few
 * microseconds of stream time.  csamd_label_reduce_last_kernel_ms returns
 * the last timed call's kernel duration in milliseconds, 0 if none. */
int csamd_label_reduce_set_timing(csamd_label_reduce_workspace *workspace, int enable);
float csamd_label_reduce_last_kernel_ms(const csamd_label_reduce_workspace *workspace);

/* ---------------------------------------------------------------------------
 * Serving loop (project extension; the reference runtime is one call per
 * query, mask_groups.cuh:159-224, and has no request loop).
 *
 * Shards stay resident in HBM as column-major (CSC) device views registered
 * with the server; their bytes count against the server's device-memory
 * budget, and registration fails past it.  A query names a shard, a feature
 * subset (ascending distinct columns), a group count G and one op:
 *   out[j][g] = sum over kept rows r with label(r) = g of
 *               v(r, features[j])      (CSAMD_SERVE_SUM,   f64)
 *               v(r, features[j])^2    (CSAMD_SERVE_SUMSQ, f64)
 *               [v(r, features[j]) != 0] (CSAMD_SERVE_NNZ,  u32)
 * over stored nonzero entries; output is feature-major [k x G].  Row labels
 * and the row mask come from device arrays [rows] when given, else from the
 * seeds: label(r) = (lowbias32(label_seed ^ (r * 0x9E3779B9)) * G) >> 32 and
 * keep(r) = (lowbias32(mask_seed ^ (r * 0x85EBCA6B)) >> 8) < keep_thr24
 * (keep_thr24 = fraction * 2^24, 2^24 keeps every row).
 *
 * Queries are admitted into output slots (slot_bytes each, counted against
 * the budget) as slots free up; admitted queries are launched together, one
 * workgroup per (query, feature) item taken from a shared counter, with up
 * to max_batches_in_flight batches running on separate streams, so queries
 * that arrive while others run overlap with them on the device.  Each
 * query's result does not depend on what it was batched with (per-(query,
 * feature) LDS accumulation in entry order; exact for integer values).
 * ------------------------------------------------------------------------- */
#define CSAMD_SERVE_SUM 0u
#define CSAMD_SERVE_SUMSQ 1u
#define CSAMD_SERVE_NNZ 2u

typedef struct CSAMD_ALIGN16 csamd_serve_query {
    unsigned int shard;                /* id from csamd_serve_add_shard */
    unsigned int op;                   /* CSAMD_SERVE_* */
    unsigned int group_count;          /* G, 1..CSAMD_SERVE_MAX_GROUPS */
    unsigned int feature_count;        /* k >= 1 */
    const unsigned int *features;      /* device [k], ascending, < cols */
    const unsigned int *row_labels;    /* device [rows] or NULL (seeded labels) */
    const unsigned char *row_keep;     /* device [rows] or NULL (seeded mask) */
    unsigned int label_seed;
    unsigned int mask_seed;
    unsigned int keep_thr24;           /* seeded mask threshold, 0..2^24 */
    unsigned int reserved;
} csamd_serve_query;
#define CSAMD_SERVE_MAX_GROUPS 4096u

typedef struct CSAMD_ALIGN16 csamd_serve_result {
    int status;                        /* 0 queued, 1 running, 2 done, 3 released, -1 unknown ticket */
    unsigned int op;
    unsigned int group_count;
    unsigned int feature_count;
    const void *out;                   /* device [k x G] (f64, or u32 for NNZ); valid until release */
    double latency_us;                 /* submit (or replay arrival) to completion seen by the host */
} csamd_serve_result;

typedef struct CSAMD_ALIGN16 csamd_serve_config {
    int device;
    unsigned int max_slots;            /* output slots (queries admitted at once) */
    unsigned int max_batch_queries;    /* queries per launch */
    unsigned int max_batches_in_flight;/* 1..4 launches running at once (one stream each) */
    unsigned long long slot_bytes;     /* output bytes per slot (k x G x 8 must fit) */
    unsigned long long budget_bytes;   /* device-memory budget: shards + slots */
} csamd_serve_config;

typedef struct CSAMD_ALIGN16 csamd_serve_workspace {
    int device;
    void *state;
} csamd_serve_workspace;

void csamd_serve_init(csamd_serve_workspace *workspace);
void csamd_serve_clear(csamd_serve_workspace *workspace);
int csamd_serve_setup(csamd_serve_workspace *workspace, const csamd_serve_config *config);
/* Register a resident CSC shard (axis CSAMD_COMPRESSED_BY_COL, values of
 * value_type; the caller keeps the device arrays alive).  Fails when it
 * would exceed the budget. */
int csamd_serve_add_shard(csamd_serve_workspace *workspace, const csamd_compressed_view *csc, unsigned int value_type,
                          unsigned int *shard_id);
/* Queue a query; *ticket identifies it.  Does not launch. */
int csamd_serve_submit(csamd_serve_workspace *workspace, const csamd_serve_query *query, unsigned long long *ticket);
/* Retire finished launches and launch queued queries that fit; never
 * blocks.  *completed (may be NULL) = queries finished by this call. */
int csamd_serve_pump(csamd_serve_workspace *workspace, unsigned int *completed);
/* Block until every submitted query is done. */
int csamd_serve_drain(csamd_serve_workspace *workspace);
int csamd_serve_get_result(csamd_serve_workspace *workspace, unsigned long long ticket, csamd_serve_result *out);
/* Free a finished query's output slot. */
int csamd_serve_release(csamd_serve_workspace *workspace, unsigned long long ticket);
/* Replay a schedule: query i is submitted when arrival_us[i] microseconds
 * have passed since the start (ascending), the loop pumps until all are
 * done, and latency_us[i] = completion seen - arrival; *wall_us = start to
 * last completion.  With tickets, results stay held (release them); with
 * tickets NULL each slot is released as soon as its query finishes. */
int csamd_serve_replay(csamd_serve_workspace *workspace, const csamd_serve_query *queries, const double *arrival_us,
                       unsigned int n, unsigned long long *tickets, double *latency_us, double *wall_us);
/* The seeded row labels / row mask of a query as device arrays (labels
 * [rows] u32, keep [rows] u8; either may be NULL).  Asynchronous. */
int csamd_serve_materialize(unsigned int rows, unsigned int group_count, unsigned int label_seed, unsigned int mask_seed,
                            unsigned int keep_thr24, unsigned int *d_labels, unsigned char *d_keep, csamd_stream_t stream);

/* ---------------------------------------------------------------------------
 * Conversion workspace (grow-only scratch + stream), shared by converters,
 * transpose and bucketing entry points that need temporary device memory.
 * ------------------------------------------------------------------------- */
typedef struct CSAMD_ALIGN16 csamd_convert_workspace {
    int device;
    csamd_stream_t stream;
    int owns_stream;
    void *scratch;
    size_t scratch_bytes;
} csamd_convert_workspace;

/* Layout autotuner over a batch of row-major partitions (project
 * extension; one launch for all partitions): per partition, the reference's
 * Blocked-ELL block-size choice among {8, 16, 32}
 * (src/convert/blocked_ell_from_compressed.cuh:157-190) and Sliced-ELL
 * slice-height choice among {32, 64}
 * (src/convert/sliced_ell_from_blocked_ell.cuh:196-221), the format bytes of
 * each (value bytes from value_type), the layout with the fewest bytes, and
 * the histogram of row lengths by log2 bucket (bucket 0: empty rows; bucket
 * b >= 1: lengths in [2^(b-1), 2^b); the last bucket takes the rest).
 * Partitions share the column count (cols <= 524,288). */
typedef struct CSAMD_ALIGN16 csamd_partition_tune_result {
    unsigned int layout;               /* CSAMD_LAYOUT_COMPRESSED / _SLICED_ELL / _BLOCKED_ELL */
    unsigned int bell_block_size;
    unsigned int bell_ell_width;       /* column blocks per row block (W) */
    unsigned int sell_slice_rows;
    unsigned int sell_slice_count;
    unsigned int reserved;
    unsigned long long sell_padded_slots;
    unsigned long long csr_bytes;      /* (rows + 1) x 4 + nnz x (4 + value bytes) */
    unsigned long long bell_bytes;     /* row blocks x W x 4 + rows x W x B x value bytes */
    unsigned long long sell_bytes;     /* (2 slices + 1) x 4 + padded slots x (4 + value bytes) */
    double bell_fill_ratio;
    double sell_score;
    unsigned int row_length_log2_hist[32];
} csamd_partition_tune_result;

int csamd_tune_partitions(const csamd_compressed_view *partitions, unsigned int partition_count, unsigned int value_type,
                          csamd_convert_workspace *workspace, csamd_partition_tune_result *out);

/* Bucketed execution partitions (src/io/csh5/execution_builders_part.hh:1035-1596; the
 * reference builds them on the host from host partitions).  Input: a device view; output:
 * device segment views (each segment one width; Sliced-ELL segments one slice), host segment
 * row offsets (in exec rows), device row maps exec <-> canonical, and the reference's packed
 * byte estimate of the segments (2-byte values).  The library owns the output memory until
 * the clear call.  Blocked-ELL: api.cuh:316-327; Sliced-ELL: api.cuh:394-406. */
typedef struct CSAMD_ALIGN16 csamd_bucketed_blocked_ell {
    unsigned int rows;
    unsigned int cols;
    unsigned int nnz;
    unsigned int segment_count;
    csamd_blocked_ell_view *segments;      /* host array [segment_count] of device views */
    unsigned int *segment_row_offsets;     /* host [segment_count + 1] */
    unsigned int *exec_to_canonical_rows;  /* device [rows] */
    unsigned int *canonical_to_exec_rows;  /* device [rows] */
    unsigned int *exec_to_canonical_cols;  /* device [cols] (identity) */
    unsigned int *canonical_to_exec_cols;  /* device [cols] (identity) */
    unsigned long long bytes;
    void *storage;
} csamd_bucketed_blocked_ell;

typedef struct CSAMD_ALIGN16 csamd_bucketed_sliced_ell {
    unsigned int rows;
    unsigned int cols;
    unsigned int nnz;
    unsigned int segment_count;
    csamd_sliced_ell_view *segments;       /* host array [segment_count] of device views */
    unsigned int *segment_row_offsets;     /* host [segment_count + 1] */
    unsigned int *exec_to_canonical_rows;  /* device [rows] */
    unsigned int *canonical_to_exec_rows;  /* device [rows] */
    unsigned int canonical_slice_count;
    unsigned int reserved;
    unsigned int *canonical_slice_row_offsets;  /* device [canonical_slice_count + 1] */
    unsigned int *canonical_slice_widths;       /* device [canonical_slice_count] */
    unsigned long long bytes;
    void *storage;
} csamd_bucketed_sliced_ell;

/* execution_builders_part.hh:1206-1246 (layout :1035-1098, fill :1100-1204) */
int csamd_build_bucketed_blocked_ell(const csamd_blocked_ell_view *src, unsigned int value_type,
                                     unsigned int requested_bucket_count, csamd_convert_workspace *workspace,
                                     csamd_bucketed_blocked_ell *out);
/* execution_builders_part.hh:1248-1275 */
int csamd_choose_bucket_count_blocked_ell(const csamd_blocked_ell_view *src, csamd_convert_workspace *workspace,
                                          unsigned int *bucket_count, unsigned long long *bucketed_bytes);
void csamd_bucketed_blocked_ell_clear(csamd_bucketed_blocked_ell *m);
/* execution_builders_part.hh:1535-1568 (layout DP :1312-1429, fill :1498-1533) */
int csamd_build_bucketed_sliced_ell(const csamd_sliced_ell_view *src, unsigned int value_type,
                                    unsigned int requested_bucket_count, csamd_convert_workspace *workspace,
                                    csamd_bucketed_sliced_ell *out);
/* execution_builders_part.hh:1570-1596 */
int csamd_choose_bucket_count_sliced_ell(const csamd_sliced_ell_view *src, csamd_convert_workspace *workspace,
                                         unsigned int *bucket_count, unsigned long long *bucketed_bytes);
void csamd_bucketed_sliced_ell_clear(csamd_bucketed_sliced_ell *m);

/* ---------------------------------------------------------------------------
 * CSPACK01 shard packs (docs/SPEC_CSPACK_V1.md:87-155): one file per shard holding its bucketed
 * execution partitions.  Container: "CSPACK01", shard_id u64, partition_count u64,
 * partition_offsets[partition_count] u64, then one blob per partition.  Blob layouts:
 * Blocked-ELL execution_builders_part.hh:2116-2132, Sliced-ELL execution_runtime_part.hh:168-195;
 * each segment is a raw payload (src/io/pack/packfile.cu:351-369 Blocked-ELL, :597-621
 * Sliced-ELL: 1-byte format tag, rows, cols, nnz, then the arrays, unpadded).  Values are
 * stored at the value type's width; files the reference reads carry 2-byte (F16) values.
 *
 * Staging goes straight to HBM: the partition's byte range is read into pinned host memory by
 * parallel readers, each piece copied to a device staging buffer as soon as it lands, and one
 * kernel moves every array from its unaligned place in the blob to an aligned place in the
 * partition's single device allocation.  Headers are checked on the host before any device
 * work; the result is the same device structure the bucketed builders produce (freed by
 * csamd_bucketed_*_clear), with `bytes` set to the segments' packed bytes.
 * ------------------------------------------------------------------------- */
#define CSAMD_CSPACK_ANY_SHARD 0xffffffffffffffffull

typedef struct CSAMD_ALIGN16 csamd_cspack_file {
    unsigned long long shard_id;
    unsigned long long partition_count;
    unsigned long long file_bytes;
    unsigned long long *partition_offsets;  /* host [partition_count] */
    unsigned long long *partition_bytes;    /* host [partition_count]: bytes up to the next blob or the file end */
    void *impl;
} csamd_cspack_file;

typedef struct CSAMD_ALIGN16 csamd_cspack_workspace {
    int device;
    unsigned int reader_threads;  /* parallel file readers while staging (0: 8) */
    csamd_stream_t stream;
    int owns_stream;
    unsigned int chunk_bytes;     /* read / copy piece size (0: 8 MiB) */
    void *private_workspace;
} csamd_cspack_workspace;

/* init / clear / setup (a null stream creates an owned non-blocking one).  The workspace keeps
 * its pinned host buffers and device staging buffer between calls. */
void csamd_cspack_init(csamd_cspack_workspace *ws);
void csamd_cspack_clear(csamd_cspack_workspace *ws);
int csamd_cspack_setup(csamd_cspack_workspace *ws, int device, csamd_stream_t stream);
/* materialize_blocked_ell_cspack (execution_runtime_part.hh:643-733): writes `path` (through a
 * temporary file renamed into place) from device-resident execution partitions. */
int csamd_cspack_write_blocked_ell(const char *path, unsigned long long shard_id, const csamd_bucketed_blocked_ell *parts,
                                   unsigned int part_count, unsigned int value_type, csamd_cspack_workspace *ws);
/* materialize_sliced_ell_cspack (execution_runtime_part.hh:735-796) */
int csamd_cspack_write_sliced_ell(const char *path, unsigned long long shard_id, const csamd_bucketed_sliced_ell *parts,
                                  unsigned int part_count, unsigned int value_type, csamd_cspack_workspace *ws);
/* The header checks of load_blocked_pack_partition_from_cspack (execution_runtime_part.hh:569-604):
 * magic, the shard id (unless CSAMD_CSPACK_ANY_SHARD) and every offset inside the file. */
int csamd_cspack_open(const char *path, unsigned long long shard_id, csamd_cspack_file *out);
void csamd_cspack_close(csamd_cspack_file *f);
/* load_blocked_pack_partition_from_cspack (execution_runtime_part.hh:569-604) with
 * load_execution_partition_blob (execution_builders_part.hh:2146-2197), staged to the device. */
int csamd_cspack_stage_blocked_ell(const csamd_cspack_file *f, unsigned long long local_partition, unsigned int value_type,
                                   csamd_cspack_workspace *ws, csamd_bucketed_blocked_ell *out);
/* load_sliced_bucketed_partition_from_cspack (execution_runtime_part.hh:606-641) with
 * load_sliced_execution_partition_blob (:197-258), staged to the device. */
int csamd_cspack_stage_sliced_ell(const csamd_cspack_file *f, unsigned long long local_partition, unsigned int value_type,
                                  csamd_cspack_workspace *ws, csamd_bucketed_sliced_ell *out);
/* fetch_cached_shard for Blocked-ELL (src/io/csh5/runtime/fetch.cc:20-36): every partition of
 * the file into out[0 .. partition_count), reading partition p + 1 while p is copied and
 * unpacked.  On failure every out[] is cleared. */
int csamd_cspack_stage_shard_blocked_ell(const csamd_cspack_file *f, unsigned int value_type, csamd_cspack_workspace *ws,
                                         csamd_bucketed_blocked_ell *out);
/* bucketed_partition_to_canonical_coo (execution_builders_part.hh:909-968), as a device CSR in
 * canonical row order: the stored values that are not +-0 at exec columns < cols, in slot order
 * within a row, columns mapped through exec_to_canonical_cols (identity when null).  The reference
 * then builds the canonical Blocked-ELL with blocked_ell_from_coo (blocked_ell_from_compressed.cuh
 * :349-434, used by load_blocked_ell_part_from_cspack, execution_runtime_part.hh:496-528);
 * csamd_blocked_ell_from_compressed on this CSR gives the same Blocked-ELL.  d_out_ptr [rows + 1]
 * is written and *nnz_out set; with d_out_idx null the call stops there (sizing), else capacity
 * must hold nnz.  ptr_ready != 0 skips the counting pass: d_out_ptr then holds a sizing call's
 * result for the same partition.  Fails when exec_to_canonical_rows is not a permutation or a
 * mapped column is >= cols; writes never leave [0, nnz) of the outputs.  Synchronizes the
 * workspace stream. */
int csamd_canonical_compressed_from_bucketed_blocked_ell(const csamd_bucketed_blocked_ell *part, unsigned int value_type,
                                                         unsigned int *d_out_ptr, int ptr_ready, unsigned int *d_out_idx,
                                                         void *d_out_val, unsigned long long capacity, unsigned long long *nnz_out,
                                                         csamd_convert_workspace *ws);
/* rebuild_canonical_sliced_partition (execution_runtime_part.hh:275-318, used by
 * load_sliced_ell_part_from_cspack, :464-494): the canonical Sliced-ELL of a bucketed Sliced-ELL
 * partition in the canonical slice layout the partition carries (canonical_slice_row_offsets /
 * widths).  Each canonical row holds its exec row's valid slots (column != invalid) in slot order,
 * then invalid slots with zero values.  d_slot_offsets [canonical_slice_count] is written and
 * *total_slots_out set; with d_col_idx null the call stops there (sizing), else capacity must hold
 * the slots.  Fails when a row's valid slots exceed its canonical width (as the reference does), when
 * exec_to_canonical_rows is not a permutation, or when the canonical slices do not cover the rows.
 * Segments may hold several slices (the reference reads only the first slice's width).
 * Synchronizes the workspace stream. */
int csamd_canonical_sliced_ell_from_bucketed_sliced_ell(const csamd_bucketed_sliced_ell *part, unsigned int value_type,
                                                        unsigned int *d_slot_offsets, unsigned int *d_col_idx, void *d_val,
                                                        unsigned long long capacity, unsigned long long *total_slots_out,
                                                        csamd_convert_workspace *ws);



void csamd_convert_init(csamd_convert_workspace *ws);
void csamd_convert_clear(csamd_convert_workspace *ws);
int csamd_convert_setup(csamd_convert_workspace *ws, int device, csamd_stream_t stream);

/* convert/blocked_ell_from_compressed.cuh:16-20 */
typedef struct CSAMD_ALIGN16 csamd_blocked_ell_tune_result {
    unsigned int block_size;
    double fill_ratio;
    size_t padded_bytes;
} csamd_blocked_ell_tune_result;

/* convert/sliced_ell_from_blocked_ell.cuh:18-26 */
typedef struct CSAMD_ALIGN16 csamd_sliced_ell_tune_result {
    unsigned int slice_rows;
    unsigned int slice_count;
    size_t padded_slots;
    size_t padded_bytes;
    double storage_ratio;
    double slice_penalty;
    double score;
} csamd_sliced_ell_tune_result;

/* Blocked-ELL shape of a device CSR (row axis): W = max unique column blocks
 * per row block, total = sum over row blocks.  Synchronous.
 * convert/blocked_ell_from_compressed.cuh:32-91 (count_blocked_ell_shape_) */
int csamd_blocked_ell_shape(const csamd_compressed_view *src, unsigned int block_size,
                            csamd_convert_workspace *ws, unsigned int *out_ell_width,
                            unsigned long long *out_total_blocks, double *out_fill_ratio);
/* convert/blocked_ell_from_compressed.cuh:157-190; padded bytes priced at
 * value_bytes per stored value (2 in the reference). */
int csamd_choose_blocked_ell_block_size(const csamd_compressed_view *src, const unsigned int *candidates,
                                        unsigned int candidate_count, unsigned int value_bytes,
                                        csamd_convert_workspace *ws, csamd_blocked_ell_tune_result *out);
/* convert/blocked_ell_from_compressed.cuh:258-347.  dst->block_size and
 * dst->ell_cols (= W * block_size from csamd_blocked_ell_shape) must be set,
 * dst->blockColIdx [row_blocks x W] and dst->val [rows x ell_cols] allocated
 * by the caller.  Writes both, sets dst->nnz (count of nonzero stored
 * values; duplicate (row, col) keeps the last value).  Synchronous. */
int csamd_blocked_ell_from_compressed(const csamd_compressed_view *src, unsigned int value_type,
                                      csamd_blocked_ell_view *dst, csamd_convert_workspace *ws);

/* Uniform Sliced-ELL layout of a device CSR/CSC (axis honoured):
 * slice_row_offsets [S+1], slice_widths [S], slice_slot_offsets [S]
 * (device, caller-allocated, S = ceil(rows / clamp(slice_rows,1,rows))).
 * Returns the total slot count (fails if it does not fit u32, as the
 * reference's slot offsets are u32).  Synchronous.
 * convert/sliced_ell_from_blocked_ell.cuh:111-142 and
 * runtime/device/sharded_device.cuh:609-620 (slot offsets) */
int csamd_sliced_ell_layout(const csamd_compressed_view *src, unsigned int slice_rows,
                            unsigned int *d_slice_row_offsets, unsigned int *d_slice_widths,
                            unsigned int *d_slice_slot_offsets, csamd_convert_workspace *ws,
                            unsigned int *out_slice_count, unsigned long long *out_total_slots);
/* Same layout for a device Blocked-ELL source (row nnz = nonzero stored
 * values at valid columns).  convert/sliced_ell_from_blocked_ell.cuh:36-62 */
int csamd_sliced_ell_layout_from_blocked_ell(const csamd_blocked_ell_view *src, unsigned int value_type,
                                             unsigned int slice_rows, unsigned int *d_slice_row_offsets,
                                             unsigned int *d_slice_widths, unsigned int *d_slice_slot_offsets,
                                             csamd_convert_workspace *ws, unsigned int *out_slice_count,
                                             unsigned long long *out_total_slots);
/* convert/sliced_ell_from_blocked_ell.cuh:196-221 (choose_sliced_ell_slice_rows)
 * over the row nnz of a device CSR (sliced_ell_from_compressed.cuh:86-94). */
int csamd_choose_sliced_ell_slice_rows(const csamd_compressed_view *src, const unsigned int *candidates,
                                       unsigned int candidate_count, csamd_convert_workspace *ws,
                                       csamd_sliced_ell_tune_result *out);
/* convert/sliced_ell_from_compressed.cuh:96-164 (CSR and CSC sources).
 * dst layout arrays must hold the csamd_sliced_ell_layout result; col_idx
 * and val [total_slots] caller-allocated.  Padding is written as
 * CSAMD_INVALID_COL / 0.  Sets dst->nnz = src->nnz.  Asynchronous. */
int csamd_sliced_ell_from_compressed(const csamd_compressed_view *src, unsigned int value_type,
                                     csamd_sliced_ell_view *dst, csamd_convert_workspace *ws);
/* convert/sliced_ell_from_blocked_ell.cuh:233-277.  Asynchronous. */
int csamd_sliced_ell_from_blocked_ell(const csamd_blocked_ell_view *src, unsigned int value_type,
                                      csamd_sliced_ell_view *dst, csamd_convert_workspace *ws);

/* CSR -> CSC transpose (or CSC -> CSR): out_ptr[uDim+1], out_idx, out_val.
 * Canonical order: ascending compressed-axis index inside every output
 * segment (what cuSPARSE csr2cscEx2 and scipy produce).
 * convert/compressed_transpose_raw.cuh:14-122 */
size_t csamd_compressed_transpose_scratch_bytes(unsigned int cDim, unsigned int uDim, unsigned int nnz);
int csamd_build_compressed_transpose_raw(unsigned int cDim, unsigned int uDim, unsigned int nnz,
                                         const unsigned int *d_cAxPtr, const unsigned int *d_uAxIdx,
                                         const csamd_half *d_val, unsigned int *d_out_uAxPtr,
                                         unsigned int *d_heads, unsigned int *d_out_cAxIdx,
                                         csamd_half *d_out_val, void *d_scan_tmp, size_t scan_bytes,
                                         csamd_stream_t stream);
int csamd_build_compressed_transpose_raw_typed(unsigned int cDim, unsigned int uDim, unsigned int nnz,
                                               const unsigned int *d_cAxPtr, const unsigned int *d_uAxIdx,
                                               const void *d_val, unsigned int value_type,
                                               unsigned int *d_out_uAxPtr, unsigned int *d_out_cAxIdx,
                                               void *d_out_val, void *d_scan_tmp, size_t scan_bytes,
                                               csamd_stream_t stream);

/* Major-axis nnz bucketing.  bucket/major_nnz_raw.cuh:12-19 (plan view) */
typedef struct CSAMD_ALIGN16 csamd_major_nnz_bucket_plan_view {
    unsigned int major_dim;
    unsigned int bucket_count;
    const unsigned int *major_order;
    const unsigned int *major_nnz_sorted;
    const unsigned int *bucket_offsets;
    const unsigned int *inverse_major_order;
} csamd_major_nnz_bucket_plan_view;

/* bucket/major_nnz_raw.cuh:27-32 */
unsigned int csamd_clamp_bucket_count(unsigned int major_dim, unsigned int requested);
/* bucket/major_nnz_raw.cuh:34-48 / 50-62 */
int csamd_major_nnz_bucket_sort_scratch_bytes(unsigned int major_dim, size_t *out_bytes);
int csamd_major_nnz_bucket_scan_scratch_bytes(unsigned int major_dim, size_t *out_bytes);
/* bucket/major_nnz_raw.cuh:64-120: nnz per major, stable ascending order,
 * equal-count offsets [bucket_count + 1]. */
int csamd_build_major_nnz_bucket_plan_raw(const unsigned int *d_major_ptr, unsigned int major_dim,
                                          unsigned int *d_major_nnz, unsigned int *d_major_nnz_sorted,
                                          unsigned int *d_major_order_in, unsigned int *d_major_order_out,
                                          unsigned int *d_bucket_offsets, unsigned int requested_bucket_count,
                                          void *d_sort_tmp, size_t sort_tmp_bytes, csamd_stream_t stream);
/* log2 nnz buckets over an ascending major_nnz_sorted: bucket of n is
 * clamp(bit_width(n) - min_log2 - 1 + 1, 0, bucket_count - 1) with
 * bucket 0 = n < 2^(min_log2); offsets [bucket_count + 1].  Project rule. */
int csamd_fill_log2_bucket_offsets_raw(const unsigned int *d_major_nnz_sorted, unsigned int major_dim,
                                       unsigned int bucket_count, unsigned int min_log2,
                                       unsigned int *d_bucket_offsets, csamd_stream_t stream);
/* bucket/major_nnz_raw.cuh:188-250 */
int csamd_rebuild_compressed_major_order_raw(const unsigned int *d_src_major_ptr, const unsigned int *d_src_minor_idx,
                                             const void *d_src_val, unsigned int value_type, unsigned int major_dim,
                                             const unsigned int *d_major_order, unsigned int *d_dst_major_ptr,
                                             unsigned int *d_dst_minor_idx, void *d_dst_val,
                                             unsigned int *d_inverse_major_order, void *d_scan_tmp,
                                             size_t scan_tmp_bytes, csamd_stream_t stream);

/* Workspace-owned major-nnz bucketing (bucket/operators/major_nnz.cuh:11-213):
 * packed grow-only device blocks for the per-major counts, the sort and
 * scan scratch and the rebuilt compressed arrays.  Result views alias the
 * workspace until its next call; work runs on the workspace stream.  The
 * fields up to d_scan_tmp are the reference's; owns_stream and
 * value_bytes_capacity (rebuilt value bytes per entry reserved: the reference
 * stores 2-byte values only) are extensions. */
typedef struct CSAMD_ALIGN16 csamd_compressed_major_bucket_workspace {
    int device;
    csamd_stream_t stream;
    unsigned int major_capacity;
    unsigned int nnz_capacity;
    unsigned int bucket_capacity;
    size_t sort_capacity;
    size_t scan_capacity;
    void *d_major_block;
    void *d_rebuilt_block;
    unsigned int *d_major_nnz;
    unsigned int *d_major_nnz_sorted;
    unsigned int *d_major_order_in;
    unsigned int *d_major_order_out;
    unsigned int *d_inverse_major_order;
    unsigned int *d_bucket_offsets;
    unsigned int *d_rebuilt_major_ptr;
    unsigned int *d_rebuilt_minor_idx;
    void *d_rebuilt_val;
    void *d_sort_tmp;
    void *d_scan_tmp;
    int owns_stream;
    unsigned int value_bytes_capacity;
} csamd_compressed_major_bucket_workspace;

/* major_nnz.cuh:40-43 */
typedef struct CSAMD_ALIGN16 csamd_compressed_major_bucket_result {
    csamd_compressed_view rebuilt;
    csamd_major_nnz_bucket_plan_view plan;
} csamd_compressed_major_bucket_result;

/* major_nnz.cuh:45-65: init / clear / setup (stream 0: an owned non-blocking stream) */
void csamd_major_bucket_init(csamd_compressed_major_bucket_workspace *ws);
void csamd_major_bucket_clear(csamd_compressed_major_bucket_workspace *ws);
int csamd_major_bucket_setup(csamd_compressed_major_bucket_workspace *ws, int device, csamd_stream_t stream);
/* major_nnz.cuh:67-143 (2-byte values); _typed reserves value_type's bytes per entry */
int csamd_major_bucket_reserve(csamd_compressed_major_bucket_workspace *ws, unsigned int major_dim, unsigned int nnz,
                               unsigned int bucket_count);
int csamd_major_bucket_reserve_typed(csamd_compressed_major_bucket_workspace *ws, unsigned int major_dim, unsigned int nnz,
                                     unsigned int bucket_count, unsigned int value_type);
/* major_nnz.cuh:145-177: stable ascending nnz order and equal-count bucket
 * offsets of src's major axis; out aliases the workspace (inverse order null). */
int csamd_major_bucket_build_plan(const csamd_compressed_view *src, unsigned int requested_bucket_count,
                                  csamd_compressed_major_bucket_workspace *ws, csamd_major_nnz_bucket_plan_view *out);
/* major_nnz.cuh:179-202: src's segments in plan order (rebuilt row i = src
 * row major_order[i]) and the inverse order; 2-byte values, or value_type's. */
int csamd_major_bucket_rebuild(const csamd_compressed_view *src, const csamd_major_nnz_bucket_plan_view *plan,
                               csamd_compressed_major_bucket_workspace *ws, csamd_compressed_major_bucket_result *out);
int csamd_major_bucket_rebuild_typed(const csamd_compressed_view *src, unsigned int value_type,
                                     const csamd_major_nnz_bucket_plan_view *plan, csamd_compressed_major_bucket_workspace *ws,
                                     csamd_compressed_major_bucket_result *out);
/* major_nnz.cuh:204-210: build_plan then rebuild */
int csamd_build_bucketed_major_view(const csamd_compressed_view *src, unsigned int requested_bucket_count,
                                    csamd_compressed_major_bucket_workspace *ws, csamd_compressed_major_bucket_result *out);
int csamd_build_bucketed_major_view_typed(const csamd_compressed_view *src, unsigned int value_type,
                                          unsigned int requested_bucket_count, csamd_compressed_major_bucket_workspace *ws,
                                          csamd_compressed_major_bucket_result *out);

/* Blocked-ELL row-block bucket plan: key = (valid slots << 32) | hash of the
 * valid block columns, stable ascending order.  Writes order and the sorted
 * widths (device, [row_block_count]).  bucket/operators/blocked_ell_bucket.cuh:77-96,149-210 */
int csamd_blocked_ell_bucket_plan_raw(const unsigned int *d_block_col_idx, unsigned int row_block_count,
                                      unsigned int width_blocks, unsigned int *d_order_out,
                                      unsigned int *d_width_sorted, unsigned long long *d_keys_sorted,
                                      csamd_convert_workspace *ws);

/* Masked Blocked-ELL -> CSR compaction with row / column remaps, and the same
 * followed by the major-nnz bucket plan and rebuild (the step after the
 * reduce: QC keep -> compact -> rebucket).
 * convert/filtered_blocked_ell_to_compressed.cuh:125-415; bucketed view:
 * bucket/operators/major_nnz.cuh:40-43,189-210.  An entry survives when its
 * row is kept and maps below output_rows, its block column is valid, col <
 * cols, its value is nonzero, its column is kept and maps below output_cols;
 * output rows keep source ell order.  Row i of the output is [ptr[i],
 * ptr[i + 1]) (the reference's scan shifts rows by one place; see the
 * integration notes).  Outputs alias workspace memory until the next call.
 * Two kept rows mapped to one output row fail the call. */
typedef struct CSAMD_ALIGN16 csamd_filtered_blocked_ell_stats {
    unsigned int kept_rows;          /* = output_rows, as the reference reports it (:362) */
    unsigned int output_rows;
    unsigned int output_cols;
    unsigned int live_nnz;
    unsigned long long kept_row_slots;  /* kept_rows x ell_cols */
    unsigned long long dead_slots;
    double live_fill_ratio;
    size_t dead_value_bytes;
} csamd_filtered_blocked_ell_stats;

typedef struct CSAMD_ALIGN16 csamd_filtered_blocked_ell_result {
    csamd_compressed_view filtered;
    csamd_compressed_major_bucket_result bucketed;
    csamd_filtered_blocked_ell_stats stats;
    int has_bucketed;
} csamd_filtered_blocked_ell_result;

typedef struct CSAMD_ALIGN16 csamd_filtered_blocked_ell_workspace {
    int device;
    csamd_stream_t stream;
    int owns_stream;
    void *private_workspace;
} csamd_filtered_blocked_ell_workspace;

/* :143-178 init / clear / setup (a null stream creates an owned non-blocking one) */
void csamd_filtered_blocked_ell_init(csamd_filtered_blocked_ell_workspace *ws);
void csamd_filtered_blocked_ell_clear(csamd_filtered_blocked_ell_workspace *ws);
int csamd_filtered_blocked_ell_setup(csamd_filtered_blocked_ell_workspace *ws, int device, csamd_stream_t stream);
/* :264-373 filter_blocked_ell_to_compressed (value_type: F16 as in the reference; F32 / BF16 / FP8 too) */
int csamd_filter_blocked_ell_to_compressed(const csamd_blocked_ell_view *src, unsigned int value_type, unsigned int output_rows,
                                           unsigned int output_cols, const unsigned char *d_keep_rows,
                                           const unsigned char *d_keep_cols, const unsigned int *d_row_remap,
                                           const unsigned int *d_col_remap, csamd_filtered_blocked_ell_workspace *ws,
                                           csamd_filtered_blocked_ell_result *out);
/* :375-412 build_bucketed_filtered_blocked_ell_major_view */
int csamd_build_bucketed_filtered_blocked_ell_major_view(const csamd_blocked_ell_view *src, unsigned int value_type,
                                                         unsigned int output_rows, unsigned int output_cols,
                                                         const unsigned char *d_keep_rows, const unsigned char *d_keep_cols,
                                                         const unsigned int *d_row_remap, const unsigned int *d_col_remap,
                                                         unsigned int requested_bucket_count,
                                                         csamd_filtered_blocked_ell_workspace *ws,
                                                         csamd_filtered_blocked_ell_result *out);

/* Repack of one shard's Blocked-ELL partitions: masked compaction to CSR
 * (kept rows counted over the shard, slots priced at each kept row's own
 * ell width), the major-nnz bucket plan and rebuild, then Blocked-ELL of the
 * requested block size over the bucketed rows.
 * repack/operators/sharded_blocked_ell.cuh:17-47 (config, stats, result,
 * workspace), :113-150 (init / clear / setup), :393-723
 * (build_repacked_shard_blocked_ell).  The reference reads the shard's
 * partitions from a sharded_device state; here the caller passes the
 * partitions' views in shard order (host array of views over device
 * arrays); keep_rows / row_remap index the shard's rows.  Row pointers as
 * in the filter above.  Outputs alias workspace memory. */
typedef struct CSAMD_ALIGN16 csamd_sharded_blocked_ell_repack_config {
    unsigned int output_rows;
    unsigned int output_cols;
    unsigned int block_size;
    unsigned int requested_bucket_count;
} csamd_sharded_blocked_ell_repack_config;

typedef struct CSAMD_ALIGN16 csamd_sharded_blocked_ell_repack_stats {
    unsigned int kept_rows;
    unsigned int output_rows;
    unsigned int output_cols;
    unsigned int live_nnz;
    unsigned long long kept_row_slots;
    unsigned long long dead_slots;
    double live_fill_ratio;
    size_t dead_value_bytes;
    unsigned int ell_width;
} csamd_sharded_blocked_ell_repack_stats;

typedef struct CSAMD_ALIGN16 csamd_sharded_blocked_ell_repack_result {
    csamd_compressed_view filtered;
    csamd_compressed_major_bucket_result bucketed;
    csamd_blocked_ell_view rebuilt;
    csamd_major_nnz_bucket_plan_view bucket_plan;
    csamd_sharded_blocked_ell_repack_stats stats;
    unsigned long shard_id;
} csamd_sharded_blocked_ell_repack_result;

typedef struct CSAMD_ALIGN16 csamd_sharded_blocked_ell_repack_workspace {
    int device;
    csamd_stream_t stream;
    int owns_stream;
    void *private_workspace;
} csamd_sharded_blocked_ell_repack_workspace;

void csamd_repack_init(csamd_sharded_blocked_ell_repack_workspace *ws);
void csamd_repack_clear(csamd_sharded_blocked_ell_repack_workspace *ws);
int csamd_repack_setup(csamd_sharded_blocked_ell_repack_workspace *ws, int device, csamd_stream_t stream);
int csamd_build_repacked_blocked_ell(const csamd_blocked_ell_view *parts, unsigned int part_count, unsigned int value_type,
                                     const csamd_sharded_blocked_ell_repack_config *cfg, const unsigned char *d_keep_rows,
                                     const unsigned char *d_keep_cols, const unsigned int *d_row_remap,
                                     const unsigned int *d_col_remap, unsigned long shard_id,
                                     csamd_sharded_blocked_ell_repack_workspace *ws, csamd_sharded_blocked_ell_repack_result *out);

/* ---------------------------------------------------------------------------
 * Host matrix structs (reference layouts) and the host-struct converters.
 * formats/compressed.cuh:24-37, formats/blocked_ell.cuh:15-26,
 * formats/sliced_ell.cuh:16-28.  Output structs are (re)allocated with one
 * packed malloc block (`storage`), exactly as sparse::allocate does, so a
 * caller's sparse::clear()/free() releases them.  flags bit 0 marks storage
 * allocated pinned with hipHostMalloc (released with hipHostFree); existing
 * memory is never registered after the fact.
 * ------------------------------------------------------------------------- */
typedef struct CSAMD_ALIGN16 csamd_host_compressed {
    unsigned int rows;
    unsigned int cols;
    unsigned int nnz;
    unsigned int axis;
    unsigned int flags;
    void *storage;
    unsigned int *majorPtr;
    unsigned int *minorIdx;
    csamd_half *val;
} csamd_host_compressed;

typedef struct CSAMD_ALIGN16 csamd_host_blocked_ell {
    unsigned int rows;
    unsigned int cols;
    unsigned int nnz;
    unsigned int block_size;
    unsigned int ell_cols;
    unsigned int flags;
    void *storage;
    unsigned int *blockColIdx;
    csamd_half *val;
} csamd_host_blocked_ell;

typedef struct CSAMD_ALIGN16 csamd_host_sliced_ell {
    unsigned int rows;
    unsigned int cols;
    unsigned int nnz;
    unsigned int slice_count;
    unsigned int flags;
    void *storage;
    unsigned int *slice_row_offsets;
    unsigned int *slice_widths;
    unsigned int *col_idx;
    csamd_half *val;
} csamd_host_sliced_ell;

/* formats/triplet.cuh:14-25 (sparse::coo): host COO, fp16 values. */
typedef struct CSAMD_ALIGN16 csamd_host_coo {
    unsigned int rows;
    unsigned int cols;
    unsigned int nnz;
    void *storage;
    unsigned int *rowIdx;
    unsigned int *colIdx;
    csamd_half *val;
} csamd_host_coo;

/* Release host storage (the formats/ headers' clear()); resets metadata. */
void csamd_host_compressed_clear(csamd_host_compressed *m);
void csamd_host_blocked_ell_clear(csamd_host_blocked_ell *m);
void csamd_host_sliced_ell_clear(csamd_host_sliced_ell *m);

/* convert/blocked_ell_from_compressed.cuh:157-190 */
int csamd_host_choose_blocked_ell_block_size(const csamd_host_compressed *src, const unsigned int *candidates,
                                             unsigned int candidate_count, csamd_blocked_ell_tune_result *out);
/* convert/blocked_ell_from_compressed.cuh:258-347 */
int csamd_host_blocked_ell_from_compressed(const csamd_host_compressed *src, unsigned int block_size,
                                           csamd_host_blocked_ell *dst);
/* convert/blocked_ell_from_compressed.cuh:449-461 */
int csamd_host_blocked_ell_from_compressed_auto(const csamd_host_compressed *src, const unsigned int *candidates,
                                                unsigned int candidate_count, csamd_host_blocked_ell *dst,
                                                csamd_blocked_ell_tune_result *picked);
/* convert/blocked_ell_from_coo_cuda.cuh:612-661: the tuner on a host COO (columns remapped
 * through feature_to_global[src->cols] when it is not NULL; padded bytes at 2-byte values). */
int csamd_choose_blocked_ell_block_size_from_coo(const csamd_host_coo *src, unsigned int cols,
                                                 const unsigned int *feature_to_global, const unsigned int *candidates,
                                                 unsigned int candidate_count, int device, csamd_blocked_ell_tune_result *out);
/* convert/blocked_ell_from_coo_cuda.cuh:663-861: host COO -> host Blocked-ELL (dst reallocated
 * with malloc).  Repeated (row, column) entries keep their last value in input order; entries
 * outside rows x cols (after the remap) are rejected. */
int csamd_blocked_ell_from_coo(const csamd_host_coo *src, unsigned int cols, const unsigned int *feature_to_global,
                               unsigned int block_size, csamd_host_blocked_ell *dst, int device);
/* convert/blocked_ell_from_coo_cuda.cuh:863-966 */
int csamd_blocked_ell_from_coo_auto(const csamd_host_coo *src, unsigned int cols, const unsigned int *feature_to_global,
                                    const unsigned int *candidates, unsigned int candidate_count, csamd_host_blocked_ell *dst,
                                    int device, csamd_blocked_ell_tune_result *picked);

/* Column / row orders against Blocked-ELL padding.
 * ingest/internal/dataset/internal/convert_layout_support_part.hh:20-86
 * (build_shard_column_maps): columns ordered by their row-block support
 * signature over the shard's Blocked-ELL partitions (views in shard order;
 * `cols` columns); host outputs [cols].  Synchronous. */
int csamd_build_shard_column_maps(const csamd_blocked_ell_view *parts, unsigned int part_count, unsigned int value_type,
                                  unsigned int cols, unsigned int *h_exec_to_canonical, unsigned int *h_canonical_to_exec,
                                  csamd_convert_workspace *ws);
/* bucket/operators/blocked_ell_bipartite_optimize.cuh:361-434: columns
 * ordered by the mean rank of their sampled rows (h_sample_row_rank [rows],
 * 0xFFFFFFFF: unranked); host outputs [cols].  :436-500: rows ordered by
 * the mean exec rank of their columns; host outputs [rows].  Rank sums are
 * exact (the reference accumulates them in float).  Synchronous. */
int csamd_build_column_order_from_sampled_coo(const csamd_host_coo *sampled, const unsigned int *h_sample_row_rank, int device,
                                              unsigned int *h_exec_to_canonical_cols, unsigned int *h_canonical_to_exec_cols);
int csamd_build_row_order_from_coo(const csamd_host_coo *src, const unsigned int *h_canonical_to_exec_cols, int device,
                                   unsigned int *h_exec_to_canonical_rows, unsigned int *h_canonical_to_exec_rows);

/* Device COO -> device CSR (the sort those converters start with; project extension, the
 * reference's compressed_from_coo_raw.cuh:97-225 has the same role): stable by (row, column),
 * optional column remap, any value type; out_ptr[rows + 1], out_idx/out_val[nnz]. */
size_t csamd_compressed_from_coo_scratch_bytes(unsigned int nnz);
int csamd_compressed_from_coo_raw(unsigned int rows, unsigned int cols, unsigned int nnz, const unsigned int *d_row_idx,
                                  const unsigned int *d_col_idx, const void *d_val, unsigned int value_type,
                                  const unsigned int *d_feature_to_global, unsigned int feature_count, unsigned int *d_out_ptr,
                                  unsigned int *d_out_idx, void *d_out_val, void *d_scratch, size_t scratch_bytes,
                                  csamd_stream_t stream);

/* convert/sliced_ell_from_compressed.cuh:86-94 */
int csamd_host_choose_sliced_ell_slice_rows(const csamd_host_compressed *src, const unsigned int *candidates,
                                            unsigned int candidate_count, csamd_sliced_ell_tune_result *out);
/* convert/sliced_ell_from_compressed.cuh:96-164 and :166-254 (CSR or CSC source) */
int csamd_host_sliced_ell_from_compressed(const csamd_host_compressed *src, unsigned int slice_rows,
                                          csamd_host_sliced_ell *dst);
/* convert/sliced_ell_from_compressed.cuh:256-284 */
int csamd_host_sliced_ell_from_compressed_auto(const csamd_host_compressed *src, const unsigned int *candidates,
                                               unsigned int candidate_count, csamd_host_sliced_ell *dst,
                                               csamd_sliced_ell_tune_result *picked);
/* convert/sliced_ell_from_blocked_ell.cuh:233-364 */
int csamd_host_sliced_ell_from_blocked_ell(const csamd_host_blocked_ell *src, unsigned int slice_rows,
                                           csamd_host_sliced_ell *dst);

/* ---------------------------------------------------------------------------
 * Device staging.  runtime/device/sharded_device.cuh:22-50 (records, cache),
 * :199-243 (allocation cache), :343-755 (upload / upload_async), :958-990
 * (release), :1336-1432 (packed shard upload).
 * One packed device allocation per partition (or per shard): the device copy
 * of the view sits at offset 0, the arrays follow 256-byte aligned.
 * ------------------------------------------------------------------------- */
typedef struct CSAMD_ALIGN16 csamd_partition_record {
    void *allocation;
    size_t allocation_bytes;
    void *storage;
    void *view;          /* device copy of the view */
    void *a0;
    void *a1;
    void *a2;
    void *a3;
    unsigned long group_begin;
    unsigned long group_end;
    int device_id;
} csamd_partition_record;

/* Per-process device allocation cache, 16 entries per device (best fit). */
typedef struct CSAMD_ALIGN16 csamd_cached_allocation {
    void *ptr;
    size_t bytes;
    int device_id;
} csamd_cached_allocation;

typedef struct CSAMD_ALIGN16 csamd_allocation_cache {
    unsigned long cache_count;
    unsigned long cache_capacity;
    csamd_cached_allocation *cache;
} csamd_allocation_cache;

void csamd_allocation_cache_init(csamd_allocation_cache *c);
void csamd_allocation_cache_clear(csamd_allocation_cache *c);
/* sharded_device.cuh:199-220: smallest cached block >= min_bytes on device, or NULL */
void *csamd_take_cached_allocation(csamd_allocation_cache *c, int device_id, size_t min_bytes, size_t *out_bytes);
/* sharded_device.cuh:222-243: keep a block (frees it when the cache is full) */
int csamd_cache_allocation(csamd_allocation_cache *c, int device_id, void *ptr, size_t bytes);

/* Upload one host partition (stream NULL = synchronous).  cache may be NULL.
 * out_view receives the host copy of the device view. */
int csamd_upload_compressed(const csamd_host_compressed *src, csamd_partition_record *record,
                            csamd_compressed_view *out_view, csamd_allocation_cache *cache, csamd_stream_t stream);
int csamd_upload_blocked_ell(const csamd_host_blocked_ell *src, csamd_partition_record *record,
                             csamd_blocked_ell_view *out_view, csamd_allocation_cache *cache, csamd_stream_t stream);
/* Computes slice_slot_offsets and slice_rows (uniform height or 0) like
 * sharded_device.cuh:609-620. */
int csamd_upload_sliced_ell(const csamd_host_sliced_ell *src, csamd_partition_record *record,
                            csamd_sliced_ell_view *out_view, csamd_allocation_cache *cache, csamd_stream_t stream);
/* sharded_device.cuh:958-990: return the allocation to the cache (or free). */
int csamd_release_record(csamd_partition_record *record, csamd_allocation_cache *cache);
/* Packed shard upload: `count` Blocked-ELL partitions into ONE allocation,
 * out_views[count] host copies (sharded_device.cuh:1336-1432). */
int csamd_upload_blocked_ell_shard(const csamd_host_blocked_ell *const *parts, unsigned int count,
                                   csamd_partition_record *record, csamd_blocked_ell_view *out_views,
                                   csamd_allocation_cache *cache, csamd_stream_t stream);
int csamd_upload_compressed_shard(const csamd_host_compressed *const *parts, unsigned int count,
                                  csamd_partition_record *record, csamd_compressed_view *out_views,
                                  csamd_allocation_cache *cache, csamd_stream_t stream);

/* ---------------------------------------------------------------------------
 * sharded<MatrixT>: partition / shard metadata of a row-partitioned matrix.
 * runtime/layout/sharded.cuh:30-50 (layout, field for field), :71-124
 * (partition_nnz / partition_aux), :172-252 (lookups), :262-344 (bytes);
 * runtime/host/sharded_host.cuh:14-396 (metadata operations).  The layout
 * does not depend on MatrixT (parts is an array of part pointers), so one C
 * struct serves every format; calls that read a part or its aux word take
 * the part format (CSAMD_FORMAT_*).  Host memory only; nothing here touches
 * a device.  parts[i] may be NULL (metadata only); a non-NULL part is a host
 * struct of the format, owned by the sharded matrix (destroyed by clear and
 * the drop calls: its storage released, the struct itself freed).
 * ------------------------------------------------------------------------- */
enum {
    CSAMD_FORMAT_COMPRESSED = 1,
    CSAMD_FORMAT_BLOCKED_ELL = 2,
    CSAMD_FORMAT_SLICED_ELL = 3
};

typedef struct CSAMD_ALIGN16 csamd_sharded {
    unsigned long rows;
    unsigned long cols;
    unsigned long nnz;
    unsigned long num_partitions;
    unsigned long partition_capacity;
    void **parts;
    unsigned long *partition_offsets;  /* [capacity + 1] */
    unsigned long *partition_rows;
    unsigned long *partition_nnz;
    unsigned long *partition_aux;
    unsigned long num_shards;
    unsigned long shard_capacity;
    unsigned long *shard_offsets;      /* [shard capacity + 1] row offsets */
    unsigned long *shard_parts;        /* [shard capacity + 1] first partition of each shard */
} csamd_sharded;

/* sharded.cuh:54-69 */
void csamd_sharded_init(csamd_sharded *m);
/* sharded_host.cuh:21-36: destroys the parts (format 0: parts are left to the caller),
 * frees the metadata, re-inits */
void csamd_sharded_clear(csamd_sharded *m, unsigned int format);
/* A zeroed host part struct of `format` (calloc), for parts a sharded matrix will own. */
void *csamd_host_part_alloc(unsigned int format);
/* formats/compressed.cuh:122-153 (allocate): rows, cols, nnz and axis set; one malloc block
 * [ptr | idx | val] (2-byte values), any previous storage released. */
int csamd_host_compressed_allocate(csamd_host_compressed *m);
/* sharded_host.cuh:38-83, :85-109 (grow-only; existing entries kept) */
int csamd_sharded_reserve_partitions(csamd_sharded *m, unsigned long capacity);
int csamd_sharded_reserve_shards(csamd_sharded *m, unsigned long capacity);
/* sharded_host.cuh:111-126, :128-141 */
void csamd_sharded_rebuild_shard_parts(csamd_sharded *m);
void csamd_sharded_rebuild_partition_offsets(csamd_sharded *m);
/* sharded_host.cuh:143-167 (rows / nnz / aux arrays may be NULL: zeros); shards = partitions.
 * Existing parts are destroyed as csamd_sharded_clear(m, format) does (format 0: left alone). */
int csamd_sharded_define_partitions(csamd_sharded *m, unsigned int format, unsigned long cols, unsigned long num_partitions,
                                    const unsigned long *partition_rows, const unsigned long *partition_nnz,
                                    const unsigned long *partition_aux);
/* sharded_host.cuh:169-176 */
int csamd_sharded_set_shards_to_partitions(csamd_sharded *m);
/* sharded_host.cuh:178-198: takes ownership of `part` (a host struct of `format`, or NULL) */
int csamd_sharded_append_partition(csamd_sharded *m, unsigned int format, void *part);
/* sharded_host.cuh:200-228: moves src's parts to the end of dst; src keeps no partitions */
int csamd_sharded_concatenate(csamd_sharded *dst, csamd_sharded *src);
/* sharded_host.cuh:230-266 */
int csamd_sharded_set_equal_shards(csamd_sharded *m, unsigned long count);
/* sharded_host.cuh:268-308: *out_offsets is malloc'ed ([*out_count + 1]); the caller frees it */
int csamd_sharded_build_shard_offsets_by_rows(const csamd_sharded *m, unsigned long target_rows_per_shard,
                                              unsigned long **out_offsets, unsigned long *out_count);
/* sharded_host.cuh:310-327: cuts must be partition boundaries, 0 first, rows last */
int csamd_sharded_reshard(csamd_sharded *m, unsigned long count, const unsigned long *offsets);
/* sharded_host.cuh:329-357, :359-389 */
int csamd_sharded_set_shards_by_nnz(csamd_sharded *m, unsigned long max_nnz);
int csamd_sharded_set_shards_by_part_bytes(csamd_sharded *m, unsigned int format, size_t max_bytes);
/* sharded_host.cuh:514-531, :533-548 (host materialization only) */
int csamd_sharded_drop_partition(csamd_sharded *m, unsigned int format, unsigned long part);
int csamd_sharded_drop_shard(csamd_sharded *m, unsigned int format, unsigned long shard);
/* sharded.cuh:101-124 (aux word of a host part: compressed axis, packed Blocked-ELL
 * (width << 16 | block size), packed Sliced-ELL (total slots << 32 | slice count)) */
unsigned long csamd_partition_aux(unsigned int format, const void *part);
/* sharded.cuh:127-139, :178-190 etc.: row -> partition / shard (count when outside) */
unsigned long csamd_sharded_find_partition(const csamd_sharded *m, unsigned long row);
unsigned long csamd_sharded_find_shard(const csamd_sharded *m, unsigned long row);
unsigned long csamd_sharded_first_partition_in_shard(const csamd_sharded *m, unsigned long shard);
unsigned long csamd_sharded_last_partition_in_shard(const csamd_sharded *m, unsigned long shard);
int csamd_sharded_row_is_partition_boundary(const csamd_sharded *m, unsigned long row);
/* sharded.cuh:262-344: host bytes of a partition (its struct + arrays; from the metadata
 * when the part is not materialized), of a shard, and of the whole matrix */
size_t csamd_sharded_partition_bytes(const csamd_sharded *m, unsigned int format, unsigned long part);
size_t csamd_sharded_shard_bytes(const csamd_sharded *m, unsigned int format, unsigned long shard);
size_t csamd_sharded_bytes(const csamd_sharded *m, unsigned int format);

/* ---------------------------------------------------------------------------
 * sharded_device<MatrixT>: device residency of a sharded matrix's partitions.
 * runtime/device/sharded_device.cuh:43-50 (layout, field for field), :119-178
 * (init / clear / reserve), :1098-1140 (device bytes, shards by device bytes),
 * :1142-1198 (upload / release partition), :1200-1432 (packed shard upload),
 * :1434-1545 (upload / release shard), :1555-1722 (stage / swap).
 * Device bytes are this library's footprint: one allocation per partition
 * (or per packed shard) with the device view at offset 0 and every array
 * 256-byte aligned.  Stream NULL = synchronous.
 * ------------------------------------------------------------------------- */
typedef struct csamd_sharded_device {
    unsigned long capacity;
    csamd_partition_record *parts;       /* [capacity] */
    unsigned long cache_count;           /* the allocation cache (csamd_allocation_cache fields) */
    unsigned long cache_capacity;
    csamd_cached_allocation *cache;
} csamd_sharded_device;

void csamd_sharded_device_init(csamd_sharded_device *s);
/* releases every resident partition, frees the cached allocations and the table */
int csamd_sharded_device_clear(csamd_sharded_device *s);
int csamd_sharded_device_reserve(csamd_sharded_device *s, unsigned long capacity);
/* device bytes one partition / shard takes once uploaded (from the metadata) */
size_t csamd_sharded_device_partition_bytes(const csamd_sharded *view, unsigned int format, unsigned long part);
size_t csamd_sharded_device_shard_bytes(const csamd_sharded *view, unsigned int format, unsigned long shard);
/* sharded_device.cuh:1111-1140: shards cut so each holds at most max_bytes of device bytes */
int csamd_sharded_set_shards_by_device_bytes(csamd_sharded *view, unsigned int format, size_t max_bytes);
/* sharded_device.cuh:1142-1165: the host part must be materialized; a partition already
 * resident on `device` is kept, one resident elsewhere is released first */
int csamd_sharded_device_upload_partition(csamd_sharded_device *s, const csamd_sharded *view, unsigned int format,
                                          unsigned long part, int device, csamd_stream_t stream);
/* sharded_device.cuh:1167-1198: a partition of a packed shard releases its whole group;
 * the allocation goes to the cache (16 entries) or is freed */
int csamd_sharded_device_release_partition(csamd_sharded_device *s, unsigned long part);
/* sharded_device.cuh:1434-1524: compressed and Blocked-ELL shards go up packed into one
 * allocation (views back to back at its start); Sliced-ELL shards one partition at a time.
 * A shard whose partitions are all resident on `device` is kept as it is. */
int csamd_sharded_device_upload_shard(csamd_sharded_device *s, const csamd_sharded *view, unsigned int format,
                                      unsigned long shard, int device, csamd_stream_t stream);
/* sharded_device.cuh:1526-1545 */
int csamd_sharded_device_release_shard(csamd_sharded_device *s, const csamd_sharded *view, unsigned long shard);
/* sharded_device.cuh:1631-1674 (stage_shard / stage_shard_async).  When every host part of
 * the shard is materialized: upload_shard.  Otherwise the parts come from `pack` (Blocked-ELL
 * only, a CSPACK01 file of the shard's bucketed execution partitions in shard order; the
 * reference's fetch_shard reads the same blobs): each partition is staged from the file
 * straight into HBM and rebuilt there as its canonical Blocked-ELL partition at the block
 * size of its aux word, which becomes the resident device partition, with no host copy.
 * drop_host_after_upload drops materialized host parts afterwards. */
int csamd_sharded_device_stage_shard(csamd_sharded_device *s, csamd_sharded *view, unsigned int format,
                                     const csamd_cspack_file *pack, unsigned long shard, int device, csamd_stream_t stream,
                                     int drop_host_after_upload);
/* sharded_device.cuh:1555-1629 (stage_partition / _async): keep a partition resident on
 * `device`, else upload its host part, else (Blocked-ELL) rebuild it from the pack of the shard
 * holding it as stage_shard does; optionally drop the host part afterwards. */
int csamd_sharded_device_stage_partition(csamd_sharded_device *s, csamd_sharded *view, unsigned int format,
                                         const csamd_cspack_file *pack, unsigned long part, int device, csamd_stream_t stream,
                                         int drop_host_after_upload);
/* sharded_device.cuh:1676-1722: stage `in_shard`, then release `out_shard` (both resident
 * for a moment) and optionally drop the outgoing shard's host parts */
int csamd_sharded_device_swap_shard(csamd_sharded_device *s, csamd_sharded *view, unsigned int format,
                                    const csamd_cspack_file *pack, unsigned long out_shard, unsigned long in_shard, int device,
                                    csamd_stream_t stream, int drop_host_after_upload, int drop_host_after_release);

/* convert/blocked_ell_from_compressed.cuh:480-551: dst (cleared first; its
 * parts are host Blocked-ELL the matrix owns) gets the rows of src (host CSR
 * partitions, compressed by row) in parts of target_block_rows x block_size
 * rows (0: one part), each converted as blocked_ell_from_compressed does
 * (on the GPU, through the host-struct converter). */
int csamd_repack_sharded_compressed_to_blocked_ell(const csamd_sharded *src, unsigned int block_size, unsigned long target_block_rows,
                                                   csamd_sharded *dst);
/* :553-600: the block size chosen over the whole matrix
 * (choose_blocked_ell_block_size), then the repack. */
int csamd_repack_sharded_compressed_to_blocked_ell_auto(const csamd_sharded *src, const unsigned int *candidates,
                                                        unsigned int candidate_count, unsigned long target_block_rows,
                                                        csamd_sharded *dst, csamd_blocked_ell_tune_result *picked);

/* Shard bucketing over the partitions of one shard resident on the device
 * (bucket/operators/sharded_major_nnz.cuh:35-135): the base workspace plus
 * the shard's part pointer tables (pinned host, device).  tables_event
 * (extension) orders rewrites of the pinned tables after the copies that
 * read them. */
typedef struct CSAMD_ALIGN16 csamd_sharded_major_bucket_workspace {
    csamd_compressed_major_bucket_workspace base;
    unsigned long part_capacity;
    void *h_part_block;
    void *d_part_block;
    const unsigned int **h_part_major_ptr;
    const unsigned int **h_part_minor_idx;
    const void **h_part_val;
    unsigned int *h_part_row_offsets;
    const unsigned int **d_part_major_ptr;
    const unsigned int **d_part_minor_idx;
    const void **d_part_val;
    unsigned int *d_part_row_offsets;
    void *tables_event;
} csamd_sharded_major_bucket_workspace;

/* sharded_major_nnz.cuh:53-57 */
typedef struct CSAMD_ALIGN16 csamd_sharded_major_bucket_result {
    csamd_compressed_view rebuilt;
    csamd_major_nnz_bucket_plan_view plan;
    unsigned long shard_id;
} csamd_sharded_major_bucket_result;

/* sharded_major_nnz.cuh:59-87 */
void csamd_sharded_major_bucket_init(csamd_sharded_major_bucket_workspace *ws);
void csamd_sharded_major_bucket_clear(csamd_sharded_major_bucket_workspace *ws);
int csamd_sharded_major_bucket_setup(csamd_sharded_major_bucket_workspace *ws, int device, csamd_stream_t stream);
/* sharded_major_nnz.cuh:89-135 */
int csamd_sharded_major_bucket_reserve_parts(csamd_sharded_major_bucket_workspace *ws, unsigned long part_count);
/* sharded_major_nnz.cuh:137-273 with major_nnz_raw.cuh:122-186,252-320: the
 * bucket plan and rebuild of shard shard_id of `view` (compressed by row)
 * from its partitions resident in `state` on the workspace's device, as
 * one matrix of the shard's rows (row i of the shard = row i -
 * partition_offsets[p] of partition p).  2-byte values, or value_type's. */
int csamd_build_bucketed_shard_major_view(const csamd_sharded_device *state, const csamd_sharded *view, unsigned long shard_id,
                                          unsigned int requested_bucket_count, csamd_sharded_major_bucket_workspace *ws,
                                          csamd_sharded_major_bucket_result *out);
int csamd_build_bucketed_shard_major_view_typed(const csamd_sharded_device *state, const csamd_sharded *view,
                                                unsigned long shard_id, unsigned int value_type,
                                                unsigned int requested_bucket_count, csamd_sharded_major_bucket_workspace *ws,
                                                csamd_sharded_major_bucket_result *out);


/* ---------------------------------------------------------------------------
 * Synthetic CSR generator for measurement and tests (not a reference path).
 * Rows: nnz ~ round(LogNormal(ln median, sigma)) capped at cols; columns
 * drawn without replacement with weights w_c (Efraimidis-Spirakis keys
 * log(u)/w_c, top k, emitted ascending); values ~ zero-truncated NB2(mean
 * col_mean[c], alpha) counts, stored as value_type (f16 or f32).
 * Deterministic in seed.
 * ------------------------------------------------------------------------- */
typedef struct CSAMD_ALIGN16 csamd_synth_params {
    unsigned int rows;
    unsigned int cols;
    unsigned int value_type;
    unsigned int seed;
    float row_nnz_median;
    float row_nnz_sigma;
    float nb_alpha;
    unsigned int reserved;
    const float *d_col_inv_weight; /* [cols]: 1 / w_c */
    const float *d_col_mean;       /* [cols] */
} csamd_synth_params;

/* Pass 1: row nnz into d_row_nnz [rows]. */
int csamd_synth_row_nnz(const csamd_synth_params *p, unsigned int *d_row_nnz, csamd_stream_t stream);
/* Exclusive scan of d_counts [n] into d_ptr [n+1] (u32; fails with an error
 * if the total overflows u32).  Synchronous; returns the total. */
/* Decode n stored values of value_type (f16, bf16, f32, fp8 e4m3) to fp32
 * with the same device decoders the kernels use.  Asynchronous. */
int csamd_decode_values(unsigned int value_type, const void *d_values, unsigned long long n, float *d_out, csamd_stream_t stream);
int csamd_exclusive_scan_u32(const unsigned int *d_counts, unsigned int n, unsigned int *d_ptr,
                             csamd_convert_workspace *ws, unsigned long long *out_total);
/* Pass 2: fill minor indices and values of every row given d_ptr [rows+1]. */
int csamd_synth_fill(const csamd_synth_params *p, const unsigned int *d_ptr, unsigned int *d_idx,
                     void *d_val, csamd_stream_t stream);
/* log1p(count / row_total * scale) transform of fp32 counts -> value_type
 * (bf16 or fp8 e4m3 with per-feature absmax scale written to d_feature_scale
 * [cols] when value_type is fp8; f32/f16/bf16 otherwise). */
int csamd_synth_log1p_normalize(unsigned int rows, unsigned int cols, const unsigned int *d_ptr,
                                const unsigned int *d_idx, const float *d_counts, float scale,
                                unsigned int value_type, void *d_val_out, float *d_feature_scale,
                                csamd_convert_workspace *ws);

#ifdef __cplusplus
}
#endif

#endif /* CELLSHARD_AMD_H */
