// Batched layout conversion of the partition pipeline (pipeline_convert.hip):
// every partition the layout tuner assigns to Sliced-ELL or Blocked-ELL is
// converted from its device CSR in a fixed number of launches over all of
// them, capturable into a graph (descriptor tables are uploaded beforehand).
#pragma once

#include "common.hh"

namespace csamd {

struct pl_state {
    void *block;                // device: descriptor tables, base arrays, nnz counters
    size_t bytes;
    unsigned int value_type;
    unsigned int n_sell, n_bell;
    unsigned int sell_slices, sell_rows, bell_row_blocks;
    unsigned int bell_nwords;   // LDS bitset words of the widest Blocked-ELL partition
    void *sell_tab, *bell_tab;  // device descriptor tables
    unsigned int *sell_slice_base, *sell_row_base, *bell_rb_base;  // device [n + 1]
    unsigned int *bell_counters;  // device [n_bell]: stored nonzeros (copied to each partition's d_nnz)
    unsigned int **bell_nnz_out;  // device [n_bell]: the callers' d_nnz
    unsigned int *error;          // device word: 1 when a row block held more block columns than its ell width
};

// Validates the SELL / BELL partitions, allocates and uploads the tables.
int pl_prepare(const csamd_pipeline_partition *parts, unsigned int n, unsigned int value_type, pl_state *st);
// Enqueues the conversions on `s` (capturable: kernels only).
int pl_enqueue(const pl_state &st, hipStream_t s);
void pl_free(pl_state *st);

} // namespace csamd
