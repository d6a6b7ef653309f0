/*
 * oracle.h -- C entry points of the CPU oracle (liboracle.so).
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a sequential/OpenMP CPU
 * restatement of the reference adaptive-tree path, used by tests/ as the
 * parity checker and by bench.py as the timed CPU baseline. The product
 * library (libadaptive_trees_hip.so) never links, loads or calls it.
 *
 * The tree entry points mirror include/adaptive_trees.h one to one, with
 * HOST coordinate pointers; buffers returned by orc_tree_buffer are host
 * memory owned by the oracle tree.
 */
#ifndef ADAPTIVE_TREES_ORACLE_H
#define ADAPTIVE_TREES_ORACLE_H

#include "../include/adaptive_trees.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_tree orc_tree;

int orc_tree_create(int real_bytes, int tree_type, const void *x, const void *y,
                    const void *z, size_t count, const at_config *config,
                    orc_tree **out);
int orc_tree_destroy(orc_tree *tree);
int orc_tree_build(orc_tree *tree);
int orc_tree_update(orc_tree *tree);
int orc_tree_compute_views(orc_tree *tree, const uint8_t *views);
int orc_tree_needs_resize(orc_tree *tree, int *needs_resize);
int orc_tree_resize_buffers(orc_tree *tree);
int orc_tree_buffer(orc_tree *tree, int buffer_id, const void **data, size_t *size,
                    size_t *capacity);
int orc_tree_scalars(orc_tree *tree, at_scalars *out);
int orc_tree_device_counters(orc_tree *tree, const int32_t **num_leaves,
                             const int32_t **num_nodes, const uint32_t **num_active);
int orc_tree_max_achieved_depth(orc_tree *tree, uint32_t *depth);
const char *orc_last_error(void);

/* Stand-alone primitives for pinning tests. */
int orc_sfc_keys(int real_bytes, const void *x, const void *y, const void *z,
                 size_t n, const at_box *box, int hilbert, uint64_t *keys);
int orc_binary_keys(int real_bytes, const void *x, const void *y, const void *z,
                    size_t n, const at_box *box, uint32_t max_depth, uint64_t *keys);
uint64_t orc_morton(uint32_t x, uint32_t y, uint32_t z);
uint64_t orc_hilbert(uint32_t x, uint32_t y, uint32_t z);
void orc_hilbert_decode(uint64_t key, uint32_t *xyz);
void orc_morton_decode(uint64_t key, uint32_t *xyz);
/* Stable ascending sort of (key, value) pairs in place. */
int orc_sort_pairs(uint64_t *keys, uint32_t *values, size_t n);
/* Near-field count of one octree node (octree_rebalance.cuh:579-634). */
uint64_t orc_near_field(const uint64_t *keys, size_t n, uint64_t node_start,
                        uint32_t level, int hilbert, uint32_t *own_count);
/* Axis schedule of the binary tree (binary_node.cuh:17-40). */
int orc_binary_geometry(const at_box *box, uint32_t max_depth, uint8_t *axis63);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
