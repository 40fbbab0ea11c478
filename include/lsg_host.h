/*
 * lsg_host.h - host-side inspection of the operator build (no device needed).
 *
 * These entry points run only the host part of lsg_create: tree, level
 * tables, NCA skeletons and fused-operator sizes. They exist so that the build
 * can be checked against a CPU restatement of the reference on machines
 * without a GPU; they perform no matrix-vector products.
 *
 *   reference                                         inspected by
 *   createUniformTree / assign_Tree_Interactions       lsg_host_info
 *   createAdaptiveTree / makeLevelRestriction_Tree     lsg_host_level_boxes,
 *     GMRES/G_FMM2DTree.hpp:776-1137                     lsg_host_leaf_boxes
 *     GMRES/G_FMM2DTree.hpp:833-941, 1356-1498
 *   evaluateM / writeMToFile (6-digit tables)          lsg_host_far_table
 *     GMRES/FarFieldInteraction.hpp:152-160, G_FMM2DTree.hpp:1252-1271
 *   getOperator / colleagueNeighborInteraction         lsg_host_near_tables
 *     GMRES/G_FMM2DTree.hpp:1773-1911
 *   getNodes_LFR / getNodes_HFR / getUserCheckPoints   lsg_host_skeleton
 *     GMRES/G_FMM2DTree.hpp:2062-2385, 3146-3571
 */
#ifndef LSG_HOST_H_
#define LSG_HOST_H_

#include <stdint.h>
#include "lsg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lsg_host lsg_host;

int lsg_host_build(const lsg_inputs* in, lsg_host** out);
/* Tree only (levels, interaction lists, cones); the table, skeleton and
 * operator queries then return empty results. */
int lsg_host_build_tree(const lsg_inputs* in, lsg_host** out);
void lsg_host_destroy(lsg_host* h);
/* A device context from a host build: the build's tree, tables and fused NCA
 * operators are uploaded as they are, with no second build. No reference
 * counterpart: the reference builds and applies the operator in one object
 * (DFMM::FMMTreeInitialisation, GMRES/GMRES.cpp:161-199). One host build can
 * serve several contexts, on one device or on several (call with each device
 * current); the host build stays valid. Errors as lsg_create
 * (lsg_create_error): LSG_ERR_ARG for a NULL or tree-only build,
 * LSG_ERR_DEVICE without a HIP device. */
int lsg_create_from_host(const lsg_host* h, lsg_ctx** out);
const char* lsg_host_error(void);
int lsg_host_info(const lsg_host* h, lsg_info* info);
/* Far-field table of the leaf level: (2*15+1) x nChebNodes^2, row-major,
 * interleaved complex; returns the number of complex values (call with
 * out == NULL to query). */
int64_t lsg_host_far_table(const lsg_host* h, double* out);
/* The 9 colleague near-field tables of the leaf level, slot-major,
 * each nChebNodes^2 x nChebNodes^2 row-major, interleaved complex. */
int64_t lsg_host_near_tables(const lsg_host* h, double* out);
/* Skeleton dump (see the layout comment in nca.hpp); returns its length. */
int64_t lsg_host_skeleton(const lsg_host* h, int* out);
/* Interaction-list pairs per level (sum over boxes of |IL|, cones summed in
 * the high-frequency levels), levels 0..nLevels. */
int lsg_host_il_pairs(const lsg_host* h, int64_t* out, int max_levels);
/* Q_pinv of the basis (NumberOfBases x nChebNodes^2, row-major). */
int64_t lsg_host_qpinv(const lsg_host* h, double* out);
/* The flattened device plan, for checking the fused operators on the CPU:
 * operator entries (interleaved complex, returns the complex count), source
 * state indices, and per destination 7 values {pass, kind (0 store,
 * 1 accumulate), rows, ncol, dst, op_off, src_off}; returns #destinations.
 * Passes run in index order; the near field runs before the first kind-1
 * pass. */
int64_t lsg_host_plan_ops(const lsg_host* h, double* out);
int64_t lsg_host_plan_src(const lsg_host* h, int* out);
int64_t lsg_host_plan_dests(const lsg_host* h, int64_t* out);
/* Leaf colleague indices (n_leaves x 9, -1 where absent) and k^2 q at nodes
 * as used by the near-field kernel. */
int64_t lsg_host_plan_nbr(const lsg_host* h, int* out);
int64_t lsg_host_plan_qk2(const lsg_host* h, double* out);
/* Per-level largest NCA rank, levels 0..nLevels. */
int lsg_host_level_ranks(const lsg_host* h, int* out, int max_levels);
/* Boxes and leaf boxes per level, levels 0..nLevels (the tree's tree[j].size()
 * and the leaves among them; createAdaptiveTree + makeLevelRestriction_Tree,
 * GMRES/G_FMM2DTree.hpp:776-831, 891-915, 1019-1137, for tree_mode 1). */
int lsg_host_level_boxes(const lsg_host* h, int* boxes, int* leaves, int max_levels);
/* The leaf order (leafNodes): per leaf its level and box index in the level;
 * 2 ints per leaf, returns the leaf count. */
int64_t lsg_host_leaf_boxes(const lsg_host* h, int* out);
/* The NCA's ACA operands per skeleton owner, in the skeleton dump's order:
 * 9 ints {level, box, cone (-1 = LFR), outgoing rows, columns, rank, incoming
 * rows, columns, rank}, rank -1 where an operand list was empty (the n_rows,
 * n_cols and ComputedRank of getNodes_*_box, G_FMM2DTree.hpp:2083-2212,
 * 2294-2358, 3322-3392, 3485-3546). Returns the int count. */
int64_t lsg_host_aca_shapes(const lsg_host* h, int* out);

/* Checks of the GPU operator build (lsg_inputs.setup_device = 1) against the
 * host one; device = 0 evaluates on the host, 1 on the GPU.
 * lsg_host_entries: K(rows, cols) column-major (nr x nc interleaved complex),
 * the far-field entry of getMatrixEntry (G_FMM2DTree.hpp:603-658).
 * lsg_host_aca: one ACA_only_nodes (ACA.hpp:253-481) of K(rows, cols) at the
 * build's tolerance; rank, row and column pivots (positions in rows / cols;
 * room for min(n1, n2) each). */
int lsg_host_entries(const lsg_host* h, const int* rows, int nr, const int* cols, int nc, int device, double* out);
int lsg_host_aca(const lsg_host* h, const int* rows, int n1, const int* cols, int n2, int device, int* rank,
                 int* row_piv, int* col_piv);

/* Fused ragged batch (no reference counterpart: the reference builds one tree
 * and solves one system per RestartedGMRES call, GMRES/GMRES.cpp:152-199,
 * 307-319). n host builds of independent scatterers (uniform trees, any
 * depth, wavenumber and contrast, the same nChebNodes) become ONE plan on the
 * current device: one state vector, each matvec pass covering every system,
 * the near field on MFMA with one table set per system. Vectors of system i
 * are N_i complex128 (lsg_ragged_size), host arrays, leaf-major as
 * everywhere. Errors of lsg_ragged_create in lsg_create_error(). */
typedef struct lsg_ragged lsg_ragged;
int lsg_ragged_create(const lsg_host* const* hosts, int n, lsg_ragged** out);
void lsg_ragged_destroy(lsg_ragged* r);
const char* lsg_ragged_error(const lsg_ragged* r);
/* N_i of system i; i = -1: the total */
int64_t lsg_ragged_size(const lsg_ragged* r, int i);
/* y_i = A_i x_i for every system with one fused matvec */
int lsg_ragged_matvec(lsg_ragged* r, const double* const* x, double* const* y);
/* GMRES(restart_len) of every system at once, each with its own Krylov
 * basis, Givens rotations and stopping point (estimate at tol/2, breakdown),
 * every cycle opening with its true FP64 residual: x_i holds the initial
 * guess and receives the solution; stats (n entries, may be NULL) per system.
 * LSG_ERR_NOCONV if a system missed tol after max_restarts cycles. */
int lsg_ragged_gmres(lsg_ragged* r, const double* const* b, double* const* x, int restart_len, int max_restarts,
                     double tol, lsg_stats* stats);
/* lsg_autotune / lsg_use_graph / lsg_time_matvec for the fused operator */
int lsg_ragged_autotune(lsg_ragged* r, int reps, int* n_changed);
int lsg_ragged_use_graph(lsg_ragged* r, int on);
int lsg_ragged_time_matvec(lsg_ragged* r, int reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* LSG_HOST_H_ */
