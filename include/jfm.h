/*
 * jfm.h -- C ABI of the MI355X-native FMM evaluation path (libjfm.so).
 *
 * Framework-neutral: plain pointers, sizes and an opaque stream handle
 * (a hipStream_t passed as void*; NULL is the legacy default stream).  No
 * exceptions cross this boundary: every function returns JFM_OK (0) or an
 * error code, and jfm_last_error() returns a thread-local message.
 *
 * Device entry points are stream-ordered, never synchronise the host and never
 * allocate: every buffer (including scratch) is owned by the caller.  Outputs
 * are written in place and must not alias inputs.
 *
 * dtype selects the working precision of a device entry point:
 *   JFM_F32 -- float32 buffers; JFM_F64 -- float64 buffers.
 *
 * Layouts (all row-major, tree order = Morton order of the topology):
 *   pos4      [N][4]       x, y, z, mass per particle in tree order
 *   centers4  [nodes][4]   centre of mass x, y, z and total mass per node
 *   coeffs    [nodes][cs]  packed real (Dehnen no-sqrt2) coefficients,
 *                          index n*n + n + m, row stride cs = jfm_coeff_stride()
 *   node ids: internal nodes [0, L-1) with the root at 0, leaf l is node L-1+l;
 *             leaf l owns tree-order particles [l*leaf_size, min(N,(l+1)*leaf_size)).
 *
 * Each entry point names the reference interface it replaces (TobiBu/jaccpot).
 */
#ifndef JFM_H
#define JFM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JFM_OK 0
#define JFM_ERR_ARG 1          /* invalid argument / unsupported size */
#define JFM_ERR_HIP 2          /* HIP runtime error (launch or API) */
#define JFM_ERR_UNSUPPORTED 3  /* order / dtype combination not compiled */
#define JFM_ERR_ALLOC 4        /* host allocation failure (topology build) */

#define JFM_F32 0
#define JFM_F64 1

/* evaluate flags */
#define JFM_EVAL_NEAR 1         /* add the P2P near field (self leaf + near list) */
#define JFM_EVAL_FAR 2          /* add the L2P far field from the leaf locals */
#define JFM_EVAL_POTENTIAL 4    /* also write potentials */
#define JFM_EVAL_COMPENSATED 8  /* blocked Kahan accumulation of the P2P sums */

#define JFM_MAX_ORDER 8

/* Thread-local message for the last failing call on this thread. */
const char* jfm_last_error(void);
/* ABI version (bumped on any signature change). */
int32_t jfm_abi_version(void);
/* Row stride (elements) of a coefficient array for order p and dtype. */
int32_t jfm_coeff_stride(int32_t p, int32_t dtype);
/* 1 if kernels for (p, dtype) are compiled into this library. */
int32_t jfm_order_supported(int32_t p, int32_t dtype);

/* ------------------------------------------------------------------------
 * Host topology (replaces yggdrax build_tree + compute_tree_mass_moments +
 * compute_tree_geometry + build_interactions_and_neighbors as called from
 * jaccpot/runtime/fmm_state.py:630-800, upward/real_tree_expansions.py:439-440,
 * upward/tree_geometry.py:65-127 and runtime/_interaction_cache.py:523,564).
 * ---------------------------------------------------------------------- */
typedef struct jfm_topology jfm_topology;

/* Build the fixed static-radix topology and its far/near lists on the host.
 * positions [n][3] and masses [n] in input order (float64).  bounds is
 * {lo_x, lo_y, lo_z, hi_x, hi_y, hi_z} or NULL for the bounding box.
 * num_threads <= 0 uses all hardware threads. */
int32_t jfm_topology_build(const double* positions, const double* masses, int64_t n, int32_t leaf_size,
                           double theta, const double* bounds, int32_t num_threads, jfm_topology** out);

/* info[0..8] = N, leaf_size, num_leaves, num_internal, num_nodes, far pairs K,
 *              near entries E, max depth, number of targets with far sources. */
int32_t jfm_topology_info(const jfm_topology* topo, int64_t* info);

/* Copy the topology arrays into caller buffers (any pointer may be NULL).
 *   perm[N]        tree slot -> input index
 *   parent[nodes]  (-1 at the root), left[I], right[I]
 *   node_range[nodes][2] inclusive tree-order particle range, depth[nodes]
 *   center[nodes][3], radius[nodes], mass[nodes]  (float64 geometry of the build)
 *   far_off[nodes+1], far_src[K]  far list by target node, sources ascending
 *   near_off[L+1], near_src[E]    near list by target leaf (leaf indices, self excluded) */
int32_t jfm_topology_export(const jfm_topology* topo, int32_t* perm, int32_t* parent, int32_t* left,
                            int32_t* right, int32_t* node_range, int32_t* depth, double* center,
                            double* radius, double* mass, int32_t* far_off, int32_t* far_src,
                            int32_t* near_off, int32_t* near_src);

/* Level schedules: m2m_nodes = internal nodes by depth, deepest level first, with
 * m2m_off[levels+1]; l2l_nodes = non-root nodes by depth, shallowest first, with
 * l2l_off[levels+1]; m2l_targets = nodes that have far sources (ascending).
 * Sizes: I, max_depth+1, nodes-1, max_depth+1, info[8]. */
int32_t jfm_topology_schedules(const jfm_topology* topo, int32_t* m2m_nodes, int32_t* m2m_off,
                               int32_t* l2l_nodes, int32_t* l2l_off, int32_t* m2l_targets);

void jfm_topology_free(jfm_topology* topo);

/* ------------------------------------------------------------------------
 * Device topology (the same yggdrax calls as above, jaccpot/runtime/
 * fmm_state.py:630-800 and runtime/_interaction_cache.py:523,564, run on the
 * GPU): the steps of jfm_topology_build as stream-ordered kernels whose
 * outputs equal the host build's bit for bit.  Positions [n][3] and masses [n]
 * are float64 device arrays in input order.  The caller sequences the steps
 * (jaccpot_amd/dtopology.py), reading back two small counts in between.
 * ---------------------------------------------------------------------- */
/* Bounds: bounds6 = {lo, hi} by exact min/max unless given != 0 (then bounds6
 * already holds them); flags |= 1 if any coordinate is not finite.  scratch
 * holds jfm_tree_bounds_scratch_doubles() doubles. */
int32_t jfm_tree_bounds_scratch_doubles(void);
int32_t jfm_tree_bounds(void* stream, const double* pos, int64_t n, int32_t given, double* scratch,
                        double* bounds6, int32_t* flags);
/* 63-bit Morton codes on the cube of bounds6 and index[i] = i. */
int32_t jfm_tree_morton(void* stream, const double* pos, int64_t n, const double* bounds6, uint64_t* codes,
                        int32_t* index);
/* Stable LSD radix sort of uint64 keys over bits [0, end_bit), optionally
 * carrying int32 values (both value pointers NULL: keys only); scratch holds
 * jfm_sort_scratch_bytes(n, with_values) bytes. */
int32_t jfm_sort_scratch_bytes(int64_t n, int32_t with_values, int64_t* bytes);
int32_t jfm_sort_u64(void* stream, int64_t n, int32_t end_bit, const uint64_t* keys_in, uint64_t* keys_out,
                     const int32_t* vals_in, int32_t* vals_out, void* scratch, int64_t scratch_bytes);
/* Binary radix tree over the leaf keys sorted_codes[l * leaf_size]: parent
 * [nodes] (-1 at the root), left/right [L-1], node_range [nodes][2]. */
int32_t jfm_tree_radix(void* stream, int64_t n, int32_t leaf_size, const uint64_t* sorted_codes, int32_t* parent,
                       int32_t* left, int32_t* right, int32_t* node_range);
/* depth [nodes] by the parent chain, *max_depth (zeroed by the caller)
 * raised to the largest; flags |= 2 if a chain is longer than any radix tree
 * can be.  Level keys for the schedules: m2m_keys [I] = (D - depth) << 32 | node,
 * l2l_keys [nodes-1] = (depth - 1) << 32 | node for nodes 1..nodes-1. */
int32_t jfm_tree_levels(void* stream, int32_t nodes, int32_t I, const int32_t* parent, int32_t* depth,
                        int32_t* max_depth, int32_t* flags, uint64_t* m2m_keys, uint64_t* l2l_keys);
/* Centres of mass: every leaf (slot order sums), then one call per internal
 * level, deepest first (level_nodes = that level's slice of the M2M schedule). */
int32_t jfm_tree_leaf_geometry(void* stream, int32_t L, int32_t I, const double* pos, const double* mass,
                               const int32_t* perm, const int32_t* node_range, double* node_mass, double* center);
int32_t jfm_tree_level_geometry(void* stream, int32_t count, const int32_t* level_nodes, const double* pos,
                                const double* mass, const int32_t* perm, const int32_t* left, const int32_t* right,
                                const int32_t* node_range, double* node_mass, double* center);
/* radius [nodes] = max distance from the centre to the node's particles;
 * scratch holds n * 3 doubles (the positions in tree order). */
int32_t jfm_tree_radius(void* stream, int32_t nodes, int64_t n, const double* pos, const int32_t* perm,
                        const int32_t* node_range, const double* center, double* scratch, double* radius);
/* One breadth-first step of the dual-tree walk in three calls.  classify:
 * codes [n_in] (what item t emits) and counts3 [3][n_in] (items, far keys,
 * near keys of item t); jfm_scan_i32 on each row of counts3 gives offsets3;
 * emit: item t writes its child items at items_out[offsets3[0][t]], its far
 * pairs (target << 32 | source node) at far_keys[far_base + offsets3[1][t]]
 * and its near pairs (target << 32 | source leaf) at near_keys[near_base +
 * offsets3[2][t]] (the caller sizes every output from the scanned totals). */
int32_t jfm_tree_walk_classify(void* stream, int32_t I, const int32_t* items_in, int64_t n_in, const double* center,
                               const double* radius, double theta, uint8_t* codes, int32_t* counts3);
int32_t jfm_tree_walk_emit(void* stream, int32_t I, const int32_t* items_in, int64_t n_in, const int32_t* left,
                           const int32_t* right, const uint8_t* codes, const int32_t* offsets3, int64_t far_base,
                           int64_t near_base, int32_t* items_out, uint64_t* far_keys, uint64_t* near_keys);
/* Exclusive prefix sum of int32 values; scratch holds jfm_scan_scratch_bytes(n). */
int32_t jfm_scan_scratch_bytes(int64_t n, int64_t* bytes);
int32_t jfm_scan_i32(void* stream, int64_t n, const int32_t* in, int32_t* out, void* scratch, int64_t scratch_bytes);
/* Sorted (row << 32 | col) keys -> CSR off [rows+1], col [K]. */
int32_t jfm_keys_to_csr(void* stream, const uint64_t* sorted_keys, int64_t K, int32_t rows, int32_t* off,
                        int32_t* col);
/* Number of significant bits of max_key (sort widths). */
int32_t jfm_key_bits(uint64_t max_key);

/* ------------------------------------------------------------------------
 * Prepare-time list surgery (the reference's per-order pair bucketing before
 * its M2L launches, runtime/_interaction_cache.py p_gears groups, and the
 * flat pair layout behind jfm_m2l_flat).
 * ---------------------------------------------------------------------- */
/* row_of [K]: the row of every entry of a CSR with off [rows+1]. */
int32_t jfm_pair_rows(void* stream, int32_t rows, const int32_t* off, int64_t K, int32_t* row_of);
/* keys [K] = (slot << 32 | k), slot = index of gear_of_pair[k] in gears3[0..2];
 * flags |= 1 if a gear is not one of them. */
int32_t jfm_gear_keys(void* stream, int64_t K, const int8_t* gear_of_pair, const int32_t* gears3, uint64_t* keys,
                      int32_t* flags);
/* out_tgt[i] = row_of[index[i]], out_src[i] = src[index[i]] for i < n. */
int32_t jfm_gather_pairs(void* stream, int64_t n, const int32_t* index, const int32_t* row_of, const int32_t* src,
                         int32_t* out_tgt, int32_t* out_src);
/* keys[i] = tgt[i] << 32 | src[i]. */
int32_t jfm_pair_keys(void* stream, int64_t n, const int32_t* tgt, const int32_t* src, uint64_t* keys);
/* Appends every row r with off[r+1] > off[r] to out (rows entries of room),
 * *count (zeroed by the caller) += appended; order unspecified. */
int32_t jfm_nonempty_rows(void* stream, int32_t rows, const int32_t* off, uint64_t* out, uint64_t* count);
/* Targets whose pairs (pair_tgt [K], grouped by target) cross a 64-pair block
 * boundary: span_tgt, first block span_w0, last block span_w1 (room for
 * (K-1)/64 entries), *count (zeroed by the caller) = number found; order
 * unspecified (jfm_m2l_flat's fix-up treats every span on its own). */
int32_t jfm_flat_spans(void* stream, int64_t K, const int32_t* pair_tgt, int32_t* span_tgt, int32_t* span_w0,
                       int32_t* span_w1, uint64_t* count);

/* Dehnen real multipole swap matrices B_U for degrees 0..p, float64, packed
 * row-major block after block (sum (2l+1)^2 entries); the constants every
 * rotation kernel is compiled with (jaccpot/operators/real_dehnen_q.py:343-470). */
int32_t jfm_tables_bu(int32_t p, double* out, int64_t capacity);

/* ------------------------------------------------------------------------
 * Device kernels.
 * ---------------------------------------------------------------------- */

/* Gather input-order positions/masses into tree-order pos4.
 * src_dtype describes positions [n][3] / masses [n]; dtype describes pos4. */
int32_t jfm_gather_sorted(void* stream, int32_t src_dtype, int32_t dtype, const void* positions,
                          const void* masses, const int32_t* perm, int32_t n, void* pos4);

/* Leaf P2M about each leaf's centre of mass; also writes the leaves' centers4.
 * Replaces _p2m_leaves_real (upward/real_tree_expansions.py:89-213) with the
 * centres of yggdrax compute_tree_mass_moments (:439-440). */
int32_t jfm_p2m(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t num_leaves,
                int32_t leaf_size, int32_t num_internal, void* multipoles, void* centers4);

/* One M2M level: for each internal node in nodes[count], combine its children
 * (centre of mass, multipole = m2m(left) + m2m(right)); node_range [nodes][2]
 * (inclusive tree-order particle ranges) weights the mean of a massless node.
 * Replaces aggregate_m2m_real_by_level (upward/real_tree_expansions.py:216-364). */
int32_t jfm_m2m_level(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t count,
                      const int32_t* left, const int32_t* right, const int32_t* node_range, void* multipoles,
                      void* centers4);

/* M2L over the far list: locals[t] = sum_{s in far(t)} m2l(M[s], c_t - c_s) for
 * every t in targets[num_targets] (written, not accumulated; other rows untouched).
 * Replaces _apply_m2l + _accumulate_m2l_* (runtime/kernels/_m2l.py:1439-1793) and
 * m2l_real_fused_pallas (pallas/m2l_real_fused.py:493-573). */
int32_t jfm_m2l(void* stream, int32_t dtype, int32_t p, const int32_t* targets, int32_t num_targets,
                const int32_t* far_off, const int32_t* far_src, const void* multipoles,
                const void* centers4, void* locals);

/* M2L at a per-pair expansion order (gear): every pair of the CSR (far_off,
 * far_src) is translated at order gear <= p from the source multipole truncated
 * to degree gear, and the result is written into (accumulate == 0, other
 * coefficients of the row zeroed) or added to (accumulate != 0) degrees <= gear
 * of the target local.  Calling it once per gear over per-gear lists implements
 * the per-pair order the reference's adaptive policy tags
 * (runtime/_adaptive_policy.py:2216-2385, far_pairs_by_gear at
 * runtime/kernels/_l2l.py:491-493). */
int32_t jfm_m2l_gear(void* stream, int32_t dtype, int32_t p, int32_t gear, int32_t accumulate,
                     const int32_t* targets, int32_t num_targets, const int32_t* far_off, const int32_t* far_src,
                     const void* multipoles, const void* centers4, void* locals);

/* M2L over a flat pair list, one lane per pair: locals[pair_tgt[k]] +=
 * m2l_gear(M[pair_src[k]], c_t - c_s) for k < num_pairs, at order gear <= p
 * (degrees <= gear accumulated).  Pairs must be grouped by target (CSR order).
 * scratch holds jfm_m2l_flat_scratch_size(gear, num_pairs) elements of the
 * working dtype (per-wave partial sums); the span list names every target
 * whose pairs cross a 64-pair boundary, with the first and last 64-pair block
 * (index k / 64) holding its pairs.  Deterministic (fixed summation order,
 * no atomics).  Same operator as jfm_m2l / jfm_m2l_gear (the reference's
 * runtime/kernels/_m2l.py:1439-1793), with a mapping that keeps every lane busy
 * when targets have few pairs. */
int64_t jfm_m2l_flat_scratch_size(int32_t gear, int32_t num_pairs);
/* Engine of the float32 M2L rotations for every later jfm_m2l / jfm_m2l_gear /
 * jfm_m2l_flat launch of the process: JFM_M2L_VALU applies the B_U blocks as
 * unrolled FMAs with the constants folded in, JFM_M2L_MFMA applies them with
 * v_mfma_f32_4x4x1_16b_f32 (each block's connected components as 4-row
 * groups; degrees >= 3) while the VALU does the Dz rotations and the z-axis
 * core.  float64 always runs on the VALU.  The same operator as the reference's
 * dense rotate -> z-core -> rotate-back (pallas/m2l_real_fused.py:227-277,
 * B_U per degree from operators/m2l_real_rot_scale.py:161-170); the engine
 * changes rounding only.  jfm_m2l_engine() returns the current engine. */
#define JFM_M2L_VALU 0
#define JFM_M2L_MFMA 1
int32_t jfm_set_m2l_engine(int32_t engine);
int32_t jfm_m2l_engine(void);
int32_t jfm_m2l_flat(void* stream, int32_t dtype, int32_t p, int32_t gear, int32_t num_pairs,
                     const int32_t* pair_tgt, const int32_t* pair_src, const void* multipoles,
                     const void* centers4, void* locals, void* scratch, int32_t num_span,
                     const int32_t* span_tgt, const int32_t* span_w0, const int32_t* span_w1);

/* One L2L level: locals[c] += l2l(locals[parent[c]], c_parent - c_c) for c in nodes[count].
 * Replaces _propagate_solidfmm_locals_by_level (runtime/kernels/_l2l.py:210-409). */
int32_t jfm_l2l_level(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t count,
                      const int32_t* parent, const void* centers4, void* locals);

/* Mixed precision, top three levels (float32 arrays + float64 shadows).
 * shadow[nodes] maps each internal node of depth <= 2 to its shadow row
 * (0..6) and every other node to -1; mult64 [7][jfm_coeff_stride(p, JFM_F64)],
 * centers64 [7][4].  jfm_m2m_top: M2M of the shadowed nodes, levels deepest
 * first (nodes / lev_off[nlev+1] on the device, lev_off_host the same offsets
 * on the host), in float64, writing shadows and rounded float32 rows.
 * jfm_l2l_top: seeds the shadows from the float32 M2L rows of top_nodes[ntop],
 * then L2L in float64 into every node of depth 1..3 (levels shallowest first),
 * adding once-rounded results to the float32 rows below the shadows.  Call it
 * after jfm_m2l and before the float32 jfm_l2l_level sweeps of depth >= 4.
 * The float64 M2M/L2L of the top levels of jaccpot's large-N lane runs in
 * its working precision; the split is this library's precision policy. */
int32_t jfm_m2m_top(void* stream, int32_t p, const int32_t* nodes, const int32_t* lev_off, int32_t nlev,
                    const int32_t* lev_off_host, const int32_t* left, const int32_t* right, const int32_t* node_range,
                    const int32_t* shadow, void* multipoles, void* centers4, void* mult64, void* centers64);
int32_t jfm_l2l_top(void* stream, int32_t p, const int32_t* nodes, const int32_t* lev_off, int32_t nlev,
                    const int32_t* lev_off_host, const int32_t* top_nodes, int32_t ntop, const int32_t* parent,
                    const int32_t* shadow, const void* centers4, const void* centers64, void* locals, void* loc64);

/* L2P far field only: writes (accumulate == 0) or adds to (accumulate != 0)
 * rows out_index[i] of acc [rows][3] (and pot) with a = -G grad_delta phi,
 * pot = -G phi, delta = centre - x, for every tree slot i.
 * Replaces _evaluate_local_expansions_for_particles (runtime/kernels/_evaluate.py:2106-2485)
 * and evaluate_local_real_with_grad (operators/real_p2m_l2p.py:494-547). */
int32_t jfm_l2p(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t leaf_size,
                int32_t num_internal, const void* locals, const void* centers4, double G, const int32_t* out_index,
                void* acc, void* pot, int32_t accumulate);

/* Combine of a deferred far field: acc[k] += -G far_acc[k] for k < 3 n and,
 * when pot is not NULL, pot[r] += -G far_pot[r] for r < n, where far_acc /
 * far_pot are the rows of a jfm_l2p run with G = -1 (unscaled) and
 * accumulate == 0.  Equal bit for bit to a jfm_l2p that accumulates onto the
 * same near-field rows; it lets L2P run before the near field is done.
 * The far + near combine of the reference's evaluation
 * (runtime/kernels/_evaluate.py:2106-2485, accelerations = far + near). */
int32_t jfm_l2p_combine(void* stream, int32_t dtype, int32_t n, double G, const void* far_acc, const void* far_pot,
                        void* acc, void* pot);

/* P2P near field only (self leaf i != j plus near leaves), leaf_size <= 256.
 * flags: JFM_EVAL_POTENTIAL, JFM_EVAL_COMPENSATED.  accumulate != 0 adds into
 * the output rows instead of overwriting them.  Replaces
 * compute_leaf_p2p_accelerations (nearfield/near_field.py:916) and
 * nearfield_fused_leaf_pallas (pallas/nearfield_fused_leaf.py:288-510). */
int32_t jfm_p2p(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t num_leaves, int32_t leaf_size,
                const int32_t* near_off, const int32_t* near_src, double eps2, double G, const int32_t* out_index,
                void* acc, void* pot, int32_t flags, int32_t accumulate);

/* P2P over work items that split long near lists (the reference's bucketed
 * near-field launches, runtime/kernels/nearfield/_kernels.py:40-338, cut
 * into equal slot ranges so no leaf's sources serialise on one workgroup).
 * Item k covers slots [item_s0[k], item_s1[k]) of leaf item_leaf[k]'s
 * sequence (self leaf, then its near leaves, leaf_size slots each); item_s0
 * and item_s1 are multiples of leaf_size (whole source leaves);
 * item_part[k] < 0: the item is its leaf's only one and writes (or adds to,
 * accumulate != 0) the output rows like jfm_p2p; otherwise it writes its
 * targets' unscaled partial sums to partials[(item_part[k] * leaf_size + i) * 4
 * + {0..3}] (num_part rows of room).  jfm_p2p_reduce then adds, for every
 * split leaf, the partials split_p0[k] .. split_p1[k]-1 in that order
 * (compensated) and writes or adds the output rows. */
int32_t jfm_p2p_items(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t num_leaves, int32_t leaf_size,
                      int32_t num_items, const int32_t* item_leaf, const int32_t* item_s0, const int32_t* item_s1,
                      const int32_t* item_part, int32_t num_part, void* partials, const int32_t* near_off,
                      const int32_t* near_src, double eps2, double G, const int32_t* out_index, void* acc, void* pot,
                      int32_t flags, int32_t accumulate);
int32_t jfm_p2p_reduce(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t leaf_size, int32_t num_split,
                       const int32_t* split_leaf, const int32_t* split_p0, const int32_t* split_p1,
                       const void* partials, double eps2, double G, const int32_t* out_index, void* acc, void* pot,
                       int32_t flags, int32_t accumulate);

/* Fused leaf evaluation: P2P near field (self leaf i != j plus near leaves) and
 * L2P far field (a = -G grad_delta phi, delta = centre - x), per flags.  Results
 * for tree slot i go to row out_index[i] of acc [rows][3] (and pot [rows]);
 * out_index NULL means identity.  Replaces compute_leaf_p2p_accelerations
 * (nearfield/near_field.py:916), nearfield_fused_leaf_pallas
 * (pallas/nearfield_fused_leaf.py:288-510) and _evaluate_local_expansions_for_particles
 * (runtime/kernels/_evaluate.py:2106-2485) with the combine/un-permute of
 * _evaluate.py:1155-1170 and fmm_evaluate.py:531. */
int32_t jfm_evaluate(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t num_leaves,
                     int32_t leaf_size, int32_t num_internal, const int32_t* near_off, const int32_t* near_src,
                     const void* locals, const void* centers4, double eps2, double G,
                     const int32_t* out_index, void* acc, void* pot, int32_t flags);

/* ------------------------------------------------------------------------
 * Adjoint (reverse-mode) kernels: exact gradients of accelerations and
 * potentials w.r.t. positions and masses at a frozen topology, replacing the
 * reference's custom_vjp rules for P2P (nearfield/grad.py:354-869), L2P
 * (operators/real_p2m_l2p.py:403-491) and M2L (pallas/m2l_real_fused.py:576-829)
 * under differentiable_accelerations (runtime/fmm_evaluate.py:733).
 * cot4 [N][4] holds (abar_x, abar_y, abar_z, phibar) per tree slot; grad4
 * [N][4] accumulates (xbar, mbar) per tree slot.  Rows named *_flip hold
 * coefficients multiplied by (-1)^n; with them the tree passes of the reverse
 * sweep reuse the forward kernels: Lbar up = jfm_m2m_level on lbar_flip,
 * Mbar_own = jfm_m2l on lbar_flip (output is mbar_own_flip), Mbar down =
 * jfm_l2l_level on mbar_flip.  Order of calls: jfm_p2p_adjoint (writes grad4),
 * jfm_leaf_adjoint_a (adds L2P terms, writes leaf rows of lbar_flip), M2M
 * levels, M2L, jfm_center_adjoint (cbar4 [nodes][4] for the nodes with far
 * pairs; other rows must be zero), L2L levels, jfm_com_chain (chain
 * [L][8]), jfm_leaf_adjoint_b (P2M adjoint + centre-of-mass chain, writes
 * gpos [rows][3], gmass [rows] at out_index). */
int32_t jfm_p2p_adjoint(void* stream, int32_t dtype, const void* pos4, const void* cot4, int32_t n,
                        int32_t num_leaves, int32_t leaf_size, const int32_t* near_off, const int32_t* near_src,
                        double eps2, double G, void* grad4);
/* jfm_p2p_adjoint over the work items of jfm_p2p_items (same item arrays):
 * single-item leaves store their grad4 rows, split leaves leave partials that
 * jfm_p2p_adjoint_reduce adds in item order and stores. */
int32_t jfm_p2p_adjoint_items(void* stream, int32_t dtype, const void* pos4, const void* cot4, int32_t n,
                              int32_t num_leaves, int32_t leaf_size, int32_t num_items, const int32_t* item_leaf,
                              const int32_t* item_s0, const int32_t* item_s1, const int32_t* item_part,
                              int32_t num_part, void* partials, const int32_t* near_off, const int32_t* near_src,
                              double eps2, double G, void* grad4);
int32_t jfm_p2p_adjoint_reduce(void* stream, int32_t dtype, const void* cot4, int32_t n, int32_t leaf_size,
                               int32_t num_split, const int32_t* split_leaf, const int32_t* split_p0,
                               const int32_t* split_p1, const void* partials, double eps2, double G, void* grad4);
int32_t jfm_leaf_adjoint_a(void* stream, int32_t dtype, int32_t p, const void* pos4, const void* cot4, int32_t n,
                           int32_t leaf_size, int32_t num_internal, const void* locals, const void* centers4,
                           double G, void* grad4, void* lbar_flip);
int32_t jfm_center_adjoint(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t num_nodes,
                           const int32_t* far_off, const int32_t* far_src, const void* multipoles,
                           const void* locals_own, const void* lbar_flip, const void* mbar_own_flip,
                           const void* centers4, void* cbar4);
/* Mutual-lane reverse pass (level-weighted): jfm_p2p_adjoint with pair (i, j)
 * scaled by level_weights[clamp(max(rung_sorted[i], rung_sorted[j]), 0,
 * num_levels - 1)] (num_levels <= 64), and jfm_center_adjoint with far pair
 * (u, v) scaled by the weight of max(node_rung[u], node_rung[v]); with them
 * and jfm_m2l_flat_weighted (the transposed M2L on the symmetric list) the
 * reverse sweep differentiates jfm_mutual_* forces (the reference's mutual
 * lane under jax.grad, mutual/force.py:962-1068). */
int32_t jfm_p2p_adjoint_weighted(void* stream, int32_t dtype, const void* pos4, const void* cot4, int32_t n,
                                 int32_t num_leaves, int32_t leaf_size, const int32_t* near_off,
                                 const int32_t* near_src, const int32_t* rung_sorted, const void* level_weights,
                                 int32_t num_levels, double eps2, double G, void* grad4);
/* jfm_p2p_adjoint_weighted over work items (as jfm_p2p_adjoint_items; the
 * split leaves' partials go through jfm_p2p_adjoint_reduce). */
int32_t jfm_p2p_adjoint_weighted_items(void* stream, int32_t dtype, const void* pos4, const void* cot4, int32_t n,
                                       int32_t num_leaves, int32_t leaf_size, int32_t num_items,
                                       const int32_t* item_leaf, const int32_t* item_s0, const int32_t* item_s1,
                                       const int32_t* item_part, int32_t num_part, void* partials,
                                       const int32_t* near_off, const int32_t* near_src, const int32_t* rung_sorted,
                                       const void* level_weights, int32_t num_levels, double eps2, double G,
                                       void* grad4);
int32_t jfm_center_adjoint_weighted(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t num_nodes,
                                    const int32_t* far_off, const int32_t* far_src, const int32_t* node_rung,
                                    const void* level_weights, int32_t num_levels, const void* multipoles,
                                    const void* locals_own, const void* lbar_flip, const void* mbar_own_flip,
                                    const void* centers4, void* cbar4);
/* jfm_center_adjoint for per-pair orders (learned gears): pair k = (u <- v) of
 * the far CSR ran at order gear_fwd[k], its mirror (v <- u) at gear_rev[k]
 * (orders <= p).  With it, Mbar_own of each gear g is jfm_m2l_gear at gear g
 * over the transposed gear-g pair lists.  Counterpart of the reference's
 * gradient through the tagged per-gear M2L (runtime/_adaptive_policy.py
 * :2216-2385 under runtime/fmm_evaluate.py:733), which JAX differentiates
 * with the tags held fixed. */
int32_t jfm_center_adjoint_gears(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t num_nodes,
                                 const int32_t* far_off, const int32_t* far_src, const int8_t* gear_fwd,
                                 const int8_t* gear_rev, const void* multipoles, const void* locals_own,
                                 const void* lbar_flip, const void* mbar_own_flip, const void* centers4,
                                 void* cbar4);
int32_t jfm_com_chain(void* stream, int32_t dtype, int32_t num_leaves, int32_t num_internal, const int32_t* parent,
                      const int32_t* node_range, const void* centers4, const void* cbar4, void* chain);
int32_t jfm_leaf_adjoint_b(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t leaf_size,
                           int32_t num_internal, const void* mbar_flip, const void* centers4, const void* chain,
                           const void* grad4, const int32_t* out_index, void* gpos, void* gmass);

/* ------------------------------------------------------------------------
 * Learned per-pair M2L order (no reference counterpart: the reference tags
 * gears with a hand-written error model, runtime/_adaptive_policy.py:2216-2384,
 * and then runs every pair at full order, runtime/kernels/_l2l.py:491-493).
 * Features [K][8] of far pairs (k -> target tgt_of[k], source far_src[k]):
 * log2(d/(r_t+r_s)), log10(m_s/m_total), log10(m_t/m_total), and
 * log10(|M_s,n|/(m_s r_s^n)) for n = 2..6; radius [nodes] float32 bounding radii.
 * Labels [K][3]: log10 of the relative force error on the target sphere of the
 * order gears3[j] translation against order 8 (degree weights n r_t^(n-1),
 * rotation-invariant degree norm; floor 1e-7; multipoles8 at order 8, float32).
 * MLP 8 -> 64 -> 64 -> 3 (ReLU), jfm_mlp_num_params() float32 parameters
 * W1[64][8] b1 W2[64][64] b2 W3[3][64] b3, used as bfloat16 (wbf16) with
 * float32 accumulation.  jfm_mlp_backward: gradient of the mean squared error
 * over the K pairs (grad_and_loss[num_params] = loss), deterministic; scratch
 * holds jfm_mlp_backward_scratch_floats(K, per_block) floats.  jfm_mlp_adam:
 * one Adam step on the float32 masters, refreshing wbf16.  jfm_mlp_forward
 * normalises the features on the fly when mean / inv_std [8] are given
 * (both NULL: X is already normalised).  jfm_select_gears:
 * per far pair of the CSR (far_off [num_nodes+1], far_src), with predictions
 * [K,3] in CSR order and w = log10(m_s / d^2): the smallest of gears3[0..2]
 * with y_g + w <= budget_t + tol_log10, where budget_t is the largest
 * y_top + w over target t's pairs; gears3[2] otherwise.  A top gear of 8 is
 * the label reference (floor label): y_top is then extrapolated,
 * y_top = y_1 + (g2 - g1) / (g1 - g0) min(y_1 - y_0, 0). */
int32_t jfm_pair_features(void* stream, int32_t p, int32_t K, const int32_t* tgt_of, const int32_t* far_src,
                          const void* multipoles, const void* centers4, const float* radius, double total_mass,
                          float* features);
int32_t jfm_pair_labels(void* stream, int32_t K, const int32_t* tgt_of, const int32_t* far_src,
                        const void* multipoles8, const void* centers4, const float* radius, const int32_t* gears3,
                        float* labels);
int32_t jfm_mlp_num_params(void);
int64_t jfm_mlp_backward_scratch_floats(int32_t K, int32_t per_block);
int32_t jfm_mlp_forward(void* stream, int32_t K, const float* X, const float* mean, const float* inv_std,
                        const uint16_t* wbf16, float* Y);
/* The same forward pass on the matrix cores (bf16 MFMA, fp32 accumulation;
 * activations rounded to bf16 between layers): the inference path. */
int32_t jfm_mlp_forward_mfma(void* stream, int32_t K, const float* X, const float* mean, const float* inv_std,
                             const uint16_t* wbf16, float* Y);
int32_t jfm_mlp_backward(void* stream, int32_t K, int32_t per_block, const float* X, const float* targets,
                         const uint16_t* wbf16, float* scratch, float* grad_and_loss);
int32_t jfm_mlp_adam(void* stream, int32_t step, float lr, float beta1, float beta2, float eps, float* master,
                     const float* grad, float* m1, float* m2, uint16_t* wbf16);
int32_t jfm_mlp_to_bf16(void* stream, const float* master, uint16_t* wbf16);
int32_t jfm_select_gears(void* stream, int32_t num_nodes, const int32_t* far_off, const int32_t* far_src,
                         const float* predictions, const void* centers4, float tol_log10, const int32_t* gears3,
                         int8_t* gears);
/* Particle-level budget of the gear choice (gear_budget="leaf"):
 * jfm_gear_node_max writes, per node, the largest top-gear error proxy of its
 * far pairs (-inf without pairs); jfm_level_path_max (one tree level, root
 * first) takes values[u] = max(values[u], values[parent[u]]);
 * jfm_level_subtree_min (internal nodes of one level, deepest first) takes
 * values[u] = min(values[left[u]], values[right[u]]); jfm_select_gears_budget
 * is jfm_select_gears with those per-target budgets.  A refinement of the
 * per-target rule for the reference's tagged gears
 * (runtime/_adaptive_policy.py:2216-2385). */
int32_t jfm_gear_node_max(void* stream, int32_t num_nodes, const int32_t* far_off, const int32_t* far_src,
                          const float* predictions, const void* centers4, const int32_t* gears3, float* node_max);
int32_t jfm_level_path_max(void* stream, const int32_t* nodes, int32_t count, const int32_t* parent, float* values);
int32_t jfm_level_subtree_min(void* stream, const int32_t* nodes, int32_t count, const int32_t* left,
                              const int32_t* right, float* values);
int32_t jfm_select_gears_budget(void* stream, int32_t num_nodes, const int32_t* far_off, const int32_t* far_src,
                                const float* predictions, const void* centers4, float tol_log10,
                                const int32_t* gears3, const float* budget, int8_t* gears);

/* Float64 direct sum at targets[nt] (tree slots) over all n particles of pos4
 * (float64), self excluded; writes acc [nt][3] and pot [nt] (pot may be NULL).
 * scratch must hold jfm_direct_scratch_doubles(nt) doubles.
 * Replaces direct_sum_gravitational_acceleration (jaccpot/autodiff.py:12-98). */
int64_t jfm_direct_scratch_doubles(int32_t nt);
int32_t jfm_direct_sum_f64(void* stream, const double* pos4, int32_t n, const int32_t* targets, int32_t nt,
                           double eps2, double G, double* scratch, double* acc, double* pot);

/* jfm_direct_sum_f64 without a scratch argument and without any allocation:
 * the per-chunk partials live in a fixed per-GPU device array (32 MiB, held
 * by each GPU the library runs on until the process ends), so at most
 * jfm_direct_sum_f64_static_max_targets() = 4096 targets per call; more is
 * refused with JFM_ERR_ARG.  Same kernels, same summation order, same result
 * bit for bit as jfm_direct_sum_f64.  The array belongs to the current device
 * (the stream must be on it); every call on a GPU reuses it, so calls that
 * overlap on one GPU (other streams, other threads) race.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc). */
int32_t jfm_direct_sum_f64_static_max_targets(void);
int32_t jfm_direct_sum_f64_static(void* stream, const double* pos4, int32_t n, const int32_t* targets, int32_t nt,
                                  double eps2, double G, double* acc, double* pot);

/* ------------------------------------------------------------------------
 * Momentum-conserving (mutual) lane with block-step level weights, the
 * reference's jaccpot/mutual (force.py:985-1068: mutual_weighted_accelerations;
 * nearfield.py: mutual_near_field_forces; farfield.py:
 * mutual_far_field_forces).  The topology's far and near lists are symmetric
 * (every unordered pair once, both directions emitted), so each pair's two
 * halves come from the same operator; interaction level of a particle pair
 * is max(rung_i, rung_j), of a cell pair max(rung_A, rung_B) with the cell
 * rung the maximum over its particles (force.py:885-960).
 *
 * jfm_mutual_rungs: rung_sorted[j] = rung[perm[j]] and node_rung[leaf node]
 * = max over the leaf (-1 if empty); node_rung [nodes].
 * jfm_mutual_rungs_level: node_rung[u] = max(node_rung[left[u]],
 * node_rung[right[u]]) for u in nodes[count] (one tree level, deepest first). */
int32_t jfm_mutual_rungs(void* stream, int32_t n, int32_t num_leaves, int32_t leaf_size, const int32_t* perm,
                         const int32_t* rung, int32_t* rung_sorted, int32_t* node_rung);
int32_t jfm_mutual_rungs_level(void* stream, const int32_t* nodes, int32_t count, const int32_t* left,
                               const int32_t* right, int32_t* node_rung);
/* Near-field forces F_i = G sum_j m_i m_j w[max(r_i, r_j)] (x_j - x_i) /
 * (|x_j - x_i|^2 + eps2)^{3/2} over the self leaf (j != i) and the near
 * leaves; each pair term is the exact negative of its mirror (j, i) to the
 * last bit, and G scales the per-target sums; rows force[out_index[i]][3]
 * (out_index NULL: tree order).  level_weights NULL: every pair weight 1
 * (rung_sorted unused).  leaf_size <= 256, num_levels <= 64. */
int32_t jfm_mutual_p2p(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t num_leaves,
                       int32_t leaf_size, const int32_t* near_off, const int32_t* near_src, const int32_t* rung_sorted,
                       const void* level_weights, int32_t num_levels, double eps2, double G, const int32_t* out_index,
                       void* force);
/* jfm_mutual_p2p over work items that cut long near lists into bounded
 * slot ranges (the reference chunks its near-pair work the same way,
 * mutual/nearfield.py:58-83 resolve_near_chunk_size, there for memory).
 * Items, partials and split arrays as jfm_p2p_items / jfm_p2p_reduce
 * (whole source leaves per item; partials rows of 4, the fourth unused);
 * the split leaves' partials are added in item order (compensated) and
 * written to their force rows, so the result is deterministic. */
int32_t jfm_mutual_p2p_items(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t num_leaves,
                             int32_t leaf_size, int32_t num_items, const int32_t* item_leaf, const int32_t* item_s0,
                             const int32_t* item_s1, const int32_t* item_part, int32_t num_part, void* partials,
                             int32_t num_split, const int32_t* split_leaf, const int32_t* split_p0,
                             const int32_t* split_p1, const int32_t* near_off, const int32_t* near_src,
                             const int32_t* rung_sorted, const void* level_weights, int32_t num_levels, double eps2,
                             double G, const int32_t* out_index, void* force);
/* jfm_m2l_flat at gear p with pair (t <- s) scaled by
 * level_weights[clamp(max(node_rung[t], node_rung[s]), 0, num_levels - 1)]
 * (level_weights in device memory, the working dtype). */
int32_t jfm_m2l_flat_weighted(void* stream, int32_t dtype, int32_t p, int32_t num_pairs, const int32_t* pair_tgt,
                              const int32_t* pair_src, const int32_t* node_rung, const void* level_weights,
                              int32_t num_levels, const void* multipoles, const void* centers4, void* locals,
                              void* scratch, int32_t num_span, const int32_t* span_tgt, const int32_t* span_w0,
                              const int32_t* span_w1);
/* Per input row r: far_force = m_r far_acc[r], acc = (near_force + far_force)
 * / m_r (0 where m_r <= 0); far_force may be NULL.  All [n][3] except
 * masses [n], in input order. */
int32_t jfm_mutual_combine(void* stream, int32_t dtype, int32_t n, const void* masses, const void* far_acc,
                           const void* near_force, void* acc, void* far_force);

/* ------------------------------------------------------------------------
 * Validation only: host (CPU) instantiation of the kernels' per-pair operator
 * math, one pair per call, for checking against the float64 oracle without a
 * GPU.  No product path uses it; it has no counterpart in the reference.
 *   op 0 m2m (in: multipole, delta = source - destination)
 *   op 1 m2l (in: multipole, delta = target - source)
 *   op 2 l2l (in: local,     delta = parent - child)
 *   op 3 p2m (in[0]: mass,   delta = particle - centre)
 *   op 4 l2p (in: local,     delta = centre - point; out = phi, dphi/ddelta[3])
 * in/out hold (p+1)^2 doubles; dtype selects the arithmetic precision. */
int32_t jfm_host_operator(int32_t op, int32_t p, int32_t dtype, const double* in, const double* delta, double* out);

#ifdef __cplusplus
}
#endif

#endif /* JFM_H */
