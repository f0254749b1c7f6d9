/* kiddo_amd.h — C ABI of the MI355X batch query engine for Kiddo's ImmutableKdTree.
 *
 * The reference exposes one query at a time through a fluent builder
 * (kiddo src/kd_tree/query/builder.rs:2318-2331 `KdTree::query`, families at :2008-2077,
 * `execute` at :1693); callers batch by looping or with rayon
 * (examples/immutable-large.rs:59-69). Each entry point below is that builder chain run
 * over a whole batch of queries on one GPU. Plain pointers and sizes only.
 *
 * Conventions
 *   - Every function returns KD_OK (0) or a negative KD_ERR_* code and never aborts;
 *     kd_last_error() describes the last failure on the calling thread.
 *   - Query and result pointers are DEVICE pointers (HBM); `stream` is a hipStream_t
 *     (NULL = the default stream). Calls are stream-ordered and asynchronous except where
 *     noted; results are ready once the stream is synchronised.
 *   - dtype: KD_F32 / KD_F64 / KD_U16 / KD_U32 coordinates (the tree's A). Queries are a
 *     row-major m x dims array of A (Kiddo's `&[A; K]` per query).
 *   - Distances are SquaredEuclidean in the output type O: O = A for f32/f64 trees, and
 *     O = f64 for u16/u32 trees (exact there: every distance is an integer below 2^53).
 *   - Radii are in metric units, i.e. SQUARED distance (Kiddo's `within(radius)` argument
 *     is `Dq::Output`, builder.rs:2033). Pass +INFINITY for "no radius".
 *   - Result order: by (distance, then lower item) for nearest_n and sorted within;
 *     by item ascending for best_n_within; Kiddo's boundary rules (inclusive `<=` unless
 *     KD_EXCLUSIVE, builder.rs:1271-1278).
 */
#ifndef KIDDO_AMD_H
#define KIDDO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum kd_status {
  KD_OK = 0,
  KD_ERR_INVALID = -1,        /* bad argument (null, size, NaN coordinate, ...)          */
  KD_ERR_UNSUPPORTED = -2,    /* dtype/dims/flag combination not provided                */
  KD_ERR_DEVICE = -3,         /* HIP runtime error                                       */
  KD_ERR_LEAF_OVERFLOW = -4,  /* a leaf would exceed 1024 entries (construction.rs:526)  */
  KD_ERR_TOO_DEEP = -5,       /* more stem levels than the engine supports               */
  KD_ERR_NO_MEMORY = -6,      /* device or host allocation failed                        */
  KD_ERR_CAPACITY = -7        /* caller-provided output capacity too small               */
};

enum kd_dtype { KD_F32 = 0, KD_F64 = 1, KD_U16 = 2, KD_U32 = 3 };

/* query flags */
enum kd_query_flags {
  KD_EXCLUSIVE = 1u << 0, /* `.exclusive_boundaries()`: strict `< radius`              */
  KD_UNSORTED = 1u << 1,  /* `.unsorted()`: within results in visit order              */
  KD_MORTON = 1u << 2,    /* traverse queries in Morton order (results keep input order) */
  KD_SCREEN_F32 = 1u << 3, /* k-NN with the FP32 collector (same results): f64 trees  */
                           /* screened in FP32, rescored in FP64; f32 trees exactly     */
  KD_GRAPH = 1u << 4        /* forests: capture the launch in a hipGraph and replay it     */
};

/* Distance metric of a query (the `D` of `nearest_n::<D>(k)` etc., src/dist/): bits 8-11
 * of the flags hold the kind, bits 12-19 the Minkowski exponent P. Every tree query entry
 * point (Cartesian and periodic, all families) and the server take it; batched forests
 * and KD_SCREEN_F32 are SquaredEuclidean only (KD_ERR_UNSUPPORTED). Radii are in the
 * metric's own units (Manhattan and Chebyshev: plain distance; Minkowski: sum of |d|^P).
 *   KD_METRIC_SQEUCLIDEAN  squared_euclidean/mod.rs: sum of d*d (the default, 0)
 *   KD_METRIC_MANHATTAN    manhattan/mod.rs:36-49: sum of |d|
 *   KD_METRIC_CHEBYSHEV    chebyshev/mod.rs:34-61: max of |d|
 *   KD_METRIC_MINKOWSKI(P) minkowski/mod.rs:55-59: sum of |d|^P, P >= 3 (P = 1, 2: use the
 *                          two above, as the reference requires; KD_ERR_INVALID) */
#define KD_METRIC_SQEUCLIDEAN 0u
#define KD_METRIC_MANHATTAN (1u << 8)
#define KD_METRIC_CHEBYSHEV (2u << 8)
#define KD_METRIC_MINKOWSKI(p) ((3u << 8) | ((uint32_t)(p) << 12))
#define KD_METRIC_MASK 0x000FFF00u

/* tree-create flags */
enum kd_tree_flags {
  KD_TREE_SHADOW_F32 = 1u << 0 /* f64 and f32 trees: keep an FP32 shadow of the leaves    */
};

/* A tree in Kiddo's ImmutableKdTree layout, exactly what a Rust shim reads through the
 * public KdTreeAccessor (kiddo src/traits/kd_tree.rs:16-46) and
 * VecOfArenas (src/leaf_strategy/vec_of_arenas.rs:60-69). All pointers are host memory. */
typedef struct kd_tree_desc {
  int32_t dtype;            /* kd_dtype                                                */
  int32_t dims;             /* K (2, 3 or 4)                                           */
  int32_t bucket;           /* B (construction bucket size)                            */
  int32_t max_stem_level;   /* KdTree::max_stem_level; -1 for a stemless tree          */
  uint64_t size;            /* number of entries                                       */
  uint64_t stems_len;       /* number of A values in `stems`                           */
  const void* stems;        /* Eytzinger stems, index 0 unused, sentinel = A::MAX/+inf */
  uint64_t leaf_count;      /* 2^(max_stem_level+1), or 1/0 for stemless trees         */
  const uint64_t* leaf_extents; /* leaf_count pairs (byte offset, entry count)         */
  uint64_t leaf_bytes_len;
  const uint8_t* leaf_bytes;    /* VecOfArenas arena (tiles 32/8/4/2/1, coords then u32 items) */
} kd_tree_desc;

typedef struct kd_layout kd_layout; /* host-side tree produced by kd_build_layout */
typedef struct kd_tree kd_tree;     /* device-resident tree                         */

typedef struct kd_tree_info {
  int32_t dtype, dims, max_stem_level, has_shadow_f32;
  uint64_t size, leaf_count, max_leaf_len;
  uint64_t device_bytes;    /* HBM held by the tree */
  double bbox_min[4], bbox_max[4];
} kd_tree_info;

/* ----------------------------------------------------------------------- misc ---- */
const char* kd_last_error(void);
int kd_version(void);

/* ------------------------------------------------------------ host construction ---- */
/* Builds the ImmutableKdTree layout on the host, as ImmutableKdTree::new_from_slice
 * (items == NULL: item = source index) or new_from_entries (explicit u32 items)
 * (kiddo src/kd_tree/construction.rs:278-303 and :403-411). points: row-major n x dims.
 * threads <= 0: all host cores. */
int kd_build_layout(int32_t dtype, int32_t dims, int32_t bucket, uint64_t n,
                    const void* points, const uint32_t* items, int32_t threads,
                    kd_layout** out);
/* Describes a built layout (pointers stay valid until kd_layout_free). */
int kd_layout_desc(const kd_layout* layout, kd_tree_desc* out);
void kd_layout_free(kd_layout* layout);

/* A KdTree with Donnelly<block_height> stems (kiddo src/stem_strategy/donnelly/core.rs:
 * stems in lines of 2^block_height slots, each a triangle of block_height levels; the
 * stem depth padded to a multiple of block_height with pivot-less levels above the real
 * root, construction.rs:431-444), re-laid out on the host in the Eytzinger layout that
 * kd_tree_create takes: the same pivot at every node of the same level, the same leaves
 * (the empty leaves under the padding levels appended), so every query answers exactly as
 * on the Donnelly tree. `donnelly` is the tree as a Rust shim reads it through
 * KdTreeAccessor (src/traits/kd_tree.rs:16-46): stems in Donnelly order, leaf_count real
 * leaves, max_stem_level including the padding. Free with kd_layout_free. */
int kd_layout_from_donnelly(const kd_tree_desc* donnelly, int32_t block_height, kd_layout** out);

/* ----------------------------------------------------------------- device trees ---- */
/* Uploads a Kiddo-layout tree to the current HIP device and re-lays it out for the
 * kernels (synchronous). Replaces holding `&ImmutableKdTree` (kiddo src/lib.rs:270-271).
 * Rejects NaN coordinates, leaves over 1024 entries and inconsistent extents. */
int kd_tree_create(const kd_tree_desc* desc, uint32_t tree_flags, kd_tree** out);

/* ImmutableKdTree::new_from_slice / new_from_entries built on the GPU from points already
 * in device memory (kiddo src/kd_tree/construction.rs:413-539, :641-691, :868-968,
 * :1089-1175): level-synchronous radix sorts, one split per node. Same stems (bit for
 * bit) and the same entries in every leaf as kd_build_layout; the order inside a leaf
 * may differ (no query observes it). points: device, row-major n x dims; items: device
 * u32[n] or NULL (item = source index). Same argument checks as kd_build_layout and
 * kd_tree_create (NaN / sentinel coordinates, leaves over 1024 entries); n < 2^31.
 * Synchronous on `stream` when it returns. */
int kd_tree_build_device(int32_t dtype, int32_t dims, int32_t bucket, uint64_t n,
                         const void* points, const uint32_t* items, uint32_t tree_flags,
                         kd_tree** out, void* stream);
/* Copies a device tree back to the host in Kiddo's layout (stems, extents, arena; the
 * buffers kd_layout_desc describes), e.g. for a Rust shim to own a KdTree built on the
 * GPU (KdTreeAccessor, src/traits/kd_tree.rs:16-46). Free with kd_layout_free. */
int kd_tree_export_layout(const kd_tree* tree, kd_layout** out);
int kd_tree_destroy(kd_tree* tree);
int kd_tree_get_info(const kd_tree* tree, kd_tree_info* out);

/* Launch tuning (no Kiddo counterpart; results never depend on it).
 * knn_variant: 0 automatic (for k <= 16, group_size lanes per query, 0 = 8 or 16, on
 * batches below 2^16 queries, or 2^19 on trees of at least 2^22 entries; else one thread
 * per query; nearest_one takes 8 lanes per query on the same batches), 1 one thread per
 * query (nearest_one too), 2 one group of group_size lanes per query
 * (group_size 0 = smallest of 4/8/16/32/64 that holds k). lds_levels: top stem
 * levels copied into LDS per block (-1 automatic, 0 none, at most 14). */
int kd_tree_tune(kd_tree* tree, int32_t knn_variant, int32_t group_size, int32_t lds_levels);
/* As kd_tree_tune, plus the minimum waves per SIMD the thread-per-query k-NN kernel is
 * compiled for (0 = default 3; 4..6 trade registers for occupancy). */
int kd_tree_tune2(kd_tree* tree, int32_t knn_variant, int32_t group_size, int32_t lds_levels,
                  int32_t waves);

/* Launch shapes of the two within passes (no Kiddo counterpart; results never depend on
 * them): 1 one thread per query, 4/8/16/32 lanes per query sharing each leaf visit
 * (SquaredEuclidean; other metrics use one thread per query), 0 automatic (the count pass:
 * one thread per query; the fill pass: 16 lanes per query when the batch averages 32 or
 * more results per query, else one thread). */
int kd_tree_tune_within(kd_tree* tree, int32_t count_group, int32_t fill_group);
/* Launch shape of best_n_within (no Kiddo counterpart; results never depend on it): 1 one
 * thread per query, 4/8/16/32 lanes per query sharing each leaf visit (SquaredEuclidean;
 * a shape with fewer lanes than n, and other metrics, use one thread per query), 0
 * automatic (for n <= 16: 16 lanes below 2^17 queries, 8 lanes for n <= 8 and 16 for
 * n <= 16 below 2^22 queries; one thread per query above). */
int kd_tree_tune_best_n_within(kd_tree* tree, int32_t group);
/* Autotuner of the within passes: times each shape of the count pass on up to 2^20 of the
 * queries (with KD_MORTON: 16 runs of their Morton order), then each shape of the fill pass
 * on as many of them as 2^29 results allow, and keeps the fastest of each in the tree (a
 * fill shape with more lanes per query only when more than 10% faster than the best with
 * fewer) (chosen2 = {count shape, fill shape}). Synchronous. */
int kd_autotune_within(kd_tree* tree, const void* queries, uint64_t m, double radius, uint32_t flags,
                       int32_t* chosen2, void* stream);

/* Autotuner: times each candidate launch shape (thread vs group per query, group size,
 * LDS stem levels, occupancy floor) on up to 2^20 of the given queries (with KD_MORTON:
 * 16 contiguous runs of their Morton order) with the given k/radius/flags, keeps the
 * fastest in the tree (as kd_tree_tune2 would) and reports it in
 * chosen4 = {variant, group_size, lds_levels, waves}. Synchronous. */
int kd_autotune_nearest_n(kd_tree* tree, const void* queries, uint64_t m, uint32_t k,
                          double radius, uint32_t flags, int32_t* chosen4, void* stream);
/* Bucket size picked by measurement for one tree built on the device from n points in HBM
 * (kd_tree_build_device arguments; no Kiddo counterpart: the bucket is ImmutableKdTree's
 * const generic B, construction.rs:278-303, and results do not depend on it): the tree is
 * built at each of the nbuckets candidates (1..1024), its k-NN launch shape autotuned and
 * nearest_n(k, radius, flags) timed on the batch's tuning sample (up to 2^20 queries,
 * Morton-coherent under KD_MORTON); *best_bucket receives the fastest, ms_out[i] (optional)
 * each candidate's kernel time. The caller then builds the tree at *best_bucket. */
int kd_autotune_bucket(int32_t dtype, int32_t dims, uint64_t n, const void* points, const uint32_t* items,
                       uint32_t tree_flags, const void* queries, uint64_t m, uint32_t k, double radius,
                       uint32_t flags, const int32_t* buckets, uint32_t nbuckets, int32_t* best_bucket,
                       float* ms_out, void* stream);

/* --------------------------------------------------------------------- queries ---- */
/* tree.query(q).nearest_one::<SquaredEuclidean>().execute()  (builder.rs:2008-2017;
 * query/nearest_one.rs:61-143). items[m], dists[m] (O). Empty tree: (+inf, 0). */
int kd_nearest_one(kd_tree* tree, const void* queries, uint64_t m, uint32_t flags,
                   uint32_t* items, void* dists, void* stream);

/* tree.query(q).nearest_one().approx().execute()  (builder.rs:2059-2065;
 * query/approx_nearest_one.rs:49-71): scans only the leaf the query descends to.
 * Float trees only (Kiddo requires Output == A). dists are in A. */
int kd_approx_nearest_one(kd_tree* tree, const void* queries, uint64_t m, uint32_t flags,
                          uint32_t* items, void* dists, void* stream);

/* tree.query(q).nearest_n(k)[.within(radius)][.exclusive_boundaries()][.unsorted()]
 * (builder.rs:2019-2042; query/nearest_n.rs:21-38, nearest_n_within.rs:127-179).
 * radius = +INFINITY for plain nearest_n. Row-major m x k outputs; row i holds counts[i]
 * results sorted by (distance, item) and is padded with item 0xFFFFFFFF, distance +inf.
 * KD_SCREEN_F32 (trees created with KD_TREE_SHADOW_F32): on f64 trees the traversal screens
 * in FP32 against a certified error bound and rescores survivors in FP64; on f32 trees
 * (k <= 16) it keeps FP32 distances only and resolves items at the end. Results are
 * identical to the call without the flag. */
int kd_nearest_n(kd_tree* tree, const void* queries, uint64_t m, uint32_t k, double radius,
                 uint32_t flags, uint32_t* items, void* dists, uint32_t* counts, void* stream);

/* kd_nearest_n that allocates nothing: no counterpart in the reference. It is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and errors as kd_nearest_n;
 * the Morton ordering of KD_MORTON runs in a scratch array the library holds in static
 * device memory (32 MiB per GPU, kept until the process ends) instead of a per-call
 * allocation. Limit: with KD_MORTON, m <= KD_NOALLOC_MAX_QUERIES (KD_ERR_CAPACITY
 * beyond); without it any m (no scratch is used). The tree's device must be current.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU (on
 * different streams, or from several threads) race; order them on one stream. */
#define KD_NOALLOC_MAX_QUERIES (1u << 20)
int kd_nearest_n_noalloc(kd_tree* tree, const void* queries, uint64_t m, uint32_t k, double radius,
                         uint32_t flags, uint32_t* items, void* dists, uint32_t* counts, void* stream);

/* The order KD_MORTON visits a batch in (no Kiddo counterpart): perm[m] receives the query
 * indices sorted by the Morton key of their coordinates over the tree's bounding box. A
 * caller that cuts a large batch into chunks (for example within results bounded by result
 * memory) can take contiguous ranges of this order, so that every chunk stays spatially
 * coherent. Device pointers; stream-ordered. */
int kd_morton_order(kd_tree* tree, const void* queries, uint64_t m, uint32_t* perm, void* stream);

/* tree.query(q).within(radius)[.unsorted()][.exclusive_boundaries()], phase 1
 * (query/within.rs:18-41, within_unsorted.rs:133-156): counts results per query and
 * writes offsets[m+1] as their exclusive prefix sum (offsets[m] = total). */
int kd_within_count(kd_tree* tree, const void* queries, uint64_t m, double radius,
                    uint32_t flags, uint64_t* offsets, void* stream);
/* Phase 2: writes each query's results into items/dists[offsets[i] .. offsets[i+1]).
 * Sorted by (distance, item) unless KD_UNSORTED (then traversal visit order). */
int kd_within_fill(kd_tree* tree, const void* queries, uint64_t m, double radius,
                   uint32_t flags, const uint64_t* offsets, uint32_t* items, void* dists,
                   void* stream);

/* tree.query(q).best_n_within(radius, n)[.exclusive_boundaries()]  (builder.rs:2044-2057;
 * query/best_n_within.rs:88-196): up to n entries within radius with the smallest items.
 * Row-major m x n outputs sorted by item ascending, padded like kd_nearest_n. */
int kd_best_n_within(kd_tree* tree, const void* queries, uint64_t m, double radius,
                     uint32_t n, uint32_t flags, uint32_t* items, void* dists,
                     uint32_t* counts, void* stream);

/* ------------------------------------------------ periodic boundary conditions ---- */
/* tree.query(q).periodic_boundary_condition(&box)...  (builder.rs:1280-1291, 1989-2003;
 * query/periodic.rs). box: K strictly positive lengths in the tree's coordinate type
 * (host memory; builder.rs:29-35 rejects a length <= 0 or NaN: KD_ERR_INVALID). Float
 * trees only (PeriodicAxis is implemented for f32/f64: KD_ERR_UNSUPPORTED otherwise).
 * Distances of nearest_n / within are minimum-image distances (per axis the smaller of
 * |q - c| and L - |q - c|); best_n_within is Cartesian-only in the reference. */

/* .nearest_one() (periodic.rs:196-244): Cartesian nearest_one of the home query, then of
 * the shifted query images (every axis -L, 0 or +L) whose lower bound does not exceed the
 * best distance, in ascending lower bound; dists in A, each being the Cartesian distance
 * to the image that produced the result. Ties: lower item. */
int kd_periodic_nearest_one(kd_tree* tree, const void* queries, uint64_t m, const void* box,
                            uint32_t flags, uint32_t* items, void* dists, void* stream);

/* .nearest_n(k)[.within(radius)][.exclusive_boundaries()][.unsorted()]
 * (periodic.rs:620-791). Outputs as kd_nearest_n. */
int kd_periodic_nearest_n(kd_tree* tree, const void* queries, uint64_t m, const void* box,
                          uint32_t k, double radius, uint32_t flags, uint32_t* items, void* dists,
                          uint32_t* counts, void* stream);

/* .within(radius)[.unsorted()][.exclusive_boundaries()], phases 1 and 2, as
 * kd_within_count / kd_within_fill (periodic.rs:620-791 with the Vec collector). */
int kd_periodic_within_count(kd_tree* tree, const void* queries, uint64_t m, const void* box,
                             double radius, uint32_t flags, uint64_t* offsets, void* stream);
int kd_periodic_within_fill(kd_tree* tree, const void* queries, uint64_t m, const void* box,
                            double radius, uint32_t flags, const uint64_t* offsets,
                            uint32_t* items, void* dists, void* stream);

/* Traversal counters of the last query call on this tree (synchronous; calls that
 * overlap on one tree share them): out4 = {leaves visited, entries scanned, stem nodes
 * read, entries re-measured in FP64 by KD_SCREEN_F32}. The first three equal what Kiddo's
 * DFS visits for the same queries (its leaf-visit hook: query/nearest_one.rs:31-32). */
int kd_query_stats(kd_tree* tree, uint64_t* out4);
/* Queries of the last call on this tree that the FP32-screened k-NN kernel (k <= 16) could
 * not settle from its candidate list (more entries within the certified rounding band of the
 * k-th distance than the list holds, e.g. many exact ties) and answered again with the plain
 * FP64 kernel (no Kiddo counterpart: measurement support). Like kd_query_stats, a count that
 * calls overlapping on one tree share; results never depend on it. */
int kd_query_redos(kd_tree* tree, uint64_t* out);

/* Timing of the main query kernel of each call with HIP events recorded on the call's
 * stream (no Kiddo counterpart: measurement support). kd_last_kernel_ms waits for the
 * last timed call and returns its kernel's elapsed milliseconds. */
int kd_tree_set_timing(kd_tree* tree, int32_t enable);
int kd_last_kernel_ms(kd_tree* tree, float* ms);

/* ------------------------------------------------------------- batched forests ---- */
/* Thousands of independent trees of uneven sizes, each with its own query set, answered by
 * one persistent launch fed by a device work queue (no Kiddo counterpart: the reference
 * would loop over trees, each query through KdTree::query, builder.rs:2318-2331).
 * Trees are given in Kiddo layout (kd_tree_desc, same dtype and dims for all). tree_flags:
 * 0, or KD_TREE_SHADOW_F32 for f32 trees (an FP32 shadow of the pooled entries, 16 B each),
 * which kd_forest_nearest_n's KD_SCREEN_F32 needs (the FP32 collector for k <= 16, thread
 * per query, exact redo pass for overflowed candidate lists; identical results; measured on
 * 4096 trees of 1k..262k points: 15.9 against 24.0 ms at k = 8, 30.3 against 41.4 ms at
 * k = 16 before the end phase read distances and items from the shadow). */
typedef struct kd_forest kd_forest;
int kd_forest_create(const kd_tree_desc* descs, uint32_t ntrees, uint32_t tree_flags, kd_forest** out);
int kd_forest_destroy(kd_forest* forest);
/* Plans a batch: query_offsets (host, ntrees+1) delimit each tree's queries in one
 * concatenated device array; work items of up to `chunk` queries (0 = 256) are ordered by
 * cost; KD_MORTON also Morton-orders the queries inside each tree's range (synchronous). */
int kd_forest_plan(kd_forest* forest, const void* queries, const uint64_t* query_offsets,
                   uint32_t chunk, uint32_t flags, void* stream);
/* nearest_n over the planned batch; outputs indexed by position in the concatenated
 * queries. KD_GRAPH captures the launch once and replays it (same pointers/arguments). */
int kd_forest_nearest_n(kd_forest* forest, const void* queries, uint32_t k, double radius,
                        uint32_t flags, uint32_t* items, void* dists, uint32_t* counts, void* stream);
/* within over the planned batch: first with items == NULL (writes offsets[m+1]), then with
 * output buffers of offsets[m] entries. */
int kd_forest_within(kd_forest* forest, const void* queries, double radius, uint32_t flags,
                     uint64_t* offsets, uint32_t* items, void* dists, void* stream);
int kd_forest_tune(kd_forest* forest, int32_t variant, int32_t group_size, int32_t lds_levels,
                   int32_t waves);
int kd_forest_autotune_nearest_n(kd_forest* forest, const void* queries, uint32_t k, double radius,
                                 uint32_t flags, uint32_t* items, void* dists, uint32_t* counts,
                                 int32_t* chosen4, void* stream);
int kd_forest_set_timing(kd_forest* forest, int32_t enable);
int kd_forest_last_kernel_ms(kd_forest* forest, float* ms);
int kd_forest_query_stats(kd_forest* forest, uint64_t* out4);
/* Bucket size picked by measurement for a batched forest (no Kiddo counterpart: the bucket
 * is ImmutableKdTree's const generic B, construction.rs:278-303, and results do not depend on
 * it). The ntrees trees (host point arrays, points[i] of sizes[i] x dims, typically a
 * sample of the forest's trees) are built at each of the nbuckets candidates, their
 * concatenated device queries planned (query_offsets: ntrees + 1 host values), the launch
 * shape autotuned and nearest_n(k, radius) timed; *best_bucket receives the fastest and
 * ms_out[i] (optional) each candidate's kernel time. k <= 64. */
int kd_forest_autotune_bucket(int32_t dtype, int32_t dims, uint32_t ntrees, const void* const* points,
                              const uint64_t* sizes, const void* queries, const uint64_t* query_offsets,
                              uint32_t k, double radius, uint32_t flags, const int32_t* buckets,
                              uint32_t nbuckets, int32_t* best_bucket, float* ms_out, void* stream);

/* ------------------------------------------------------------------- serving ---- */
/* A continuous-batching request loop over resident trees (no Kiddo counterpart; it serves
 * the same families as KdTree::query). Requests carry HOST query arrays and HOST result
 * buffers and may be submitted from any thread; each is cut into parts of at most
 * `max_part` queries (kd_server_set_part; 16384 by default). The loop admits a part when
 * one of `slots` streams is free and the device arena (`arena_bytes`, mirrored in pinned
 * host memory) has room for its queries and results, taking first the oldest part of a
 * request with the fewest parts on the device (a large request does not hold every slot
 * while smaller ones wait); it coalesces pending parts of the same kind (tree, family, k,
 * radius, flags), one per request, into the same launch, up to `max_coalesce` queries.
 * Each slot has a worker thread that stages, runs and copies back its launches, so the
 * server runs 1 + `slots` threads; launches on different slots overlap on the device. A
 * request completes when its last part does; results never depend on the batching.
 * Served trees should not have timing enabled (the slots' calls share the tree's events). */
enum kd_family {
  KD_FAMILY_NEAREST_ONE = 1,   /* items[m], dists[m]                                   */
  KD_FAMILY_NEAREST_N = 2,     /* items[m*k], dists[m*k], counts[m]; radius as kd_nearest_n */
  KD_FAMILY_BEST_N_WITHIN = 3, /* items[m*k], dists[m*k], counts[m]; k is n             */
  KD_FAMILY_APPROX = 4         /* items[m], dists[m]                                   */
};
typedef struct kd_request {
  int32_t tree;       /* index into the server's trees */
  int32_t family;     /* kd_family */
  uint32_t k;
  uint32_t flags;     /* KD_EXCLUSIVE, KD_MORTON, KD_SCREEN_F32 */
  double radius;      /* +INFINITY for none */
  const void* queries; /* host, m x dims of the tree's dtype; read when a part is launched */
  uint64_t m;
  uint32_t* items;    /* host outputs, valid once kd_server_wait returns */
  void* dists;
  uint32_t* counts;
} kd_request;
typedef struct kd_server kd_server;
int kd_server_create(kd_tree* const* trees, uint32_t ntrees, uint64_t arena_bytes, uint32_t slots,
                     uint64_t max_coalesce, kd_server** out);
/* Non-blocking; the request's buffers must stay valid until kd_server_wait returns. */
int kd_server_submit(kd_server* server, const kd_request* request, uint64_t* ticket);
/* Blocks until the request is answered; returns its status and its latency (submit to
 * results in host memory). Each ticket is waited for once. */
int kd_server_wait(kd_server* server, uint64_t ticket, double* latency_ms);
/* Requests above max_part queries are cut into parts of max_part (0: never cut). Applies
 * to requests submitted afterwards. */
int kd_server_set_part(kd_server* server, uint64_t max_part);
/* out6 = {requests done, launches, queries launched, parts coalesced into another part's
 * launch, admissions deferred for memory, peak arena bytes in use}. */
int kd_server_stats(kd_server* server, uint64_t* out6);
int kd_server_destroy(kd_server* server);

#ifdef __cplusplus
}
#endif

#endif /* KIDDO_AMD_H */
