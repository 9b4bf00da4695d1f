/* jagua_amd.h — C ABI of the MI355X (gfx950) collision-query engine.
 *
 * Drop-in boundary for the collision-query hot path of jagua-rs: the CDEngine / Hazard /
 * HazardFilter / HazardCollector / SPSurrogate API, re-exposed as plain C so a Rust crate
 * (or any FFI) can bind it with `extern "C"`.  Every entry point cites the reference
 * interface it replaces (paths relative to jagua-rs/src/ and lbf/ in JeroenGar/jagua-rs).
 *
 * Conventions
 *  - Return value: jg_status, 0 = JG_OK, negative = error (no entry point aborts).
 *  - Geometry is IEEE binary32, the reference's `f32`; vertex arrays are xy-interleaved.
 *  - Handles are opaque and not thread-safe; queries read the engine, updates write it.
 *  - Query entry points take DEVICE pointers (resident in HBM) and a HIP stream passed as
 *    `void*` (a hipStream_t; NULL = the default stream).  Outputs are valid after the
 *    stream is synchronised.  jg_*_host variants take host pointers and synchronise.
 *  - A 2x3 transform is the reference's Transformation matrix rows
 *    (m00 m01 m02 m10 m11 m12), transformation.rs:131-139; points map as
 *    x' = m00*x + m01*y + m02*1, y' = m10*x + m11*y + m12*1 (point.rs:26-31).
 */
#ifndef JAGUA_AMD_H
#define JAGUA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t jg_status;
#define JG_OK 0
#define JG_ERR_INVALID_ARG (-1)
#define JG_ERR_POLYGON (-2)          /* SPolygon::new bail! (simple_polygon.rs:42-51) */
#define JG_ERR_NO_EXTERIOR (-3)      /* CDEngine::new expect (cd_engine.rs:44-48) */
#define JG_ERR_NOT_REGISTERED (-4)   /* deregister expect (cd_engine.rs:78,91) */
#define JG_ERR_SURROGATE (-5)        /* surrogate missing / ff_poles slice (sp_surrogate.rs:56) */
#define JG_ERR_HIP (-6)              /* a HIP runtime call failed */
#define JG_ERR_DEPTH (-7)            /* quadtree depth above JG_MAX_QT_DEPTH */
#define JG_ERR_CAPACITY (-8)         /* an output buffer is too small */
#define JG_ERR_INTERNAL (-9)

/* Per-query status written by the detect entry points. */
#define JG_Q_CLEAR 0                 /* no collision */
#define JG_Q_COLLISION 1             /* collision */
#define JG_Q_PANIC_FILTERED_ENTIRE 2 /* the reference reaches unreachable!() here:
                                        a filtered Entire hazard in the virtual root
                                        (cd_engine.rs:176-181); reported, not aborted */

#define JG_MAX_QT_DEPTH 12

/* Query composition flags (detect). Default (0) = SURROGATE|POLY. */
#define JG_QUERY_SURROGATE 1u        /* CDEngine::detect_surrogate_collision (cd_engine.rs:204-223) */
#define JG_QUERY_POLY 2u             /* CDEngine::detect_poly_collision (cd_engine.rs:160-197) */

typedef struct jg_spolygon jg_spolygon;   /* SPolygon (+ SPSurrogate), simple_polygon.rs:24-37 */
typedef struct jg_cde jg_cde;             /* CDEngine, cd_engine.rs:21-31 */
typedef struct jg_snapshot jg_snapshot;   /* CDESnapshot, cd_engine.rs:356-359 */
typedef struct jg_plan jg_plan;           /* batched persistent query plan (no reference counterpart) */

typedef struct {                          /* SPSurrogateConfig, sp_surrogate.rs:120-131 */
    uint32_t n_pole_limits_n[3];
    float n_pole_limits_cov[3];
    uint32_t n_ff_poles;
    uint32_t n_ff_piers;
} jg_surrogate_config;

typedef struct {                          /* CDEConfig, cd_engine.rs:345-353 */
    uint32_t quadtree_depth;
    uint32_t cd_threshold;
    jg_surrogate_config item_surrogate_config;
} jg_cde_config;

typedef enum {                            /* HazardEntity, hazards/hazard.rs:38-51 */
    JG_ENTITY_PLACED_ITEM = 0,            /* a = item id, b = placed-item key, dt = DTransformation */
    JG_ENTITY_EXTERIOR = 1,
    JG_ENTITY_HOLE = 2,                   /* a = idx */
    JG_ENTITY_INFERIOR_QUALITY_ZONE = 3   /* a = quality, b = idx */
} jg_entity_kind;

typedef struct {
    uint32_t kind;
    uint32_t _pad;
    uint64_t a;
    uint64_t b;
    float dt[3];                          /* rotation, tx, ty */
    uint32_t _pad2;
} jg_hazard_entity;

typedef struct {                          /* HazKey (slotmap key), hazard.rs:8-11 */
    uint32_t idx;
    uint32_t version;
} jg_hazkey;

/* HazardFilter (hazards/filter.rs:7-9) */
#define JG_FILTER_NONE 0u                 /* NoFilter (filter.rs:57-65) */
#define JG_FILTER_KEYS 1u                 /* HazKeyFilter / HazKey self-filter (filter.rs:13-55) */
#define JG_FILTER_HASHED 2u               /* per-query pseudo-random subset: hazard slot s is
                                             irrelevant for query q iff
                                             jg_filter_hash(seed, q, s) < threshold */
typedef struct {
    uint32_t kind;
    uint32_t n_keys;
    const jg_hazkey* keys;                /* host pointer, JG_FILTER_KEYS */
    uint32_t seed;
    uint32_t threshold;
} jg_hazard_filter;

/* The hashed filter predicate and the order-independent set hash, specified bit-exactly so
 * any host can reproduce them:
 *   fmix32(h)      = murmur3 finaliser (h^=h>>16; h*=0x85ebca6b; h^=h>>13; h*=0xc2b2ae35; h^=h>>16)
 *   jg_filter_hash = fmix32(seed ^ fmix32(lo32(q) ^ fmix32(hi32(q) + 0x9e3779b9)) ^ (s * 0x85ebca77))
 *   mix64(z)       = splitmix64 finaliser
 *   set hash       = sum over slots s in the set of mix64(s)   (mod 2^64)
 */

/* ---------------------------------------------------------------- SPolygon / SPSurrogate */

/* SPolygon::new (simple_polygon.rs:41-71): CCW normalisation, area, diameter, bbox, POI. */
jg_status jg_spolygon_new(const float* xy, uint32_t n_vertices, jg_spolygon** out);
/* SPolygon::generate_surrogate (simple_polygon.rs:73-80) -> SPSurrogate::new (sp_surrogate.rs:33-53). */
jg_status jg_spolygon_generate_surrogate(jg_spolygon* poly, const jg_surrogate_config* cfg);
/* Transformable::transform_clone with DTransformation::compose (placed_item.rs:27-36,
 * simple_polygon.rs:195-224): vertices, POI and surrogate transformed, bbox regenerated. */
jg_status jg_spolygon_transform_clone(const jg_spolygon* poly, const float dt[3], jg_spolygon** out);

/* Many SPolygon::new (simple_polygon.rs:41-71) at once, each followed by
 * generate_surrogate (sp_surrogate.rs:33-53) when cfg is not NULL, with the pole searches
 * (fail_fast/pole.rs:12-80) and, when cfg->n_ff_piers > 0, the pier searches
 * (fail_fast/piers.rs:25-98) run on the GPU: out[i] is bit-identical to jg_spolygon_new on
 * xy[2*offsets[i] .. 2*offsets[i+1]) followed by jg_spolygon_generate_surrogate.  offsets
 * has n_polys + 1 entries.  A polygon whose search outgrows the device buffers (more than
 * 8192 cells on one level, more poles than the largest pole limit, a pier ray crossing its
 * boundary more than 32 times) or that the reference would reject during pier generation is
 * finished by the host routine instead; their number goes to *n_host_completed (may be
 * NULL).  On error no handle is returned and the message names the polygon. */
jg_status jg_spolygon_new_batch(const float* xy, const uint32_t* offsets, uint32_t n_polys,
                                const jg_surrogate_config* cfg, jg_spolygon** out, uint32_t* n_host_completed);
void jg_spolygon_free(jg_spolygon* poly);

typedef struct {
    uint32_t n_vertices;
    uint32_t n_poles;                     /* all surrogate poles (0 without surrogate) */
    uint32_t n_piers;
    uint32_t n_ff_poles;                  /* SPSurrogate::ff_poles() length */
    float area, diameter;
    float bbox[4];                        /* x_min y_min x_max y_max */
    float poi[3];                         /* cx cy r */
    float convex_hull_area;
} jg_spolygon_info;
jg_status jg_spolygon_get_info(const jg_spolygon* poly, jg_spolygon_info* out);
jg_status jg_spolygon_get_vertices(const jg_spolygon* poly, float* xy_out);       /* [n][2] */
jg_status jg_spolygon_get_poles(const jg_spolygon* poly, float* out);             /* [n_poles][3] */
jg_status jg_spolygon_get_piers(const jg_spolygon* poly, float* out);             /* [n_piers][4] */
/* SPolygon::centroid (simple_polygon.rs:173-192). */
jg_status jg_spolygon_centroid(const jg_spolygon* poly, float out[2]);

/* DTransformation::compose (d_transformation.rs:45-47 -> transformation.rs:91-101,131-139):
 * dt[i] = (rotation, tx, ty) -> mat6[i]; sin/cos come from the host libm (as in Rust std). */
jg_status jg_transforms_from_dt(const float* dt, uint64_t n, float* mat6_out);

/* The transformation chain of the instance import and export: out = decompose(rot_transl(b) .
 * X . I) with X = rot_transl(a), or its inverse when invert_a (transformation.rs:60-88,
 * 132-210, all products in f32 in written order; decompose takes atan2f).  invert_a = 1 is
 * ext_to_int_transformation(ext = b, pre = a) (io/import.rs:222-233); invert_a = 0 is
 * int_to_ext_transformation(int = b, pre = a) (io/export.rs:34-45).  Transformations are
 * (rotation in radians, tx, ty). */
jg_status jg_dt_chain(const float a[3], uint32_t invert_a, const float b[3], float out[3]);

/* ---------------------------------------------------------------- CDEngine */

/* CDEngine::new (cd_engine.rs:34-56).  shapes are copied; dynamic[i] is Hazard::dynamic. */
jg_status jg_cde_new(const float bbox[4], const jg_cde_config* cfg, const jg_hazard_entity* entities,
                     const jg_spolygon* const* shapes, const uint8_t* dynamic, uint32_t n, jg_cde** out);
void jg_cde_free(jg_cde* cde);
/* CDEngine::register_hazard (cd_engine.rs:59-69); the new key is written to *key_out (may be NULL). */
jg_status jg_cde_register_hazard(jg_cde* cde, const jg_hazard_entity* entity, const jg_spolygon* shape,
                                 uint8_t dynamic, jg_hazkey* key_out);
/* CDEngine::deregister_hazard_by_entity / _by_key (cd_engine.rs:72-96). */
jg_status jg_cde_deregister_hazard_by_entity(jg_cde* cde, const jg_hazard_entity* entity);
jg_status jg_cde_deregister_hazard_by_key(jg_cde* cde, jg_hazkey key);
/* CDEngine::save / restore (cd_engine.rs:98-149). */
jg_status jg_cde_save(const jg_cde* cde, jg_snapshot** out);
jg_status jg_cde_restore(jg_cde* cde, const jg_snapshot* snap);
void jg_snapshot_free(jg_snapshot* snap);
/* CDEngine::bbox (cd_engine.rs:329-331) and the exterior key (cd_engine.rs:29-30). */
jg_status jg_cde_bbox(const jg_cde* cde, float out[4]);
jg_status jg_cde_hkey_exterior(const jg_cde* cde, jg_hazkey* out);

/* Sizes of the device image: node records (the host node pool, free blocks included), entry
 * and inline-edge records (arena use, including the slack of relocated ranges), slots and
 * hazard vertices; device_bytes is the image's layout size with its headroom.  A full upload
 * (flatten_ms, upload_ms) rebuilds the image in breadth-first order; a patch (patch_ms,
 * patch_bytes) rewrites only the nodes and slots changed since the last upload. */
typedef struct {
    uint64_t n_nodes, n_entries, n_edge_refs, n_hazard_vertices, n_slots;
    uint64_t device_bytes;
    double flatten_ms, upload_ms;
    double patch_ms;
    uint64_t patch_bytes, n_patches, n_full_uploads;
} jg_cde_stats;
/* Brings the engine's HBM image up to date when it changed since the last upload (queries do
 * this implicitly): in place, when the changes fit the image's headroom, else by a full
 * rebuild.  Must not run while a launch that reads this engine is in flight on another
 * thread; earlier launches are waited for. */
jg_status jg_cde_sync(jg_cde* cde);
jg_status jg_cde_get_stats(const jg_cde* cde, jg_cde_stats* out);
/* The flattened tree in BFS order (same record layout as the oracle's tree dump):
 * node records [8 x u32]: level, n_children (0|4), n_hazards, n_active_edges, bbox (4 x f32 bits);
 * hazard records [8 x u32]: slot idx, presence (1 Partial, 2 Entire), n_edges, 0, ff_bbox.
 * Call with NULL buffers to obtain *n_nodes and *n_haz. */
jg_status jg_cde_dump(const jg_cde* cde, uint32_t* nodes, uint64_t nodes_cap, uint32_t* haz, uint64_t haz_cap,
                      uint64_t* n_nodes, uint64_t* n_haz);

/* ---------------------------------------------------------------- collision queries */

/* One collision query per transform i, as lbf samples them (lbf/src/opt/search.rs:41-57) and
 * ci_bench composes them (lbf/benches/ci_bench.rs:107-120):
 *   status[i] = detect_surrogate_collision(item.surrogate, T_i, filter)  (cd_engine.rs:204-223)
 *            || detect_poly_collision(item transformed by T_i, filter)   (cd_engine.rs:160-197)
 * item must carry a surrogate when JG_QUERY_SURROGATE is set.
 * d_mat6: device [n][6] f32; d_status: device [n] u8 (JG_Q_*). */
jg_status jg_detect_batch(jg_cde* cde, const jg_spolygon* item, const float* d_mat6, uint64_t n,
                          const jg_hazard_filter* filter, uint32_t flags, uint8_t* d_status, void* stream);

/* One hazard-collection query per transform i (ci_bench.rs:56-69):
 *   set_i = collect_surrogate_collisions (cd_engine.rs:296-310)
 *         ∪ collect_poly_collisions       (cd_engine.rs:261-289)
 * BasicHazardCollector (collector.rs:44-67) is returned as, per query: d_count[i] (set size),
 * d_hash[i] (set hash, see above) and the first `capacity` slot indices in d_keys[i*capacity..]
 * in ascending order (d_keys may be NULL).  Sets larger than `capacity` are still counted
 * and hashed exactly. */
jg_status jg_collect_batch(jg_cde* cde, const jg_spolygon* item, const float* d_mat6, uint64_t n,
                           uint32_t capacity, uint32_t* d_count, uint64_t* d_hash, uint32_t* d_keys,
                           void* stream);

/* Host-pointer conveniences of the two calls above (copy in, run, copy out, synchronise). */
jg_status jg_detect_batch_host(jg_cde* cde, const jg_spolygon* item, const float* mat6, uint64_t n,
                               const jg_hazard_filter* filter, uint32_t flags, uint8_t* status);
jg_status jg_collect_batch_host(jg_cde* cde, const jg_spolygon* item, const float* mat6, uint64_t n,
                                uint32_t capacity, uint32_t* count, uint64_t* hash, uint32_t* keys);

/* ---------------------------------------------------------------- batched persistent form */

/* A set of queries: one engine, one item, a contiguous range of transforms in d_mat6 and
 * of outputs.  Many sets (many layouts, many items) run in ONE persistent launch fed by a
 * device work queue.  No reference counterpart: it is lbf's sampled placement evaluation
 * (search.rs:41-101) batched across items and layouts. */
typedef struct {
    jg_cde* cde;
    const jg_spolygon* item;
    uint64_t first;                       /* index of the set's first transform / output */
    uint64_t count;
    jg_hazard_filter filter;
} jg_query_set;

typedef struct {
    uint32_t chunk;                       /* queries per work unit: 16, 32, 64 (0 = 64) */
    uint32_t workgroups;                  /* persistent one-wavefront workgroups (0 = 16 per compute unit) */
    uint32_t lanes_per_query;             /* 0 = per item: its vertex count rounded up to a power
                                             of two, 8..32 */
    uint32_t flags;                       /* JG_QUERY_* */
    uint32_t stages;                      /* 0 or 2: surrogate and polygon stages as two kernels,
                                             survivors passed in groups of one polygon pass;
                                             1: one kernel runs both stages of a work unit */
} jg_plan_config;

jg_status jg_plan_create(const jg_query_set* sets, uint32_t n_sets, const jg_plan_config* cfg, jg_plan** out);
void jg_plan_free(jg_plan* plan);
/* d_mat6 [total][6] and d_status [total] are device arrays indexed by jg_query_set::first + i. */
jg_status jg_plan_detect(jg_plan* plan, const float* d_mat6, uint8_t* d_status, void* stream);
/* The same launch captured once in a hipGraph and replayed (launch overhead amortised). */
jg_status jg_plan_detect_graph(jg_plan* plan, const float* d_mat6, uint8_t* d_status, uint32_t replays,
                               void* stream);
/* The hazard set of every query of the plan, as jg_collect_batch computes it for one set
 * (collect_surrogate_collisions ∪ collect_poly_collisions, cd_engine.rs:261-310, composed as in
 * lbf/benches/ci_bench.rs:56-69), for all sets in one persistent launch.  Outputs are indexed
 * like d_status: d_count[i], d_hash[i] and, when d_keys is not NULL, the first `capacity` slots
 * ascending in d_keys[i*capacity..] (the rest 0xffffffff).  The collect APIs of the reference
 * take no filter; a set's JG_FILTER_KEYS filter is read as a pre-seeded collector instead (the
 * collector is the filter, filter.rs:69-76): every query's set starts with those keys and the
 * result is the seed together with the hits.  A JG_FILTER_HASHED filter is rejected.  Every
 * item needs its surrogate (its ff poles are collected). */
jg_status jg_plan_collect(jg_plan* plan, const float* d_m6, uint32_t capacity, uint32_t* d_count, uint64_t* d_hash,
                          uint32_t* d_keys, void* stream);
/* Schedule autotuner (no reference counterpart: the reference evaluates one query at a
 * time).  Times the plan's launch on `stream` with HIP events for candidate schedules —
 * stage layout {two kernels, one kernel}, lanes per query in the polygon stage {per item, 8,
 * 16, 32, 64}, workgroups {4, 8, 16, 32} per compute unit and queries per work unit {16, 32,
 * 64}: all 120 combinations plus the current one (median of `reps` timed launches each, after
 * one warm-up), and keeps the fastest in the plan.  The plan's flags are kept.  Every schedule computes the same verdicts;
 * d_status holds them afterwards.  `best` and `best_ms` may be NULL. */
jg_status jg_plan_autotune(jg_plan* plan, const float* d_mat6, uint8_t* d_status, uint32_t reps, void* stream,
                           jg_plan_config* best, float* best_ms);
/* The schedule a plan currently runs with (after defaults are resolved: lanes_per_query 0
 * means per item). */
jg_status jg_plan_get_config(const jg_plan* plan, jg_plan_config* out);

/* The batched detection of jg_plan_create + jg_plan_detect in one call that allocates no
 * memory.  No reference counterpart: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The sets, schedule, verdicts and kernels are jg_plan_detect's.  Instead of a plan's own
 * buffers it uses one fixed device array of JG_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends, for the launch table (set descriptors, work units, claim counters, survivor
 * groups), the query items' records and JG_FILTER_KEYS masks; the caller passes the
 * transforms and the status array.  Requirements and limits:
 *  - every engine's image must be current on the device (jg_cde_sync after the last update;
 *    JG_ERR_INVALID_ARG otherwise), since uploading an image allocates;
 *  - a call whose table outgrows the scratch returns JG_ERR_CAPACITY (about 8 bytes per query
 *    plus 16 per survivor pass: roughly 2M queries of configuration-2 items fit);
 *  - every call on one GPU shares that GPU's scratch: calls that overlap on one GPU (other
 *    streams or threads, before the previous call's kernels finished) race.  Calls in order
 *    on one stream are safe: the staging copy is ordered after the previous call's kernels.
 * The staging copy is from pageable host memory, so the call returns once it is staged. */
#define JG_NOALLOC_SCRATCH_BYTES (32u << 20)
jg_status jg_detect_sets_noalloc(const jg_query_set* sets, uint32_t n_sets, const jg_plan_config* cfg,
                                 const float* d_m6, uint8_t* d_status, void* stream);

/* ---------------------------------------------------------------- placement search */

/* lbf's sampled placement search (lbf/src/opt/search.rs:17-106), resident on the GPU for many
 * independent (engine, item) searches at once.  No reference counterpart as an API: the
 * reference evaluates one sample at a time and tightens its sampling box after every
 * improvement.  Here every round draws samples_per_round transforms per search on the device,
 * evaluates them with the batched collision query (surrogate || polygon, bit-exact), and one
 * wave per search keeps the clear sample of lowest LBFLoss (x_max * 10 + y_max of the placed
 * shape's bbox, lbf/src/opt/loss.rs:29-41; ties to the first sample) if it beats the best so
 * far.  Within a round that is exactly the sequential rule applied to the same samples.
 *   uniform rounds: rotation uniform over the search's rotation table, translation uniform in
 *                   the sampling box, initially the CDE bbox; after each round x_max of the box
 *                   is tightened to best cost / 10 (LBFLoss::tighten_sample_bbox);
 *   local rounds:   only for searches that found a placement: translation normal around the
 *                   best with standard deviation decaying from ls_sd_transl[0] to [1] times
 *                   the larger side of the CDE bbox over the rounds (ls_sampler.rs:8-12,
 *                   80-95); rotation: discrete tables keep the best's rotation, continuous
 *                   tables step the best's angle by a normal with deviation ls_sd_rot[0..1].
 * Rotations come from a host-built table (sin/cos with libm, as DTransformation::compose), so
 * the device never evaluates trigonometry: each sample's matrix equals the host composition of
 * its (rotation, tx, ty).  Continuous rotation uses JG_SEARCH_CONT_ROTATIONS angles
 * 2 pi k / JG_SEARCH_CONT_ROTATIONS.  Random numbers are a counter-based hash of
 * (seed, search, round, sample): a run is reproducible. */
#define JG_SEARCH_CONT_ROTATIONS 4096u

typedef struct {
    uint32_t samples_per_round;           /* per search and round, 1 .. 65536 */
    uint32_t uniform_rounds;
    uint32_t local_rounds;
    uint32_t pad;
    uint64_t seed;
    float ls_sd_transl[2];                /* 0.01, 0.0005 in lbf */
    float ls_sd_rot[2];                   /* 2 and 0.5 degrees in lbf, radians */
} jg_search_config;

typedef struct {
    jg_cde* cde;
    const jg_spolygon* item;              /* with its surrogate */
    const float* rotations;               /* allowed rotations (radians), n_rotations of them */
    uint32_t n_rotations;                 /* 0 with continuous = 0: no rotation */
    uint32_t continuous;                  /* 1: continuous rotation (RotationRange::Continuous) */
    jg_hazard_filter filter;
} jg_search_item;

typedef struct {
    uint32_t found;
    float rotation, tx, ty;               /* best DTransformation */
    float cost, x_max, y_max;             /* its LBFLoss */
    uint32_t samples;                     /* samples drawn for this search */
} jg_search_result;

typedef struct jg_search jg_search;

jg_status jg_search_create(const jg_search_item* items, uint32_t n_items, const jg_search_config* cfg,
                           jg_search** out);
void jg_search_free(jg_search* s);
/* All rounds of all searches on `stream`; graph = 1 captures the run in a hipGraph once and
 * replays it on later calls.  Results are valid after the stream synchronises. */
jg_status jg_search_run(jg_search* s, uint32_t graph, void* stream);
/* Copies the results of the last run (synchronises). */
jg_status jg_search_results(jg_search* s, jg_search_result* out);
/* The last run's samples: dt[(round * n_items + item) * samples_per_round + j][3] (rotation,
 * tx, ty) and their statuses (JG_Q_*), host arrays of rounds * n_items * samples_per_round
 * entries (synchronises). */
jg_status jg_search_samples(jg_search* s, float* dt, uint8_t* status);

/* ---------------------------------------------------------------- overlap gradient */

/* Pole-pair overlap quantity (an extension with no reference counterpart; its
 * primitives are Circle::separation_distance for circles and for edges, circle.rs:184-229):
 *   Q = sum over candidate item pairs (a,b), poles i of a, j of b, with |c_i - c_j| < r_i + r_j,
 *         of (sqrt((r_i + r_j)^2) - |c_i - c_j|)^2
 *     + sum over poles i with D(c_i) < r_i of (r_i - D(c_i))^2, where D is the signed distance to
 *         the container boundary (the container edges): the distance to the nearest edge,
 *         positive inside (ray-crossing parity over the edges) and negative outside, so that a
 *         pole whose centre has left the container is pushed back by r_i + |D|
 * where c_i = R(theta_a) p_i + (x_a, y_a) for item a's pose (x, y, theta) and local pole
 * (p_i, r_i).  Forward and backward are FP32 kernels; the backward recomputes the pole
 * transforms instead of storing them.  Partial sums are reduced in a fixed order
 * (deterministic run to run). */
typedef struct jg_overlap jg_overlap;
/* poles [total][3] (px, py, r) and pole_off [n_items + 1] (host); pairs [n_pairs][2] item
 * indices (host); cedges [n_cedges][4] container edges x1 y1 x2 y2 (host). */
jg_status jg_overlap_create(const float* poles, const uint32_t* pole_off, uint32_t n_items, const uint32_t* pairs,
                            uint64_t n_pairs, const float* cedges, uint32_t n_cedges, jg_overlap** out);
/* Replaces the candidate pairs (host [n_pairs][2]). */
jg_status jg_overlap_set_pairs(jg_overlap* ov, const uint32_t* pairs, uint64_t n_pairs);
/* Forward: d_pose device [n_items][3] (x, y, theta; may be NULL when n_items is 0); d_q device f32[1]. */
jg_status jg_overlap_fwd(jg_overlap* ov, const float* d_pose, float* d_q, void* stream);
/* Backward: dQ/d(x, y, theta) per item into d_grad device [n_items][3]. */
jg_status jg_overlap_bwd(jg_overlap* ov, const float* d_pose, float* d_grad, void* stream);
/* Forward and backward in one pass over the pairs: Q into d_q and the gradient into d_grad,
 * bit for bit those of jg_overlap_fwd and jg_overlap_bwd (one evaluation of an optimiser
 * step). */
jg_status jg_overlap_fwd_bwd(jg_overlap* ov, const float* d_pose, float* d_q, float* d_grad, void* stream);
void jg_overlap_free(jg_overlap* ov);
/* Candidate pairs from the engine's quadtree, found on the GPU, replacing the overlap's pairs:
 * item k is the hazard of slot item_slots[k] in cde, with bounding circle centre (x, y) of its
 * pose d_pose[k] (device) and radius rad[k] (host).  For each item i the quadtree nodes meeting
 * its circle's box are visited down to their leaves; a hazard present there that is an item
 * k > i whose circle overlaps i's gives the pair (i, k).  Every pair of items whose shapes
 * overlap is found (both are present in the leaves covering the overlap).  The pairs are in a
 * fixed order (by i, then by slot), and so is the adjacency the backward pass gathers with.
 * Synchronises `stream` (the pair count is read back); *n_pairs_out may be NULL. */
jg_status jg_overlap_pairs_from_cde(jg_overlap* ov, jg_cde* cde, const uint32_t* item_slots, const float* d_pose,
                                    const float* rad, uint64_t* n_pairs_out, void* stream);
/* The overlap's current candidate pairs copied to the host ([n][2]); with pairs_out NULL only
 * their number. */
jg_status jg_overlap_get_pairs(const jg_overlap* ov, uint32_t* pairs_out, uint64_t cap, uint64_t* n_out);
/* Candidate pairs for the overlap quantity: items whose bounding circles (centre (x, y),
 * radius rad) overlap, found with a uniform grid over `bbox` of 2^depth cells per side (the
 * leaf cells of a quadtree of that depth).  Writes up to `cap` pairs (a < b) and their count. */
jg_status jg_overlap_candidate_pairs(const float* pose, const float* rad, uint32_t n_items, const float bbox[4],
                                     uint32_t depth, uint32_t* pairs_out, uint64_t cap, uint64_t* n_out);

/* ---------------------------------------------------------------- graphs */

/* Captures the launches the calls between jg_graph_begin and jg_graph_end make on `stream`
 * (not the null stream) into one HIP graph, replayed by jg_graph_launch: for instance a
 * step of jg_plan_detect, jg_plan_collect and jg_overlap_fwd / jg_overlap_bwd.  No reference
 * counterpart.  Only launches may happen while capturing: call each captured entry point once
 * beforehand (it builds its tables and buffers then), and update no engine the captured calls
 * use while the graph is in use.  A call that would allocate or copy fails the capture, and
 * jg_graph_end reports the error. */
typedef struct jg_graph jg_graph;
jg_status jg_graph_begin(void* stream);
jg_status jg_graph_end(void* stream, jg_graph** out);
jg_status jg_graph_launch(const jg_graph* graph, void* stream);
void jg_graph_free(jg_graph* graph);

/* Version / build info. */
const char* jg_version(void);
/* Message of the last error on this thread ("" when none). */
const char* jg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* JAGUA_AMD_H */
