/*
 * d2hip.h -- C ABI of the MI355X contact-detection library (libd2hip.so).
 *
 * This is the drop-in boundary for Delta2's data-parallel contact-detection
 * path: the surrogate-tree discrete narrow phase, the linear-interpolation
 * CCD with its TOI -> step rule, and contact generation over broad-phase
 * pair lists.  Every entry point names the reference interface it replaces
 * (paths relative to the reference's src/).  Plain pointers and sizes only;
 * no exceptions cross the ABI; every call returns a status code (D2_OK == 0)
 * and d2_last_error() holds the message of the last failure.
 *
 * Flat record formats (all little-endian, densely packed):
 *
 *   body record, 40 doubles (model/particle.h:9-64, model/state.h:6-36)
 *     [0..13]  current state: rotation quaternion w,x,y,z | translation xyz |
 *              velocity xyz | angular momentum xyz | time
 *     [14..27] future state, same layout
 *     [28..36] body-frame inverse inertia, 3x3 row-major (already / density)
 *     [37] geo_eps   [38] is_static (0/1)   [39] last_time_step_size
 *
 *   node record, 8 int32 (model/surrogate_components.h:11-20)
 *     node_id, bucket_id, is_inner, depth, parent_id, child_id_first,
 *     num_children, num_tris (of the node's bucket)
 *
 *   bucket record, 88 doubles, structure of arrays over the 8 slots
 *     (model/surrogate_components.h:66-88)
 *     Ax[8] Ay[8] Az[8] Bx[8] By[8] Bz[8] Cx[8] Cy[8] Cz[8] eps[8] inner_eps[8]
 *     Node i of a tree owns bucket i; unused slots are zero.
 *
 *   contact, 8 int32 + 11 doubles (collision_detection/contact.h:9-33,96-108)
 *     ints : pair, node_a, node_b, kind, ia, ib, level, 0
 *            kind 0: triangle pair (ia of a, ib of b) of the node pair
 *            kind 1: CCD leaf pair, vertex ia of a vs triangle ib of b
 *            kind 2: CCD leaf pair, vertex ib of b vs triangle ia of a
 *            kind 3: CCD leaf pair, edge ia of a vs edge ib of b
 *     reals: A xyz, B xyz, eps_a, eps_b, eps_inner_a, eps_inner_b, toc
 *   Contacts of a pair are returned in the reference's order (breadth-first
 *   level, node-pair order, then triangle/sub-test order).
 *
 * Threading: one context per host thread; a context owns one HIP stream
 * (default priority) and all its device buffers.  Outputs stay valid until
 * the next call on the same context.
 */
#ifndef D2HIP_H
#define D2HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum d2_status {
  D2_OK = 0,
  D2_ERR_ARG = 1,      /* invalid argument (reference: std::invalid_argument) */
  D2_ERR_HIP = 2,      /* HIP runtime failure */
  D2_ERR_CAPACITY = 3, /* a work list exceeded its capacity (see d2_set_capacity) */
  D2_ERR_STATE = 4,    /* call order: e.g. run before upload */
  D2_ERR_RUNTIME = 5   /* reference: std::runtime_error (e.g. no surrogate triangle) */
};

enum d2_precision {
  D2_FP64 = 0, /* every test in FP64: identical results to the reference path */
  D2_FP32 = 1  /* FP32 culling with error-inflated margins, FP64 for every kept result */
};

typedef struct d2_ctx d2_ctx;

/* ---- context ----------------------------------------------------------- */
/* No reference counterpart (the reference path works on host objects): a
 * context owns one HIP stream at default priority, the resident inputs and
 * the work lists of one device.  d2_ctx_create fails (D2_ERR_HIP) without a
 * visible GPU; there is no CPU fallback. */
int d2_ctx_create(int device, int precision, d2_ctx** out);
int d2_ctx_destroy(d2_ctx* ctx);
const char* d2_last_error(const d2_ctx* ctx);
int d2_set_precision(d2_ctx* ctx, int precision);
/* Work-list capacities.  By default they start at 1<<22 node pairs per list,
 * 1<<24 items per level, 1<<22 contacts and 1<<20 CCD soup items, and grow
 * on demand (a run that overflows is repeated with capacities sized from its
 * recorded counts).  d2_set_capacity fixes every capacity at `entries` (CCD
 * soup items at min(entries, 1<<20)); an overflow then fails with
 * D2_ERR_CAPACITY. */
int d2_set_capacity(d2_ctx* ctx, int64_t entries);
/* Turns growth on demand on or off (on by default; d2_set_capacity turns it off). */
int d2_set_auto_grow(d2_ctx* ctx, int on);

/* Launch shapes (no reference counterpart: the reference runs on the CPU).
 * chunk_blocks: blocks of the chunked list kernels (1-4096; a list is split
 * into that many contiguous ranges); stream_blocks: blocks of the
 * grid-stride kernels (1-16384); post_blocks: 256-thread blocks of the
 * per-pair post-processing (1-16384).  Defaults: 4, 8 and 16 per CU.
 * Results do not depend on them; d2_autotune picks them by timing. */
typedef struct d2_tune {
  int chunk_blocks;
  int stream_blocks;
  int post_blocks;
  int gate_phases; /* phase mask of the learned gate when it is on (1 CCD, 2 discrete, 3 both; 0 none) */
} d2_tune;
int d2_set_tune(d2_ctx* ctx, const d2_tune* tune);
int d2_get_tune(const d2_ctx* ctx, d2_tune* tune);
/* Autotuner: coordinate descent over the three block counts (and, with the
 * learned gate on, over its phase mask), timing whole
 * runs on the resident inputs (kind 0: d2_narrow_run(filter), 1: d2_ccd_run,
 * 2: d2_step_run(dt, filter); graph or eager as set), median of reps runs per
 * candidate.  Leaves the best shape set and returns it with its time (ms). */
int d2_autotune(d2_ctx* ctx, int kind, float dt, int filter, int reps, d2_tune* best, double* best_ms);

/* Graph mode (no reference counterpart).  When on, each kind of run (narrow,
 * CCD, step) is captured into a HIP graph and replayed on later calls as one
 * launch: every phase runs the deepest level count seen so far (the first run
 * of a kind is eager, to see it); a run whose traversal goes deeper is
 * repeated eagerly and the next capture holds more levels.  Uploads that
 * change sizes or launch parameters recapture.  Results are identical to
 * eager runs. */
int d2_set_graph(d2_ctx* ctx, int on);
/* Graph segments captured (and instantiated) and graph launches since the
 * context was created (no reference counterpart): a sequence whose pair
 * list, clusters or level grids change between steps recaptures; this tells
 * how often. */
int d2_graph_stats(const d2_ctx* ctx, int64_t* captures, int64_t* launches);

/* ---- surrogate trees ---------------------------------------------------- */
/* Builds one surrogate tree on the host from a triangle mesh, as
 * model::SurrogateTree(V, F, opt) does (model/surrogate_tree.cpp:9-18,39-112,
 * model/construct_surrogate_tree.h:8-185, model/fit_surrogate_triangle.h:246-298).
 * rng is the PCG32 state {state, inc} of the shared Options generator
 * (common/cli_options.h:46); it is advanced in place.  Writes at most
 * max_nodes node/bucket records and the node count to *n_nodes_out. */
int d2_tree_build(const double* V, int n_vertices, const int32_t* F, int n_faces, uint64_t* rng,
                  int max_nodes, int32_t* nodes_out, double* buckets_out, int* n_nodes_out);
/* PCG32 seeding as pcg32_srandom_r (common/pcg_basic.cpp:42-49). */
int d2_pcg_seed(uint64_t initstate, uint64_t initseq, uint64_t* rng_out);

/* Uploads n_trees concatenated trees; tree t owns node records
 * [tree_node_offset[t], tree_node_offset[t+1]).  Replaces MeshData storage
 * read through SurrogateTree::getNode/getBucket (model/surrogate_tree.cpp:20-27).
 * Body records name trees by index, so an upload forgets the bodies, scene
 * groups, pair list and clusters: runs fail with D2_ERR_STATE until they are
 * uploaded again. */
int d2_trees_upload(d2_ctx* ctx, int n_trees, const int32_t* tree_node_offset, const int32_t* nodes,
                    const double* buckets);

/* ---- bodies ------------------------------------------------------------- */
/* Uploads n body records; tree_of_body[i] selects the tree of body i.
 * Replaces the Particle fields the path reads (model/particle.h:9-64).
 * Last states are reset to the current states (d2_bodies_last_upload sets
 * them).  With the same body count as before, the pair list, its clusters and
 * the scene groups stay; with another count they are forgotten (runs fail
 * with D2_ERR_STATE until a pair list is uploaded or built again). */
int d2_bodies_upload(d2_ctx* ctx, int n_bodies, const int32_t* tree_of_body, const double* records);
int d2_bodies_download(d2_ctx* ctx, double* records);
/* future = current.extrapolate(step[i]) for every non-null body
 * (Particle::projectFutureState, model/particle.cpp:64-66; state.cpp:90-102). */
int d2_project_future(d2_ctx* ctx, const double* step_per_body);

/* ---- pair lists --------------------------------------------------------- */
/* Uploads a broad-phase pair list (collision_detection/broad_phase.h:5-7:
 * geometry ids of body a and b) with the CCD interval of each pair
 * (separate_clusters.cpp:490: min of the two last_time_step_size; may be
 * NULL for the discrete phase). */
int d2_pairs_upload(d2_ctx* ctx, int n_pairs, const int32_t* pairs, const double* max_time);

/* ---- the path ----------------------------------------------------------- */
/* Discrete narrow phase over the resident pairs: per pair
 * compareTreesFull<double,8,8>(a, b, ContactDetectionComparison)
 * (collision_detection/full_tree_comparison.h:13-67, comparison_dist.h:205-240,
 * strategies/contact_detection/contact_detection_comparison.cpp:12-14), then
 * filterContacts(contacts, 0.0) when filter != 0 (contact.h:44-69), as
 * strategies/PDE/PDE_explicit.cpp:87-88 does.  Uses the bodies' future
 * states.  *n_contacts_out receives the total contact count. */
int d2_narrow_run(d2_ctx* ctx, int filter, int64_t* n_contacts_out);

/* CCD over the resident pairs: per pair
 * compareTreesFullContinuous<double,8,8>(a, b, max_time[p], min_time)
 * (full_tree_comparison.h:69-219, continuous_comparison.h:12-98,565-910) and
 * the TOI -> step rule of separate_clusters.cpp:496-564.  Per pair results:
 * step scale (min_scaling_seen, 1.0 when no contact) and min_time_out.
 * The future state is projected per pair on the device; the bodies'
 * stored future states are not modified. */
int d2_ccd_run(d2_ctx* ctx, int64_t* n_contacts_out);

/* Per-cluster step of fineCollisionClustersWithTimeStepSelection + the
 * clamp of TimeStepSelectionDynamicContinuous::selectTimeStep
 * (separate_clusters.cpp:462-599, time_step_selection_dynamic_continuous.cpp:17-21)
 * from the last d2_ccd_run.  cluster_of_pair[p] in [0, n_clusters); dt is
 * Options::time_step_size (a float, common/cli_options.h:11). */
int d2_cluster_steps(d2_ctx* ctx, const int32_t* cluster_of_pair, int n_clusters, float dt, double* cluster_step_out);

/* Cluster of every resident pair (the connected components of
 * separate_clusters.cpp:355-459, computed by the caller); needed by
 * d2_step_run.  cluster_of_pair[p] in [0, n_clusters). */
int d2_clusters_upload(d2_ctx* ctx, const int32_t* cluster_of_pair, int n_clusters);

/* One contact-detection step over the resident pairs, as
 * fineCollisionClustersWithTimeStepSelection and PDEExplicit::step sequence
 * it (separate_clusters.cpp:462-599, strategies/PDE/PDE_explicit.cpp:49-95),
 * all on the device:
 *   1. CCD with max_time = min(last_time_step_size of a, b) and the
 *      TOI -> step rule (separate_clusters.cpp:479-566);
 *   2. the clamped cluster steps (separate_clusters.cpp:570-598,
 *      time_step_selection_dynamic_continuous.cpp:17-21), written to
 *      cluster_step_out[n_clusters] when it is not NULL;
 *   3. future = current.extrapolate(cluster step) for every non-static body
 *      in a pair; a body in no pair is a cluster of its own and steps with
 *      its clamped last step (PDE_explicit.cpp:49-56, separate_clusters.cpp:
 *      594-596; static bodies keep their future state);
 *   4. the discrete narrow phase and, when filter != 0, filterContacts.
 * Afterwards d2_result_contacts returns the discrete contacts and
 * d2_result_pairs the CCD values of step 1.  The bodies' future states are
 * overwritten (as in the reference); current states and last steps are not. */
int d2_step_run(d2_ctx* ctx, float dt, int filter, double* cluster_step_out, int64_t* n_contacts_out);

/* ---- the contact solver (consumer of the filtered contacts) ------------- */
/* Per-body mass (> 0) and friction coefficient (>= 0), the solver's view of
 * Particle::getMass / friction_coeff (model/particle.h:17, :47); n_bodies
 * must equal the uploaded count.  The inverse inertia is the body record's. */
int d2_solver_bodies(d2_ctx* ctx, int n_bodies, const double* mass, const double* friction);
/* strategy::SequentialImpulses::solve (strategies/contact_force/
 * sequential_impulses.cpp:104-638, warm start sequential_impulses_warm_start.
 * cpp:12-54) for every cluster of the last d2_step_run, which must have run
 * with the filter: each cluster's hits are its pairs' filtered contacts,
 * pairs in index order (PDE_explicit.cpp:77-95), solved at the cluster's step
 * with external force mass * gravity (contact_force_strategy.cpp:10-13);
 * bodies in no pair are clusters without contacts (gravity only).  Solved
 * clusters replace their dynamic bodies' future states; a cluster the
 * reference refuses (a contact closer than 1e-6 with a step above 1e-4)
 * keeps them and gets status 0.  cluster_status (n_clusters, may be NULL)
 * receives 1 / 0 per cluster, *n_failed the refusals.  The warm store (the
 * impulse of each body pair's last contact) persists across calls.  Every
 * dynamic body's pairs must lie in one cluster (D2_ERR_ARG otherwise). */
int d2_solve(d2_ctx* ctx, const double* gravity, int32_t* cluster_status, int* n_failed);
/* Per contact of the solved step's filtered result, in d2_result_contacts'
 * order, 11 doubles: the contact normal normalized((B - A) / 2), the normal
 * impulse, the position-offset impulse, the friction impulse (on body a) and
 * the tangential relative velocity v_ab - (v_ab . n) n of the last friction
 * sweep (sequential_impulses.cpp:124, :197, :455). */
int d2_solve_results(d2_ctx* ctx, double* per_contact);
/* PDEExplicit::step's tail for the dynamic bodies (PDE_explicit.cpp:117-131):
 * current state <- future state, last step <- the body's step of the last
 * d2_step_run, future state <- extrapolate(current, step).  Static bodies
 * keep theirs (the reference's last step of a static body depends on the
 * order its clusters are stepped in). */
int d2_commit_step(d2_ctx* ctx);
/* Empties the solver's warm store. */
int d2_solver_reset(d2_ctx* ctx);

/* ---- broad phase and clusters -------------------------------------------- */
/* Per-tree mesh bounds the broad phase reads: the float AABB (lo xyz, hi xyz)
 * and the bounding radius of MeshData::initBoudingData (model/mesh.cpp:44-51,
 * getAABB :132-153, getSphereBounds :155-158, getBoundingRadius :160-162).
 * d2_trees_upload computes them from the leaf triangles (every vertex of a
 * mesh whose vertices all belong to a face); the upload overrides them. */
int d2_tree_bounds_upload(d2_ctx* ctx, const float* aabb6, const double* radius);
int d2_tree_bounds_download(d2_ctx* ctx, float* aabb6, double* radius);
/* Particle::last_state of every body (14 doubles each, state layout; the
 * individual-time broad phase and the cluster roll-back read it).  Until it is
 * uploaded a body's last state equals its current state. */
int d2_bodies_last_upload(d2_ctx* ctx, const double* last14);
/* Independent scenes in one context (no reference counterpart): bodies of
 * different groups never pair.  NULL puts every body in one group. */
int d2_bodies_groups_upload(d2_ctx* ctx, const int32_t* group_of_body);

/* Broad phase on the device: broadPhaseEmbree (collision_detection/
 * broad_phase_embree.cpp:248-289) without Embree.  mode 0 is
 * broadPhaseEmbree(particles, t) (boundsFunc :60-84, collideFunc :143-176);
 * mode 1 is broadPhaseEmbree(particles) (boundsFuncIndividualTime :86-112,
 * collideFuncIndividualTime :178-226; t unused).  Float bounding boxes that
 * overlap (touching included) and pass the sphere test of
 * sphereSphereFirstIntersect (continuous_comparison.h:12-36) give a pair
 * (a, b), a < b, not both static, same group.  The pairs become the resident
 * pair list (sorted by a, then b; Embree's order is not deterministic) with
 * max_time = min(last_time_step_size of a, b) (separate_clusters.cpp:490).
 * D2_ERR_RUNTIME when State::interpolate would throw (state.cpp:56-58). */
int d2_broad_phase(d2_ctx* ctx, int mode, double t, int64_t* n_pairs_out);
/* Copies the resident pair list (and its max_time when not NULL). */
int d2_pairs_download(d2_ctx* ctx, int32_t* pairs, double* max_time);

/* Clusters of the resident pairs on the device: separateCollisionClusters
 * (collision_detection/separate_clusters.cpp:355-459) without sleeping
 * bodies.  Connected components of the pairs between two non-static bodies,
 * numbered as libigl's connected_components numbers them (by smallest body
 * index), components without a non-static body dropped; a pair belongs to the
 * cluster of its non-static body.  Every non-static member is rolled back to
 * its cluster's earliest current time (Particle::rollBackState,
 * model/particle.cpp:96-98).  Afterwards d2_step_run uses these clusters; a
 * cluster without pairs steps with its first particle's clamped last step.
 * D2_ERR_ARG for a pair of two static bodies (it belongs to no cluster). */
int d2_clusters_build(d2_ctx* ctx, int* n_clusters_out);
/* Any output may be NULL: cluster of each resident pair, of each body (-1
 * for static bodies), and per cluster its first particle, earliest current
 * time and smallest last_time_step_size (Cluster::min_current_time /
 * step_size before time-step selection). */
int d2_clusters_download(d2_ctx* ctx, int32_t* cluster_of_pair, int32_t* cluster_of_body, int32_t* first,
                         double* min_time, double* step);

/* ---- learned pair gate (no reference counterpart) ------------------------
 * A 24 -> 64 -> 64 -> 2 MLP (ReLU; bf16 weights, FP32 accumulation on MFMA)
 * proposes per pair (y0) its step scale and (y1) a lower bound on its time of
 * impact as a fraction of the CCD interval.  With the gate on, a pair with
 * y0 >= s_gate and y1 >= t_gate is checked by a sound separation certificate
 * (vertex sets separated by more than the search distance plus rounding and
 * float-barycentric margins, in every frame the sub-tests use); certified
 * pairs skip the traversal and take the plain path's result for such pairs
 * (no contact, scale 1, min_time = max_time); all other pairs, including
 * failed certificates, run the plain path.  Results are identical either way.
 * Parameters: FP32 master weights, layout W1[24][64] b1[64] W2[64][64] b2[64]
 * W3[64][2] b3[2] (row = input), d2_mlp_num_params() floats. */
int d2_mlp_num_params(void);
int d2_mlp_set_params(d2_ctx* ctx, const float* params);
int d2_mlp_get_params(d2_ctx* ctx, float* params);
/* Inference (the MFMA kernel the gate uses) of n feature rows X[n][24] -> Y[n][2]. */
int d2_mlp_infer(d2_ctx* ctx, const float* X, int n, float* Y);
/* Mean loss and its gradient over n samples (one batch; the training kernels),
 * loss per sample sum_o lw_o * huber_delta(y_o - t_o). */
int d2_mlp_loss_grad(d2_ctx* ctx, const float* X, const float* T, int n, float lw0, float lw1, float delta,
                     float* loss, float* grad);
/* Adam training on the device (forward, backward and update kernels),
 * minibatches of `batch` shuffled with `seed`; mean loss per epoch out. */
int d2_mlp_train(d2_ctx* ctx, const float* X, const float* T, int n, int epochs, int batch, float lr, float lw0,
                 float lw1, float delta, uint64_t seed, float* epoch_loss);
/* The 24 features of every resident pair (relative pose and motion in body
 * a's frame, lengths in units of r_a + r_b, times in units of the interval). */
int d2_pair_features(d2_ctx* ctx, float* X);
/* The gate for later runs (needs parameters): on = 0 off, or a mask of the
 * phases it applies to (1: CCD, 2: discrete, 3: both). */
int d2_set_learned(d2_ctx* ctx, int on, float t_gate, float s_gate);

/* Results of the last run: CSR offsets (n_pairs+1), contact records (the
 * Contact<double> vectors the reference fills, collision_detection/contact.h:
 * 10-28: A, B, eps_a, eps_b, plus pair, node and sub-test ids), and per-pair
 * values (CCD: scale, min_time_out; discrete: unused).
 * Any pointer may be NULL.  For a filtered run the call computes the
 * filtered CSR offsets on the context's stream before copying. */
int d2_result_contacts(d2_ctx* ctx, int32_t* offsets, int32_t* ints, double* reals);
int d2_result_pairs(d2_ctx* ctx, double* per_pair);

/* The plugin point's bucket comparison:
 * strategy::ContactDetectionStrategy::findContactsBucketComparison
 * (strategies/contact_detection/contact_detection_strategy.h:11) as
 * ContactDetectionComparison implements it (contact_detection_comparison.cpp:
 * 12-14 -> collision_detection/comparison_dist.h:205-240).  node_pairs[k] =
 * (node of body_a's tree, node of body_b's tree) -- the DeferredCompare list
 * of one BFS level; every triangle of the first bucket is compared with every
 * triangle of the second at the bodies' future states, in FP64.  Hits are the
 * contact records (layout above; pair 0, kind 0, level 0) in the reference's
 * order (node pair, then a's triangle, then b's); pair_used_out[k] = hits of
 * node pairs 0..k (the reference appends to `hits` and records hits.size(), so
 * a caller appending to a non-empty vector adds its size).  *n_hits_out gets
 * the hit count; when it exceeds max_hits nothing but pair_used_out and the
 * count is written and the call fails with D2_ERR_CAPACITY (retry with more
 * room).  D2_ERR_ARG for a node index outside its tree. */
int d2_find_contacts_bucket(d2_ctx* ctx, int body_a, int body_b, int n_node_pairs, const int32_t* node_pairs,
                            int32_t* pair_used_out, int64_t max_hits, int32_t* ints_out, double* reals_out,
                            int64_t* n_hits_out);
/* d2_find_contacts_bucket without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as d2_find_contacts_bucket, on a fixed
 * scratch array that each GPU holds once for the life of the process: at most
 * 32768 node pairs per call (more: D2_ERR_CAPACITY before any device work)
 * and at most 131072 hits (more: D2_ERR_CAPACITY with *n_hits_out and
 * pair_used_out written, as for max_hits).  Every call on one GPU uses that
 * GPU's one array, from every context: calls that overlap on one GPU (two
 * threads, two contexts' streams) race and must be serialised by the
 * caller.  The call synchronises the context's stream before it returns. */
int d2_find_contacts_bucket_noalloc(d2_ctx* ctx, int body_a, int body_b, int n_node_pairs, const int32_t* node_pairs,
                                    int32_t* pair_used_out, int64_t max_hits, int32_t* ints_out, double* reals_out,
                                    int64_t* n_hits_out);
/* filterContacts(unfiltered, margin) (collision_detection/contact.h:44-69) on
 * the device over n contacts given as reals (11 per contact, layout above;
 * the filter reads A, B, eps_a and eps_b).  The reference sorts `unfiltered`
 * in place with libstdc++ std::sort by (eps_a + eps_b, |A - B|) and keeps a
 * contact unless its midpoint lies within eps_a + eps_b + margin of an
 * earlier kept one: sorted_out[t] = the input index at sorted position t
 * (that permutation, n entries), kept_out[0..*n_kept_out) = the kept input
 * indices in sorted order (room for n).  Uses the context's run buffers: the
 * last run's results are no longer readable afterwards. */
int d2_filter_contacts(d2_ctx* ctx, int n, const double* reals, double margin, int32_t* sorted_out, int32_t* kept_out,
                       int* n_kept_out);

/* Convenience: upload pairs and run, as one call (host buffers in/out). */
int d2_narrow(d2_ctx* ctx, int n_pairs, const int32_t* pairs, int filter, int64_t* n_contacts_out);
int d2_ccd(d2_ctx* ctx, int n_pairs, const int32_t* pairs, const double* max_time, int64_t* n_contacts_out);

/* ---- instrumentation ---------------------------------------------------- */
/* When on, every kernel launch is bracketed by HIP events on the context's
 * stream; d2_kernel_stats returns per-kernel launch counts and summed
 * milliseconds (index order: d2_kernel_name(i)). */
int d2_set_profiling(d2_ctx* ctx, int on);
int d2_num_kernels(void);
const char* d2_kernel_name(int i);
int d2_kernel_stats(d2_ctx* ctx, int i, int64_t* launches, double* total_ms, double* items);
int d2_reset_stats(d2_ctx* ctx);
/* Counters of the last run (read back when it finished, also after a
 * failure): 8 global counters (contacts, overflow flags: 1 CCD phase,
 * 2 discrete phase, triangle-pair items, output contacts, then the learned
 * gate's proposed and certified pairs of the CCD and of the discrete phase), then
 * per level 0..16 eight sizes (soup node pairs, connected node pairs, items,
 * point and edge sub-tests, 3 reserved) of the discrete phase (of the CCD
 * phase for d2_ccd_run), then the same table of the last CCD phase (its
 * contact count in the last field of level 0).  Copies min(n, total) values;
 * returns the total. */
int d2_run_info(const d2_ctx* ctx, int64_t* info, int n);
/* Subtree pruning of the discrete traversal (no reference counterpart; on by
 * default): a node pair whose leaf triangles are provably farther apart than
 * any leaf contact could be (separating-axis gap of the oriented leaf boxes
 * against the search distance plus the nodes' largest leaf eps and
 * float-barycentric slack) is not tested and not expanded.  The reference
 * finds no contact below such a node pair, so every result is unchanged;
 * only the work per level shrinks.  Accounting runs (d2_set_touch_stats)
 * walk the reference's node pairs without it. */
int d2_set_pruning(d2_ctx* ctx, int on);
/* Algorithmic-byte accounting (no reference counterpart: instrumentation).
 * When on, runs are eager and every level also records which (body, node)
 * records its traversal tests (a discrete node pair whose parent missed is
 * not tested).  d2_touch_stats copies, for the last run, per phase (0
 * discrete, 1 CCD) and level 0..16 four counts -- distinct inner nodes,
 * distinct leaf nodes, the triangles of those leaves, distinct bodies --
 * then the same four counts over the run's union of levels and phases
 * (144 values).  Copies min(n, total) values; returns the total.  Turning
 * the accounting on forgets the last counts; turning it off keeps them. */
int d2_set_touch_stats(d2_ctx* ctx, int on);
int d2_touch_stats(const d2_ctx* ctx, int64_t* out, int n);
/* Synchronise the context's stream. */
int d2_sync(d2_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* D2HIP_H */
