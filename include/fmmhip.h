/*
 * fmmhip.h - C ABI of the MI355X FMM evaluator (libfmmhip.so).
 *
 * This boundary replaces the FMM evaluation core of FastMultipole.jl:
 *   - Tree(systems, target; ...)                     src/tree.jl:3-122
 *   - build_interaction_lists + sort_by_target        src/interaction_list.jl:3-19, 183-213
 *   - fmm!(targets, target_tree, sources, source_tree, ..., m2l_list, direct_list, ...)
 *                                                     src/fmm.jl:914-1176
 *   - direct!(targets, sources; ...)                  src/direct.jl:20-75
 * for systems whose body_to_multipole! / direct! are one of the built-in point kernels
 * (fmmhip_kernel below).
 *
 * Data layout across the ABI follows the reference's buffers (src/compatibility.jl:5-20,
 * 155-172): column-major double matrices, i.e. one record per body.
 *   target positions : 3 x N   [x, y, z]
 *   source buffer    : data_per_body x N  [x, y, z, radius, strength...]
 *                      (5 rows for POINT_SOURCE, 7 for the other point kernels, 14 / 17 for panels)
 *   target results   : 16 x N  [x, y, z, phi, grad(3), hessian(9)] in the ORIGINAL body
 *                      order (the reference's target buffer after unsorting through
 *                      sort_index, src/compatibility.jl:324-338). Hessian entries are in
 *                      the reference's linear order (vxx, vxy, vxz, vyx, ... ).
 * Indices returned by the getters are 0-based; the reference's are these plus one.
 *
 * Errors never unwind across the ABI: every call returns an fmmhip_status and
 * fmmhip_last_error() describes the last failure. Each context owns one HIP
 * stream and is not thread-safe; use one context per host thread.
 */
#ifndef FMMHIP_H
#define FMMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FMMHIP_OK = 0,
    FMMHIP_EINVAL = 1,   /* bad argument (reference: @assert / ArgumentError) */
    FMMHIP_ENOMEM = 2,   /* device allocation failed */
    FMMHIP_EDEVICE = 3,  /* HIP runtime or kernel launch error, or no GPU */
    FMMHIP_ETOL = 4,     /* error tolerance not met (reference: error_success == false), results delivered;
                            returned by the request server (fmmhip_server_wait); the direct calls return
                            FMMHIP_OK and report fmmhip_stats.error_success, as fmm! returns error_success */
    FMMHIP_EDEPTH = 5    /* octree deeper than 42 levels (coincident bodies) */
} fmmhip_status;

/* Built-in kernels (the system's body_to_multipole! / direct! tag):
 *   POINT_SOURCE     Point{Source}, test/gravitational.jl:80-104; rows [x,y,z,radius,q]
 *   POINT_VORTEX     Point{Vortex}, test/vortex.jl:83-128;         rows [x,y,z,sigma,Gx,Gy,Gz]
 *   POINT_DIPOLE     Point{Dipole} expansion (src/bodytomultipole.jl:431-463) with the direct
 *                    kernel phi = q.d/(4 pi r^3);                   rows [x,y,z,radius,qx,qy,qz]
 *   VORTEX_GAUSSIAN  vortex particle with a Gaussian core of width sigma (row 4): the
 *                    singular kernel times erf(rho/sqrt2) - sqrt(2/pi) rho exp(-rho^2/2),
 *                    rho = r/sigma; expansions as POINT_VORTEX;     rows [x,y,z,sigma,Gx,Gy,Gz]
 *   PANEL_*          constant-strength planar panels, Panel{Source} / Panel{Dipole}: nearfield
 *                    induced() of test/panels.jl, expansions body_to_multipole_panel!
 *                    (src/bodytomultipole.jl:643-831); triangles: rows [centroid(3), radius,
 *                    strength, v1(3), v2(3), v3(3)] (14 rows); quadrilaterals add v4(3) (17 rows);
 *                    vertices counter-clockwise about the normal; FP64 nearfield in both precisions
 *   POINT_SOURCE_VORTEX  Point{SourceVortex} (src/bodytomultipole.jl:491-496): a point source and a point
 *                    vortex on one body; expansions the source's, then the vortex's; nearfield the
 *                    POINT_SOURCE pair terms (potential, gradient, and the gradient's derivative, which
 *                    the far field carries) plus the POINT_VORTEX ones (velocity, its gradient), each
 *                    written as the switches ask; rows [x,y,z,sigma,q,Gx,Gy,Gz]
 *   PANEL_SOURCE_DIPOLE_*  Panel{SourceDipole}: a source and a dipole strength on one panel; expansions the
 *                    source's then the dipole's (src/bodytomultipole.jl:717-720); nearfield the two closed
 *                    forms of test/panels.jl:323-429; rows [centroid(3), radius, source strength, dipole
 *                    strength, v1(3), v2(3), v3(3)] (15 rows), quadrilaterals add v4(3) (18 rows)
 * Results use the expansion evaluation's convention: slot "gradient" = -grad(phi) for the
 * scalar kernels, the velocity for the vortex kernels; slot "hessian" = d(gradient slot). */
typedef enum {
    FMMHIP_POINT_SOURCE = 0,
    FMMHIP_POINT_VORTEX = 1,
    FMMHIP_POINT_DIPOLE = 2,
    FMMHIP_VORTEX_GAUSSIAN = 3,
    FMMHIP_PANEL_SOURCE_TRI = 4,
    FMMHIP_PANEL_DIPOLE_TRI = 5,
    FMMHIP_PANEL_SOURCE_QUAD = 6,
    FMMHIP_PANEL_DIPOLE_QUAD = 7,
    FMMHIP_POINT_SOURCE_VORTEX = 8,
    FMMHIP_PANEL_SOURCE_DIPOLE_TRI = 9,
    FMMHIP_PANEL_SOURCE_DIPOLE_QUAD = 10
} fmmhip_kernel;
typedef enum { FMMHIP_FP64 = 0, FMMHIP_MIXED = 1 } fmmhip_precision;
typedef enum { FMMHIP_BARBA = 0, FMMHIP_SELFTUNING = 1, FMMHIP_SELFTUNING_TREESTOP = 2 } fmmhip_list_method;
typedef enum {
    FMMHIP_ERR_NONE = 0,
    FMMHIP_ERR_ROTATED_COEFFICIENTS_GRADIENT = 1, /* RotatedCoefficients{Absolute,Relative}Gradient */
    FMMHIP_ERR_POWER_POTENTIAL = 2,               /* Power{Absolute,Relative}Potential */
    FMMHIP_ERR_POWER_GRADIENT = 3                 /* Power{Absolute,Relative}Gradient */
} fmmhip_error_kind;

/* Keyword arguments of fmm! (src/fmm.jl:862-923); defaults from fmmhip_options_default(). */
typedef struct {
    int32_t expansion_order;      /* 5 */
    int32_t leaf_size_source;     /* 20 */
    int32_t leaf_size_target;     /* same as source */
    double multipole_acceptance;  /* 0.4 */
    int32_t list_method;          /* FMMHIP_SELFTUNING_TREESTOP */
    int32_t shrink_recenter;      /* 1 */
    int32_t error_kind;           /* FMMHIP_ERR_NONE */
    int32_t error_relative;       /* 0: Absolute, 1: Relative (runs estimate_influence!) */
    double error_tolerance;       /* epsilon */
    int32_t error_bonus;          /* BE, 1 */
    int32_t scalar_potential;     /* 0 */
    int32_t gradient;             /* 1 */
    int32_t hessian;              /* 0 */
    int32_t farfield, nearfield, self_induced;  /* 1, 1, 1 */
    int32_t precision;            /* FMMHIP_FP64 */
} fmmhip_options;

typedef struct {
    int32_t pmax;                 /* max expansion order used by M2L (+1 when a tolerance was missed) */
    int32_t error_success;        /* reference's error_success */
    int64_t n_target_branches, n_source_branches;
    int64_t n_m2l, n_direct;      /* interaction list lengths */
    int64_t n_p2p_pairs;          /* body-body pairs evaluated by the nearfield */
    double ms_tree, ms_lists, ms_p2p, ms_upward, ms_m2l, ms_downward, ms_total;  /* device time per phase */
    double ms_estimate;           /* estimate_influence pre-pass (relative tolerances only) */
    double ms_gather;             /* scatter of sorted results back to the caller's order */
} fmmhip_stats;

typedef struct fmmhip_ctx fmmhip_ctx;
typedef struct fmmhip_eval fmmhip_eval;

void fmmhip_options_default(fmmhip_options* opt);

/* Context: one HIP device and one stream. device < 0 selects the current device. */
fmmhip_status fmmhip_ctx_create(int device, fmmhip_ctx** out);
void fmmhip_ctx_destroy(fmmhip_ctx* ctx);
const char* fmmhip_last_error(const fmmhip_ctx* ctx);
fmmhip_status fmmhip_sync(fmmhip_ctx* ctx);
/* Launch shape of later evaluations on this context (no counterpart in the reference; set by the
 * project's autotuner, core.autotune): workgroup slots per compute unit the persistent nearfield leaves
 * free for the far-field stream (0..3, default 1). */
fmmhip_status fmmhip_ctx_set_p2p_free_slots(fmmhip_ctx* ctx, int32_t slots);

/* An evaluation keeps trees, lists, expansions and buffers resident in HBM between
 * calls (the reference's Cache + Tree objects, src/containers.jl:264-292). */
fmmhip_status fmmhip_eval_create(fmmhip_ctx* ctx, fmmhip_eval** out);
void fmmhip_eval_destroy(fmmhip_eval* ev);

/* Hand over one target system and one source system (fmm!(targets, sources)).
 * on_device = 0: host pointers, copied in. on_device = 1: device pointers (HBM),
 * read in place during the following calls; they must stay valid, and positions read in
 * place must not change between fmmhip_eval_build_trees and the evaluation that follows
 * (the trees are built from them, and the influence estimate of a relative tolerance
 * reuses the source tree's octant keys within that evaluation; fmmhip_fmm builds and
 * evaluates in one call). */
fmmhip_status fmmhip_eval_set_systems(fmmhip_eval* ev, int32_t kernel, int64_t n_targets, const double* target_xyz,
                                      int64_t n_sources, const double* source_buffer, int32_t on_device);

/* Several independent systems of one kernel, each its own target and source (fmm!(system) for each of
 * them), evaluated together: one tree per system (its own root cube, the tree the reference builds for
 * that system alone), interaction lists within each system only, and every pass of all systems in the
 * same launches. No counterpart in the reference, whose callers loop over fmm!; for batches of many
 * small systems. counts: n_systems sizes, each >= 1; source_buffer: the systems' columns one after
 * another (data_per_body x sum(counts), column-major). The stages below, fmmhip_fmm and
 * fmmhip_eval_get_targets then work on the batch; results follow the same concatenated order, and the
 * stats (pmax, error_success, counts, times) cover the whole batch. FastGaussSeidel takes one system. */
fmmhip_status fmmhip_eval_set_batch(fmmhip_eval* ev, int32_t kernel, int64_t n_systems, const int64_t* counts,
                                     const double* source_buffer, int32_t on_device);
/* Several systems in one tree, each a target and a source of all of them (fmm!((sys1, sys2, ...)),
 * src/fmm.jl:862-884 with the tuple trees of src/tree.jl:3-122): one tree over the bodies of every system,
 * whose leaf test counts each system against its leaf size (src/tree.jl:492-569; opt->leaf_size_source for
 * each, or fmmhip_eval_set_tuple_leaf_sizes); SelfTuningTreeStop or Barba lists (SelfTuning lists are
 * FMMHIP_EINVAL: their decisions weigh each system's count too). kernels: any FMMHIP_POINT_* or FMMHIP_PANEL_* kernel
 * per system. buffer: R x sum(counts) column-major, R the largest data_per_body of the kernels (7 for point
 * systems, 14 with triangles, 17 with quadrilaterals), each system's rows [x, y, z, radius, strength
 * rows...] first and zeros after them, the systems one after another in the order of kernels. The
 * nearfield and P2M run once per source system with its kernel; results follow the same concatenated
 * order. Not combined with batches or FastGaussSeidel. */
fmmhip_status fmmhip_eval_set_tuple(fmmhip_eval* ev, int64_t n_systems, const int32_t* kernels, const int64_t* counts,
                                     const double* buffer, int32_t on_device);
/* Leaf sizes per system for the tuple set last (the reference's SVector leaf_size, src/tree.jl:492-569):
 * a branch splits by the reference's test on each system's count against its own leaf size (SelfTuningTreeStop
 * and Barba trees; without this call every system takes opt->leaf_size_source). Kept until the next
 * fmmhip_eval_set_tuple. */
fmmhip_status fmmhip_eval_set_tuple_leaf_sizes(fmmhip_eval* ev, int64_t n_systems, const int64_t* leaf_sizes);
/* After an evaluation: per system of the batch, the expansion order and error_success its own fmm! call
 * reports (pmax: the largest order its M2L interactions used, plus one where the tolerance was not met).
 * Arrays of n_systems entries (one entry for an evaluation of a single system); either may be null.
 * No counterpart in the reference (it belongs to fmmhip_eval_set_batch). */
fmmhip_status fmmhip_eval_get_batch_stats(fmmhip_eval* ev, int32_t* pmax, int32_t* error_success);

/* Stages of fmm!: Tree() for both systems (src/tree.jl:3-122) */
fmmhip_status fmmhip_eval_build_trees(fmmhip_eval* ev, const fmmhip_options* opt);
/* build_interaction_lists + sort_by_target (src/fmm.jl:904-907) */
fmmhip_status fmmhip_eval_build_lists(fmmhip_eval* ev, const fmmhip_options* opt);
/* the fmm! core: estimate_influence!, nearfield, upward, horizontal, downward passes (src/fmm.jl:914-1152) */
fmmhip_status fmmhip_eval_run(fmmhip_eval* ev, const fmmhip_options* opt, fmmhip_stats* stats);
/* all three stages (fmm!(targets, sources; kwargs...), src/fmm.jl:862-884) */
fmmhip_status fmmhip_fmm(fmmhip_eval* ev, const fmmhip_options* opt, fmmhip_stats* stats);

/* Results: 16 x n_targets in original body order. on_device selects a device or host destination. */
fmmhip_status fmmhip_eval_get_targets(fmmhip_eval* ev, double* out, int32_t on_device);

/* Tree and list inspection (host copies). which = 0 target tree, 1 source tree. */
int64_t fmmhip_eval_n_branches(fmmhip_eval* ev, int32_t which);
int64_t fmmhip_eval_n_levels(fmmhip_eval* ev, int32_t which);
int64_t fmmhip_eval_n_leaves(fmmhip_eval* ev, int32_t which);
/* ints: n x 6 [n_bodies, first_body, n_branches, first_child, parent, leaf]
 * reals: n x 17 [source_center(3), target_center(3), source_radius, target_radius,
 *                source_box(3), target_box(3), max_influence, 0, 0]          (Branch, src/containers.jl:150-164) */
fmmhip_status fmmhip_eval_get_branches(fmmhip_eval* ev, int32_t which, int64_t* ints, double* reals);
fmmhip_status fmmhip_eval_get_levels(fmmhip_eval* ev, int32_t which, int64_t* first_last);
fmmhip_status fmmhip_eval_get_leaf_index(fmmhip_eval* ev, int32_t which, int64_t* out);
fmmhip_status fmmhip_eval_get_sort_index(fmmhip_eval* ev, int32_t which, int64_t* out);
/* which = 0 m2l list, 1 direct list; n x 2 int32 (target, source) sorted by target */
int64_t fmmhip_eval_n_list(fmmhip_eval* ev, int32_t which);
fmmhip_status fmmhip_eval_get_list(fmmhip_eval* ev, int32_t which, int32_t* out);
/* expansion order chosen for each M2L interaction, in m2l-list order */
fmmhip_status fmmhip_eval_get_m2l_orders(fmmhip_eval* ev, int32_t* out);
/* expansions of a tree, [branch][H][phi/chi][re/im] doubles (src/tree.jl:1334-1339 layout per branch) */
fmmhip_status fmmhip_eval_get_expansions(fmmhip_eval* ev, int32_t which, double* out);

/* direct!(targets, sources; scalar_potential, gradient, hessian): O(N^2) on the GPU (src/direct.jl:43-75).
 * target_out: 16 x n_targets, accumulated (+=) like the reference's set_* accessors. */
fmmhip_status fmmhip_direct(fmmhip_ctx* ctx, int32_t kernel, int64_t n_targets, const double* target_xyz,
                            int64_t n_sources, const double* source_buffer, int32_t scalar_potential,
                            int32_t gradient, int32_t hessian, int32_t precision, double* target_out,
                            int32_t on_device);

/* Allocation-free direct!: no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as fmmhip_direct with on_device = 1 (target_xyz, source_buffer and
 * target_out are device pointers); it calls no hipMalloc, hipFree or allocating routine. Few
 * targets against many sources are split into source chunks whose sums go to a static workspace of
 * the library (52 MiB per device, held until the process ends); the call fails with FMMHIP_EINVAL if
 * a size would need more (it cannot with the library's chunk choice). Calls that overlap on one
 * device share that workspace and race: serialise them. */
fmmhip_status fmmhip_direct_noalloc(fmmhip_ctx* ctx, int32_t kernel, int64_t n_targets, const double* target_xyz,
                                    int64_t n_sources, const double* source_buffer, int32_t scalar_potential,
                                    int32_t gradient, int32_t hessian, int32_t precision, double* target_out);

/* ---- FastGaussSeidel (src/solve.jl) for one system that is both target and source
 * (fmmhip_eval_set_systems with target_xyz = NULL): point sources or source panels.
 * influence: the value matched per body, FMMHIP_INFLUENCE_POTENTIAL (the scalar potential,
 * the reference's test influence!, test/solve_test.jl:11-13) or FMMHIP_INFLUENCE_NORMAL_GRADIENT
 * (panels: the gradient along the panel's unit normal). */
enum { FMMHIP_INFLUENCE_POTENTIAL = 0, FMMHIP_INFLUENCE_NORMAL_GRADIENT = 1 };

/* FastGaussSeidel(system; expansion_order, multipole_acceptance, leaf_size, shrink_recenter)
 * (src/solve.jl:538-645): trees, Barba lists, per-leaf self and non-self influence matrices in HBM.
 * Uses opt->expansion_order, leaf_size_source, multipole_acceptance, shrink_recenter; leaves hold at
 * most 4096 bodies (FMMHIP_EINVAL otherwise). Self blocks of up to 101 bodies are inverted in LDS, larger
 * ones in device memory. */
fmmhip_status fmmhip_fgs_setup(fmmhip_eval* ev, const fmmhip_options* opt, int32_t influence);
/* solve!(system, solver; max_iterations, tolerance) (src/solve.jl:797-946). external: the external
 * influence per body (the reference's target_influence_to_buffer! values), original order, host or
 * device memory. Strengths start from the source buffer given to fmmhip_eval_set_systems;
 * strengths_out (host, original order, may be NULL) receives the solved strengths; the target
 * buffer (potential, gradient: farfield of the last iteration plus the nearfield of the solved
 * strengths) is read with fmmhip_eval_get_targets. sweeps: Gauss-Seidel sweeps performed;
 * mse_history (host, max_iterations entries, may be NULL): mean squared residual over all bodies
 * at each iteration. */
fmmhip_status fmmhip_fgs_solve(fmmhip_eval* ev, const double* external, int32_t on_device, int32_t max_iterations,
                               double tolerance, double* strengths_out, int32_t* sweeps, double* mse_history);
/* inspection: per leaf (leaf_index order) [branch, n_bodies, non-self rows, first sorted body];
 * matrices column-major (which = 0 self n x n, 1 non-self rows x n, src/solve.jl:3-51);
 * original index of the body of every non-self row */
int64_t fmmhip_fgs_n_leaves(fmmhip_eval* ev);
fmmhip_status fmmhip_fgs_leaf_info(fmmhip_eval* ev, int64_t* out);
fmmhip_status fmmhip_fgs_get_matrix(fmmhip_eval* ev, int32_t which, int64_t leaf, double* out);
fmmhip_status fmmhip_fgs_nonself_rows(fmmhip_eval* ev, int64_t leaf, int64_t* out);

/* ---- Request server: many fmm!(system) evaluations and FastGaussSeidel solves of uneven
 * sizes, submitted from any host thread, run concurrently on one device. The reference has
 * no counterpart (a Julia caller loops over fmm!(system) / solve!(system, solver),
 * src/fmm.jl:862-884, src/solve.jl:797-946); each request does exactly what that call does
 * for one system that is both target and source.
 * n_slots worker threads (1..16) each own a context (stream) and an evaluation whose device
 * buffers persist between requests. A free worker takes the oldest waiting request whose
 * estimated HBM footprint fits beside the running ones within hbm_budget_bytes (or the oldest
 * one when nothing runs). */
enum { FMMHIP_REQUEST_FMM = 0, FMMHIP_REQUEST_SOLVE = 1 };
typedef struct {
    int32_t kind;                 /* FMMHIP_REQUEST_FMM or FMMHIP_REQUEST_SOLVE */
    int32_t kernel;               /* fmmhip_kernel; SOLVE: point sources or source panels */
    int64_t n;                    /* bodies */
    const double* source_buffer;  /* host, data_per_body x n; copied by submit */
    fmmhip_options options;       /* fmm! keywords; SOLVE uses the ones fmmhip_fgs_setup reads */
    int32_t influence;            /* SOLVE: FMMHIP_INFLUENCE_* */
    int32_t max_iterations;       /* SOLVE */
    double tolerance;             /* SOLVE */
    const double* external;       /* SOLVE: host, n external influences; copied by submit */
} fmmhip_request;
typedef struct {
    int32_t status;               /* the request's fmmhip_status; FMMHIP_ETOL: results delivered, but the FMM's
                                     error tolerance was missed (error_success == 0) or the solve stopped at
                                     max_iterations above its residual tolerance (converged == 0) */
    int32_t sweeps;               /* SOLVE: Gauss-Seidel sweeps */
    int32_t pmax;                 /* FMM: fmmhip_stats.pmax */
    int32_t error_success;        /* FMM: fmmhip_stats.error_success (1 for a solve) */
    int32_t converged;            /* SOLVE: residual tolerance met (1 for an FMM) */
    double queue_ms, run_ms, latency_ms;  /* submit -> start, start -> end, submit -> end */
    double start_ms, end_ms;      /* start and end, in ms since the server was created */
    double footprint_bytes;       /* the admission estimate */
} fmmhip_request_stats;
typedef struct {
    int64_t completed, queued;
    int32_t running;
    double p50_ms, p99_ms;        /* latency percentiles over completed requests */
    double mean_queue_ms;
    double span_ms;               /* first submit -> last completion */
} fmmhip_server_stats_t;
typedef struct fmmhip_server fmmhip_server;

fmmhip_status fmmhip_server_create(int device, int32_t n_slots, double hbm_budget_bytes, fmmhip_server** out);
/* waits for running requests; requests not yet taken are dropped */
void fmmhip_server_destroy(fmmhip_server* server);
/* FMM requests with the same kernel and options that wait together run as one batch evaluation
 * (fmmhip_eval_set_batch): a worker takes its share of the waiting requests (the queue spread over it
 * and the idle workers), up to max_requests requests and max_bodies bodies (defaults 64 and 2 Mi);
 * max_requests = 1 runs every request alone. Results and statuses are those of the requests' own
 * calls. No counterpart in the reference. */
fmmhip_status fmmhip_server_set_batching(fmmhip_server* server, int32_t max_requests, int64_t max_bodies);
fmmhip_status fmmhip_server_submit(fmmhip_server* server, const fmmhip_request* request, int64_t* ticket);
/* blocks until the request ends and returns its status; target_out (16 x n, original order) and
 * strengths_out (SOLVE, n) are host buffers or NULL, filled when the status is FMMHIP_OK or
 * FMMHIP_ETOL. A ticket is waited on once. */
fmmhip_status fmmhip_server_wait(fmmhip_server* server, int64_t ticket, double* target_out, double* strengths_out,
                                 fmmhip_request_stats* stats);
fmmhip_status fmmhip_server_stats(fmmhip_server* server, fmmhip_server_stats_t* out);

#ifdef __cplusplus
}
#endif
#endif /* FMMHIP_H */
