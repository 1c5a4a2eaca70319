/*
 * pisurf: MI355X (gfx950) implementation of the ParticleIsosurface field-evaluation and
 * vertex-placement path, behind a plain C ABI.
 *
 * The reference has no FFI layer: its seams are the C++ classes HashGrid
 * (include/hash_grid.h:9-45), Evaluator (include/evaluator.h:13-108) and SurfReconstructor
 * (include/surface_reconstructor.h:21-93). Each entry point below replaces one of their
 * methods (cited per function); the C++ mirror of those classes sits on top of this ABI
 * (particleisosurface_amd/csrc/host_facade.h).
 *
 * Conventions
 *  - every function returns PISURF_OK (0) or a negative PISURF_E_* code; no exception or
 *    abort crosses the ABI; pisurf_last_error() gives the message of the last failure;
 *  - host buffers are caller-owned, device buffers are context-owned;
 *  - a context is used by one host thread at a time (any thread; calls must not overlap) and
 *    launches on one HIP stream of default priority;
 *  - positions are 3 packed floats per particle (layout of std::vector<Eigen::Vector3f>),
 *    3x3 matrices are 9 floats/doubles column-major (layout of Eigen::Matrix3f);
 *  - value outputs are double: FP32 results are exact in double, FP64 results are native.
 */
#ifndef PISURF_H
#define PISURF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PISURF_OK 0
#define PISURF_E_INVALID (-1)
#define PISURF_E_OOM (-2)
#define PISURF_E_HIP (-3)
#define PISURF_E_UNSUPPORTED (-4)
#define PISURF_E_STATE (-5)

/* arithmetic of the field path */
#define PISURF_FP32 0        /* float storage and sums */
#define PISURF_FP32_KAHAN 1  /* float storage and terms, sums accumulated in FP64, FP64 eigensolve for clamped spectra */
#define PISURF_FP64 2        /* double everywhere after the float positions */

/* leaf vertex placement */
#define PISURF_QEF_OFF 0      /* shipped behaviour: leaf vertex = centre (src/surface_reconstructor.cpp:258,270) */
#define PISURF_QEF_REFERENCE 1/* TNode::NodeCalcNode: float 4x4 inverse (src/iso_method_ours.cpp:101-174) */
#define PISURF_QEF_SVD64 2    /* FP64 truncated-SVD pseudo-inverse about the cell centre */

typedef struct pisurf_ctx pisurf_ctx;

typedef struct pisurf_params {
    float radius;              /* constant particle radius r */
    float smooth_factor;       /* box margin factor of checkEmptyAndCalcCurv (2.0) */
    float iso_factor;          /* poly6 iso factor (1.9) */
    int32_t use_ani;           /* USE_ANI (1) */
    int32_t use_poly6;         /* USE_POLY6 (0: cubic B-spline) */
    int32_t calc_normals;      /* CALC_P_NORMAL (1) */
    int32_t oversample;        /* _OVERSAMPLE_QEF; only 2 is supported */
    int32_t precision;         /* PISURF_FP32 / PISURF_FP32_KAHAN / PISURF_FP64 */
    int32_t qef_mode;          /* PISURF_QEF_* */
    float qef_svd_tol;         /* relative singular-value cutoff of PISURF_QEF_SVD64 */
    int32_t depth_max_override;/* 0: derived D_max (resizeRootBoxConstR) */
    int32_t mesh_order;        /* 0: extraction on the device, dual cells in point-key order;
                                  1: host traversal in the reference's order (byte-identical OBJ) */
    int32_t sample_dedupe;     /* 1 (default): node samples shared by neighbouring nodes of a level
                                  (bitwise-identical coordinates) are evaluated once; 0: 27 per node */
    int32_t frames_in_flight;  /* pisurf_run_frames: frames meshed at once, each on a context and
                                  stream of its own (0: the library's choice, 1: one after another) */
    int32_t level_graph;       /* 1: the launches of the octree levels above D_min are captured as a
                                  HIP graph, instantiated once and updated in place per frame
                                  (default 0: launched one by one, which measured 0.1 ms faster) */
    int32_t reserved[1];
} pisurf_params;

typedef struct pisurf_grid_info {
    float bounding[6];         /* HashGrid::Bounding */
    uint32_t xyz_cell_num[3];  /* HashGrid::XYZCellNum */
    int64_t cell_num;          /* HashGrid::CellNum */
    float cell_size;           /* HashGrid::CellSize */
    int64_t n_invalid;         /* particles whose cell hash is -1 */
} pisurf_grid_info;

typedef struct pisurf_tree_info {
    int32_t depth_max, depth_min;
    float root_center[3];
    float root_half;
    int64_t n_nodes;
    int32_t n_levels;
    int64_t n_leaves;
} pisurf_tree_info;

typedef struct pisurf_stats {
    int64_t field_evals;       /* SingleEval-equivalents since the last reset (the reference's count:
                                  deduplicated node samples count once per node) */
    int64_t grad_evals;        /* SingleEvalWithGrad-equivalents */
    int64_t qef_solves;
    int64_t candidates;        /* candidate particles visited by all field evaluations */
    int64_t sample_evals;      /* evaluations of the node-sampling kernel (27 per sampled node) */
    int64_t sample_candidates; /* candidates visited by the node-sampling kernel */
    int64_t sample_launches;   /* launches of the per-point node-sampling kernel */
    int64_t fp64_eigensolves;  /* clamped spectra re-solved in FP64 (PISURF_FP32_KAHAN) */
    double ms_grid, ms_aniso, ms_normals, ms_tree, ms_extract_host, ms_bisect, ms_total;
    double ms_sample_kernel;   /* summed device time of the per-point node-sampling kernel (HIP events) */
    int64_t n_dual_cells, n_verts, n_tris, n_nodes;
    int64_t sample_unique;     /* sample evaluations performed after deduplication (<= sample_evals) */
    /* the run kernel of deduplicated sampling (distinct points grouped by hash cell, candidates
     * staged in LDS); its evaluations and candidates are included in sample_unique /
     * sample_candidates */
    int64_t run_evals, run_candidates, run_launches;
    double ms_run_kernel;
} pisurf_stats;

/* ---- context ---- */
int pisurf_create(pisurf_ctx** out, int device);
int pisurf_destroy(pisurf_ctx* ctx);
const char* pisurf_last_error(const pisurf_ctx* ctx);
int pisurf_default_params(pisurf_params* p);
int pisurf_set_params(pisurf_ctx* ctx, const pisurf_params* p);
/* SurfReconstructor ctor copies the particles (src/surface_reconstructor.cpp:13-28).
 * on_device != 0: xyz is a device pointer (HBM-resident input); it is copied. */
int pisurf_upload_particles(pisurf_ctx* ctx, const float* xyz, int64_t n, int32_t on_device);

/* ---- H1-H5: HashGrid (src/hash_grid.cpp) ---- */
/* HashGrid::HashGrid const-radius ctor (src/hash_grid.cpp:4-34). bbox NULL: the tight
 * particle bbox of SurfReconstructor::loadRootBox (src/surface_reconstructor.cpp:30-49). */
int pisurf_hashgrid_build(pisurf_ctx* ctx, const float* bbox, float inf_factor, pisurf_grid_info* out);
/* IndexList (src/hash_grid.cpp:71-85) and the cell windows StartList/EndList
 * (src/hash_grid.cpp:103-122) as dense arrays of cell_num entries (start = end for empty cells). */
int pisurf_hashgrid_download(pisurf_ctx* ctx, int32_t* index_list, int32_t* cell_start, int32_t* cell_end);
/* HashGrid::GetPIdxList for nq points (src/hash_grid.cpp:196-214): offsets[nq+1], ids[cap];
 * *total receives the number of ids (ids beyond cap are not written). */
int pisurf_neighbors(pisurf_ctx* ctx, const float* q, int64_t nq, int64_t* offsets, int32_t* ids, int64_t cap,
                     int64_t* total);
/* No counterpart in the reference: pisurf_neighbors for callers that cannot allocate device or
 * host memory during the call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). q (nq x 3 floats), offsets (nq + 1 int64) and ids (cap int32) are DEVICE
 * pointers; the per-query counts, 32-bit offsets and scan workspace live in a fixed device array
 * of the library (one per GPU, held until the process ends, about 8.4 MB), so nq is at most
 * PISURF_NOALLOC_MAX_QUERIES. *total (host, may be NULL) receives the number of ids; when it
 * exceeds cap nothing is written and PISURF_E_INVALID is returned. Calls that overlap on one GPU
 * race on the shared array; the call returns after its results are written. */
#define PISURF_NOALLOC_MAX_QUERIES (1 << 20)
int pisurf_neighbors_noalloc(pisurf_ctx* ctx, const float* q, int64_t nq, int64_t* offsets, int32_t* ids,
                             int64_t cap, int64_t* total);
/* HashGrid::GetInBoxParticles (src/hash_grid.cpp:164-194) */
int pisurf_box_particles(pisurf_ctx* ctx, const float* box1, const float* box2, int32_t* ids, int64_t cap,
                         int64_t* total);

/* ---- H6-H12: Evaluator (src/evaluator.cpp) ---- */
/* Evaluator::compute_Gs_xMeans (src/evaluator.cpp:656-749) */
int pisurf_anisotropy(pisurf_ctx* ctx);
/* GlobalxMeans / GlobalGs / GlobalDeterminant (include/evaluator.h:72-74), particle order */
int pisurf_download_aniso(pisurf_ctx* ctx, double* xmean, double* G9, double* det);
/* CalculateMaxScalarConstR + RecommendIsoValueConstR (src/evaluator.cpp:272-291,323-335) */
int pisurf_iso_constants(pisurf_ctx* ctx, double* max_scalar, double* iso_value);
/* CalcParticlesNormal (src/evaluator.cpp:363-375); normals may be NULL */
int pisurf_particle_normals(pisurf_ctx* ctx, double* normals);
/* SingleEval (src/evaluator.cpp:60-87) at n points; grad != NULL gives SingleEvalWithGrad
 * (src/evaluator.cpp:89-125) */
int pisurf_eval_points(pisurf_ctx* ctx, const float* q, int64_t n, double* f, double* grad);

/* ---- H13-H15: SurfReconstructor / TNode ---- */
/* One batch of nodes of one level: checkEmptyAndCalcCurv + GenerateSampling + GridEval +
 * classification (src/surface_reconstructor.cpp:152-273; src/iso_method_ours.cpp:52-99;
 * src/evaluator.cpp:127-252). samples (n x 27 x 4) and grads (n x 27 x 3) may be NULL. */
int pisurf_sweep_level(pisurf_ctx* ctx, const float* centers, const float* halfs, const int32_t* depths, int64_t n,
                       uint8_t* type, double* f, float* curv, int32_t* signchange, double* samples, double* grads);
/* genIsoOurs state 0 (src/surface_reconstructor.cpp:275-407) with resizeRootBoxConstR
 * (:97-122): host level loop over batched device launches. */
int pisurf_build_tree(pisurf_ctx* ctx, pisurf_tree_info* out);
/* nodes in creation order (level by level, children of a parent consecutive) */
int pisurf_tree_download(pisurf_ctx* ctx, float* center3, float* half, int32_t* depth, uint8_t* type, uint64_t* id,
                         double* node4, int32_t* child);

/* ---- UniformGrid::Run (src/marching.cpp:93-186): the field sampled on a dense lattice ----
 * Box: the tight particle box, per axis resized to ceil(len / r) * r + 4 r about its centre
 * (resizeRootBoxConstR, src/marching.cpp:41-56); dims[a] = int(len' / r) + 1 points, spacing r.
 * scalars (NULL: geometry only): dims[0] * dims[1] * dims[2] SingleEval values, x fastest
 * (gridSampling, src/marching.cpp:72-91). dims / box / spacing may each be NULL. */
int pisurf_uniform_grid(pisurf_ctx* ctx, int32_t dims[3], float box[6], float* spacing, float* scalars);

/* ---- H16: TNode::NodeCalcNode (src/iso_method_ours.cpp:101-174) for n nodes ---- */
int pisurf_qef_solve(pisurf_ctx* ctx, const float* centers, const float* halfs, const double* samples,
                     const double* grads, int64_t n, int32_t mode, double* node4);

/* ---- H17: VisitorExtract::calc_vertices bisection (src/visitorextract.cpp:47-84) ----
 * ends: n x 2 x 4 (x, y, z, f) edge endpoints; out_xyz n x 3; n_evals may be NULL. */
int pisurf_bisect_edges(pisurf_ctx* ctx, const double* ends, int64_t n, double* out_xyz, int32_t* n_evals);

/* ---- frame: extraction and mesh (genIsoOurs state 1, src/surface_reconstructor.cpp:290-324) ---- */
/* traverse_node<trav_vert> with VisitorExtract::on_vert (include/traverse.h:12-312,
 * src/visitorextract.cpp:229-257): the dual grids of the built octree in the reference's traversal
 * order, each as 8 node indices (creation order, as pisurf_tree_download lists them) already
 * permuted by the procedure table's `trans`, and its sign index (bit i set when corner i of the
 * traversal has f > 0; the procedure is table[sign_index]). *total receives the number of dual
 * grids; at most cap are written. */
int pisurf_dual_cells(pisurf_ctx* ctx, int32_t* nodes8, uint8_t* sign_index, int64_t cap, int64_t* total);
int pisurf_extract(pisurf_ctx* ctx, int64_t* n_verts, int64_t* n_tris);
int pisurf_mesh_download(pisurf_ctx* ctx, double* verts, int32_t* tris);
/* SurfReconstructor::Run (src/surface_reconstructor.cpp:409-479) on the uploaded particles */
int pisurf_run_frame(pisurf_ctx* ctx, pisurf_tree_info* tree);
/* runOurs' frame loop (src/main.cpp:330-405: one SurfReconstructor per frame in DATA_PATHES)
 * in one call. Frame k has n[k] particles at xyz[k] (host, or device with on_device != 0) and
 * radius radius[k]; the other parameters are the context's. After each frame `on_mesh` receives
 * the mesh (vertices x,y,z as double, 1-based triangles, valid only during the call); `per_frame`
 * (nullable, nframes entries) receives each frame's statistics. Nothing carries over between
 * frames, and every mesh is the one a fresh context gives.
 * frames_in_flight = 1: the frames run one after another on this context; its device buffers are
 * kept and grow to the largest frame, and frame k+1's particles are copied to the device (pinned
 * staging, a second stream) while frame k is meshed.
 * frames_in_flight = K > 1 (0 takes 2): K frames are meshed at once, frame k on lane k mod K;
 * lane 0 is this context and the others are contexts (each with its own stream and buffers) that
 * this context creates on first use and keeps. One lane's kernels fill the time another spends
 * waiting for a count or launching small kernels. Callbacks come one at a time and in frame order,
 * from the lanes' threads (lane 0's is the caller's); while one runs, other lanes, this context
 * among them, may still be meshing, so a callback must not call the library on this context.
 * When a frame fails, the frames before it have been delivered, none after it is, and the error
 * names the frame.
 * Afterwards pisurf_get_stats reports the last frame meshed on this context. */
typedef void (*pisurf_mesh_cb)(void* user, int32_t frame, const pisurf_tree_info* tree, const double* verts,
                               int64_t n_verts, const int32_t* tris, int64_t n_tris);
int pisurf_run_frames(pisurf_ctx* ctx, int32_t nframes, const float* const* xyz, const int64_t* n,
                      const float* radius, int32_t on_device, pisurf_mesh_cb on_mesh, void* user,
                      pisurf_stats* per_frame);
/* The same frames meshed together: every frame gets a context of its own on this context's
 * stream, and the frames go through each step together, so a step costs one host synchronisation
 * for the whole batch. The octree levels of all frames advance in lockstep, and at each level one
 * launch evaluates the distinct sample points and leaf-depth centres of every frame, its
 * workgroups taking (frame, run of cells) items from a device work queue. Arguments and callback
 * as pisurf_run_frames. No counterpart in the reference (its frames run one after another).
 * Currently disabled: returns PISURF_E_UNSUPPORTED after checking its arguments (its first
 * device run stopped making progress, cause not yet found); use pisurf_run_frames. */
int pisurf_run_batch(pisurf_ctx* ctx, int32_t nframes, const float* const* xyz, const int64_t* n,
                     const float* radius, int32_t on_device, pisurf_mesh_cb on_mesh, void* user,
                     pisurf_stats* per_frame);
int pisurf_get_stats(pisurf_ctx* ctx, pisurf_stats* out);
int pisurf_reset_stats(pisurf_ctx* ctx);
/* device stream the library launches on (hipStream_t) */
void* pisurf_stream(pisurf_ctx* ctx);
/* Shape autotuning of the run launches (node samples, leaf-depth centres, particle normals): the
 * first launch of a (precision, launch kind, points-per-cell bucket) times the candidate
 * (threads per workgroup, cells per run) shapes on its own data and keeps the fastest for the
 * process; every shape gives the same results. enabled = 0 takes the heuristic shape instead
 * (also PISURF_RUN_TUNING=0 in the environment); clear = 1 forgets the kept shapes. Returns the
 * number of kept shapes. No counterpart in the reference (it has no device launches). */
int pisurf_run_tuning(int32_t enabled, int32_t clear);

#ifdef __cplusplus
}
#endif
#endif
