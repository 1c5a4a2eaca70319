/*
 * CPU oracle for the ParticleIsosurface field-evaluation / vertex-placement path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker and the timed CPU
 * baseline; the product (libpisurf) never links, loads or calls it.
 *
 * It is a clean-room restatement of the reference's CPU algorithm (HashGrid, Evaluator,
 * SurfReconstructor, TNode, VisitorExtract, Mesh), keeping the reference's data
 * structures (std::map cell tables, per-query candidate vectors, OpenMP per-particle /
 * per-node loops) so its cost profile matches. Two arithmetic modes:
 *   fp64 = 0  FP32-faithful: float storage and the reference's float->double promotion points
 *   fp64 = 1  the same algorithm in double (binning from the float positions, so
 *             candidate sets are identical to mode 0)
 * All value outputs are returned as double (float results are exact in double).
 */
#ifndef PISURF_ORACLE_H
#define PISURF_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

typedef struct {
    float radius;          /* particle radius r */
    float smooth_factor;   /* SurfReconstructor smooth factor (box margin r*smooth), 2.0 */
    float iso_factor;      /* 1.9, used only in poly6 mode */
    int use_ani;           /* 1 */
    int use_poly6;         /* 0 = cubic B-spline (shipped default) */
    int calc_normals;      /* 1 */
    int oversample;        /* 2 -> 27 samples per node */
    int fp64;              /* 0 FP32-faithful, 1 FP64 */
    int qef_mode;          /* 0 off (shipped), 1 reference 4x4 inverse, 2 truncated-SVD pseudo-inverse */
    float qef_svd_tol;     /* relative singular-value cutoff for qef_mode 2 */
    int depth_max_override;/* 0 = derived from the bounding box */
    int threads;           /* OpenMP threads, 0 = all */
} orc_params;

orc_ctx* orc_create(const float* xyz, int64_t n, const orc_params* p);
void orc_destroy(orc_ctx* c);

/* H1-H3 hash grid. bbox may be NULL: the tight particle bbox (loadRootBox) is used. */
int orc_hashgrid_build(orc_ctx* c, const float* bbox, float inf_factor, float* out_bounding6,
                       uint32_t* out_cellnum3, int64_t* out_cellnum, float* out_cellsize,
                       int64_t* out_n_invalid);
int orc_hashgrid_index_list(orc_ctx* c, int32_t* index_list);
/* non-empty cells: key, StartList, EndList (count returned; arrays may be NULL to query) */
int64_t orc_hashgrid_cells(orc_ctx* c, int64_t* keys, int32_t* start, int32_t* end);
/* H4: 27-cell candidates (reference order) of nq points; returns total (offsets nq+1) */
int64_t orc_neighbors(orc_ctx* c, const float* q, int64_t nq, int64_t* offsets, int32_t* ids, int64_t cap);
/* H5: box query */
int64_t orc_box_particles(orc_ctx* c, const float* box1, const float* box2, int32_t* ids, int64_t cap);

/* H8-H9 */
int orc_anisotropy(orc_ctx* c);
int orc_get_aniso(orc_ctx* c, double* xmean, double* G9_colmajor, double* det);
int orc_set_aniso(orc_ctx* c, const double* xmean, const double* G9_colmajor, const double* det);
/* H6 */
int orc_iso_constants(orc_ctx* c, double* max_scalar, double* iso);
int orc_get_constants(orc_ctx* c, double* influnce2, double* sigma);
/* H12 */
int orc_particle_normals(orc_ctx* c, double* normals);
/* H11 (+ H12 form when grad != NULL) */
int orc_eval_points(orc_ctx* c, const float* q, int64_t n, double* f, double* grad);
/* H13: 27 samples of one node box */
int orc_grid_eval(orc_ctx* c, const float* center, float half, double* samples, double* grads,
                  int* signchange);
/* H14: classification of a batch of nodes (one level) */
int orc_classify_nodes(orc_ctx* c, const float* centers, const float* halfs, const int32_t* depths,
                       int64_t n, uint8_t* type, double* f, float* curv, int32_t* signchange);
/* H15: whole octree */
int orc_build_tree(orc_ctx* c);
int orc_tree_info(orc_ctx* c, int* depth_max, int* depth_min, float* root_center3, float* root_half,
                  int64_t* n_nodes);
/* node table in creation (level) order: center, half, depth, type, id, node xyzf, first child */
int orc_tree_nodes(orc_ctx* c, float* center3, float* half, int32_t* depth, uint8_t* type,
                   uint64_t* id, double* node4, int32_t* child);
/* H16 for one node with given samples/grads (27x4, 27x3) */
int orc_qef(orc_ctx* c, const float* center, float half, const double* samples, const double* grads,
            int mode, double* node4);
/* H17: ends n x 2 x 4 (x,y,z,f) -> vertex xyz */
int orc_bisect_edges(orc_ctx* c, const double* ends, int64_t n, double* out_xyz, int32_t* n_evals);
/* Extraction + mesh (dual cells, bisection, mesh assembly) */
int orc_extract(orc_ctx* c);
int orc_mesh_counts(orc_ctx* c, int64_t* n_verts, int64_t* n_tris, int64_t* n_dual);
int orc_mesh_get(orc_ctx* c, double* verts, int32_t* tris);
int orc_dual_cells(orc_ctx* c, int32_t* nodes8, uint8_t* sign_index);
int orc_write_obj(orc_ctx* c, const char* path);
/* Whole frame: grid, anisotropy, constants, normals, tree, extraction (times in seconds) */
int orc_run(orc_ctx* c, double* phase_times7);
int64_t orc_eval_count(orc_ctx* c);
int orc_num_threads(void);
/* the oracle's symmetric eigensolvers, for pinning against a library eigensolver:
 * a = row-major n x n (n = 3 or 4), lam[n], Q[n*n] row-major (eigenvectors in columns) */
int orc_test_eig(const double* a, int n, int fp64, double* lam, double* Q);
/* compute_G_Yus tail on a given covariance (6 upper entries): G (9, column-major), det */
int orc_test_g_from_cov(const double* cov6, int fp64, double* G9, double* det);

#ifdef __cplusplus
}
#endif
#endif
