/*
 * bemfmm.h — C ABI of the MI355X-native FMM-accelerated GMRES BVP engine.
 *
 * Drop-in for the matvec/solve path of the reference's BEM_BVP (repository
 * tomoakihirakawa/BEM_for_Nonlinear_Waves): the caller flattens its Network
 * (nodes, boundary triangles, per-(point,face) DOFs) into plain arrays, and the
 * library performs what BEM_BVP::solveGMRES does after that point: coordinate
 * scaling, sources, octree, interaction lists, near-field quadrature, the
 * rigid-mode diagonal, and FMM-accelerated GMRES.
 *
 * Model of the existing device C ABI in the reference: bem/metal_m2l/metal_m2l.h:23-129
 * and bem/metal_nearfield/metal_nearfield.h:33-141 (opaque context, int return
 * codes, host-owned inputs copied at call time).
 *
 * Conventions
 *   - Return codes: 0 = OK, negative = error (see BEMFMM_E_*); bemfmm_last_error()
 *     returns the message of the last failing call on a context.
 *   - All pointers without a _device suffix are host memory owned by the caller and
 *     are only read or written during the call.
 *   - A context is used by one host thread at a time; several contexts may share a
 *     device, each with its own HIP stream.
 *   - DOF kinds: 0 = Dirichlet (unknown phi_n, known phi), 1 = Neumann (unknown phi,
 *     known phi_n), 2 = CORNER Neumann-face DOF (unknown phi; its row is the
 *     constraint phi = phi_D, with phi_D passed in phi[]).
 *     (isDirichletID_BEM / isNeumannID_BEM, bem/BEM_setBoundaryTypes.hpp:29-40)
 *     3 = ROBIN (unknown phi, phi_n = kappa phi + g with kappa from bemfmm_set_robin and g
 *     passed as its phi_n; the frequency-domain free surface of bem/main_freq_domain.cpp:411-426,
 *     543-555, where g = 0 and kappa = omega^2 / g_gravity).
 */
#ifndef BEMFMM_H
#define BEMFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BEMFMM_OK 0
#define BEMFMM_E_INVALID (-1)      /* invalid argument or geometry */
#define BEMFMM_E_DEVICE (-2)       /* no usable GPU / HIP runtime error */
#define BEMFMM_E_OOM (-3)          /* device allocation failed */
#define BEMFMM_E_NONFINITE (-4)    /* non-finite input or result */
#define BEMFMM_E_NOT_CONVERGED (-5) /* GMRES stopped above tolerance */
#define BEMFMM_E_STATE (-6)        /* call out of order (e.g. solve before prepare) */

#define BEMFMM_PRECISION_FP64 0     /* every stage in FP64 */
#define BEMFMM_PRECISION_FP32_REFINE 1 /* FP32 near field and M2L inside FP64 iterative refinement */

#define BEMFMM_PRECOND_NONE 0      /* reference default (main_time_domain.cpp:38) */
#define BEMFMM_PRECOND_JACOBI 1    /* diagonal scaling (BEM_solveBVP_GMRES_FMM.hpp:2095-2104) */
#define BEMFMM_PRECOND_SCHWARZ 2   /* overlapping-block Schwarz on the near field (SCHWARZ, GMRES_FMM.hpp:899-1013;
                                      SchwarzAdditive_CRS, basic_linear_systems.hpp:2870-3003), restricted combination */
#define BEMFMM_PRECOND_ILU0 3      /* ILU(0) (ILU, ILU0_CRS, basic_linear_systems.hpp:2337-2606) of the reference's sparse
                                      matrix (buildSparseMatrixForILU, GMRES_FMM.hpp:215-662; neighbourhood from
                                      ilu_neighborhood / ilu_kring); factorised on the host, sweeps on the GPU by level */
#define BEMFMM_PRECOND_MILU0 4     /* MILU(0): ILU(0) with the dropped fill-in times milu_omega moved to the diagonal
                                      (MILU, ILU0_CRS(modified = true), basic_linear_systems.hpp:2379-2520) */
#define BEMFMM_PRECOND_ILUT 5      /* ILUT (ILUT_CRS, basic_linear_systems.hpp:2610-2807) on the same sparse matrix */

typedef struct bemfmm_ctx bemfmm_ctx;

typedef struct bemfmm_opts {
  int p;              /* expansion order N (reference: Buckets<...,10>, BEM_solveBVP.hpp:563) */
  double theta;       /* mac_theta, default 0.25 (main_time_domain.cpp:63) */
  int leaf_max;       /* fmm_bucket_max_points, default 50 (main_time_domain.cpp:54) */
  int max_level;      /* fmm_max_level, default 7 (main_time_domain.cpp:55) */
  int q;              /* p2m_quadrature_points, default 6 (main_time_domain.cpp:62) */
  int precision;      /* BEMFMM_PRECISION_* */
  int preconditioner; /* BEMFMM_PRECOND_* */
  double refine_eta;  /* FP32_REFINE: each inner FP32 solve reduces the FP64 residual by this factor (default 3e-7,
                         above the ~1.5e-7 per step that the FP32 operator reaches) */
  int schwarz_max_core;  /* schwarz_max_core_size, default 64 (main_time_domain.cpp:47) */
  int schwarz_max_block; /* schwarz_max_block_size, default 128, at most 128 (main_time_domain.cpp:48) */
  int schwarz_overlap;   /* schwarz_overlap_k: rings of neighbours added to a core, default 1 (main_time_domain.cpp:46) */
  int m2l_tile_pairs;    /* M2L box pairs per matrix-core tile: 16, 32 or 64; 0 = chosen by timing 16 and 32 on the
                            plan during prepare (16 for plans under 20,000 pairs) */
  int dof_reorder;       /* 1 (default): number DOFs internally in Morton order of their points, as solveGMRES
                            does (BEM_solveBVP_GMRES_FMM.hpp:3164-3246); 0: keep the caller's numbering.
                            Every array of this ABI stays in the caller's numbering either way. */
  int use_graphs;        /* 0 (default): launch the matvec's kernels one by one; 1: replay it from a HIP graph captured
                            at the first matvec. With graphs, the ROCm 7 runtime kept the buffers a context's graphs
                            had used allocated after a re-prepare or teardown freed them (measured on the MI355X:
                            4.1 GB per re-prepare at 80k nodes), and the eager launches were as fast at 209k nodes
                            (2.549 against 2.569 ms per FP32 matvec) */
  int m2l_kernel;        /* M2L of the FP32 operator (FP32_REFINE): 0 (default) = rotation where compiled for p
                            (8, 10), else matrix cores; 1 = dense matrix-core tiles; 2 = rotation (error if
                            not compiled for p); 3 = plane waves (lib_plane_wave_m2l.hpp:293-490, rotations to all 6
                            directions): per offset the chain rotate / multipole-to-exponential / diagonal shift /
                            exponential-to-local / rotate back composed into one operator, applied by the dense
                            matrix-core tiles. The FP64 operator uses the dense SimpleM2L tiles, except with 4 =
                            plane waves for both operators. */
  int overlap_near;      /* 1 (default): the near-field columns of each matvec run on a second stream beside the
                            FMM passes (they need only the input vector); 0: after L2L, one GEMV */
  int refine_recycle;    /* FP32_REFINE: 1 (default): each inner FP32 solve after the first minimises over the
                            earlier steps' Krylov directions too (they share the operator M^-1 A_32) and extends
                            that basis, up to `restart` + 1 vectors; 0: every step restarts GMRES from its residual */
  int ilu_neighborhood;  /* BEMFMM_PRECOND_ILU0 rows: 0 (default) = BUCKETS (faces sharing the point's cell of a
                            uniform grid of clamp(cbrt(faces), 8, 64) cells per axis, plus the point's faces),
                            1 = K-RING (the point's faces and ilu_kring rings of edge neighbours);
                            ilu_neighborhood_type, main_time_domain.cpp:39 */
  int ilu_kring;         /* rings for K-RING, default 1 (ilu_kring_num, main_time_domain.cpp:40) */
  int schwarz_variant;   /* BEMFMM_PRECOND_SCHWARZ: 0 (default) = blocks grown on the mesh DOF graph, entries from
                            the FMM near field, restricted combination (each DOF from its own core's block);
                            1 = the reference's SCHWARZ: blocks on the graph of the ILU(0) sparse matrix (cores by
                            breadth-first search to depth schwarz_core_k, GMRES_FMM.hpp:901-1010), entries from that
                            matrix, combination averaged over the covering blocks (SchwarzAdditive_CRS::solve,
                            basic_linear_systems.hpp:2960-3003) */
  int schwarz_core_k;    /* schwarz_variant 1: core search depth, default 1 (main_time_domain.cpp:45) */
  double milu_omega;     /* MILU(0) diagonal compensation, default 1 (main_time_domain.cpp:41) */
  double ilut_drop_tol;  /* ILUT: entries below drop_tol * |row of A|_1 are dropped, default 1e-3 (:42) */
  int ilut_max_entries;  /* ILUT: largest entries kept per L row and per U row, default 50 (:43) */
  double ilut_pivot_min; /* ILUT: smaller diagonals are set to +-ilut_pivot_min, default 1e-12 (:44) */
  int plane_wave_digits; /* m2l_kernel 3 / 4: quadrature of the exponential expansions, 10 (default; 1751
                            exponentials, operators within ~1e-11 of SimpleM2L) or 6 (558, ~1e-7); the published
                            rules of Greengard & Rokhlin 1997 (lib_plane_wave_quadrature.hpp) */
  int tree_on_gpu;       /* 1 (default): the octree, the near / M2L lists and the target boxes are built on the GPU
                            (level-synchronous radix sorts on path keys); 0: on the host. Both give the same plan
                            (generateTree, lib_spatial_partitioning.hpp:1079-1115; moments.hpp:968-997) */
  int element_order;     /* 1 (default): linear triangles; 2: true quadratic faces (isTrueQuadraticElement,
                            BEM_solveBVP_GMRES_FMM.hpp:1311-1657): bemfmm_set_geometry then takes six nodes and six
                            DOF entries per face (vertices p0 p1 p2, then the midpoint nodes of edges p0p1, p1p2, p2p0;
                            a midpoint DOF of -1 has its weight redistributed to the edge's end vertices), the
                            midpoint nodes are targets with their own BIE rows (:111-144); preconditioners NONE,
                            JACOBI or SCHWARZ (schwarz_variant 0) */
  int near_rows_per_wave; /* near-field GEMV launch shape: rows each wavefront streams at once, 2, 4 or 8; 0
                             (default): chosen by timing the three at the first prepare of a problem class (FP64 and
                             FP32 operators separately; no counterpart in the reference) */
  int p2m_loads;          /* P2M launch shape: column loads in flight per thread, 4, 8 or 16; 0 (default): timed
                             like near_rows_per_wave */
  int inner_p;            /* FP32_REFINE with the rotation M2L: expansion order of the FP32 operator's M2L, below p
                             (5, 6 or 8); 0 (default): p. The inner solves then run on the expansions truncated at
                             inner_p (P2M, M2M, L2L and L2P keep p, whose leading terms are the inner_p ones); the
                             FP64 residual of every refinement step keeps p, so the result is that of the p operator
                             (the reference switches its single-precision M2L to double near convergence instead,
                             BEM_solveBVP_GMRES_FMM.hpp:2306-2319) */
} bemfmm_opts;

typedef struct bemfmm_result {
  int iters;              /* GMRES iterations (matvecs inside Arnoldi) */
  int n_matvec;           /* all matvecs including residual checks */
  int converged;          /* 1 if |M^-1 (b - A x)|_2 <= tol */
  double residual;        /* last true residual (absolute, scaled system) */
  double time_setup;      /* seconds: plan + near-field quadrature + alpha */
  double time_solve;      /* seconds: RHS + GMRES */
  int refinement_steps;   /* FP32_REFINE: outer FP64 refinement steps */
} bemfmm_result;

typedef struct bemfmm_stats {
  int n_nodes, n_faces, n_dof;
  int n_levels, n_boxes, n_leaves;
  int64_t n_m2l_pairs;
  int64_t near_nnz;       /* sum over targets of near DOF columns */
  int n_m2l_ops;          /* distinct M2L offsets */
  double L;               /* coordinate scale (bounding-box diagonal) */
  double device_bytes;    /* resident device memory of the plan */
  /* seconds per setup stage: geometry + octree + target boxes, interaction lists, near-field structure, P2M
     blocks, translation groups and operators, uploads + P2M matrices, near-field quadrature, rigid-mode matvec */
  double t_tree, t_lists, t_near_struct, t_p2m, t_ops, t_upload, t_near_quad, t_alpha;
  int n_precond_blocks;         /* Schwarz: number of blocks */
  int precond_fallback_blocks;  /* Schwarz: blocks whose elimination failed (diagonal used instead) */
  double precond_avg_block;     /* Schwarz: mean block size */
  int n_ops_uploaded;           /* translation operators uploaded by the last prepare (the rest were resident) */
  int m2l_tile_pairs;           /* M2L pairs per tile in use (autotuned unless set in the options) */
  int m2l_rotation;             /* 1 if the FP32 operator's M2L runs by rotation */
  int64_t precond_nnz;          /* ILU(0)/MILU(0)/ILUT (the factors) and the reference's Schwarz (the matrix): entries */
  int precond_levels;           /* ILU(0)/MILU(0)/ILUT: dependency levels of the forward plus the backward sweep */
  int m2l_plane_wave;           /* 1: the FP32 operator's M2L through plane waves; 2: both operators' */
  int tree_on_gpu;              /* 1 if the last prepare built the tree and lists on the GPU */
  double t_tune;                /* seconds: M2L tile autotuning (after t_alpha, the rigid-mode matvec) */
  double t_precond;             /* seconds: preconditioner build */
  double t_fp32;                /* seconds: FP32 copies of the operator (FP32_REFINE) and launch-shape tuning */
  int near_rows_per_wave[2];    /* near GEMV launch shape in use: FP64 operator, FP32 operator */
  int p2m_loads[2];             /* P2M launch shape in use: FP64 operator, FP32 operator */
  int inner_p;                  /* expansion order of the FP32 operator's M2L when below p, else 0 */
} bemfmm_stats;

/* Availability (metal_m2l.h: metal_m2l_is_available / metal_m2l_device_name). */
int bemfmm_is_available(void);
const char* bemfmm_device_name(void);

void bemfmm_default_opts(bemfmm_opts* opts);

/* Context lifecycle (replaces the BEM_BVP members B_poles/targets/preconditioner). */
bemfmm_ctx* bemfmm_ctx_create(int device, const bemfmm_opts* opts);
void bemfmm_ctx_destroy(bemfmm_ctx* ctx);
const char* bemfmm_last_error(const bemfmm_ctx* ctx);

/* Geometry and DOF model: replaces WATERS[0] (Network) + setNodeFaceIndices (f2Index).
 * X: n_nodes*3 coordinates (unscaled); tri: n_faces*3 node ids, outward (out of the fluid);
 * face_dof: n_faces*3 DOF id of each face corner (pf2Index(p, f)); dof_kind/dof_point: per DOF. */
int bemfmm_set_geometry(bemfmm_ctx* ctx, int n_nodes, const double* X, int n_faces, const int32_t* tri,
                        const int32_t* face_dof, int n_dof, const int32_t* dof_kind, const int32_t* dof_point);

/* Known boundary values per DOF (phiOnFace / phinOnFace): phi for Dirichlet DOFs (and phi_D for
 * CORNER Neumann DOFs), phi_n for Neumann and CORNER Neumann DOFs; other entries are ignored. */
int bemfmm_set_bc(bemfmm_ctx* ctx, const double* phi, const double* phin);

/* Robin coefficient kappa per DOF (n_dof, physical units, used on ROBIN DOFs only): replaces the
 * postprocess hook that sets phin = kappa * phi on the is_robin DOFs of every matvec
 * (main_freq_domain.cpp:543-555). kappa is part of the operator and of its preconditioner, so a
 * call on a prepared context requires bemfmm_prepare again (the operator table stays resident).
 * Without a call kappa is 0. */
int bemfmm_set_robin(bemfmm_ctx* ctx, const double* kappa);

/* Setup (solveGMRES up to GMRES, BEM_solveBVP_GMRES_FMM.hpp:3076-3450 + initializeFMM,
 * moments.hpp:1239-1541): scaling, sources, tree, lists, operators, near-field quadrature,
 * rigid-mode diagonal alpha, preconditioner. */
int bemfmm_prepare(bemfmm_ctx* ctx);

/* y = A x for the scaled system (return_A_dot_v without preconditioner,
 * BEM_solveBVP_GMRES_FMM.hpp:2124-2155 -> compute_Ax_minus_b :34-158). n_dof entries. */
int bemfmm_matvec(bemfmm_ctx* ctx, const double* x, double* y);
/* Same on device pointers (inputs resident in HBM); ordered on the context's stream. */
int bemfmm_matvec_device(bemfmm_ctx* ctx, const double* x_device, double* y_device);
/* y = A_32 x: the same operator with the near field, P2M and M2M/M2L/L2L products in FP32
 * (matrices rounded once from FP64; FP32 MFMA accumulation per translation), FP64 vectors and
 * FP64 sums across boxes. Counterpart of the reference's single-precision Metal M2L
 * (bem/metal_m2l/metal_m2l.h:23-129, used by BEM_solveBVP_GMRES_FMM.hpp:2246-2320).
 * Requires a context created with BEMFMM_PRECISION_FP32_REFINE (else BEMFMM_E_STATE). */
int bemfmm_matvec_fp32(bemfmm_ctx* ctx, const double* x, double* y);
int bemfmm_matvec_fp32_device(bemfmm_ctx* ctx, const double* x_device, double* y_device);
/* y = A x on device pointers, as bemfmm_matvec_device, without allocating anything: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc). The per-call work memory (phi / phi_n with the box locals, multipoles,
 * per-node sums) is taken from one array of 64 MiB per GPU that the library holds from load until the process
 * ends, shared by every call on that GPU and left holding the last call's values: calls that overlap on one GPU
 * race, so a caller waits for the context's stream (bemfmm_sync) before another context on the same GPU makes
 * such a call. The FP64 operator, launched kernel by kernel on the context's stream (no graph); its M2L
 * accumulates into the locals directly (no per-pair staging). Returns BEMFMM_E_INVALID if the problem needs
 * more than 64 MiB (about (2 n_dof + 4 C n_boxes + 2 n_nodes) doubles: a 200k-node tank at p = 10 needs 49). */
int bemfmm_matvec_device_noalloc(bemfmm_ctx* ctx, const double* x_device, double* y_device);
/* Synchronise the context's stream. */
int bemfmm_sync(bemfmm_ctx* ctx);

/* b = -(A x - b)|_{x=0} with CORNER rows = phi_D (solveSystemGMRES, BEM_solveBVP_GMRES_FMM.hpp:2025-2047). */
int bemfmm_rhs(bemfmm_ctx* ctx, double* b);
/* Rigid-mode diagonal coefficients per DOF (updateGeometricProperties, :2674-2716). */
int bemfmm_alpha(bemfmm_ctx* ctx, double* alpha);

/* Solve (solveSystemGMRES + run_gmres, :2007-2355). x0/x: n_dof, physical units (phi_n unscaled
 * on Dirichlet DOFs as stored by storePhiPhin, :3462-3489). Returns BEMFMM_E_NOT_CONVERGED if the
 * true residual stays above tol (x still holds the last iterate). */
int bemfmm_solve(bemfmm_ctx* ctx, const double* x0, double* x, double tol, int max_iter, int restart,
                 bemfmm_result* result);

int bemfmm_get_stats(const bemfmm_ctx* ctx, bemfmm_stats* stats);

/* GMRES residual estimate |g_n| per iteration of the last solve (the reference prints it per
 * inner iteration, BEM_solveBVP_GMRES_FMM.hpp:2297-2300). Copies up to cap values; returns the count. */
int bemfmm_residual_history(const bemfmm_ctx* ctx, double* out, int cap);

/* z = M^-1 r with the context's preconditioner (none: copy; Jacobi: diagonal scaling; Schwarz: restricted
 * additive Schwarz over overlapping DOF blocks) — the reference's preconditioner solve
 * (ilu->solve / schwarz_preconditioner->solve, BEM_solveBVP_GMRES_FMM.hpp:2124-2171). n_dof entries. */
int bemfmm_precond_apply(bemfmm_ctx* ctx, const double* r, double* z);

/* Internal DOF order of a prepared context: order[i] = caller DOF at internal position i (n_dof entries).
 * With dof_reorder this is the Morton order (21-bit keys in the point bounding box, stable radix sort on
 * the GPU, equal keys kept in caller order); a caller may renumber its f2Index by it as the reference does. */
int bemfmm_dof_order(const bemfmm_ctx* ctx, int32_t* order);

/* Schwarz blocks of a prepared context: DOFs idx[ptr[b] .. ptr[b+1]) of block b, the first ncore[b] of them
 * its core (every DOF lies in exactly one core). Any pointer may be NULL. Returns the number of blocks
 * (0 without the Schwarz preconditioner); ptr needs n_blocks+1 entries, idx ptr[n_blocks]. */
int bemfmm_schwarz_blocks(const bemfmm_ctx* ctx, int64_t* ptr, int32_t* idx, int32_t* ncore);

/* 64-bit digest of a prepared context's tree stage: level offsets, parents, leaf flags, box centres, the
 * depth-first face order, box face ranges, target boxes and the near / M2L lists. No counterpart in the
 * reference: equal digests mean equal trees and lists (host and GPU builds, two prepares of one mesh). */
int bemfmm_tree_digest(const bemfmm_ctx* ctx, uint64_t* digest);

/* FP64 residual |M^-1 (b - A x)|_2 at the start of each refinement step of the last FP32_REFINE solve
 * (the last entry is the final residual). The reference's mixed-precision switch
 * (BEM_solveBVP_GMRES_FMM.hpp:2246-2320) prints the true residual at its one switch point.
 * Copies up to cap values; returns the count. */
int bemfmm_refinement_history(const bemfmm_ctx* ctx, double* out, int cap);

/* Kernel timing of the last matvec (milliseconds, HIP events on the context's stream):
 * [0] scatter, [1] P2M, [2] M2M, [3] M2L translations, [4] M2L per-box reduction, [5] L2L,
 * [6] near+L2P, [7] assemble. Enables event recording on the first call (and then returns zeros);
 * while it is on, matvecs launch their kernels one by one instead of replaying the graph.
 * ms8 = NULL turns it off. */
int bemfmm_matvec_profile(bemfmm_ctx* ctx, double* ms8);

#ifdef __cplusplus
}
#endif

#endif /* BEMFMM_H */
