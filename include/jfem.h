/*
 * jfem.h -- C ABI of libjfem_hip.so, the MI355X (gfx950) elasticity library.
 *
 * The library is the device side of JuliaFEM's linear-elasticity path.  Each
 * entry point replaces one piece of the reference (paths relative to the
 * reference repository root); the Julia and Python bindings live in the
 * package and are shown in INTEGRATION.md.
 *
 * Conventions
 *   - Plain C, no exceptions cross the boundary.  Every int-returning call
 *     returns JFM_OK (0) or a negative JFM_E* code; jfm_last_error(ctx)
 *     then describes the failure (valid until the next call on that ctx).
 *   - Host buffers belong to the caller; the library copies in or out and
 *     keeps no host pointer after a call returns.  Calls that return host
 *     data synchronise the context's stream first.
 *   - Node coordinates X are 3 x n_nodes column-major (Julia's layout; the
 *     same bytes as an n_nodes x 3 row-major C array).  Connectivity is
 *     npe x n_el column-major int64 with index_base 1 (Julia) or 0.
 *   - DOF of (node a, component c) is 3a + c (0-based), i.e. the
 *     reference's gdofs = 3(node-1)+c (src/assembly/problems.jl:466-478).
 *   - One context = one device + one HIP stream; a context is not
 *     thread-safe, distinct contexts are independent.
 *   - Results are deterministic: no floating-point atomics anywhere; every
 *     sum has a fixed order (ascending element id for nodal sums, fixed
 *     reduction trees for dot products), so repeated calls give identical bits.
 *   - PCG iterations run in batches captured as HIP graphs; setting the
 *     environment variable JFEM_GRAPHS=0 launches the same kernels one by
 *     one instead (for tools that do not follow graph launches).
 */
#ifndef JFEM_H
#define JFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jfm_ctx jfm_ctx;

enum {
    JFM_OK = 0,
    JFM_EINVAL = -1,      /* bad argument */
    JFM_EHIP = -2,        /* HIP runtime error */
    JFM_ENOTSPD = -3,     /* non-positive pivot / inverted element */
    JFM_ENOCONV = -4,     /* iteration limit reached */
    JFM_ESTATE = -5,      /* call out of order (e.g. solve before mesh) */
    JFM_ENOMEM = -6,      /* device allocation failed */
    JFM_EUNSUPPORTED = -7 /* element type not supported */
};

/* K.u operator variants */
enum {
    JFM_OP_CSR = 0,      /* assembled K (3x3-block compressed CSR) */
    JFM_OP_EBE = 1,      /* stored element matrices, deterministic nodal gather */
    JFM_OP_MATFREE = 2,  /* element geometric factors, recompute, nodal gather */
    JFM_OP_TANGENT = 3   /* neo-Hookean tangent K_T(u) at the state set by the
                            nonlinear solver (matrix-free, nodal gather) */
};
enum { JFM_FP64 = 0, JFM_FP32 = 1 };

/* solve methods */
enum {
    JFM_SOLVE_PCG = 0,       /* FP64 preconditioned CG to rtol */
    JFM_SOLVE_MIXED_IR = 1   /* FP32 inner PCG inside FP64 iterative refinement */
};
/* preconditioners */
enum {
    JFM_PREC_BLOCK_JACOBI = 0, /* inverse 3x3 node diagonal blocks */
    JFM_PREC_LEARNED = 1,      /* network-predicted SPD blocks M = C (L L' + eps I) C' */
    JFM_PREC_USER = 2          /* blocks given with jfm_precond_set_blocks */
};

typedef struct {
    int32_t method;       /* JFM_SOLVE_* */
    int32_t op;           /* JFM_OP_* used for the (inner) iteration */
    int32_t op_outer;     /* JFM_OP_* used for the FP64 outer residual (mixed only) */
    int32_t precond;      /* JFM_PREC_* */
    double rtol;          /* target ||f - K u|| / ||f|| over free DOFs */
    double inner_rtol;    /* inner relative tolerance (mixed only) */
    int32_t maxit;        /* iteration cap (inner iterations summed for mixed) */
    int32_t max_outer;    /* refinement-step cap (mixed only) */
    int32_t check_every;  /* iterations per device batch between host checks */
    int32_t learned_probe;/* learned only: PCG iterations run with each preconditioner
                           * before the solve; the learned blocks are used only if
                           * they capture more energy b.x_k (else block Jacobi,
                           * stats.fallback = 1); 0 = no probe */
    int32_t inner_maxit;  /* cap on the iterations of one inner solve (mixed); 0 = none */
    double first_rtol;    /* mixed only: FP64 PCG to this relative residual first; 0 = none */
    int32_t coarse_nodes; /* > 0: add the two-level correction Q A_c^-1 Q' (rigid-body
                           * modes on node aggregates of about this many nodes,
                           * A_c = Q' K Q rebuilt per solve) to the block
                           * preconditioner; 0 = single level */
} jfm_solve_opts;

typedef struct {
    int32_t iterations;   /* PCG iterations (inner iterations summed for mixed, probe included) */
    int32_t outer;        /* refinement steps (mixed), 1 for plain PCG */
    int32_t converged;    /* 1 when the FP64 true residual met rtol */
    int32_t fallback;     /* 1 when the learned preconditioner was rejected */
    double relres;        /* final FP64 true relative residual */
    double solve_ms;      /* device time of the solve (HIP events) */
    double setup_ms;      /* preconditioner setup time inside the call */
} jfm_solve_stats;

typedef struct {
    int64_t n_nodes, n_el, n_dofs, nnz, nnzb, n_faces;
    int32_t npe, affine;
    int64_t device_bytes;  /* bytes currently allocated by the context */
    /* patch form of the matrix-free operator (0 until it is first used):
     * patches of 480 elements, their (patch, node) entries, the nodes shared
     * between patches and their partial sums; the bytes per connectivity
     * entry the patch kernel reads (2: offsets from each patch's lowest node
     * id, when every patch's ids span at most 65,536; else 4); the factor
     * rows per element its FP32 form reads (9: sqrt(mu w detJ / 5) invJ,
     * affine Tet10 with one Poisson's ratio for all elements; else 11) */
    int64_t n_patches, patch_nodes, shared_nodes, shared_slots, patch_conn_bytes, patch_fp32_factor_rows;
} jfm_info;

/* ---- context --------------------------------------------------------- */
int jfm_create(int device, jfm_ctx **out);
void jfm_destroy(jfm_ctx *ctx);
const char *jfm_last_error(const jfm_ctx *ctx);
int jfm_get_info(jfm_ctx *ctx, jfm_info *info);

/* ---- problem upload (ext/JuliaFEMCUDAExt.jl:86-217 initialize_gpu_data!) --
 * Replaces the host-side renumbering and CuArray uploads.  npe = 10 (Tet10),
 * 4 (Tet4) or 8 (Hex8).  Builds the node->element incidence in ascending
 * element id (ext/JuliaFEMCUDAExt.jl:179-193), the 3x3-block sparsity
 * pattern and the per-element geometric factors.  Fails with JFM_ENOTSPD if
 * any element has detJ <= 0 at a Gauss point (src/basis/math.jl:202 keeps the
 * signed determinant). */
int jfm_upload_mesh(jfm_ctx *ctx, int64_t n_nodes, const double *X, int64_t n_el, int32_t npe,
                    const int64_t *conn, const double *E, const double *nu, int32_t index_base);

/* Dirichlet DOFs and surface traction (src/problems_dirichlet.jl:60-89,
 * src/problems_elasticity.jl:454-502).  fixed/u_prescribed have n_dofs
 * entries (u_prescribed may be NULL = zero).  face_conn is
 * nodes_per_face x n_faces column-major (6 = Tri6, 3 = Tri3, 4 = Quad4),
 * traction is 3 x n_faces (force per unit area, constant on a face). */
int jfm_upload_bcs(jfm_ctx *ctx, const uint8_t *fixed, const double *u_prescribed,
                   int64_t n_faces, int32_t nodes_per_face, const int64_t *face_conn,
                   const double *traction, int32_t index_base);

/* ---- assembly (assemble!(problem, time), src/assembly/assembly.jl:31-70 ->
 * src/problems_elasticity.jl:65-81, 203-451; sparse(COO), src/sparse/sparse.jl:54)
 * CSR sparsity of K over all DOFs, identical to SparseArrays.sparse() of the
 * reference's COO triplets (duplicates summed, explicit zeros kept, columns
 * sorted).  K is structurally symmetric, so this is also its CSC pattern.
 * Pass rowptr/colind = NULL to query nnz only. */
int jfm_csr_pattern(jfm_ctx *ctx, int64_t *nnz_out, int64_t *rowptr, int64_t *colind, int32_t base);

/* Element stiffness at Gauss points and deterministic scatter into the CSR
 * values (summed in ascending element id), plus f = f_ext - K u for the
 * displacement u (u = NULL means zero).  vals (nnz) and f (n_dofs) may be
 * NULL to keep the results on the device only. */
int jfm_assemble(jfm_ctx *ctx, const double *u, double *vals, double *f);

/* The same for Elasticity(finite_strain = true) (src/problems_elasticity.jl:
 * 203-451 with props.finite_strain; St. Venant-Kirchhoff S = D E of the
 * Green-Lagrange strain, BL(F) :255-289): the material tangent Km(u) into
 * vals, with geometric != 0 the geometric stiffness Kg(u) (:378-404, the
 * reference's assembly.Kg) into vals_kg, and f = f_ext - f_int(u), f_int =
 * sum w BL' S (:406-409, 432-433).  Same pattern as jfm_csr_pattern.  The
 * context's CSR operator (JFM_OP_CSR) then holds K = Km + Kg (Km when
 * geometric = 0), the matrix the reference solves, until the next assembly.
 * Any output may be NULL. */
int jfm_assemble_tl(jfm_ctx *ctx, int32_t geometric, const double *u, double *vals, double *vals_kg, double *f);

/* ---- operator (compute_residual_gpu!, ext/JuliaFEMCUDAExt.jl:442-478) --
 * y = K x on the free rows, y = 0 on the rows of fixed DOFs; x is used as
 * given on every DOF.  Host pointers, FP64 in and out; precision selects the
 * arithmetic of the operator. */
int jfm_apply(jfm_ctx *ctx, int32_t op, int32_t precision, const double *x, double *y);

/* K.u that allocates nothing: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops with fixed memory).  x and y are DEVICE pointers (n_dofs doubles);
 * the work is enqueued on `stream` (a hipStream_t; NULL = the context's
 * stream) and the call returns without synchronising.  op is
 * JFM_OP_MATFREE (two-phase form) or JFM_OP_EBE (after the element matrices
 * exist, i.e. after one jfm_apply or jfm_assemble with it).  The element
 * results (3 npe n_el values) and, in FP32, the converted vectors live in
 * one 64 MiB device array per GPU held until the process ends: meshes that
 * need more (e.g. more than 279,620 Tet10 elements in FP64) get JFM_EINVAL.
 * Every call on a GPU shares that array, so calls that overlap on one GPU
 * race.  Same result as jfm_apply up to the order of the nodal sums. */
int jfm_apply_noalloc(jfm_ctx *ctx, int32_t op, int32_t precision, const double *x, double *y, void *stream);

/* Average device times (ms) over `reps` back-to-back applications on
 * device-resident vectors, timed with HIP events on the context's stream:
 * avg_ms[0] = whole K.u, avg_ms[1] = its first kernel (SpMV; the element
 * phase of EBE; the patch kernel of MATFREE), avg_ms[2] = the second (the
 * nodal gather of EBE; the shared-node kernel of MATFREE; 0 for CSR); each as
 * reps back-to-back launches between two events. */
int jfm_time_apply(jfm_ctx *ctx, int32_t op, int32_t precision, int32_t reps, double *avg_ms);

/* The coarse matrix of the two-level preconditioner, before inversion (no
 * reference counterpart; for checking the coarse level against Q' K Q): the
 * node aggregates for coarse_nodes and A_c = Q' K Q (op JFM_OP_MATFREE or
 * JFM_OP_CSR: the linear K; JFM_OP_TANGENT: the neo-Hookean tangent at the
 * stored state), with Q the orthonormalised rigid-body modes of each
 * aggregate (six columns per aggregate, in aggregate order; a dropped mode
 * has a zero row and column and a unit diagonal).  *nc receives the order;
 * Ac (nc x nc, row major) and node_agg (n_nodes) are written when not NULL.
 * The next solve rebuilds the coarse level. */
int jfm_coarse_matrix(jfm_ctx *ctx, int32_t op, int32_t coarse_nodes, int64_t *nc, double *Ac, int32_t *node_agg);

/* ---- solve (solve!(solver), src/solvers.jl:192-216, 273-344; and
 * solve_backend!, ext/JuliaFEMCUDAExt.jl:886-916) ---------------------
 * Solves K u = f on the free DOFs with u = u_prescribed on fixed DOFs.
 * f = NULL uses the assembled load vector.  u (n_dofs) receives the
 * displacement; la (n_dofs, may be NULL) receives the reactions
 * f - K u on fixed DOFs (the reference's Lagrange multipliers la[B]). */
int jfm_solve(jfm_ctx *ctx, const jfm_solve_opts *opts, const double *f, double *u, double *la,
              jfm_solve_stats *stats);

/* User 3x3 preconditioner blocks (n_nodes x 9, row-major, each SPD on the
 * free DOFs; rows/columns of fixed DOFs are zeroed by the library), used by
 * jfm_solve with precond = JFM_PREC_USER until the mesh or BCs change. */
int jfm_precond_set_blocks(jfm_ctx *ctx, const double *M);

/* Convergence history of the last jfm_solve: per refinement step (or PCG
 * restart) the cumulative iteration count and the FP64 true relative
 * residual after it.  Returns the number of entries (<= max_entries). */
int jfm_solve_history(jfm_ctx *ctx, int32_t *iterations, double *relres, int32_t max_entries);

/* ---- nonlinear path (run!(solver::Solver{Nonlinear}), src/solvers.jl:575-620;
 * solve_newton_krylov_gpu!, ext/JuliaFEMCUDAExt.jl:685-854) -----------------
 * Compressible neo-Hookean, total Lagrangian (src/materials/neo_hookean.jl:
 * 129-231; src/problems_elasticity.jl:255-289, 378-404), Lame parameters from
 * the uploaded E and nu.  Load is applied in equal steps of the uploaded
 * traction; each step runs Newton to ||lf f_ext - f_int(u)|| <= rtol ||lf f_ext||
 * on the free DOFs; each tangent solve is PCG (FP64, or FP32 inside FP64
 * refinement) to an Eisenstat-Walker tolerance. */
typedef struct {
    int32_t load_steps;   /* equal load increments (>= 1) */
    int32_t max_newton;   /* Newton iterations per load step */
    double rtol;          /* Newton tolerance on the relative residual */
    int32_t method;       /* JFM_SOLVE_PCG or JFM_SOLVE_MIXED_IR for the tangent solves */
    int32_t precond;      /* JFM_PREC_BLOCK_JACOBI or JFM_PREC_LEARNED (diagonal of K_T) */
    double eta_max;       /* first and largest linear tolerance of Eisenstat-Walker
                             choice 2 (0: 0.1) */
    int32_t maxit_linear; /* PCG iteration cap per tangent solve */
    int32_t learned_probe;/* as in jfm_solve_opts, at the first Newton iterate of each load
                             step; the step's later iterates keep the outcome (accepted
                             blocks are recomputed from each new tangent diagonal) */
    double inner_rtol;    /* JFM_SOLVE_MIXED_IR: FP32 inner tolerance (0: 1e-2) */
    int32_t inner_maxit;  /* JFM_SOLVE_MIXED_IR: FP32 iterations per inner solve (0: no cap) */
    int32_t check_every;  /* PCG iterations between host checks (0: 32) */
    int32_t coarse_nodes; /* as in jfm_solve_opts; A_c = Q' K_T Q at the first Newton
                             iterate of each load step */
} jfm_newton_opts;

typedef struct {
    int32_t newton_iterations; /* summed over load steps */
    int32_t linear_iterations; /* PCG iterations summed over all tangent solves */
    int32_t converged;         /* 1 when every load step met rtol */
    int32_t fallbacks;         /* tangent solves where the learned blocks were rejected */
    double relres;             /* final relative residual of the last load step */
    double solve_ms;           /* device time */
} jfm_newton_stats;

/* R = lf f_ext - f_int(u) on the free DOFs (0 on fixed ones) for the
 * displacement u (n_dofs); *rnorm = |R|.  R may be NULL. */
int jfm_nl_residual(jfm_ctx *ctx, const double *u, double load_factor, double *R, double *rnorm);

/* Load-stepped Newton from u = u_prescribed; u (n_dofs) receives the final
 * displacement. */
int jfm_solve_nonlinear(jfm_ctx *ctx, const jfm_newton_opts *opts, double *u, jfm_newton_stats *stats);

/* ---- postprocessing (postprocess!(problem::Problem{Elasticity}, time, Val{:strain/:stress}),
 * src/problems_elasticity.jl:520-594) ------------------------------------------
 * Small-strain strain and linear-elastic stress of displacement u (n_dofs) at
 * every Gauss point, ABAQUS order 11, 22, 33, 12, 23, 13 with tensor shear:
 * strain_gp / stress_gp hold n_el x n_gauss x 6 doubles (n_gauss = 4 Tet10,
 * 1 Tet4, 8 Hex8).  strain_nodal / stress_nodal (n_nodes x 6) receive the
 * least-squares nodal fits (consistent mass solve, as lsq_fit).  Any output
 * may be NULL. */
int jfm_postprocess(jfm_ctx *ctx, const double *u, double *strain_gp, double *stress_gp, double *strain_nodal,
                    double *stress_nodal);

/* ---- steady heat conduction (Problem{Heat}, src/problems_heat.jl:36-113) --
 * On the uploaded mesh (element type and geometry; E and nu unused):
 *     K T = f,  K = sum w detJ k dN'dN  +  sum_faces w detJ h N N',
 *               f = sum w detJ q N      +  sum_faces w detJ (g + h T_ext) N,
 * with T = T_fixed on nodes where fixed_node is set.  k, q (may be NULL):
 * per element; flux, h, T_ext (each may be NULL; h and T_ext together):
 * per boundary face (n_faces x nodes_per_face connectivity, index_base).
 * Block-Jacobi PCG in FP64 to rtol; T receives n_nodes temperatures.
 * Does not change the elasticity boundary conditions. */
int jfm_heat_solve(jfm_ctx *ctx, const double *k, const double *q, int64_t n_faces, int32_t nodes_per_face,
                   const int64_t *face_conn, const double *flux, const double *h, const double *T_ext,
                   const uint8_t *fixed_node, const double *T_fixed, int32_t index_base, double rtol, int32_t maxit,
                   double *T, jfm_solve_stats *stats);

/* ---- learned preconditioner (no reference counterpart) ---------------
 * Per-node MLP 24 -> 64 -> 64 -> 6 (GELU), BF16 weights, FP32 accumulation.
 * The weight blob is the packed layout described in jfem_learned.h of the
 * package sources; jfm_learned_blob_bytes() gives its size. */
size_t jfm_learned_blob_bytes(void);
int32_t jfm_learned_n_params(void);   /* trainable parameters, gradient order W1 b1 W2 b2 W3 b3 */
int jfm_learned_set_weights(jfm_ctx *ctx, const void *blob, size_t bytes);
int jfm_learned_get_weights(jfm_ctx *ctx, void *blob, size_t bytes);
/* The 24 node features (n_nodes x 24, FP32) for the uploaded mesh and BCs. */
int jfm_learned_features(jfm_ctx *ctx, float *feat);
/* Network outputs (n_nodes x 6) and preconditioner blocks (n_nodes x 9,
 * row-major 3x3; may be NULL) for the current weights. */
int jfm_learned_outputs(jfm_ctx *ctx, float *out, double *M);
/* Training step pieces: mean over n_pairs (residual r, error e) pairs
 * (each n_dofs, row p at r + p*n_dofs, zero on fixed DOFs, e = K^-1 r) of
 *     loss = 1 - (z.r)^2 / ((z.Kz)(e.r)),  z = M r (+ Q A_c^-1 Q' r),
 * and its gradient with respect to the parameters (jfm_learned_n_params()
 * floats), computed by the network's forward and backward kernels.  With
 * coarse_nodes > 0 the blocks are trained as the fine level of the two-level
 * preconditioner of jfm_solve_opts.coarse_nodes. */
int jfm_learned_grad(jfm_ctx *ctx, int32_t n_pairs, const double *r, const double *e, int32_t coarse_nodes,
                     double *loss, float *grad);

#ifdef __cplusplus
}
#endif
#endif /* JFEM_H */
