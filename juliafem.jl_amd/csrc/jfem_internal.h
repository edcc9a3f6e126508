// jfem_internal.h -- shared declarations of libjfem_hip.so (not installed).
//
// Data layout in HBM (n = 3N DOFs, DOF 3a+c):
//   X        N x 3 f64            node coordinates
//   conn     npe x n_el i32 (SoA) conn[i*n_el + e] = node of local i
//   gf       GF x n_el (SoA)      per element (affine) or per Gauss point:
//                                 invJ (9, row-major), lambda*w*detJ, mu*w*detJ
//   inc      node -> (e*npe + i)  incidences in ascending element id
//   brow/bcol                     3x3-block pattern (node adjacency), rows sorted
//   vals     9*nnzb               CSR values in the scalar CSR order: the block
//                                 row a occupies vals[9*brow[a] ...]; scalar row
//                                 3a+c starts at 9*brow[a] + c*3*nb, entry
//                                 (3a+c, 3*bcol[s]+d) at + 3*(s-brow[a]) + d
//   Ke       n_el x NPAIR x 9     element matrix blocks (i <= j), row-major 3x3
//   ye       (n_el*npe) x 3       element results of the two-phase operators,
//                                 in node-sorted incidence order (slot k of
//                                 inc), so each node's terms are contiguous
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/jfem.h"

namespace jfem {

// ------------------------------------------------------------------ traits --
// Reference elements, Gauss rules and dN/dxi at the Gauss points, as
// constexpr tables so that fully unrolled loops fold them into immediates.
//   Tet10: src/basis/lagrange_generated.jl:267-282, 4-pt rule with the
//          15-digit literals of src/integration/gauss_points.jl:173-182
//   Tet4:  :229-256, 1-pt rule (gauss_points.jl:166-170)
//   Hex8:  :285-310, 2x2x2 rule (gauss_points.jl:207-219)
template <int NPE> struct Elem;

template <> struct Elem<10> {
    static constexpr int NPE = 10, NGP = 4, NPAIR = 55;
    __host__ __device__ static constexpr double w(int) { return 1.0 / 24.0; }
    __host__ __device__ static constexpr double xi(int q, int c) {
        return (q == 3) ? 0.138196601125011 : ((q == c) ? 0.585410196624968 : 0.138196601125011);
    }
    __host__ __device__ static constexpr double dN(int q, int i, int c) {
        const double u = xi(q, 0), v = xi(q, 1), s = xi(q, 2);
        switch (i) {
        case 0: return -3.0 + 4.0 * u + 4.0 * v + 4.0 * s;
        case 1: return c == 0 ? -1.0 + 4.0 * u : 0.0;
        case 2: return c == 1 ? -1.0 + 4.0 * v : 0.0;
        case 3: return c == 2 ? -1.0 + 4.0 * s : 0.0;
        case 4: return c == 0 ? 4.0 - 8.0 * u - 4.0 * v - 4.0 * s : -4.0 * u;
        case 5: return c == 0 ? 4.0 * v : (c == 1 ? 4.0 * u : 0.0);
        case 6: return c == 1 ? 4.0 - 8.0 * v - 4.0 * u - 4.0 * s : -4.0 * v;
        case 7: return c == 2 ? 4.0 - 8.0 * s - 4.0 * u - 4.0 * v : -4.0 * s;
        case 8: return c == 0 ? 4.0 * s : (c == 2 ? 4.0 * u : 0.0);
        default: return c == 1 ? 4.0 * s : (c == 2 ? 4.0 * v : 0.0);
        }
    }
    __host__ __device__ static constexpr double N(int q, int i) {
        const double u = xi(q, 0), v = xi(q, 1), s = xi(q, 2), l = 1.0 - u - v - s;
        switch (i) {
        case 0: return l * (2.0 * l - 1.0);
        case 1: return u * (2.0 * u - 1.0);
        case 2: return v * (2.0 * v - 1.0);
        case 3: return s * (2.0 * s - 1.0);
        case 4: return 4.0 * u * l;
        case 5: return 4.0 * u * v;
        case 6: return 4.0 * v * l;
        case 7: return 4.0 * s * l;
        case 8: return 4.0 * u * s;
        default: return 4.0 * v * s;
        }
    }
};

template <> struct Elem<4> {
    static constexpr int NPE = 4, NGP = 1, NPAIR = 10;
    __host__ __device__ static constexpr double w(int) { return 1.0 / 6.0; }
    __host__ __device__ static constexpr double dN(int, int i, int c) {
        return i == 0 ? -1.0 : ((i - 1) == c ? 1.0 : 0.0);
    }
    __host__ __device__ static constexpr double N(int, int) { return 0.25; }  // at the centroid
};

template <> struct Elem<8> {
    static constexpr int NPE = 8, NGP = 8, NPAIR = 36;
    // 1/sqrt(3) to double precision
    static constexpr double A = 0.57735026918962576451;
    __host__ __device__ static constexpr double w(int) { return 1.0; }
    __host__ __device__ static constexpr double ref(int i, int c) {
        // corners (-1,-1,-1),(1,-1,-1),(1,1,-1),(-1,1,-1),(-1,-1,1),(1,-1,1),(1,1,1),(-1,1,1)
        return c == 0 ? ((i == 1 || i == 2 || i == 5 || i == 6) ? 1.0 : -1.0)
             : c == 1 ? ((i == 2 || i == 3 || i == 6 || i == 7) ? 1.0 : -1.0)
                      : (i >= 4 ? 1.0 : -1.0);
    }
    // Gauss point q: x fastest (gauss_points.jl:207-219 ordering)
    __host__ __device__ static constexpr double xi(int q, int c) {
        return ((q >> c) & 1) ? A : -A;
    }
    __host__ __device__ static constexpr double dN(int q, int i, int c) {
        const double a = 1.0 + ref(i, 0) * xi(q, 0), b = 1.0 + ref(i, 1) * xi(q, 1),
                     d = 1.0 + ref(i, 2) * xi(q, 2);
        return c == 0 ? 0.125 * ref(i, 0) * b * d : (c == 1 ? 0.125 * a * ref(i, 1) * d : 0.125 * a * b * ref(i, 2));
    }
    __host__ __device__ static constexpr double N(int q, int i) {
        return 0.125 * (1.0 + ref(i, 0) * xi(q, 0)) * (1.0 + ref(i, 1) * xi(q, 1)) * (1.0 + ref(i, 2) * xi(q, 2));
    }
};

constexpr int GF = 11;  // invJ(9), lambda*w*detJ, mu*w*detJ

// --------------------------------------------------------- PCG scalars ----
// Iteration-parity slots: kernels of iteration i read slot i&1 and the last
// kernel of the iteration writes slot (i+1)&1, so no kernel reads a value that
// another block of the same kernel may be writing.
struct PcgScalars {
    double rz[2];       // r.z entering the iteration of that parity
    double alpha, beta, pq, rr;   // last values (diagnostics)
    double bb;          // b.b over free DOFs
    double tol2;        // rtol^2 * bb
    double rtol2;       // set by the host
    double tmp0, tmp1;  // results of one-off reductions
    double rr0;         // r.r at the PCG start
    int done[2];        // 0 running, DONE_* otherwise, by parity
    int iter, maxit;
};
// done codes
enum { DONE_RUNNING = 0, DONE_CONVERGED = 1, DONE_BREAKDOWN = 2, DONE_MAXIT = 3 };
// what a one-block finalize kernel does with the summed partials
enum { FIN_NONE = 0, FIN_INIT = 3, FIN_DOT = 4, FIN_DOT2 = 5, FIN_BB = 6 };
// partials layout (doubles): [PQ slot0 | PQ slot1 | RZ slot0 (2) | RZ slot1 (2) | misc (2)]
constexpr int kMaxPart = 1024;      // producer grids are capped at this many blocks
inline constexpr int part_pq(int slot) { return slot * kMaxPart; }
inline constexpr int part_rz(int slot) { return 2 * kMaxPart + slot * 2 * kMaxPart; }
constexpr int kPartMisc = 6 * kMaxPart;
constexpr int kPartCoarse = 8 * kMaxPart;  // w.zc partials of the coarse product
constexpr int kPartTotal = 9 * kMaxPart;

// --------------------------------------------------------------- context --
struct Workspace {        // PCG vectors of one precision
    void *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *b = nullptr;
    void *Minv = nullptr; // N x 9 preconditioner blocks
    bool ok = false;
};

}  // namespace jfem

struct jfm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t bytes = 0;
    std::vector<void *> allocs;

    // mesh
    int64_t N = 0, n_el = 0, ndof = 0;
    int npe = 0, ngp = 0, npair = 0;
    bool affine = false, has_mesh = false;
    bool nu_uniform = false;      // every element has Poisson's ratio nu0
    double nu0 = 0.0;
    std::vector<int32_t> h_conn;  // n_el x npe (element-major), 0-based
    double *X = nullptr;
    int32_t *conn = nullptr;      // SoA
    double *E = nullptr, *nu = nullptr;
    double *gf64 = nullptr;
    float *gf32 = nullptr;
    int ngf = 0;                  // factor sets per element (1 if affine else ngp)
    int32_t *inc_ptr = nullptr, *inc = nullptr;
    int32_t *inc_slot = nullptr;  // SoA npe x n_el: slot in inc of incidence (e, i)

    // pattern and assembly
    int64_t nnzb = 0;
    std::vector<int32_t> h_brow, h_bcol;
    int32_t *brow = nullptr, *bcol = nullptr, *bdiag = nullptr, *blk_row = nullptr;
    int32_t *cptr = nullptr, *contrib = nullptr;
    double *vals64 = nullptr;
    float *vals32 = nullptr;
    double *Ke = nullptr;      // EBE element matrices: entry q = 9 p + m at [((q / 2) n_el + e) 2 + q % 2]
    float *Ke32 = nullptr;     // the same in FP32, chunks of 4: [((q / 4) n_el + e) 4 + q % 4]
    // non-null when Ke and Ke32 are stored in the patch order (position k
    // holds element ke_perm[k]; ke_perm is the patch structure's eperm)
    const int32_t *ke_perm = nullptr;
    bool assembled = false;

    // boundary conditions and loads
    bool has_bcs = false;
    uint8_t *fixed = nullptr;
    double *g = nullptr;          // prescribed values (n)
    bool g_nonzero = false;
    int64_t n_faces = 0;
    int nfn = 0;
    int32_t *face_conn = nullptr;  // SoA nfn x n_faces
    double *face_t = nullptr;      // n_faces x 3
    int32_t *finc_ptr = nullptr, *finc = nullptr;
    double *fe = nullptr;          // n_faces x nfn x 3
    double *f_ext = nullptr;
    double *load = nullptr;        // f = f_ext - K u from the last assembly
    bool has_load = false;

    // operator scratch
    double *ye64 = nullptr;
    float *ye32 = nullptr;
    double *vx64 = nullptr, *vy64 = nullptr;  // apply/time buffers
    float *vx32 = nullptr, *vy32 = nullptr;

    // solver
    jfem::Workspace ws64, ws32;
    double *u64 = nullptr;         // outer solution (mixed IR)
    double *rr64 = nullptr;        // outer residual (mixed IR)
    jfem::PcgScalars *S = nullptr;
    jfem::PcgScalars *S_host = nullptr;  // pinned
    double *partials = nullptr;          // kPartTotal doubles
    unsigned *counter = nullptr;         // small integer scratch (element checks)
    // captured PCG iteration batches, keyed by (op, precision, preconditioner)
    struct Graph { int op = -1, prec = -1, batch = 0, coarse = 0; const void *minv = nullptr; hipGraphExec_t exec = nullptr; };
    std::vector<Graph> graphs;
    std::vector<int32_t> hist_it;          // convergence history of the last solve
    std::vector<double> hist_res;

    // learned preconditioner
    void *mlp_w = nullptr;         // weight blob on the device
    std::vector<unsigned char> mlp_host;
    bool has_mlp = false;
    bool learned_ready = false;    // Mlearn matches the current mesh/BCs/weights
    bool mlp_dev = false;          // mlp_w holds mlp_host
    // the learned blocks formed on stream2 beside other work (jfem_learned_launch):
    // pending until jfem_learned_prepare joins them (ev_l0/ev_l1 bracket the work)
    bool learned_pending = false;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_l0 = nullptr, ev_l1 = nullptr;
    double *Mlearn64 = nullptr;    // N x 9 learned SPD blocks
    double *xprobe = nullptr;      // the learned-selection probe's winning iterate
    float *Mlearn32 = nullptr;
    float *feat = nullptr;         // N x 24 node features
    double *efeat = nullptr;       // 3 x n_el per-element feature inputs
    float *lout = nullptr;         // N x 6 network outputs
    // nonlinear state (tangent operator)
    double *ustate64 = nullptr;    // displacement the tangent is taken at
    float *ustate32 = nullptr;
    double *Dtan = nullptr;        // N x 9 tangent diagonal blocks (features)
    bool has_state = false;
    bool minv_tangent = false;     // ws*.Minv hold the tangent block Jacobi
    double *nl_u = nullptr, *nl_r = nullptr, *nl_du = nullptr, *nl_zero = nullptr;
    // two-level coarse space (coarse.hip)
    std::vector<double> h_X;       // host copy of the coordinates (aggregation)
    int coarse_built_nodes = 0;    // coarse_nodes the aggregation was built for
    int coarse_gen = 0;            // bumped per aggregation (keys captured graphs)
    bool coarse_ready = false;     // Ahat holds a coarse level
    int coarse_tag = -1;           // operator it was built from: 0 K, 1 the tangent
    int64_t n_agg = 0, nc = 0;
    int ncolor = 0;
    int32_t *node_agg = nullptr, *agg_ptr = nullptr, *agg_nodes = nullptr, *agg_color = nullptr;
    int32_t *partner = nullptr;    // ncolor x n_agg: the coupled aggregate of each colour
    int32_t *dropped = nullptr;    // nc: mode dropped (dependent on its aggregate)
    int32_t *ecol = nullptr, *ncol = nullptr;        // per colour: probed elements, nodes
    hipGraphExec_t coarse_graph[2] = {nullptr, nullptr};  // probe sequence for K, K_T
    int coarse_graph_gen[2] = {0, 0};
    hipGraphExec_t coarse_inv_graph = nullptr;  // the inverse and fold sequence
    int coarse_inv_gen = 0;
    std::vector<int32_t> h_ecol_ptr, h_ncol_ptr, h_cseg;
    int32_t *segb = nullptr, *sega = nullptr;        // colour node lists by aggregate
    int32_t *csub_ptr = nullptr, *csub = nullptr;    // colour node lists: their slots in the colour's elements
    double *cd64 = nullptr;        // N x 3 offsets from the aggregate centroid
    float *cd32 = nullptr;
    double *Rinv = nullptr;        // n_agg x 36
    double *Ac = nullptr;          // nc x nc (inverted through Ac2, a second copy)
    double *Ac2 = nullptr;
    int32_t *apos = nullptr;       // node -> its position in agg_nodes
    double *rA = nullptr;          // 3 x N: the masked residual in aggregate order (FP64 or FP32 view)
    double *cdA64 = nullptr;       // cd in aggregate order (agg_nodes), for the restriction
    float *cdA32 = nullptr;
    int64_t ldc = 0;               // row stride of Ahat32: nc rounded up to 256, zero padded
    float *Ahat32 = nullptr;       // nc x ldc: R^-1 A_c^-1 R^-T
    float *wc32 = nullptr, *zc32 = nullptr;  // wc32 has ldc entries, zero past nc
    double *wc64 = nullptr;        // 2 x nc: Q' m r by iteration parity (slot 0 at the PCG start)
    double *gjP = nullptr;         // Gauss-Jordan pivot inverses (two, by step parity)
    // A_c element by element (linear operator): listed elements, their
    // aggregates and entry positions, items, pair ranges, unordered pairs,
    // and the entries (36 doubles each)
    int32_t *gal_el = nullptr, *gal_gptr = nullptr, *gal_glist = nullptr, *gal_gq = nullptr, *gal_gpos = nullptr;
    int32_t *gal_item = nullptr, *gal_pptr = nullptr, *gal_upair = nullptr;
    double *gal_C = nullptr;
    int gal_nitems = 0, gal_nupair = 0;
    // patch operator (patch.hip): elements in Morton order, 480 per patch
    bool patch_ready = false;
    int npatch = 0, npn = 0, nbnode = 0, nslot = 0;
    int32_t *eperm = nullptr;                 // patch order -> element id
    int32_t *connP = nullptr;                 // npe x n_el, patch order
    uint16_t *connP16 = nullptr;              // the same as offsets from pbase[patch] (null: ids span too far)
    int32_t *pbase = nullptr;                 // lowest node id of every patch
    double *gfP64 = nullptr;                  // geometric factors, patch order
    float *gfP32 = nullptr;
    float *gfQ32 = nullptr;                   // affine Tet10, uniform nu: sqrt(cm) invJ (9 rows), patch order
    float rho32 = 0.0f;                       // lambda / mu for gfQ32
    int32_t *pn_ptr = nullptr, *pn_out = nullptr;  // patch nodes: range per patch, target
    uint16_t *pn_cnt = nullptr;               // patch node: incidences in the patch
    int32_t *rstart = nullptr, *rbase = nullptr;  // rounds of 256 patch nodes
    uint16_t *linT = nullptr;                 // patch-local incidences, per round transposed
    int32_t *bptr = nullptr, *bnode = nullptr;  // shared nodes, their partial slots
    double *bpart = nullptr;                  // 3 x slots (FP32 or FP64 view)
    double *ppart = nullptr;                  // per-patch partials of y.dotv
    int bmaxcnt = 0;                          // most slots of one shared node
    int32_t *qsrc = nullptr;                  // N, aggregate order: -1 or (first slot << 5 | slots)
    int qsrc_gen = -1;                        // coarse_gen qsrc was built for
    double *pv6 = nullptr, *qv6 = nullptr;  // probe vectors, 6 x 3N (zero outside the probed rows)
    double *ye6 = nullptr;                  // probe incidence slots, 6 x 3 npe n_el (zero outside the probed elements)
    // heat conduction (heat.hip): temperature in component 0
    double *heat_k = nullptr;      // n_el conductivities
    uint8_t *heat_fixed = nullptr; // 3N row mask of the heat system
    double *heat_g = nullptr;      // 3N prescribed temperatures (component 0)
    double *heat_Minv = nullptr;   // N x 9 Jacobi blocks
    int64_t hf_n = 0;              // boundary faces with flux / convection
    int hf_nfn = 0;
    int32_t *hf_conn = nullptr, *hf_ptr = nullptr, *hf_inc = nullptr;
    double *hf_h = nullptr, *hf_g = nullptr, *hf_text = nullptr, *hf_fe = nullptr;
    double *Muser64 = nullptr;     // N x 9 user blocks (JFM_PREC_USER)
    float *Muser32 = nullptr;
    bool has_user = false;
};

// api.hip helpers used by learned.hip
int jfem_ensure_workspaces(jfm_ctx *ctx);                                   // vectors + block Jacobi
int jfem_apply(jfm_ctx *ctx, int op, const double *x, double *y);          // masked FP64 K.x (device)
int jfem_dot(jfm_ctx *ctx, const double *a, const double *b, int64_t n, double *out);
// noalloc.hip: looks up the scratch array's device address (once per device)
int jfem_noalloc_init(jfm_ctx *ctx);

// -------------------------------------------------------------- helpers ----
int jfm_fail(jfm_ctx *ctx, int code, const std::string &msg);
int jfm_hip_fail(jfm_ctx *ctx, hipError_t e, const char *what);
#define JFM_HIP(ctx, call)                                   \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return jfm_hip_fail(ctx, e_, #call); \
    } while (0)

template <typename T> T *jfm_alloc(jfm_ctx *ctx, int64_t count);
void jfm_free(jfm_ctx *ctx, void *p);

// ------------------------------------------------------- kernel launchers --
// fem_kernels.hip
namespace jfem {
hipError_t launch_geom_factors(jfm_ctx *ctx, int *d_bad);
hipError_t launch_element_matrices(jfm_ctx *ctx);
hipError_t launch_assemble_values(jfm_ctx *ctx);
// finite-strain element matrices (which: 0 Km, 1 Kg, 2 Km + Kg), element-major
// into Ke and, when ye != null, f_int(u) into the incidence slots
hipError_t launch_tl_element_matrices(jfm_ctx *ctx, const double *u, int which, double *Ke, double *ye);
// Values per element of the EBE matrices: 9 NPAIR rounded up to a multiple
// of the chunk width V (2 doubles or 4 floats).
inline int64_t ke_stride(int npair, int V) { return (9 * (int64_t)npair + V - 1) / V * V; }
// CSR values from element-major packed blocks Ke[(e * NPAIR + p) * 9 + m]
hipError_t launch_assemble_values_ke(jfm_ctx *ctx, const double *Ke);
hipError_t launch_face_loads(jfm_ctx *ctx);
hipError_t launch_block_jacobi(jfm_ctx *ctx, double *Minv64, float *Minv32);
// Two-phase operators; `part` (may be null) receives one partial of
// dot(y, dotv) per block of the gather; returns the partial count in *npart.
template <typename T>
hipError_t launch_elem_only(jfm_ctx *ctx, int op, const T *x, const int *done);
template <typename T>
hipError_t launch_elem_into(jfm_ctx *ctx, int op, const T *x, T *ye);
template <typename T>
hipError_t launch_gather(jfm_ctx *ctx, const T *ye, const uint8_t *fixed, T *y, const T *dotv, double *part,
                         const int *done, int *npart);
hipError_t launch_face_gather(jfm_ctx *ctx);
template <typename T>
hipError_t launch_two_phase(jfm_ctx *ctx, int op, const T *x, T *y, const T *dotv, double *part,
                            const int *done, int *npart);
// nonlinear.hip (neo-Hookean): element phase of f_int(u) (tangent = false)
// or K_T(u) v (tangent = true) into the incidence slots
template <typename T>
hipError_t launch_nh_elem(jfm_ctx *ctx, bool tangent, const T *gf, const T *u, const T *v, T *ye, const int *done,
                          const int32_t *elist = nullptr, int n_items = 0, int nv = 1, int64_t vs = 0, int64_t ys = 0);
// FP64 masked K x (or K_T x) over an element subset and a node subset only,
// for nv vectors at once (x_j = x + j xs, y_j = y + j xs, slots ye + j ys)
hipError_t launch_subset_apply(jfm_ctx *ctx, bool tangent, const double *x, double *y, double *ye, int nv, int64_t xs,
                               int64_t ys, const int32_t *elist, int ne, const int32_t *nlist, int nn);
hipError_t launch_nh_block_jacobi(jfm_ctx *ctx, const double *u, double *M64, float *M32, double *D);
// patch.hip: fused K.u over element patches (JFM_OP_MATFREE, and JFM_OP_EBE
// when the element matrices are stored in the patch order)
bool jfem_patch_enabled();  // false when JFEM_PATCH=0 (two-phase operator)
int jfem_patch_setup(jfm_ctx *ctx);
void jfem_patch_release(jfm_ctx *ctx);
// whether `op` runs in the patch form on this context
inline bool jfem_patch_form(const jfm_ctx *ctx, int op);
template <typename T>
hipError_t launch_patch_phase(jfm_ctx *ctx, int op, const T *x, T *y, const T *dotv, const int *done,
                              double *pout = nullptr);
template <typename T>
hipError_t launch_patch_combine(jfm_ctx *ctx, T *y, const T *dotv, double *part, const int *done, int *npart);
template <typename T>
hipError_t launch_patch_apply(jfm_ctx *ctx, int op, const T *x, T *y, const T *dotv, double *part, const int *done,
                              int *npart);
// The two-level PCG may leave the shared nodes' sums to its update kernel,
// which visits the nodes in aggregate order: qsrc[k] is -1 for the k-th node
// (agg_nodes[k]) when the patch kernel writes it to y, else its first
// partial slot << 5 | its slot count.  jfem_patch_qsrc builds it for the
// current aggregation (a no-op when the form is not available);
// jfem_patch_defer_shared tells whether the PCG's K.p of `op` can use it
// (JFEM_DEFER_SHARED=0 turns it off).
int jfem_patch_qsrc(jfm_ctx *ctx);
bool jfem_patch_defer_shared(const jfm_ctx *ctx, int op);
// coarse.hip
// With learned_beside, the learned blocks start on the second stream
// (jfem_learned_launch) once A_c is formed, beside its inversion.
int jfem_coarse_setup(jfm_ctx *ctx, int op, int coarse_nodes, double *ms, bool learned_beside = false);
int jfem_coarse_matrix(jfm_ctx *ctx, int op, int coarse_nodes, int64_t *nc, double *Ac, int32_t *node_agg);
bool jfem_use_graphs();  // false when JFEM_GRAPHS=0 (profilers)
void jfem_coarse_release(jfm_ctx *ctx);
template <typename T>
hipError_t launch_coarse_correct(jfm_ctx *ctx, const T *r, T *z, T *p, double *part, const int *done, int *npart);
// the coarse product of a PCG iteration: zc = Ahat Q' m r, with the w.zc
// partials (npart_wz of them) at partials + kPartCoarse
template <typename T>
hipError_t launch_coarse_product(jfm_ctx *ctx, const T *r, const int *done, int *npart_wz);
// the same from m o r in aggregate order, as the update kernel writes it to
// coarse_rA<T>(ctx)
template <typename T>
hipError_t launch_coarse_product_agg(jfm_ctx *ctx, const T *rA, const int *done, int *npart_wz);
template <typename T> T *coarse_rA(jfm_ctx *ctx);
// The iteration's coarse right-hand side w = Q' m r of the new residual,
// formed inside the update kernel, one block per aggregate
// (launch_pcg_update(..., restrict_q = true)), in FP64 at wc64 + (slot ^ 1)
// nc and in FP32 at wc32, where the GEMV reads it.  Used when
// jfem_coarse_fused_restriction(ctx) (JFEM_FUSED_RESTRICTION=0 turns it off).
bool jfem_coarse_fused_restriction(const jfm_ctx *ctx);
hipError_t launch_coarse_gemv_fused(jfm_ctx *ctx, int slot, int *npart_wz);
// p = z + m o Q zc + beta p, beta from the (r.z, r.r) and w.zc partials
template <typename T>
hipError_t launch_pcg_direction_coarse(jfm_ctx *ctx, int slot, int npart_rz, int npart_wz, T *p, const T *z, T *x);
// heat.hip (kOpHeat: conduction + convection, FP64, row mask ctx->fixed)
constexpr int kOpHeat = 17;
hipError_t launch_heat_apply(jfm_ctx *ctx, const double *x, double *y, const double *dotv, double *part, const int *done,
                             int *npart);
hipError_t launch_heat_load(jfm_ctx *ctx, const double *src, double *f);
hipError_t launch_heat_jacobi(jfm_ctx *ctx, double *Minv);
// postprocess.hip (kOpMass: consistent scalar mass, three components, no row mask)
constexpr int kOpMass = 16;
hipError_t launch_mass_elem(jfm_ctx *ctx, const double *x, const int *done);
hipError_t launch_post_gp(jfm_ctx *ctx, const double *u, double *strain, double *stress);
hipError_t launch_post_rhs(jfm_ctx *ctx, const double *u, int field, int comp0, double *b);
hipError_t launch_mass_diag(jfm_ctx *ctx, double *Minv);
// linalg_kernels.hip
template <typename T>
hipError_t launch_bsr_spmv(jfm_ctx *ctx, const T *vals, const T *x, T *y, const T *dotv, double *part,
                           const int *done, int *npart);
template <typename T>
hipError_t launch_pcg_init(jfm_ctx *ctx, T *r, T *z, T *p, const T *Minv, int *npart);
// With shared_q (and restrict_q), q holds only what the patch kernel writes
// and the update sums the shared nodes' partial slots itself (qsrc).
template <typename T>
hipError_t launch_pcg_update(jfm_ctx *ctx, int slot, int npart_pq, T *x, T *r, T *z, const T *p, const T *q,
                             const T *Minv, int *npart, T *rA = nullptr, bool restrict_q = false,
                             bool shared_q = false);
template <typename T>
hipError_t launch_pcg_direction(jfm_ctx *ctx, int slot, int npart_rz, T *p, const T *z);
hipError_t launch_finalize(jfm_ctx *ctx, const double *part, int npart, int nv, int fin);
hipError_t launch_zero_u32(jfm_ctx *ctx, unsigned *p, int64_t n);
template <typename TI, typename TO>
hipError_t launch_convert(jfm_ctx *ctx, const TI *in, TO *out, int64_t n, double scale);
// r = b - Ku (fixed rows 0) and one partial of r.r per block -> partials misc
hipError_t launch_residual(jfm_ctx *ctx, const double *b, const double *Ku, const uint8_t *fixed, double *r,
                           int64_t n, int *npart);
hipError_t launch_axpy_scaled(jfm_ctx *ctx, double *u, const float *d, double scale, int64_t n);
hipError_t launch_add(jfm_ctx *ctx, const double *a, const double *b, double *out, int64_t n);  // out = a + b
template <typename T>
hipError_t launch_dot(jfm_ctx *ctx, const T *a, const T *b, int64_t n, int *npart);
inline bool jfem_patch_form(const jfm_ctx *ctx, int op) {
    return ctx->patch_ready && (op == JFM_OP_MATFREE || (op == JFM_OP_EBE && ctx->ke_perm));
}

}  // namespace jfem
