/*
 * cpm_gpu.h -- C ABI of the MI355X closest-point-method engine (libcpmgpu.so).
 *
 * The reference has no FFI: its boundary for the solve is the C++ class
 * cpm::Solver (src/Solver.h:19-53) and, for the per-tube setup, the classes
 * Tube (src/Tube.h:26-205), Interpolation (src/Interpolation.h:16-43) and
 * FDMatrices (src/FDMatrices.h).  Each entry point below names the reference
 * interface it replaces.  All functions are extern "C", take plain pointers
 * and sizes, never throw, and return CPM_OK (0) or a negative status; the
 * message of the last failure on the calling thread is cpm_last_error().
 *
 * Host buffers are owned by the caller and copied in/out (no retention).
 * Device memory belongs to the handles.  One HIP stream per context; a
 * context is used by one host thread at a time.
 */
#ifndef CPM_GPU_H
#define CPM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum cpm_status {
    CPM_OK = 0,
    CPM_ERR_ARG = -1,       /* invalid argument / inconsistent sizes        */
    CPM_ERR_HIP = -2,       /* HIP runtime failure                           */
    CPM_ERR_NOMEM = -3,     /* device allocation failed or budget exceeded   */
    CPM_ERR_GEOMETRY = -4,  /* tube construction failed (stencil off tube)   */
    CPM_ERR_STATE = -5,     /* handle used in the wrong state                */
    CPM_BUSY = 1            /* cpm_server_submit: not admitted, no room yet  */
};

/* Preconditioner, src/Defines.h:13-18 */
enum cpm_precond { CPM_PRECOND_IDENTITY = 0, CPM_PRECOND_DIAGONAL = 1, CPM_PRECOND_JACOBI = 2 };

/* Precision of the Krylov iterations. */
enum cpm_precision {
    CPM_FP64 = 0,          /* reference arithmetic: FP64 BiCGStab               */
    CPM_FP32_REFINE = 1    /* FP64 iterative refinement around a continuous FP32 BiCGStab      */
};

/* Per-solve termination status (cpm_solve_info.status). */
enum cpm_solve_status {
    CPM_SOLVE_CONVERGED = 1,
    CPM_SOLVE_MAXIT = 2,
    CPM_SOLVE_ZERO_RHS = 4,
    CPM_SOLVE_BREAKDOWN = 5,  /* NaN/Inf in a recurrence scalar */
    CPM_SOLVE_NOT_WRITTEN = 6 /* serving: writing out the request's results failed (the cpm_server_run
                                 call that retired it returned an error); its d_x are not valid */
};

typedef struct cpm_ctx cpm_ctx;
typedef struct cpm_server cpm_server;
typedef struct cpm_op cpm_op;
typedef struct cpm_tube cpm_tube;

typedef struct cpm_solve_opts {
    double tol;            /* relative residual ||b-Ax|| <= tol ||b||; reference 1e-10 (src/Solver.cpp:110) */
    int max_iters;         /* reference 100000 (src/Solver.cpp:110)                                         */
    int precision;         /* enum cpm_precision                                                            */
    double update_delta;   /* FP32_REFINE: FP32 residual drop that triggers an FP64 refinement step (<=0: 1e-2) */
    int reserved;          /* must be 0                                                                     */
} cpm_solve_opts;

typedef struct cpm_solve_info {
    int status;            /* enum cpm_solve_status                              */
    int iters;             /* BiCGStab iterations                                */
    int restarts;          /* rho-breakdown restarts (src/BiCGStab.h:62-71)       */
    int updates;           /* FP64 refinement steps (FP32_REFINE)                 */
    double rel_residual;   /* ||b-Ax||/||b||: FP64 recurrence residual (FP64), true FP64 residual (FP32_REFINE) */
    int fallback_iters;    /* FP32_REFINE: FP64 iterations of the re-solve from zero that a system gets when
                              its refinement steps stop reducing the true residual (0 = none).  Once an
                              operator's re-solves make FP64 the cheaper choice for it (its fallback
                              fraction and iteration counts so far), later FP32_REFINE solves of it run
                              in FP64 from the start: iters = 0 and fallback_iters = the FP64 iterations
                              (environment CPMGPU_FP64_ROUTING=0: FP32 is tried every time)               */
} cpm_solve_info;

/* Defaults of the reference: tol 1e-10, max 100000, FP64. */
void cpm_solve_opts_default(cpm_solve_opts* opts);

const char* cpm_last_error(void);

/* ---- context ---------------------------------------------------------- */
int cpm_ctx_create(int device, cpm_ctx** out);
int cpm_ctx_destroy(cpm_ctx* ctx);
int cpm_ctx_synchronize(cpm_ctx* ctx);
/* Bytes of device memory held by all handles of this context. */
int64_t cpm_ctx_device_bytes(const cpm_ctx* ctx);

/* ---- operator from host CSR -------------------------------------------
 * Replaces cpm::Solver::Init(m, n, L, E, precond, identity_rows, nD)
 * (src/Solver.h:22-23, src/Solver.cpp:8-68 -> GeneralizedBiCGStabSystem::
 * SetComponents, src/GeneralizedBiCGStabSystem.h:37-53).  L and E are D x D
 * row-major CSR with int32 indices -- exactly Eigen::SparseMatrix<double,
 * RowMajor>'s outerIndexPtr()/innerIndexPtr()/valuePtr() -- and columns
 * ascending inside each row.  identity_rows has D bytes (0/1).  The operator is
 *   y = m x + n (Ldiag x + (L - Ldiag) E x),  identity rows y = x,
 *   last nD rows y = (m - n) x
 * (GeneralizedBiCGStabSystem::Multiply, src/GeneralizedBiCGStabSystem.h:191-211). */
int cpm_op_create_csr(cpm_ctx* ctx, int64_t D, int nD,
                      const int32_t* L_rowptr, const int32_t* L_col, const double* L_val,
                      const int32_t* E_rowptr, const int32_t* E_col, const double* E_val,
                      const uint8_t* identity_rows, double m, double n,
                      int precond, int precond_iters, cpm_op** out);
int cpm_op_destroy(cpm_op* op);
int64_t cpm_op_rows(const cpm_op* op);
/* Execution format of an operator: 0 general ELL, 1 tensor-product stencils
 * (tube-built operators; an internal row order, invisible to callers).  No
 * counterpart in the reference, which stores E as a general sparse matrix
 * (src/Interpolation.cpp:204-225). */
int cpm_op_format(const cpm_op* op);
/* LDS tiles of a tensor-product operator's E gather, into out[6]:
 * shared-stencil tiles (0: none), their items (one per thread), their items
 * of plain rows sharing a stencil block, rows in their lists; row tiles with
 * a list, rows in the row tiles' lists.  No counterpart in the reference. */
int cpm_op_tile_stats(const cpm_op* op, int64_t* out);
/* Which E gather single-operator launches of `precision` (CPM_FP64 or
 * CPM_FP32_REFINE) with nrhs (1, 2, 4, 8) interleaved right-hand sides use:
 * 0 direct reads of x (kernel k_egather), 1 row tiles (k_egather),
 * 2 shared-stencil tiles (k_egather_grp); -1 bad arguments.  No counterpart
 * in the reference. */
int cpm_op_gather_kind(const cpm_op* op, int precision, int nrhs);
/* Time stepping re-uses the operator with new (m, n) (problems/HeatWithIBC.cpp:166-167). */
int cpm_op_set_mn(cpm_op* op, double m, double n);
/* Replaces cpm::Solver::SetPrecond(precond_iters) (src/Solver.h:25). */
int cpm_op_set_precond(cpm_op* op, int precond, int precond_iters);

/* y = A(m, n) x with host buffers of D doubles.
 * Replaces cpm::Solver::Multiply (src/Solver.cpp:122-138). */
int cpm_multiply(cpm_op* op, const double* x, double* y, double m, double n);
/* Same with device pointers (D doubles each), on the context stream. */
int cpm_multiply_device(cpm_op* op, const double* d_x, double* d_y, double m, double n);

/* Solve A x = rhs from x0 = 0 (host buffers).  Replaces cpm::Solver::Solve
 * (src/Solver.cpp:69-114 -> BiCGStab::Solve, src/BiCGStab.h:30-108). */
int cpm_solve(cpm_op* op, const double* rhs, double* x, const cpm_solve_opts* opts, cpm_solve_info* info);
/* Same with device pointers; no host transfer of vectors. */
int cpm_solve_device(cpm_op* op, const double* d_rhs, double* d_x, const cpm_solve_opts* opts,
                     cpm_solve_info* info);

/* Solve without device allocation.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same as
 * cpm_solve_device (one right-hand side, device pointers, FP64 or
 * FP32_REFINE with its FP64 re-solve) with every workspace -- Krylov vectors,
 * segment and pack tables, states' partial sums -- carved from a fixed device
 * array of CPM_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it
 * runs on until the process ends; nothing is allocated or freed.  An operator
 * needs about 88 B per row (up to ~760,000 rows); larger ones are rejected
 * with CPM_ERR_ARG before anything is launched.  Shared-stencil tiles are
 * used only if an earlier solve of the operator built them (results are the
 * same either way).  Every call on one GPU shares that GPU's array and leaves
 * its contents behind: calls that overlap on one GPU (other threads, other
 * contexts) race, and the caller must serialise them. */
#define CPM_NOALLOC_SCRATCH_BYTES (64ll << 20)
int cpm_solve_device_noalloc(cpm_op* op, const double* d_rhs, double* d_x, const cpm_solve_opts* opts,
                             cpm_solve_info* info);

/* Batched solve: n_inst independent systems, instance i solving
 * ops[i] x_i = rhs_i, all advanced together (one launch per kernel covers every
 * instance; converged instances are masked).  rhs/x are host pointer arrays,
 * each of cpm_op_rows(ops[i]) doubles.  Instances that share an operator are
 * grouped into multi-RHS sweeps that read the operator once per iteration.
 * The reference solves these one after another (e.g.
 * examples/DiffusionCurvesCodimZero/DiffusionCurvesCodimZero.cpp:49-52). */
int cpm_solve_batch(cpm_ctx* ctx, int n_inst, cpm_op* const* ops, const double* const* rhs,
                    double* const* x, const cpm_solve_opts* opts, cpm_solve_info* infos);
int cpm_solve_batch_device(cpm_ctx* ctx, int n_inst, cpm_op* const* ops, const double* const* d_rhs,
                           double* const* d_x, const cpm_solve_opts* opts, cpm_solve_info* infos);
/* Streamed solves of one operator (continuous batching, FP32_REFINE only): the
 * n_rhs systems op x_i = rhs_i flow through `lanes` (1, 2, 4, 8, or 16, 32 and
 * 64 as 2, 4 and 8 groups of 8 in every launch) interleaved lanes; a lane whose
 * system has stopped takes the next
 * right-hand side at the next check (every 16 iterations), so a slow system
 * does not hold the others' lanes.  Device pointers, D doubles each.  No
 * counterpart in the reference, which solves one system per Solver::Solve. */
int cpm_solve_stream_device(cpm_op* op, int n_rhs, const double* const* d_rhs, double* const* d_x, int lanes,
                            const cpm_solve_opts* opts, cpm_solve_info* infos);

/* ---- serving: continuous batching across operators --------------------
 * No counterpart in the reference, which runs one Solver::Solve at a time
 * (src/Solver.cpp:69-114).  A server keeps a pool of FP32+refine workspace
 * (pool_bytes: 80 B per row and lane; opts->precision = CPM_FP64 selects an
 * FP64 server instead: the reference's arithmetic, 88 B per row and lane, no
 * refinement and no re-solves; opts = NULL: FP32+refine) on the context's device and iterates
 * every admitted request together: a request is an operator with 1..lanes
 * right-hand sides (device vectors, caller row order, kept valid until the
 * request completes).  cpm_server_submit admits a request when a slot of
 * its size is free and returns CPM_BUSY otherwise (submit it again after
 * cpm_server_run has retired others).  cpm_server_run queues `iterations`
 * iterations over all admitted requests, then writes the solutions of the
 * requests that stopped into their d_x and reports them: done_tags[k] and
 * done_infos[k * lanes + j] for j < that request's n_rhs (at most max_done
 * per call; the rest wait for the next call), n_active = requests admitted
 * and not yet reported.  Each lane's arithmetic depends only on its own
 * request, so results are bit-identical alone and in any mix.  Operators
 * with trailing nD rows or the Jacobi preconditioner are rejected; an
 * operator may come from another context on the same device (e.g. a build
 * thread's, cpm_problem_create_on) and must stay alive until its request is
 * reported.  Systems whose refinement stalls are re-solved in FP64 when their
 * request leaves (cpm_solve_info.fallback_iters); cpm_server_submit grows the
 * context's workspaces for that re-solve (about 90 B per row and lane) before
 * it admits a request, so retiring requests allocates nothing, and returns
 * CPM_ERR_NOMEM if they cannot grow.  If writing out a request's results fails,
 * cpm_server_run still retires every stopped request, reports that one with
 * status CPM_SOLVE_NOT_WRITTEN and returns the error.  One host thread per server. */
int cpm_server_create(cpm_ctx* ctx, int64_t pool_bytes, int lanes, const cpm_solve_opts* opts, cpm_server** out);
int cpm_server_destroy(cpm_server* server);
int cpm_server_submit(cpm_server* server, cpm_op* op, int n_rhs, const double* const* d_rhs, double* const* d_x,
                      int64_t tag);
int cpm_server_run(cpm_server* server, int iterations, int max_done, int64_t* done_tags, cpm_solve_info* done_infos,
                   int* n_done, int* n_active);

/* ---- tube (closest-point computational band) -------------------------
 * Replaces Tube(SurfaceSpecifier&, interp_deg, dx) (src/Tube.h:45,
 * src/Tube.cpp:88-461): closest points on the GPU, the banded index set in
 * the reference's breadth-first order (node i here is node i there),
 * neighbours (src/Tube.cpp:620-695) and interpolation base points
 * (:464-531).  surface_type / params / bbox follow SurfaceSpecifier
 * (src/SurfaceSpecifier.cpp:65-256); nparams = 0 selects the reference
 * defaults.  "Polyline" takes nverts vertices (x,y,z) as consecutive segments.
 * Supported on the GPU: Sphere (radius[, centre]), Hemisphere, Torus (R, r),
 * Ellipsoid (a, b, c), Circle3D (r, centre, normal), Point, Polyline, and
 * triangle meshes through cpm_tube_create_mesh; dim must be 3. */
typedef struct cpm_tube_info {
    int64_t nodes;          /* nNodes()                                  */
    int64_t n;              /* grid has 2^n + 1 nodes per axis           */
    int64_t N1d;
    int64_t nnz_L;          /* entries of L (diagonal + tube neighbours) */
    int64_t levels;         /* breadth-first levels                      */
    double xstart;          /* grid origin (same on every axis)          */
    double dx;
    double tube_radius;
    double tube_radius_factor;
    double build_ms;        /* wall time of cpm_tube_create              */
    int dim;
    int p;                  /* interpolation degree                      */
} cpm_tube_info;

int cpm_tube_create(cpm_ctx* ctx, const char* surface_type, const double* params, int nparams, double bbox_lo,
                    double bbox_hi, int dim, const double* verts, int64_t nverts, int p, double dx,
                    cpm_tube** out);
/* Tube around a closed triangle mesh: SurfaceSpecifier(".obj") with surface
 * type "Triangulation" (src/SurfaceSpecifier.cpp:298-363) and the closest
 * point of Surface::cpTri (src/Surface.cpp:1751-1779), which the reference
 * answers with fcpw's BVH (:2537-2592).  verts: nverts x 3, faces: nfaces x 3
 * vertex indices.  bbox_lo >= bbox_hi takes the reference's box
 * [min(0, coordinates), max(0, coordinates)].  Closest points are the exact
 * FP64 minimisers (lowest triangle index on ties), from a BVH built here. */
int cpm_tube_create_mesh(cpm_ctx* ctx, const double* verts, int64_t nverts, const int32_t* faces, int64_t nfaces,
                         double bbox_lo, double bbox_hi, int p, double dx, cpm_tube** out);
/* ---- mesh and curve files ------------------------------------------------
 * Wavefront OBJ as SimplePolygonMesh::readMeshFromObjFile reads it
 * (src/SimplePolygonMesh.cpp:139-213): "v" positions, "f" faces of 1-based
 * position indices ("p/t/n" forms: texture and normal indices ignored), "l"
 * polylines split into 2-vertex edges; other records ignored.  Call with
 * verts = faces = NULL for the counts, then with arrays of at least
 * *nverts x 3 and *nfaces x *face_size entries (indices returned 0-based).
 * face_size is 3 (SurfaceSpecifier's "Triangulation") or 2 (polyline edges,
 * "Polyline", src/SurfaceSpecifier.cpp:298-343).  Rejected with
 * CPM_ERR_ARG: files mixing triangles and edges, polygons other than
 * triangles, vertices without exactly 3 coordinates, non-positive or
 * out-of-range indices. */
int cpm_obj_read(const char* path, double* verts, int64_t* nverts, int32_t* faces, int64_t* nfaces, int* face_size);
/* Consecutive polyline edges (edge k starts where edge k-1 ends) as the
 * vertex chain that "Polyline" surfaces and cpm_ibc_desc take; a closed
 * polyline (SurfaceSpecifier.cpp:334-341: first and last vertices equal, or
 * the last edge touching vertex 0) ends with its first vertex again.
 * chain = NULL returns the count in *nchain. */
int cpm_polyline_chain(const double* verts, int64_t nverts, const int32_t* edges, int64_t nedges, double* chain,
                       int64_t* nchain);
/* Tube of the surface an OBJ file holds: SurfaceSpecifier(file) + Tube
 * (src/SurfaceSpecifier.cpp:298-363): triangles as cpm_tube_create_mesh,
 * edges as a "Polyline" surface; bbox_lo >= bbox_hi takes the reference's
 * box [min(0, coordinates), max(0, coordinates)]. */
int cpm_tube_create_obj(cpm_ctx* ctx, const char* path, double bbox_lo, double bbox_hi, int p, double dx,
                        cpm_tube** out);
int cpm_tube_destroy(cpm_tube* tube);
int cpm_tube_get_info(const cpm_tube* tube, cpm_tube_info* info);
/* Node data in tube order (any pointer may be NULL): linear grid index
 * (Grid.cpp:123-144), coordinates x (3 per node), closest points, distance,
 * boundary tag, 6 neighbours (+x,-x,+y,-y,+z,-z; -1 = none, src/Node.h:23),
 * base point (tube index). */
int cpm_tube_export(const cpm_tube* tube, int64_t* rec, double* x, double* cpx, double* dist, int32_t* bdy,
                    int32_t* nbr, int32_t* bpt);
/* E (nodes x nodes, (p+1)^3 entries per row, ascending columns) -- the matrix of
 * Interpolation(t).BuildInterpolationMatrix (src/Interpolation.cpp:8-121). */
int cpm_tube_export_E(const cpm_tube* tube, int32_t* rowptr, int32_t* col, double* val);
/* L (nodes x nodes, nnz_L entries) -- FDMatrices::BuildLaplacianMatrix (src/FDMatrices.cpp:11-49). */
int cpm_tube_export_L(const cpm_tube* tube, int32_t* rowptr, int32_t* col, double* val);
/* Interpolation rows at arbitrary points inside the tube, nq x nodes with
 * (p+1)^3 entries per row: Interpolation(t, xq) (src/Interpolation.cpp:23-35),
 * e.g. the plotting matrix of Helpers::GetPlotInterpMatrix (src/Helpers.cpp:97-103). */
int cpm_tube_interp_matrix(const cpm_tube* tube, const double* xq, int64_t nq, int32_t* rowptr, int32_t* col,
                           double* val);
/* Operator of the plain closest-point problem built on the GPU from the tube:
 * L and E of Helpers::SetupMatrices (src/Helpers.cpp:86-93) and Poisson::
 * Initialize without boundaries (problems/Poisson.cpp:233-237). */
int cpm_op_create_from_tube(cpm_tube* tube, double m, double n, int precond, int precond_iters, cpm_op** out);

/* ---- problems with interior boundary conditions -------------------------
 * Replaces PoissonWithIBC(surface_specs, ibc_specs, dx, p)
 * (problems/PoissonWithIBC.cpp:8-47,162-192): IBC subsets
 * (src/IBCSubsets.cpp:11-116), stencil directions from the Jacobian of the
 * closest point (src/IBCSubset.cpp:190-268, src/Geometry.cpp:27-98), and the
 * IBC-modified Lfull / Efull with identity rows
 * (src/IBCMatrixManipulations.cpp:12-71,118-251,423-463), all on the GPU.
 * Rows: [tube N | ghost DOFs G | 2nd-order Dirichlet rows K]. */
typedef struct cpm_ibc_desc {
    const char* type;          /* "Polyline", "Circle3D" or "Point"                          */
    const double* params;      /* SurfaceSpecifier parameters (Circle3D: r, centre, normal;  */
    int nparams;               /*  Point: x, y, z)                                           */
    const double* verts;       /* Polyline: nverts x 3 vertices, consecutive segments        */
    int64_t nverts;
    int boundary_order;        /* IBCMetaData::boundary_order (1 or 2)                       */
    int boundary_type[2];      /* 0 Dirichlet, 1 Neumann (IBCMetaData::boundary_type)         */
    double on_ibc_tolerance;   /* IBCMetaData::on_ibc_tolerance                              */
    double dof_radius_scale;   /* IBCMetaData::ibc_DOFSubset_radius_scale (<= 0: 1.0)        */
} cpm_ibc_desc;

typedef struct cpm_problem cpm_problem;

typedef struct cpm_problem_info {
    int64_t rows;                 /* D = N + G + K                                  */
    int64_t tube_nodes;           /* N                                              */
    int64_t ibc_dofs;             /* G  (IBCSubsets::TotalIBCDOFs)                   */
    int64_t dirichlet_rows;       /* K  (IBCSubsets::Num2ndOrderDirichletRows)       */
    int64_t identity_rows_start;  /* first 2nd-order Dirichlet row (0 when none)    */
    int64_t tagging_size0;        /* tagging subset size of curve 0                 */
    int64_t dof_size0;            /* DOF subset size of curve 0                     */
    int n_ibc;
    double build_ms;
} cpm_problem_info;

int cpm_problem_create(cpm_tube* tube, int n_ibc, const cpm_ibc_desc* ibcs, cpm_problem** out);
/* The same, built with ctx's stream and workspaces (a context on the tube's device), so that a
 * second host thread can build problems on a resident tube while the tube's context keeps
 * solving; operators made from the problem belong to ctx.  No counterpart in the reference. */
int cpm_problem_create_on(cpm_ctx* ctx, cpm_tube* tube, int n_ibc, const cpm_ibc_desc* ibcs, cpm_problem** out);
int cpm_problem_destroy(cpm_problem* prob);
/* Rows of curve c in [N | G | K] (IBCSubsets::IBCStartingColumns, src/IBCSubsets.cpp:74-116):
 * rows[0], rows[1] = first ghost row and ghost-row count (the DOF subset), rows[2], rows[3] = first
 * row and count of the rows that take the curve's Dirichlet values (GetRHS kind 2; -1, 0 for
 * Neumann curves). */
int cpm_problem_curve_rows(const cpm_problem* prob, int c, int64_t* rows);
int cpm_problem_get_info(const cpm_problem* prob, cpm_problem_info* info);
/* Evaluation point of every row's right-hand side (PoissonWithIBC::GetRHS,
 * problems/PoissonWithIBC.cpp:50-90): kind 0 = f at cp(x), 1 = f at cp_c,
 * 2 = boundary value at cp_c. */
int cpm_problem_rhs_points(const cpm_problem* prob, double* pts, int32_t* kind);
/* Tagging (which = 0) or DOF (which = 1) subset of curve c in subset order:
 * tube index, closest point on the curve, stencil direction. */
int cpm_problem_subset(const cpm_problem* prob, int c, int which, int32_t* set_of_sub, double* cpc, double* dirs);
/* Curve parameter of n points (3 doubles each, host memory) on curve c:
 * Surface::IBCParameter (src/Surface.cpp:2731-2788), reached through
 * PoissonWithIBC::IBCParameter / IBCSubsets::IBCParameter (src/IBCSubsets.h:43-46)
 * for boundary data such as u = sin(t) (examples/LaplaceTorusTorusKnot/
 * LaplaceTorusTorusKnot.cpp:35).  vertex_param NULL: the normalised arc length
 * in [0, 1] of the point's closest point along the curve's parameterisation
 * points (a polyline's vertices; 100 points of a Circle3D or an Arc), the
 * reference's rule for polylines; 0 for a Point.  vertex_param non-NULL (one
 * value per polyline vertex): the value of the vertex nearest to the point,
 * the reference's rule for curves given by parameterised samples (its
 * "Torus Knot", Surface.cpp:2737-2751).  Computed on the GPU.
 * CPM_ERR_GEOMETRY if a closest point lies on no edge (the reference throws). */
int cpm_problem_ibc_parameter(cpm_problem* prob, int c, const double* pts, int64_t n, const double* vertex_param,
                              double* out);
/* Number of nodes in that subset (IBCSubset::size / IBCDOFSubset::size, src/IBCSubset.h, src/IBCDOFSubset.h). */
int cpm_problem_subset_size(const cpm_problem* prob, int c, int which, int64_t* n);
/* Lfull / Efull as CSR (ascending columns, explicit zeros dropped) and the
 * identity rows.  Call with NULL arrays to get nnz_L / nnz_E. */
int cpm_problem_export(const cpm_problem* prob, int32_t* L_rowptr, int32_t* L_col, double* L_val, int32_t* E_rowptr,
                       int32_t* E_col, double* E_val, uint8_t* identity_rows, int64_t* nnz_L, int64_t* nnz_E);
/* Operator m I + n (Ld + Loff E) of the problem (cpm::Solver::Init with the
 * problem's Lfull, Efull, identity rows). */
int cpm_op_create_from_problem(cpm_problem* prob, double m, double n, int precond, int precond_iters, cpm_op** out);

/* Plotting rows of a problem: Interpolation(tube, xq) (as
 * cpm_tube_interp_matrix) with the stencil entries across a curve moved to
 * its ghost columns, IBCMatrixManipulations::AddIBCPlotInterpolation
 * (src/IBCMatrixManipulations.cpp:74-106) with the plot points' directions of
 * PlotPointsCPDiff (:254-416), curves not oriented.  nq x D CSR with (p+1)^3
 * entries per row (p <= 3), ascending columns; rowptr has nq + 1 entries. */
int cpm_problem_plot_matrix(cpm_problem* prob, const double* xq, int64_t nq, int32_t* rowptr, int32_t* col,
                            double* val);

/* Geodesic distance from the problem's curves by the heat method,
 * GeodesicDistance::ComputeDistance (problems/GeodesicDistance.cpp:11-67):
 * heat flow from the curves, Solver(1, -dt) with jacobi_sweeps Jacobi sweeps
 * (the reference uses dt = dx^2 and 2000 sweeps, :112-115); three rounds of
 * normalised tangential gradient (centred differences, FDMatrices.cpp:64-92,
 * extended by the tube's E), divergence and Poisson solve Solver(0, 1),
 * shifted to minimum 0 (:36-60, :167-313).  d_phi receives phi on the
 * problem's D rows (device memory; the reference then applies its plotting
 * interpolation, :62-66).  Rows other than tube rows of the Poisson
 * right-hand sides are 0.  infos (may be NULL): 4 entries, the heat solve and
 * the three Poisson solves. */
int cpm_geodesic_distance(cpm_problem* prob, double dt, int jacobi_sweeps, const cpm_solve_opts* opts, double* d_phi,
                          cpm_solve_info* infos);

/* ---- time stepping ------------------------------------------------------
 * One step of HeatWithIBC::TimeStep (problems/HeatWithIBC.cpp:80-103) on
 * device vectors of the operator's D rows (the problem's row order):
 *   method 0, explicit:         u <- E (u + dt L u), then identity rows <- g
 *   method 1, backward Euler:   rhs = u,             identity rows <- g, u <- A^-1 rhs
 *   method 2, Crank-Nicolson:   rhs = A(1, dt/2) u,  identity rows <- g, u <- A^-1 rhs
 * For methods 1 and 2 `op` is the problem operator with (m, n) = (1, -dt) or
 * (1, -dt/2), as HeatWithIBC::Initialize builds it (:160-166); for method 0 its
 * L and E are applied and (m, n) are ignored.  d_g holds this step's Dirichlet
 * values on the identity rows (SetDirichletValue, :50-78); other entries are
 * not read.  info may be NULL (method 0 solves nothing). */
int cpm_heat_step_device(cpm_op* op, int method, double dt, const double* d_g, double* d_u,
                         const cpm_solve_opts* opts, cpm_solve_info* info);

/* Gray-Scott reaction-diffusion, ReactionDiffusionWithIBC::TimeStep
 * (problems/ReactionDiffusionWithIBC.cpp:79-104), with its constants as
 * defaults (cpm_rd_params_default: F 0.054, k 0.063, nu_u (dx/3)^2,
 * nu_v nu_u/3 (:81-84), dt 0.1 dx^2 / max(nu) as
 * examples/ReactionDiffusionFish/ReactionDiffusionFish.cpp:61-63).
 *   method 0, explicit (the reference): u <- E (u + dt nu_u L u + dt Ru(u, v)),
 *             v likewise, then identity rows <- g
 *   method 1, implicit diffusion (no counterpart in the reference):
 *             (1 - dt nu_u L) u' = u + dt Ru(u, v) with the identity rows of
 *             the right-hand side set to g, solved as the problem's operator
 *             with (m, n) = (1, -dt nu); the 2 n_fields solves run as one batch.
 * Ru = -u v^2 + F (1 - u), Rv = u v^2 - (F + k) v.  op's (m, n) are not used.
 * d_u[f], d_v[f] (f < n_fields): device fields of D rows, updated in place;
 * d_g_u / d_g_v: Dirichlet values read on the identity rows.  infos (may be
 * NULL) receives 2 n_fields entries, u then v of each field. */
typedef struct {
    double F, k;        /* feed and kill rates */
    double nu_u, nu_v;  /* diffusion coefficients */
    double dt;
    int method;         /* 0 explicit, 1 implicit diffusion */
    int reserved;
} cpm_rd_params;
void cpm_rd_params_default(double dx, cpm_rd_params* p);
int cpm_rd_step_device(cpm_op* op, const cpm_rd_params* p, int n_fields, double* const* d_u, double* const* d_v,
                       const double* d_g_u, const double* d_g_v, const cpm_solve_opts* opts, cpm_solve_info* infos);

/* ---- measurement ---------------------------------------------------------
 * No counterpart in the reference.  Times the operator's E-gather kernel (the
 * dominant kernel of every BiCGStab iteration) on the context stream with HIP
 * events: `reps` back-to-back launches over `nrhs` interleaved right-hand
 * sides (1, 2, 4 or 8) in the given precision, rotating over copies of the
 * input and output vectors that together exceed the last-level cache.  Returns the average launch
 * duration and the launch's algorithmic HBM bytes: E columns + values once,
 * x read once, tmp written once, one flag byte per row. */
int cpm_profile_egather(cpm_op* op, int precision, int nrhs, int reps, double* avg_ms, double* bytes_per_launch);
/* The same with `groups` (1..64) copies of the operator in every launch, as
 * the launches of streamed solves with 16 or 32 lanes (2 or 4 groups of 8)
 * and of batches over several operators carry them; bytes cover all groups.
 * No counterpart in the reference. */
int cpm_profile_egather_groups(cpm_op* op, int precision, int nrhs, int groups, int reps, double* avg_ms,
                               double* bytes_per_launch);

/* Plain device-to-device copy of `bytes` (multiple of 16), `reps` launches:
 * the reference point for HBM bandwidth and for scaling hardware traffic
 * counters (kernel `k_copy16`).  No counterpart in the reference. */
int cpm_profile_copy(cpm_ctx* ctx, int64_t bytes, int reps, double* avg_ms);

/* Device scratch for callers that keep vectors resident (bench, serving). */
int cpm_device_alloc(cpm_ctx* ctx, int64_t bytes, void** d_ptr);
int cpm_device_free(cpm_ctx* ctx, void* d_ptr);
int cpm_memcpy_h2d(cpm_ctx* ctx, void* d_dst, const void* h_src, int64_t bytes);
int cpm_memcpy_d2h(cpm_ctx* ctx, void* h_dst, const void* d_src, int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* CPM_GPU_H */
