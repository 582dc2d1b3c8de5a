/*
 * bpx.h — C ABI of the MI355X (gfx950) BK/BP hex hot path.
 *
 * Drop-in boundary for paranumal/benchparanumal's OCCA path. Two levels:
 *
 *  1. Kernel level: one entry point per reference kernel, with the
 *     reference kernel's positional arguments (device pointers, counts,
 *     scalars), asynchronous on the platform's stream. They replace
 *     platform_t::buildKernel + occa::kernel launches.
 *  2. Object level: platform / mesh / ogs / BP solver handles whose calls
 *     mirror mesh_t, ogs::ogs_t, linAlg_t, linearSolver_t (cg) and bpN_t.
 *
 * Conventions (mirroring include/types.h:32-69 and include/memory.hpp):
 *  - dfloat is double (suffix-free entry points) or float (_f32);
 *    dlong is int32; hlong (global ids) is int64.
 *  - Every pointer named o_* or documented "device" is device memory owned
 *    by the caller; the library never frees caller memory and never
 *    allocates inside a kernel-level call.
 *  - Every call returns BPX_OK (0) or a negative bpx_status; the message of
 *    the last failure on the calling thread is bpx_last_error(). Nothing
 *    throws across the ABI (the reference's LIBP_ABORT, libs/core/
 *    exception.cpp:75-79, becomes an error code).
 *  - Calls are asynchronous on the platform stream unless they return a
 *    host scalar (then they synchronise that stream, as linAlg_t::norm2 does
 *    at libs/core/linAlg.cpp:61-62).
 */
#ifndef BPX_H
#define BPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  BPX_OK = 0,
  BPX_ERR_INVALID_ARGUMENT = -1,
  BPX_ERR_UNSUPPORTED = -2,
  BPX_ERR_HIP = -3,
  BPX_ERR_OUT_OF_MEMORY = -4,
  BPX_ERR_NEGATIVE_JACOBIAN = -5, /* libs/mesh/meshGeometricFactorsHex3D.cpp:69-70,118-119 */
  BPX_ERR_OGS = -6,               /* libs/ogs/ogsOperator.cpp:487-488, ogsSetup.cpp:827-830 */
  BPX_ERR_BREAKDOWN = -7,         /* non-finite p.Ap or r.r in CG */
  BPX_ERR_NO_DEVICE = -8
} bpx_status;

typedef struct bpx_platform_s* bpx_platform; /* replaces platform_t (include/platform.hpp:61-185) */
typedef struct bpx_mesh_s* bpx_mesh;         /* replaces mesh_t (include/mesh.hpp) */
typedef struct bpx_ogs_s* bpx_ogs;           /* replaces ogs::ogs_t (include/ogs.hpp:216-341) */
typedef struct bpx_bp_s* bpx_bp;             /* replaces bpN_t (BP/BP5/bp5.hpp:43-74) */
typedef struct bpx_batch_s* bpx_batch;       /* no reference counterpart: many BP5 problems */
typedef struct bpx_cg_s* bpx_cg;             /* replaces cg (include/linearSolver.hpp:69-81) */

const char* bpx_last_error(void);
int bpx_version(void);
/* 1 when this build has HIP device code for gfx950 and a device is visible. */
int bpx_device_count(int* count);

/* Synthetic inputs (no reference counterpart): out[i] is N(0,1) draw number
 * counter0+i of the SplitMix64 + Box-Muller stream `seed` (host memory). */
int bpx_random_normal(uint64_t seed, int64_t counter0, int64_t n, double* out);

/* ------------------------------------------------------------ platform --
 * platform_t(settings) (include/platform.hpp:72-82): picks the device and
 * owns the stream, linAlg scratch and pinned scalars. */
int bpx_platform_create(int device, bpx_platform* out);
int bpx_platform_destroy(bpx_platform p);
/* Run on a caller stream (hipStream_t) instead of the platform's own. */
int bpx_platform_set_stream(bpx_platform p, void* hip_stream);
void* bpx_platform_stream(bpx_platform p);
/* platform_t::finish (include/platform.hpp:179-181) */
int bpx_platform_finish(bpx_platform p);

/* ------------------------------------------------ kernel level: Ax ------
 * Same positional arguments as the OKL kernels. N is the polynomial
 * degree (1..15), the reference's compile-time p_N. Unused reference
 * arguments (vgeo, S, ...) are accepted and ignored, as in the OKL. */

/* okl/bp5AxHex3D.okl:415-427 */
int bpx_bp5_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                   const int* GlobalToLocal, const double* wJ, const double* vgeo,
                   const double* ggeo, const double* gllw, const double* D, const double* S,
                   const double* MM, double lambda, const double* q, double* Aq);
/* okl/bp5AxAffineHex3D.okl:419-431 */
int bpx_bp5_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                          const int* GlobalToLocal, const double* wJ, const double* vgeo,
                          const double* ggeo, const double* gllw, const double* D,
                          const double* S, const double* MM, double lambda, const double* q,
                          double* Aq);
/* No reference counterpart (the reference applies Ax to one vector per
 * launch; okl/bp5AxHex3D.okl:415-427 is the single-vector form): BP5 Ax on K
 * right-hand sides at once. The right-hand sides are stored one after
 * another: q + r*ldq (global, indexed through GlobalToLocal) and Aq +
 * r*ldAq (local, e*Np + n), r = 0..K-1. The form:
 *  - chunk = 0 (default): the pencil form up to N = 9 (and
 *    Nelements*Np*8 < 2^31), else the sheet form (non-affine) or the sweep
 *    form with its default chunk;
 *  - chunk = -2, the pencil form: each element's geometric factors loaded
 *    once into registers, the K right-hand sides applied one after another,
 *    every contraction along a pencil one lane holds. Agrees with
 *    bpx_bp5_ax_hex to rounding (the weak divergence's r and s parts are
 *    summed apart), not bit for bit;
 *  - chunk = -1, the sheet form (non-affine): one workgroup per element and
 *    group of right-hand sides, whose sheets share each slab of geometric
 *    factors through LDS;
 *  - chunk = 1, 2 or 4 (dividing K), the sweep form: each element swept
 *    K/chunk times with chunk right-hand sides in registers.
 * The sheet and sweep forms equal bpx_bp5_ax_hex bit for bit (to rounding
 * at N = 15, where that one runs on the matrix cores). */
int bpx_bp5_ax_hex_krhs(bpx_platform p, int N, int K, int chunk, int Nelements,
                        const int* elementList, const int* GlobalToLocal, const double* wJ,
                        const double* vgeo, const double* ggeo, const double* gllw,
                        const double* D, const double* S, const double* MM, double lambda,
                        const double* q, int64_t ldq, double* Aq, int64_t ldAq);
int bpx_bp5_ax_affine_hex_krhs(bpx_platform p, int N, int K, int chunk, int Nelements,
                               const int* elementList, const int* GlobalToLocal,
                               const double* wJ, const double* vgeo, const double* ggeo,
                               const double* gllw, const double* D, const double* S,
                               const double* MM, double lambda, const double* q, int64_t ldq,
                               double* Aq, int64_t ldAq);
/* No reference counterpart (single-vector FP32 form: okl/bp5AxHex3D.okl:415-427
 * with dfloat = float): the k-RHS Ax above in FP32 (the same forms and
 * chunk values), stored factors rounded to float. Each right-hand side
 * agrees with bpx_bp5_ax_hex_f32 on it (bit for bit in the sheet and sweep
 * forms, to rounding in the pencil form). */
int bpx_bp5_ax_hex_krhs_f32(bpx_platform p, int N, int K, int chunk, int Nelements,
                            const int* elementList, const int* GlobalToLocal, const float* wJ,
                            const float* vgeo, const float* ggeo, const float* gllw,
                            const float* D, const float* S, const float* MM, float lambda,
                            const float* q, int64_t ldq, float* Aq, int64_t ldAq);
int bpx_bp5_ax_affine_hex_krhs_f32(bpx_platform p, int N, int K, int chunk, int Nelements,
                                   const int* elementList, const int* GlobalToLocal,
                                   const float* wJ, const float* vgeo, const float* ggeo,
                                   const float* gllw, const float* D, const float* S,
                                   const float* MM, float lambda, const float* q, int64_t ldq,
                                   float* Aq, int64_t ldAq);
/* okl/bp6AxHex3D.okl:555-567 (3 fields, local layout [e][f][n]) */
int bpx_bp6_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                   const int* GlobalToLocal, const double* wJ, const double* vgeo,
                   const double* ggeo, const double* gllw, const double* D, const double* S,
                   const double* MM, double lambda, const double* q, double* Aq);
/* okl/bp6AxAffineHex3D.okl */
int bpx_bp6_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                          const int* GlobalToLocal, const double* wJ, const double* vgeo,
                          const double* ggeo, const double* gllw, const double* D,
                          const double* S, const double* MM, double lambda, const double* q,
                          double* Aq);
/* okl/bp1AxHex3D.okl:61-68 (I = cubInterp [cubNq x Nq]) */
int bpx_bp1_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                   const int* GlobalToLocal, const double* cubwJ, const double* I,
                   const double* MM, const double* q, double* Aq);
/* okl/bp1AxAffineHex3D.okl:59-66 (cubwJ = per-element J) */
int bpx_bp1_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                          const int* GlobalToLocal, const double* cubwJ, const double* I,
                          const double* MM, const double* q, double* Aq);
/* okl/bp2AxHex3D.okl (3 fields of BP1) */
int bpx_bp2_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                   const int* GlobalToLocal, const double* cubwJ, const double* I,
                   const double* MM, const double* q, double* Aq);
int bpx_bp2_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                          const int* GlobalToLocal, const double* cubwJ, const double* I,
                          const double* MM, const double* q, double* Aq);
/* okl/bp3AxHex3D.okl:369-381 (D = cubD [cubNq x cubNq], I = cubInterp) */
int bpx_bp3_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                   const int* GlobalToLocal, const double* cubwJ, const double* cubggeo,
                   const double* D, const double* I, const double* invV, const double* S,
                   const double* MM, double lambda, const double* q, double* Aq);
/* okl/bp3AxAffineHex3D.okl:668-680 (D = modal cubD [Nq x Nq], invV) */
int bpx_bp3_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                          const int* GlobalToLocal, const double* wJ, const double* ggeo,
                          const double* D, const double* I, const double* invV,
                          const double* S, const double* MM, double lambda, const double* q,
                          double* Aq);
/* okl/bp4AxHex3D.okl (3 fields of BP3) */
int bpx_bp4_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                   const int* GlobalToLocal, const double* cubwJ, const double* cubggeo,
                   const double* D, const double* I, const double* invV, const double* S,
                   const double* MM, double lambda, const double* q, double* Aq);
int bpx_bp4_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList,
                          const int* GlobalToLocal, const double* wJ, const double* ggeo,
                          const double* D, const double* I, const double* invV,
                          const double* S, const double* MM, double lambda, const double* q,
                          double* Aq);

/* FP32 operator (no reference counterpart: the reference builds its
 * kernels with dfloat=double only, include/types.h:32-40). Same arguments
 * with float data; geometric factors and D/I matrices rounded to FP32. */
int bpx_bp5_ax_hex_f32(bpx_platform p, int N, int Nelements, const int* elementList,
                       const int* GlobalToLocal, const float* wJ, const float* vgeo,
                       const float* ggeo, const float* gllw, const float* D, const float* S,
                       const float* MM, float lambda, const float* q, float* Aq);
int bpx_bp3_ax_hex_f32(bpx_platform p, int N, int Nelements, const int* elementList,
                       const int* GlobalToLocal, const float* cubwJ, const float* cubggeo,
                       const float* D, const float* I, const float* invV, const float* S,
                       const float* MM, float lambda, const float* q, float* Aq);

/* okl/rhsHex3D.okl:44-85 / :87-128. Nfields 1 or 3. */
int bpx_rhs_hex(bpx_platform p, int N, int Nfields, int affine, int Nelements,
                const double* wJ, const double* gllw, const double* MM, const double* x,
                const double* y, const double* z, double lambda, double* rhs);

/* ---------------------------------------------- kernel level: ogs -------
 * libs/ogs/okl/ogsKernels.okl:86-92 (gather, Add) and :127-133 (scatter).
 * K interleaved vectors (entry k of node n at k + n*K). blockStarts are
 * the <=512-nnz row blocks of libs/ogs/ogsOperator.cpp:474-546. */
int bpx_ogs_gather(bpx_platform p, int Nblocks, int K, const int* blockStarts,
                   const int* gatherStarts, const int* gatherIds, const double* q,
                   double* gatherq);
int bpx_ogs_gather_f32(bpx_platform p, int Nblocks, int K, const int* blockStarts,
                       const int* gatherStarts, const int* gatherIds, const float* q,
                       float* gatherq);
int bpx_ogs_scatter(bpx_platform p, int Nblocks, int K, const int* blockStarts,
                    const int* scatterStarts, const int* scatterIds, const double* gatherq,
                    double* q);
/* libs/ogs/okl/ogsKernels.okl:32-81 (gather then scatter back in place) */
int bpx_ogs_gather_scatter(bpx_platform p, int Nblocks, int K, const int* blockStarts,
                           const int* gatherStarts, const int* gatherIds, double* q);

/* ------------------------------------------- kernel level: linAlg -------
 * libs/core/linAlg.cpp:38-89 (set, axpy, norm2, innerProd). The
 * reductions are deterministic: the same inputs give the same bits on
 * every call. */
int bpx_linalg_set(bpx_platform p, int N, double alpha, double* x);
int bpx_linalg_axpy(bpx_platform p, int N, double alpha, const double* x, double beta,
                    double* y);
/* Host-scalar results synchronise the stream (libs/core/linAlg.cpp:61-62). */
int bpx_linalg_inner_prod(bpx_platform p, int N, const double* x, const double* y,
                          double* result);
int bpx_linalg_norm2(bpx_platform p, int N, const double* x, double* result);
/* cg::UpdateCG (libs/core/linearSolverCG.cpp:131-156): r -= alpha*Ap,
 * x += alpha*p, returns r.r. */
int bpx_cg_update(bpx_platform p, int N, const double* Ap, const double* o_p, double alpha,
                  double* r, double* x, double* rdotr);

/* ------------------------------------------------------------------ cg --
 * cg(platform, N, Nhalo) (libs/core/linearSolverCG.cpp:33-60) and
 * linearSolver_t::Solve(solver_t&, o_x, o_rhs, tol, MAXIT, verbose)
 * (include/linearSolver.hpp:49-54) on any operator: solver_t::Operator
 * (include/solver.hpp:47-49) is the callback `op`, called on the host
 * thread with device vectors of N + Nhalo entries; it must queue its work
 * on the platform's stream (bpx_platform_stream) and return 0, or a
 * nonzero code that ends the solve with that code. The cg object owns its
 * work vectors (allocated here, as the reference's constructor does).
 * Scalars stay on the device; an iteration never waits for the host. o_r
 * holds the right-hand side on entry and the residual on exit, o_x the
 * initial guess and the solution. rnorm_hist (host, may be NULL) gets
 * MAXIT + 1 residual norms (entry 0 = initial). */
typedef int (*bpx_operator_fn)(void* ctx, const double* o_q, double* o_Aq);
int bpx_cg_create(bpx_platform p, int N, int Nhalo, bpx_cg* out);
int bpx_cg_destroy(bpx_cg c);
int bpx_cg_solve(bpx_cg c, bpx_operator_fn op, void* ctx, double* o_x, double* o_r, double tol,
                 int maxit, int verbose, int* iterations, double* rnorm_hist);

/* bpx_cg_solve without a bpx_cg object, allocating nothing. No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP stream capture, tight serving loops, harnesses that
 * forbid hipMalloc). The work vectors, partial sums, state and histories
 * come from a fixed device array of about 33 MiB that each GPU the library
 * runs on holds until the process ends; nothing is allocated (no hipMalloc,
 * hipFree or pinned memory, no host heap). Same arguments and results as
 * bpx_cg_solve (verbose aside), with the work queued on hip_stream (NULL:
 * the null stream of the current device) and the operator callback
 * receiving device vectors of N entries that it must fill on that stream.
 * Limits: 0 <= N <= BPX_CG_NOALLOC_MAX_N and 0 <= maxit <=
 * BPX_CG_NOALLOC_MAX_ITER; larger values return BPX_ERR_INVALID_ARGUMENT.
 * Every call on one GPU shares that GPU's array: the calls of one process
 * are serialised per GPU (the call returns once the solve is done), and a
 * call made from inside the operator callback of another is refused. */
#define BPX_CG_NOALLOC_MAX_N 2097152
#define BPX_CG_NOALLOC_MAX_ITER 65535
int bpx_cg_solve_noalloc(void* hip_stream, int N, bpx_operator_fn op, void* ctx, double* o_x,
                         double* o_r, double tol, int maxit, int* iterations, double* rnorm_hist);

/* ---------------------------------------------------------------- mesh --
 * mesh_t(platform, settings, comm) -> Setup for a hex box
 * (libs/mesh/meshSetup.cpp:31-85, meshSetupBoxHex3D.cpp:31-131).
 * lo/dim: box corner and size (reference: lo=-1, dim=2). perturb: seeded
 * interior-vertex displacement as a fraction of h (0 = the reference's
 * undistorted box). cubature: also build the Gauss (p+2) factors
 * (meshCubatureSetupHex3D.cpp). Host-side setup; copies to the device on
 * first use. elementType: the reference's codes (meshSetElementType.cpp):
 * 12 (or 0) hexahedra, 4 quadrilaterals (meshSetupBoxQuad2D.cpp), 3
 * triangles (meshSetupBoxTri2D.cpp, two per square), 6 tetrahedra
 * (meshSetupBoxTet3D.cpp, six per cube); the 2D boxes ignore nz, lo[2] and
 * dim[2]. Triangles and tetrahedra are straight-sided (affine per element
 * whatever `affine` says); cubature factors exist for hex and quad. */
typedef struct {
  int N, nx, ny, nz;
  int affine;
  int cubature;
  double lo[3], dim[3];
  double perturb;
  uint64_t seed;
  int elementType;
} bpx_mesh_opts;

int bpx_mesh_create(const bpx_mesh_opts* opts, bpx_mesh* out);
/* The same mesh set up on the device (the setup path of
 * libs/mesh/meshPhysicalNodesHex3D.cpp:31-116,
 * meshGeometricFactorsHex3D.cpp:31-153 and meshCubatureSetupHex3D.cpp:97-257
 * run as GPU kernels): the vertices and 1D operators are formed on the host,
 * the nodes, global ids and geometric and cubature factors on the device, with
 * the same bits as bpx_mesh_create's host arrays. bpx_ogs_setup_masked and
 * bpx_bp_create on such a mesh build their maps on the device too. The
 * per-node host arrays are copied down on first host use (bpx_mesh_array,
 * bpx_mesh_plot_fields). The mesh keeps using p's device and stream. */
int bpx_mesh_create_dev(bpx_platform p, const bpx_mesh_opts* opts, bpx_mesh* out);
int bpx_mesh_destroy(bpx_mesh m);
/* counts: N Nq Np cubNq cubNp Nelements elementType dim Nverts (int64 each) */
int bpx_mesh_info(bpx_mesh m, int64_t* out, int capacity);
/* Copy a host array of the mesh (names: gllz gllw D MM cubr cubw cubInterp
 * cubD invV x y z wJ ggeo cubwJ cubggeo EX EY EZ globalIds; triangles and
 * tetrahedra also rN sN tN (reference nodes), Dsim (the dim packed
 * derivative matrices) and Ssim (the packed stiffness matrices)) into dst.
 * *count gets the element count; dtype 0 = f64, 2 = i64. dst may be NULL
 * to query the size. */
int bpx_mesh_array(bpx_mesh m, const char* name, void* dst, int64_t* count, int* dtype);

/* ----------------------------------------------------------------- ogs --
 * ogs::ogs_t::Setup (include/ogs.hpp:221-232, libs/ogs/ogsSetup.cpp:87-290)
 * on one rank, Signed, unique: ids[N] (0 = masked), then the T maps and
 * GlobalToLocal (ogsSetup.cpp:833-849). Host-side; no device needed. */
int bpx_ogs_setup(int N, const int64_t* ids, bpx_ogs* out);
/* ogs_t::Setup on the device (GPU sort, scans and bisection in place of
 * libs/ogs/ogsSetup.cpp:87-290 and ogsOperator.cpp:474-546 on the host):
 * o_ids is a device array of N ids; the maps equal bpx_ogs_setup's. Host
 * copies of the maps are made on first bpx_ogs_array request. */
int bpx_ogs_setup_dev(bpx_platform p, int N, const int64_t* o_ids, bpx_ogs* out);
/* mesh_t::MaskedGatherScatterSetup(Nfields) (meshMaskedGatherScatterSetup.
 * cpp:31-77): masked ids of the box, then ogs_t::Setup. */
int bpx_ogs_setup_masked(bpx_mesh m, int Nfields, bpx_ogs* out);
int bpx_ogs_destroy(bpx_ogs o);
/* counts: N Ngather nnzT NrowBlocksT (int64 each) */
int bpx_ogs_info(bpx_ogs o, int64_t* out, int capacity);
/* names: rowStartsT colIdsT blockRowStartsT GlobalToLocal maskedIds */
int bpx_ogs_array(bpx_ogs o, const char* name, void* dst, int64_t* count, int* dtype);
/* ogs_t::Gather(o_gv, o_v, K, Add, Trans) and Scatter(o_v, o_gv, K, NoTrans)
 * (libs/ogs/ogs.cpp:205-301, :394-492) on device vectors. */
int bpx_ogs_gather_dev(bpx_platform p, bpx_ogs o, int K, const double* v, double* gv);
int bpx_ogs_scatter_dev(bpx_platform p, bpx_ogs o, int K, const double* gv, double* v);
/* ogs_t::GatherScatter(o_v, K, Add, Sym) (libs/ogs/ogs.cpp:494-581) in
 * place: every copy of a node gets the sum over its copies; local entries
 * in no row (masked) keep their value. */
int bpx_ogs_gather_scatter_dev(bpx_platform p, bpx_ogs o, int K, double* v);

/* ------------------------------------------------------------------ BP --
 * bpN_t(platform, settings, mesh, lambda) (BP/BP5/src/bp5Setup.cpp:29-95)
 * for bp in 1..6: masked ogs, GlobalToLocal, operator data on the device. */
int bpx_bp_create(bpx_platform p, bpx_mesh m, int bp, double lambda, bpx_bp* out);
int bpx_bp_destroy(bpx_bp b);
/* counts: bp Nfields Ngather NLocal Nelements N Nnodes (int64 each; Nnodes =
 * points of the node lattice, (nx*N+1)(ny*N+1)(nz*N+1)) */
int bpx_bp_info(bpx_bp b, int64_t* out, int capacity);
/* No reference counterpart: launch choices the setup timed on this
 * problem, all giving identical results (times in us, -1 = not timed).
 * out[0..2]: gather workgroup order (0 strided, 1 XCD-blocked), time of
 * order 0, of order 1; out[3..5]: FP64 CG operator fused (1: the Ax writes
 * single-copy rows, the gather sums the shared rows) or unfused (0: Ax,
 * then the full gather; -1 = no fused form for this problem), time fused,
 * time unfused; out[6..8]: the same for the FP32 operator (timed when it is
 * first used); out[9]: 1 if the FP32 operator forms its geometric factors
 * from the elements' trilinear maps, 0 if it reads stored factors rounded
 * to FP32, -1 if it has not been built yet. Returns the number of values. */
int bpx_bp_tuning(bpx_bp b, double* out, int capacity);
/* No reference counterpart: copy one of the problem's device arrays to host
 * memory dst (NULL: query *count). Names: rowStartsT colIdsT blockRowStartsT
 * GlobalToLocal (int32, the masked ogs), wJ ggeo (f64, operator factors),
 * solvePerm solveGlobalToLocal solveRowStarts solveColIds solveBlocks
 * sharedRowStarts sharedColIds sharedBlocks (int32, the fused BP5 operator's
 * solve-space maps; empty when it has none). dtype as bpx_mesh_array, with
 * 1 = int32. */
int bpx_bp_array(bpx_bp b, const char* name, void* dst, int64_t* count, int* dtype);
/* No reference counterpart: pin launch choices instead of the timed ones
 * (for reproducible timings; results do not depend on them): gather order
 * 0/1, FP64 operator fused 0/1, FP32 operator fused 0/1; -1 keeps a
 * choice. */
int bpx_bp_set_tuning(bpx_bp b, int gatherMode, int fused, int fused32);
/* No reference counterpart: the geometry of the FP32 BP5 operator (the
 * inner operator of bpx_bp_solve_mixed). mode 0 (default): formed in the
 * kernel from each element's trilinear map when every node of the mesh is
 * its element's trilinear map of the 8 vertices (checked on the host; the
 * box meshes of bpx_mesh_create are), else stored factors; mode 1: the FP64
 * geometric factors rounded to FP32 and streamed (the operator of the
 * reference's kernels with dfloat = float). Other operators always use
 * stored factors. The FP32 operator is rebuilt on its next use. */
int bpx_bp_set_fp32_geometry(bpx_bp b, int mode);
/* No reference counterpart: where the BP5 Ax at N = 15 in FP64 (one field)
 * runs its r and s contractions: 0 = auto (the matrix cores; default), 1 =
 * vector ALU, 2 = matrix cores (BPX_ERR_UNSUPPORTED for other problems).
 * Covers the plain and the fused (CG) form; the two engines agree to
 * rounding, not bit for bit. Other problems run on the vector ALU. */
int bpx_bp_set_ax_engine(bpx_bp b, int engine);
/* No reference counterpart (launch choice): which kernel runs the BP1-BP4
 * Ax of a problem object on non-affine meshes at N <= 7: 0 = auto (the
 * pencil kernel; default), 1 = pencil (several fields or elements per
 * workgroup, every contraction along a line one lane holds, in even-odd
 * form), 2 = one element per wavefront, 3 = pencil with row-by-row
 * contractions (the form the kernel-level entry points run). The forms
 * agree to rounding, not bit for bit. Other problems have one kernel. */
int bpx_bp_set_cub_form(bpx_bp b, int form);
/* No reference counterpart (launch choice, same results): mask = 1 (FP64
 * solves), 2 (the FP32 iterations of the mixed solve) or 3 (both): those CG
 * solves that use the fused BP5 operator do the direction update of the
 * elements' interior rows (x += alpha p, p = r + beta p) inside that
 * operator's Ax, and the separate update kernel covers the shared rows
 * only; elsewhere (mask 0: everywhere) one update kernel covers all rows.
 * Default 1: at 24^3, p = 7 the fold saves 2% of an FP64 iteration and
 * costs 2% of an FP32 one, whose Ax is bound by instruction issue. */
int bpx_bp_set_cg_fold(bpx_bp b, int mask);
/* No reference counterpart (measurement): the CG operator as the solvers
 * apply it (their row order, the tuned fused or unfused form) in FP64
 * (precision 64) or FP32 (32), `launches` times back to back after 3
 * warm-ups: the Ax timed by HIP events its own dispatch stamps, the gather
 * from the Ax's end to an event after it. out[0] = mean Ax launch in us,
 * out[1] = mean gather launch in us, out[2] = 1 if the fused form ran.
 * Synchronises. */
int bpx_bp_time_operator(bpx_bp b, int precision, int launches, double* out);
/* No reference counterpart (measurement): the BP5 k-RHS Ax (K right-hand
 * sides, chunk as bpx_bp5_ax_hex_krhs) on o_q (device, K vectors of Ngather
 * one after another; NULL: zeros, which draw less power and so run at a
 * higher clock than data do) and the single-vector Ax on o_q's first
 * vector, on the problem's arrays, each launched `launches` times back to
 * back after 2 warm-ups and timed by HIP events its own dispatch stamps.
 * out[0] = mean k-RHS launch in us, out[1] = mean single-vector launch in
 * us. Synchronises. */
int bpx_bp_time_krhs(bpx_bp b, int K, int chunk, int launches, const double* o_q, double* out);
/* bpN_t::Operator(o_q, o_Aq) (BP/BP5/src/bp5Operator.cpp:29-85): Ax over the
 * local elements, then the T gather. Device vectors of length Ngather. */
int bpx_bp_operator(bpx_bp b, const double* o_q, double* o_Aq);
/* The operator as CG applies it: cg::Solve's Operator followed by
 * innerProd(p, Ap) (libs/core/linearSolverCG.cpp:99-106) in one pass. For
 * BP5 the Ax kernel writes nodes that are the only copy of their global row
 * straight into o_Aq and the gather sums only the shared rows; o_Aq equals
 * bpx_bp_operator's bit for bit. *qAq (host) = q.Aq. Synchronises. */
int bpx_bp_operator_dot(bpx_bp b, const double* o_q, double* o_Aq, double* qAq);
/* Ax alone into the local vector (BK/BK5/src/bk5Run.cpp:59-71). */
int bpx_bp_ax_local(bpx_bp b, const double* o_q, double* o_AqL);
/* FP32 operator on the same mesh (D, I and factors rounded to float; BP5
 * non-affine forms its geometric factors in the kernel from each element's
 * trilinear map, agreeing with the FP64 factors to float rounding). */
int bpx_bp_operator_f32(bpx_bp b, const float* o_q, float* o_Aq);
/* No reference counterpart: bpN_t::Operator (BP/BP5/src/bp5Operator.cpp:29-85)
 * on K right-hand sides (BP5 only), stored one after another: o_q and o_Aq
 * hold K vectors of Ngather. k-RHS Ax, then one gather per right-hand side. */
int bpx_bp_operator_krhs(bpx_bp b, int K, int chunk, const double* o_q, double* o_Aq);
/* No reference counterpart (the reference's BP5 coefficient is constant,
 * okl/bp5AxHex3D.okl:415-569): variable coefficient -div(kappa grad u) +
 * lambda u for BP5. o_kappa (device, Nnodes values, lattice order
 * X + (nx*N+1)*(Y + (ny*N+1)*Z)) scales the geometric factors G at every
 * node; bpx_bp_operator and bpx_bp_solve then use it. NULL restores the
 * constant coefficient. Call again after changing kappa. */
int bpx_bp_set_kappa(bpx_bp b, const double* o_kappa);
/* No reference counterpart: adjoint gradient for the variable coefficient.
 * With A(kappa) u = b and A lam = dL/du (A is symmetric), dL/dkappa at
 * lattice point p is -sum over the local copies of p of (D lam).G(D u);
 * o_u, o_lam: device, Ngather; o_grad: device, Nnodes. Requires
 * bpx_bp_set_kappa first. */
int bpx_bp_kappa_gradient(bpx_bp b, const double* o_u, const double* o_lam, double* o_grad);
/* mesh_t::PlotFields (libs/mesh/meshPlotFields3D.cpp:31-121): a VTU file
 * (ascii) of the mesh's local nodes with Nfields values each, QL (host,
 * layout [element][field][node]); each element is N^3 sub-cubes of 6
 * tetrahedra. Host only. */
int bpx_mesh_plot_fields(bpx_mesh m, int Nfields, const double* QL, const char* filename);
/* The solution written as bpN_t::Run does with an output file name
 * (BP/BP5/src/bp5Run.cpp:220-238): o_u (device, Ngather, reference row
 * order) scattered to the local nodes (Dirichlet nodes 0), then
 * bpx_mesh_plot_fields. Synchronises. */
int bpx_bp_plot_solution(bpx_bp b, const double* o_u, const char* filename);
/* No reference counterpart (verification): o_u (device, Ngather) against
 * the manufactured solution of the forcing, u = sin(pi x) sin(pi y)
 * sin(pi z) for BP3/BP5 (-lap u + lambda u = (3 pi^2 + lambda) u) and
 * 3 pi^2 u for BP1 (forced with lambda = 0), over the local nodes with the
 * GLL quadrature: out[0] = max |u_h - u|, out[1] = ||u_h - u||_L2,
 * out[2] = ||u_h - u||_L2 / ||u||_L2. One-field problems only.
 * Synchronises. */
int bpx_bp_error_norms(bpx_bp b, const double* o_u, double* out);
/* b = gather(forcing) as bpN_t::Run does (BP/BP5/src/bp5Run.cpp:51-64). */
int bpx_bp_rhs_forcing(bpx_bp b, double* o_b);
/* cg::Solve (libs/core/linearSolverCG.cpp:62-129) with device-resident
 * scalars: no host wait inside an iteration. o_r holds b on entry and the
 * residual on exit; o_x the initial guess / solution. rnorm_hist (host,
 * may be NULL) gets maxit+1 residual norms (entry 0 = initial). Returns
 * BPX_OK and *iterations. */
int bpx_bp_solve(bpx_bp b, double* o_x, double* o_r, double tol, int maxit, int verbose,
                 int* iterations, double* rnorm_hist);
/* No reference counterpart: CG in FP32 (FP32 operator and vectors, FP64
 * dot accumulation) with an FP64 defect correction every `interval`
 * iterations (x += c; r = b - A x in FP64; the FP32 recurrence continues
 * from float(r)) until ||b - A x|| / ||b|| < tol. o_x: initial guess and
 * solution (device, FP64). stats (host, may be NULL): [defect corrections,
 * FP32 iterations, final FP64 relative residual]. */
int bpx_bp_solve_mixed(bpx_bp b, double* o_x, const double* o_b, double tol, int interval,
                       int maxit, double* stats);
/* No reference counterpart (measurement): bpx_bp_solve_mixed with the FP32
 * CG operator timed inside the solve: each application's Ax by HIP events
 * its own dispatch stamps (start and end), its gather from that end to an
 * event after it (the first 8192 applications). out (host): [mean Ax us, mean gather us, applications
 * timed, fused 0/1]. */
int bpx_bp_time_operator_in_solve(bpx_bp b, double* o_x, const double* o_b, double tol, int interval,
                                  int maxit, double* out);

/* ---------------------------------------------------------------- batch --
 * No reference counterpart. A ragged batch of BP5 problems (box meshes of
 * any size and degree, each set up as bp5_t does, BP/BP5/src/bp5Setup.cpp:
 * 29-95), solved together by one CG per problem (the recurrence of
 * cg::Solve, libs/core/linearSolverCG.cpp:62-129) whose scalars stay on the
 * device; blocks of iterations are captured once in a hipGraph and
 * replayed until every problem has stopped. Problems of one degree share one
 * Ax launch; the setup times the workgroup shapes of each degree's Ax (all
 * give identical results) and keeps the fastest. Vectors are the problems'
 * global vectors concatenated in order (bpx_batch_offsets). */
int bpx_batch_create(bpx_platform p, int nprob, const bpx_mesh_opts* opts, double lambda,
                     bpx_batch* out);
int bpx_batch_destroy(bpx_batch b);
/* counts: nprob Ntot ngroups nchunks nsupers NLocal replays-of-last-solve */
int bpx_batch_info(bpx_batch b, int64_t* out, int capacity);
/* off[0..nprob]: problem i owns [off[i], off[i+1]) of the vectors */
int bpx_batch_offsets(bpx_batch b, int64_t* off);
/* per degree group, 6 values: N, chosen shape, elements per workgroup,
 * time (us) of shapes 0, 1, 2 (-1 = not available). Returns the count. */
int bpx_batch_tuning(bpx_batch b, double* out, int capacity);
/* x_i = A_i^{-1} b_i to ||r_i|| <= max(tol ||b_i||, tol) or maxit
 * iterations, x from 0. iterations, rnorm: host arrays of nprob (may be
 * NULL). BPX_ERR_BREAKDOWN if any problem broke down. */
int bpx_batch_solve(bpx_batch b, double* o_x, const double* o_b, double tol, int maxit,
                    int* iterations, double* rnorm);
/* No reference counterpart (the reference solves in FP64 only): every
 * problem solved by FP32 CG with FP64 defect corrections, the recurrence
 * of bpx_bp_solve_mixed (FP32 vectors and operator with the geometric
 * factors rounded to FP32, FP64 dot products; every `interval` FP32
 * iterations x += c and r = b - A x in FP64, where ||b - A x|| <= tol ||b||
 * is tested), x from 0, problems in lockstep segments replayed from one
 * hipGraph. Host arrays of nprob (may be NULL): FP32 iterations, FP64
 * corrections, final ||b - A x|| / ||b||. tol > 0, interval >= 1, maxit >= 1;
 * BPX_ERR_BREAKDOWN if a residual became non-finite. */
int bpx_batch_solve_mixed(bpx_batch b, double* o_x, const double* o_b, double tol, int interval,
                          int maxit, int* iterations, int* updates, double* relres);

#ifdef __cplusplus
}
#endif
#endif /* BPX_H */
