/*
 * ufc_hip.h -- C-ABI of the MI355X (gfx950) static-stiffness library.
 *
 * Drop-in for UFC's hot path: C3D8 element stiffness (L4_PH), fixed-pattern
 * CSR assembly with penalty BCs and face loads (L5_RT), CSR SpMV and
 * Jacobi-PCG in FP64, FP32 and FP32-in-FP64-refinement form plus the adjoint
 * solve and element sensitivities (L2_NM / new), and the J2 material update
 * (L4_PH Material).  Each entry point takes UFC's call-bundle shape
 * (desc, state, algo, ctx, args) as C structs of plain pointers and sizes;
 * there are no torch or HIP types in any signature.
 *
 * Citations are to the reference (zning8251-jpg/ufc-fem-kernel), paths
 * relative to its root, "file:line" of the interface each entry replaces.
 *
 * Conventions
 *   - Array arguments are plain pointers to host memory or to device memory
 *     of the context's GPU (both are accepted; copies use unified addressing).
 *     Mesh arrays of ufc_mesh_desc must be host memory.
 *   - Index arrays carry an explicit index_base: 1 for Fortran callers, 0 for C.
 *   - Equations are numbered eq = 3(node-1)+d, d = 1..3, 1-based
 *     (RT_Asm_DofMap.f90:199-205); every node has 3 equations, clamped ones
 *     included (Dirichlet conditions are diagonal penalties).
 *   - Voigt order [xx, yy, zz, xy, yz, xz] with engineering shear strains
 *     (PH_Elem_C3D8.f90:550-573).
 *   - Every call returns ufc_status.status_code (also the return value); it
 *     never aborts.  Calls on one context are not re-entrant; calls are
 *     blocking (results are complete when the call returns).
 */
#ifndef UFC_HIP_H
#define UFC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: IF_STATUS_* of L1_IF/Error/IF_Err_Brg.f90:95-109 ---- */
#define UFC_STATUS_OK 0
#define UFC_STATUS_INVALID 1
#define UFC_STATUS_ERROR 2
#define UFC_STATUS_CONVERGED 3
#define UFC_STATUS_NOT_CONVERGED 4
#define UFC_STATUS_WARN 6
#define UFC_STATUS_MEM_ERROR 7
#define UFC_STATUS_FATAL 10
#define UFC_STATUS_UNSUPPORTED 11

/* ---- precision selectors ---- */
#define UFC_PREC_FP64 0  /* FP64 element integration, values, vectors        */
#define UFC_PREC_FP32 1  /* FP32 element integration, values, vectors        */
#define UFC_PREC_MIXED 2 /* FP32 values + inner PCG, FP64 outer refinement   */

/* ---- material selectors for assembly ---- */
#define UFC_MAT_ELASTIC 0 /* isotropic, E per element, one nu               */
#define UFC_MAT_J2 1      /* per-Gauss-point tangent from the J2 state      */

/* ---- face load kinds ---- */
#define UFC_LOAD_TRACTION 0 /* traction vector x area/4 per face node       */
#define UFC_LOAD_PRESSURE 1 /* consistent pressure along the outward normal */

/* Mirrors ErrorStatusType (L1_IF/Error/IF_Err_Def.f90:121-136): code,
 * severity, category, message(512), source(64), line number. */
typedef struct ufc_status {
  int32_t status_code;
  int32_t severity;
  int32_t category;
  int32_t line_number;
  char message[512];
  char source[64];
} ufc_status;

/* Opaque context: owns every device buffer (the NM_Solver_Ctx work arrays,
 * NM_Solv_Def.f90:122-132, and the CSR matrix the caller owns on the Fortran
 * side, RT_Solv_Sparse.f90:67-70).  One GPU per context. */
typedef struct ufc_ctx ufc_ctx;

/* Mesh description (L3 mesh pool data the assembly reads; element coordinates
 * as in elem_coords_cache(3,8,iel), PH_Elem_Domain.f90:167-168). */
typedef struct ufc_mesh_desc {
  int32_t n_nodes;
  int32_t n_elems;
  int32_t index_base;    /* base of conn                                    */
  int32_t reserved;
  const double* coords;  /* (3, n_nodes) column-major                       */
  const int32_t* conn;   /* (8, n_elems) Abaqus C3D8 node order             */
} ufc_mesh_desc;

/* ====================================================================== */
/* Context                                                                */
/* ====================================================================== */

/* Creates a context on HIP device `device` (no counterpart: the reference
 * keeps its state in the g_ufc_global singleton, L0_Global). */
int ufc_ctx_create(int32_t device, ufc_ctx** ctx, ufc_status* st);
int ufc_ctx_destroy(ufc_ctx* ctx, ufc_status* st);

/* Sizes after ufc_rt_asm_build_pattern. */
typedef struct ufc_rt_asm_info {
  int64_t n_dof;
  int64_t nnz;
  int32_t n_nodes;
  int32_t n_elems;
  int32_t max_row_nnz;
  int32_t nb_index_bytes; /* neighbour index bytes per node or neighbour    */
                          /* the FP32 SpMV view reads: 0 (symmetric block   */
                          /* diagonals), 1 (stencil pattern), 2, 4          */
  int32_t fp32_view;      /* UFC_FP32_VIEW_*: by default the fastest the    */
                          /* mesh allows, in the order SYMDIAG, PATTERN,    */
                          /* OFF16, ID32                                     */
  int32_t sym_tile;       /* nodes per tile of the SYMDIAG layout (one per  */
                          /* thread of the FP32 SpMV's workgroup): 256, or  */
                          /* 512 on meshes of at least 800,000 nodes; 0     */
                          /* when the mesh has no SYMDIAG view              */
} ufc_rt_asm_info;
int ufc_ctx_info(ufc_ctx* ctx, ufc_rt_asm_info* info, ufc_status* st);

/* Index views of the FP32 operator (the FP32 SpMV inside FP32 and MIXED
 * solves and ufc_nm_spmv in FP32).  All apply the same difference-form K:
 *   SYMDIAG: when the offsets m - n > 0 of all neighbour pairs take at most
 *            13 distinct values (a lexicographically numbered structured
 *            box), each off-diagonal block is stored once, for its upper
 *            node, and applied as K(n,m) and K(n,m)^T; a structured box in
 *            any other node numbering gets the view in its lexicographic
 *            numbering (ufc_mesh_box_numbering), the FP32 inner solve then
 *            running in that numbering;
 *   PATTERN: at most 256 distinct neighbour-offset lists, all offsets below
 *            2^15: one pattern byte per node;
 *   OFF16:   16-bit offset per neighbour;  ID32: node id per neighbour. */
#define UFC_FP32_VIEW_ID32 0
#define UFC_FP32_VIEW_OFF16 1
#define UFC_FP32_VIEW_PATTERN 2
#define UFC_FP32_VIEW_SYMDIAG 3

/* ====================================================================== */
/* L4_PH: element kernels                                                 */
/* ====================================================================== */

/* C3D8 variants: the PH_ELEM_C3D8_VARIANT_* values (PH_Elem_C3D8.f90:111-116)
 * as PH_Elem_C3D8_StiffMatrixByVariant (:400-417) computes them.  0, 1 and 5
 * are the standard 2x2x2 element; 3 (the reference's HYBRID) is the
 * selective (B-bar) element.  REDUCED uses the hourglass base vectors
 * xi.eta.zeta, eta.zeta, xi.zeta, xi.eta projected off the linear fields
 * (the reference's 4th vector is the linear mode -zeta, which stiffens rigid
 * rotations); INCOMPAT uses the modes 1 - xi_k^2 of each displacement
 * component with J0-scaled gradients (the reference's three scalar modes
 * fail the patch test). */
#define UFC_VARIANT_STANDARD 1
#define UFC_VARIANT_REDUCED 2   /* C3D8R: 1 point + hourglass stiffness    */
#define UFC_VARIANT_SELECTIVE 3 /* B-bar (reference: VARIANT_HYBRID)       */
#define UFC_VARIANT_INCOMPAT 4  /* C3D8I: incompatible modes, condensed    */
#define UFC_VARIANT_27PT 6      /* 3x3x3 Gauss                             */

typedef struct ufc_ph_ke_algo {
  int32_t precision; /* UFC_PREC_FP64 or UFC_PREC_FP32 (Ke written in it)  */
  int32_t variant;   /* UFC_VARIANT_* (0: standard)                         */
} ufc_ph_ke_algo;

/* PH_Elem_Eval_Ke_Arg (PH_Elem_Def.f90:218-222) for a batch of C3D8. */
typedef struct ufc_ph_ke_args {
  int32_t n_elems;
  int32_t reserved;
  const double* coords; /* (3, 8, n_elems)                                  */
  const double* young;  /* (n_elems): mat_props_in(1)                        */
  double nu;            /* mat_props_in(2)                                   */
  void* ke;             /* (24, 24, n_elems) column-major, FP64 or FP32 out  */
} ufc_ph_ke_args;

/* Batched Ke = sum_gp detJ w B^T D B, 2x2x2 Gauss: replaces
 * PH_Elem_Eval_Ke(desc,algo,arg,status) (PH_Elem_Eval.f90:35-81) ->
 * Compute_Ke_C3D8 (PH_ElemKeDispatch.f90:207-237) ->
 * PH_Elem_C3D8_StiffMatrix (PH_Elem_C3D8.f90:759-780); algo->variant
 * selects PH_Elem_C3D8_StiffMatrixByVariant's element (:400-417). */
int ufc_ph_c3d8_eval_ke(ufc_ctx* ctx, const ufc_ph_ke_algo* algo, ufc_ph_ke_args* args,
                        ufc_status* st);

/* ufc_ph_c3d8_eval_ke without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  args->coords, young and
 * ke must be device memory of the context's GPU (host arrays return
 * UFC_STATUS_INVALID: staging them would allocate).  The incompatible-mode
 * element's condensation factor (24 x 9 doubles per element) goes to a
 * static 16 MiB device array that each GPU holds until the process ends, so
 * a UFC_VARIANT_INCOMPAT batch is limited to 9709 elements per call
 * (UFC_STATUS_INVALID beyond); the other variants need no scratch and take
 * any n_elems.  Results are bitwise those of ufc_ph_c3d8_eval_ke.  Calls that
 * overlap on one GPU (several contexts or threads) race on the array. */
int ufc_ph_c3d8_eval_ke_noalloc(ufc_ctx* ctx, const ufc_ph_ke_algo* algo, ufc_ph_ke_args* args,
                                ufc_status* st);

/* Adjoint element sensitivities for E_e = E_e(rho_e):
 *   grad[e] = -dEdrho[e] * lambda_e^T K_e(E=1) u_e,
 * recomputing K_e(E=1) instead of storing it (after an elastic assembly
 * with another variant: that element, as lambda_e^T K_e u_e / E_e with the
 * assembly's E_e and nu, which must then be non-zero).  No counterpart in the
 * reference (it has no sensitivity kernel); consumes the output of
 * ufc_nm_adjoint_solve (NM_Adjoint_Solve, NM_Solv_GMRESTranspose.f90:229-291). */
typedef struct ufc_ph_sens_args {
  const double* u;      /* (n_dof) forward solution                          */
  const double* lambda; /* (n_dof) adjoint solution                          */
  const double* dedrho; /* (n_elems) dE_e/drho_e                             */
  double nu;
  double* grad;         /* (n_elems) out                                     */
} ufc_ph_sens_args;
int ufc_ph_c3d8_sensitivity(ufc_ctx* ctx, const ufc_ph_sens_args* args, ufc_status* st);

/* Isotropic hardening laws of PH_J2_ComputeHardening /
 * PH_J2_ComputeHardeningTangent (PH_Mat_Plast_J2_Iso_Core.f90:549-605), the
 * PH_MAT_J2_HARD_* values (:33-36). */
#define UFC_J2_HARD_LINEAR 1 /* sigma_y = sigma_y0 + H peeq  (0 selects it too) */
#define UFC_J2_HARD_SWIFT 2  /* sigma_y = K (eps0 + peeq)^n                    */
#define UFC_J2_HARD_VOCE 3   /* sigma_y = sigma_y0 + Q (1 - exp(-delta peeq))   */
#define UFC_J2_HARD_AF 4     /* Armstrong-Frederick backstress: not supported   */

/* J2 props (PH_J2_Props, PH_Mat_Plast_J2_Iso_Core.f90:52-91: PH_J2_Cfg_Elastic,
 * _Yield, _Harden, _Control).  The radial return runs the reference's local
 * Newton on R(dg) = q_trial - 3G dg - sigma_y(peeq_n + dg) (:355-456) from
 * dg = f_trial / (3G + H'(peeq_n)), to |R| < 1e-10 sigma_y0, for at most
 * max_local_iter iterations; no convergence is a local failure (the update
 * returns UFC_STATUS_ERROR, the reference's pnewdt = 0.25 cut-back signal,
 * :419-427).  The tangent uses H'(peeq_n + dg).  The reference's
 * Armstrong-Frederick backstress (:449-453, C_af, gamma_af) is updated but
 * never subtracted from the trial deviator (:254-317), so it cannot change a
 * result; UFC_J2_HARD_AF or a non-zero af_c / af_gamma returns
 * UFC_STATUS_UNSUPPORTED from ufc_ph_j2_init rather than pretend to model
 * kinematic hardening. */
typedef struct ufc_ph_j2_props {
  double young;
  double nu;
  double sigma_y0;
  double hardening;       /* H of UFC_J2_HARD_LINEAR                        */
  int32_t hardening_type; /* UFC_J2_HARD_*; 0 = UFC_J2_HARD_LINEAR          */
  int32_t max_local_iter; /* local Newton iterations; <= 0: 25              */
                          /* (PH_MAT_J2_MAX_LOCAL_ITER, :41)                */
  double swift_k;         /* K_swift  (> 0 for UFC_J2_HARD_SWIFT)           */
  double swift_n;         /* n_swift  (>= 0)                                */
  double swift_eps0;      /* eps0_swift (> 0)                               */
  double voce_q;          /* sigma_inf of UFC_J2_HARD_VOCE                  */
  double voce_delta;      /* delta_voce (>= 0)                              */
  double af_c;            /* C_af: must be 0                                */
  double af_gamma;        /* gamma_af: must be 0                            */
} ufc_ph_j2_props;

/* Allocates the per-Gauss-point FP64 state of every element (13 doubles per
 * point: stress(6), eq. plastic strain, plastic strain(6); the statev layout
 * of PH_Mat_Plast_J2_UMAT_Core.f90:502-532 with stress in front) in a
 * committed and a trial copy, both zero. */
int ufc_ph_j2_init(ufc_ctx* ctx, const ufc_ph_j2_props* props, ufc_status* st);

/* Strain increments B(u - u_n) at every Gauss point, then
 * PH_J2_ComputeStress (PH_Mat_Plast_J2_Iso_Core.f90:133-207) from the
 * committed state into the trial state, with the consistent tangent kept on
 * the device for assembly (UFC_MAT_J2).  u, u_n: (n_dof).  n_yield, if not
 * NULL, receives the number of Gauss points on the plastic branch. */
int ufc_ph_j2_update(ufc_ctx* ctx, const double* u, const double* u_n, int64_t* n_yield,
                     ufc_status* st);
/* trial -> committed (end of a converged increment, RT_Step_Exec.f90:459-470) */
int ufc_ph_j2_commit(ufc_ctx* ctx, ufc_status* st);
/* copy the committed state out / in (checkpoint; RT_Out_Restart.f90) */
int ufc_ph_j2_get_state(ufc_ctx* ctx, double* state, ufc_status* st);
int ufc_ph_j2_set_state(ufc_ctx* ctx, const double* state, ufc_status* st);

/* ====================================================================== */
/* L5_RT: pattern, assembly, boundary conditions, loads, residual          */
/* ====================================================================== */

/* Uploads the mesh and builds the merged, column-sorted CSR pattern
 * (RT_CSR_FromTripletMerged, RT_Solv_Sparse.f90:336-447) directly from node
 * adjacency, plus the precomputed element-to-slot map that lets assembly
 * write every CSR value once without atomics.  Replaces the full-build path of
 * RT_Asm_GlobalStiffness (RT_Asm_Solv.f90:1902-1965).  Any hex mesh whose
 * nodes have at most 256 neighbours (UFC_STATUS_UNSUPPORTED otherwise); a
 * node may belong to any number of elements (nodes beyond 8 elements or 32
 * neighbours are assembled by general kernels, one node per workgroup). */
int ufc_rt_asm_build_pattern(ufc_ctx* ctx, const ufc_mesh_desc* mesh, ufc_rt_asm_info* info,
                             ufc_status* st);

/* Exports row_ptr (n_dof+1) and col_idx (nnz) in the caller's index base,
 * the RT_CSRMatrix rowPtr/colInd (NM_Assem_Sparse.f90:45-56). */
int ufc_rt_asm_get_pattern(ufc_ctx* ctx, int32_t index_base, int32_t* row_ptr,
                           int32_t* col_idx, ufc_status* st);

typedef struct ufc_rt_asm_algo {
  int32_t precision; /* FP64 values; FP32 values; MIXED = FP32 values        */
  int32_t material;  /* UFC_MAT_ELASTIC or UFC_MAT_J2                        */
  int32_t variant;   /* UFC_VARIANT_* for UFC_MAT_ELASTIC (0: standard); the  */
                     /* FP64 refinement residual, ufc_rt_residual and         */
                     /* ufc_ph_c3d8_sensitivity use the same element. J2:     */
                     /* standard only (others: UFC_STATUS_UNSUPPORTED).       */
  int32_t reserved;
} ufc_rt_asm_algo;

typedef struct ufc_rt_asm_args {
  const double* young; /* (n_elems) for UFC_MAT_ELASTIC, else ignored        */
  double nu;
} ufc_rt_asm_args;

/* K = sum_e L_e^T K_e L_e into the fixed pattern, then the registered
 * penalties on the diagonal.  Replaces the reuse path of
 * RT_Asm_GlobalStiffness (RT_Asm_Solv.f90:1840-1900) with
 * RT_Asm_ScatterKe_CSR_Atomic (:3467-3493). */
int ufc_rt_asm_assemble_k(ufc_ctx* ctx, const ufc_rt_asm_algo* algo,
                          const ufc_rt_asm_args* args, ufc_status* st);

/* Copies the values of precision `precision` (FP64 -> double, FP32 -> float)
 * in CSR order.  FP64: K itself.  FP32: K in the difference form the FP32
 * SpMV applies -- every node's 3x3 self block is zero apart from the penalty
 * diagonal; the self blocks of K follow from K 1 = 0 per block row,
 * K(n,n) = -sum_{m != n} K(n,m). */
int ufc_rt_asm_get_values(ufc_ctx* ctx, int32_t precision, void* values, ufc_status* st);

/* Dirichlet conditions as diagonal penalties, RT_Asm_ApplyBC _Idx path
 * (RT_Asm_Solv.f90:318-344): K(eq,eq) += penalty, F(eq) += penalty*value.
 * The set is kept and re-applied by every later ufc_rt_asm_assemble_k;
 * calling again replaces the set (penalties on the current K are swapped). */
typedef struct ufc_rt_bc_args {
  int32_t n_bc;
  int32_t index_base;
  const int32_t* eq;    /* (n_bc) equation ids                               */
  const double* value;  /* (n_bc) prescribed values, NULL = 0                */
  double penalty;       /* 1e12 in the reference                             */
} ufc_rt_bc_args;
int ufc_rt_asm_apply_bc(ufc_ctx* ctx, const ufc_rt_bc_args* args, ufc_status* st);

/* Face loads into the load vector F (RT_Asm_GlobalLoad surface branch,
 * RT_Asm_Solv.f90:1614-1647; lumping of LoadBC_DistributeLoad_ToSurface,
 * MD_LBC_Apply.f90:258-293; consistent pressure of
 * PH_Elem_C3D8_FormFacePressure, PH_Elem_C3D8.f90:2360-2541, positive
 * pressure along the outward normal).  Faces are numbered 1..6 as in
 * PH_ELEM_C3D8_FACE_NODES (PH_Elem_C3D8.f90:120-121) when index_base is 1. */
typedef struct ufc_rt_load_args {
  int32_t n_faces;
  int32_t kind;         /* UFC_LOAD_TRACTION or UFC_LOAD_PRESSURE            */
  int32_t index_base;   /* base of elem and face                             */
  int32_t accumulate;   /* 0: F is cleared first                             */
  const int32_t* elem;  /* (n_faces)                                         */
  const int32_t* face;  /* (n_faces)                                         */
  double magnitude;     /* pressure (UFC_LOAD_PRESSURE)                      */
  double traction[3];   /* traction vector (UFC_LOAD_TRACTION)               */
} ufc_rt_load_args;
int ufc_rt_asm_load(ufc_ctx* ctx, const ufc_rt_load_args* args, ufc_status* st);

/* Replaces the FP64 values (nnz, CSR order, as ufc_rt_asm_get_values returns
 * them, penalties of the current boundary conditions included): a caller's
 * own operator on the pattern, e.g. a non-symmetric tangent
 * (RT_CSRMatrix%values, NM_Assem_Sparse.f90:45-56).  Only FP64 solves, the
 * FP64 SpMVs and ufc_nm_adjoint_solve then apply; MIXED / FP32 solves and
 * ufc_rt_residual, whose operators come from the elements, return
 * UFC_STATUS_INVALID until the next ufc_rt_asm_assemble_k. */
int ufc_rt_asm_set_values(ufc_ctx* ctx, int32_t precision, const double* values, ufc_status* st);

/* Selects the FP32 view (UFC_FP32_VIEW_*, or -1 for the fastest the mesh
 * allows, the default after ufc_rt_asm_build_pattern); assembled FP32
 * values are re-laid out at once.  UFC_STATUS_UNSUPPORTED if the mesh does
 * not allow the view.  No counterpart in the reference, whose SpMV has one
 * storage format: for tests and measurements of the individual views. */
int ufc_rt_asm_set_fp32_view(ufc_ctx* ctx, int32_t view, ufc_status* st);

/* Selects the layout the FP64 SpMV reads (no counterpart in the reference,
 * whose SpMV has one storage format): UFC_FP64_VIEW_CSR (default), the CSR
 * with a column index per value as NM_CSR_MatVec_Optimized
 * (NM_Mtx_Sparse.f90:341-365) reads it, 12 bytes per value; or
 * UFC_FP64_VIEW_BLOCK, the same values with one 4-byte neighbour id per 3x3
 * block, 8.4 bytes per value.  Both give the same bits.  Kept for the mesh's
 * lifetime (FP64 SpMVs, PCG, adjoint). */
#define UFC_FP64_VIEW_CSR 0
#define UFC_FP64_VIEW_BLOCK 1
int ufc_rt_asm_set_fp64_view(ufc_ctx* ctx, int32_t view, ufc_status* st);

/* Sets / reads the load vector F (n_dof); penalty terms of apply_bc are
 * included when read. */
int ufc_rt_set_rhs(ufc_ctx* ctx, const double* f, ufc_status* st);
int ufc_rt_get_rhs(ufc_ctx* ctx, double* f, ufc_status* st);

/* Residual R = F - F_int(u) - P u on penalty rows (F includes P ubar):
 * RT_Asm_ComputeResidual (RT_Asm_Solv.f90:1311-1426).  F_int comes from the
 * J2 trial stresses when the J2 state exists, else from the elastic field
 * given at the last assembly.  norms[0] = ||R||_2, norms[1] = ||F_int||_2. */
int ufc_rt_residual(ufc_ctx* ctx, const double* u, double* r, double* norms, ufc_status* st);

/* ====================================================================== */
/* L2_NM: SpMV and solvers                                                */
/* ====================================================================== */

/* y = K x with the assembled values of `precision` (FP64: scalar CSR as
 * NM_CSR_MatVec_Optimized, NM_Mtx_Sparse.f90:341-365; FP32: node-block index
 * view of the same CSR).  x, y: (n_dof) FP64. */
int ufc_nm_spmv(ufc_ctx* ctx, int32_t precision, const double* x, double* y, ufc_status* st);

/* NM_Solver_Algo (NM_Solv_Def.f90:86-100) with the fields the solver reads
 * (NM_Solv_Core.f90:392-399 uses maxiter/rtol). */
typedef struct ufc_nm_solver_algo {
  int32_t precision;   /* UFC_PREC_FP64, UFC_PREC_FP32 or UFC_PREC_MIXED    */
  int32_t max_iter;    /* PCG iterations (per inner solve for MIXED)        */
  double rtol;         /* ||r||/||b|| target (outer, FP64, for MIXED)        */
  double inner_rtol;   /* MIXED: inner FP32 PCG tolerance                    */
  int32_t max_outer;   /* MIXED: refinement steps                            */
  int32_t check_every; /* 0: the GPU reports each finished iteration into   */
                       /* pinned host memory and the host keeps 3 queued;    */
                       /* > 0: the host polls after every check_every        */
                       /* iterations; either way the device decides the      */
                       /* stopping iteration                                 */
  double replace_rtol; /* MIXED: inside each inner solve the FP32 residual  */
                       /* is replaced by the FP64 b - K u every time it has  */
                       /* dropped by replace_rtol (<= 0: 0.1), and the inner */
                       /* PCG keeps its search direction across refinement   */
                       /* steps; >= 1: no replacement, each step restarts    */
  int32_t krylov;      /* ufc_nm_adjoint_solve: UFC_KRYLOV_PCG (K symmetric) */
                       /* or UFC_KRYLOV_GMRES (any K, FP64)                  */
  int32_t restart;     /* GMRES restart length (<= 0: 30, the reference's)   */
} ufc_nm_solver_algo;
#define UFC_KRYLOV_PCG 0
#define UFC_KRYLOV_GMRES 1

/* NM_Solver_State (NM_Solv_Def.f90:102-120). */
typedef struct ufc_nm_solver_state {
  double* x;          /* (n_dof) in: x0 / out: solution                      */
  const double* b;    /* (n_dof) right-hand side; NULL = the load vector F   */
  int32_t niter;      /* out: PCG iterations (sum over inner solves)         */
  int32_t n_outer;    /* out: refinement steps (MIXED)                       */
  double rnorm;       /* out: ||r||/||b|| at exit (MIXED: FP64 true residual) */
  int32_t converged;  /* out                                                 */
  int32_t x0_zero;    /* in: 1 = ignore x on entry and start from 0          */
} ufc_nm_solver_state;

/* Jacobi-preconditioned CG on the assembled K: NM_Solver_PCG
 * (NM_Solv_Core.f90:363-423) with NM_Solver_Jacobi_Precond (:131-147).
 * MIXED wraps FP32 PCG (FP64 dot accumulation) in FP64 iterative refinement:
 * r = F - K u recomputed element by element in FP64, u += d in FP64.
 * Returns UFC_STATUS_OK when converged, UFC_STATUS_INVALID when max_iter is
 * reached (the reference's code, NM_Solv_Core.f90:419-422), UFC_STATUS_ERROR
 * on a breakdown (p.Ap <= 0 or non-finite). */
int ufc_nm_pcg(ufc_ctx* ctx, const ufc_nm_solver_algo* algo, ufc_nm_solver_state* state,
               ufc_status* st);

/* K^T lambda = dJdu, replacing NM_Adjoint_Solve (NM_Solv_GMRESTranspose.f90:
 * 229-291).  algo->krylov = UFC_KRYLOV_PCG: the same Jacobi-PCG as ufc_nm_pcg
 * (K symmetric, so K^T = K; the reference's is_symmetric path,
 * NM_CG_Solve_Transpose, :118-224, unpreconditioned).  UFC_KRYLOV_GMRES: the
 * is_symmetric = .FALSE. path, GMRES(restart) on K^T with right Jacobi
 * preconditioning over the FP64 values (algo->precision must be FP64; the
 * reference's NM_GMRES_Solve_Transpose forms K^T and then runs a damped
 * Richardson placeholder, :337-393); state->n_outer receives the number of
 * restarts.  lambda_0 = 0; state->x receives lambda; state->b is ignored. */
int ufc_nm_adjoint_solve(ufc_ctx* ctx, const ufc_nm_solver_algo* algo, const double* dJdu,
                         ufc_nm_solver_state* state, ufc_status* st);

/* y = K^T x with the FP64 values (precision must be UFC_PREC_FP64), read in
 * place through the mirrored block of each neighbour: replaces
 * NM_CSR_Transpose + SpMV (NM_Mtx_SpMVCSRTranspose.f90:46-145). */
int ufc_nm_spmv_t(ufc_ctx* ctx, int32_t precision, const double* x, double* y, ufc_status* st);

/* Device time in ms of the last assembly, the last solve and the last J2
 * update on this context (HIP events on the context's stream), and the PCG
 * iteration counts. */
typedef struct ufc_timing {
  double assemble_ms;
  double solve_ms;
  double spmv_ms_avg;  /* mean SpMV kernel time of the last solve (events)   */
  int32_t spmv_calls;
  int32_t n_replace;   /* MIXED: FP64 residual replacements inside the inner */
                       /* solves of the last solve                           */
  double j2_update_ms; /* device time of the last ufc_ph_j2_update's kernel  */
} ufc_timing;
int ufc_ctx_timing(ufc_ctx* ctx, ufc_timing* t, ufc_status* st);
/* per_kernel = N > 0 brackets every N-th SpMV launch with a HIP event pair
 * so that spmv_ms_avg is the mean kernel duration over the sampled launches
 * (an event record costs a few microseconds of queue time, so N > 1 keeps a
 * timed solve close to an untimed one); 0 switches it off.  No counterpart
 * in the reference, whose timers are CPU_TIME around whole phases
 * (RT_Asm_Global.f90:354,399). */
int ufc_ctx_enable_timing(ufc_ctx* ctx, int32_t per_kernel, ufc_status* st);

/* ====================================================================== */
/* Mesh and field generators for the structured benchmark inputs          */
/* (no counterpart in the reference, which reads meshes from input decks) */
/* ====================================================================== */

/* Box [0,lx]x[0,ly]x[0,lz] of nx*ny*nz C3D8, node id i + (nx+1)(j + (ny+1)k),
 * elements in the same order; conn in index_base.  jitter > 0 moves interior
 * nodes by U[-jitter, jitter] x h per coordinate (seeded). */
int ufc_mesh_box(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                 double jitter, uint64_t seed, int32_t index_base, double* coords,
                 int32_t* conn, ufc_status* st);
/* Lexicographic numbering of a mesh that is a structured box of C3D8 under
 * any node numbering (host only, no GPU): perm[n] (0-based) = i + NX (j + NY
 * k) for node n's lattice position, recovered from the element edges;
 * UFC_STATUS_UNSUPPORTED when the mesh is not such a box.  The library uses
 * it for the symmetric FP32 view of meshes whose own numbering is not
 * lexicographic; a caller may use it to renumber its mesh. */
int ufc_mesh_box_numbering(int32_t n_nodes, int32_t n_elems, int32_t index_base,
                           const int32_t* conn, int32_t* perm, ufc_status* st);
/* Per-element fields from a counter-based generator keyed by (seed, element):
 * kind 0: lognormal exp(a + b z); kind 1: uniform a + (b - a) u. */
int ufc_field(int32_t kind, int64_t n, uint64_t seed, double a, double b, double* out,
              ufc_status* st);

/* ====================================================================== */
/* L5_RT: restart                                                          */
/* ====================================================================== */

/* The step data of MD_RestartData as RT_Out_RestartSave writes it
 * (L5_RT/Output/RT_Out_Restart.f90:189-306; saved per converged increment by
 * RT_Step_Exec.f90:551-558). */
typedef struct ufc_rt_restart_data {
  double time;           /* step time at the end of the increment          */
  double lambda;         /* load factor                                    */
  double residual_norm;  /* last Newton residual norm                      */
  int32_t increment;     /* converged increments so far                    */
  int32_t iterations;    /* Newton iterations of the last increment        */
  int32_t converged;     /* 1 / 0                                          */
  int32_t n_dof;         /* see below; set by a load                       */
  double* u;             /* (n_dof) host or device; v = a = 0 (static)     */
} ufc_rt_restart_data;
/* With a context, u holds the context's n_dof values (host or device) and
 * data->n_dof is ignored on entry.  ctx may be NULL when with_state is 0:
 * u is then a host array of data->n_dof doubles (save) or with room for
 * data->n_dof, which must equal the file's (load). */

/* Writes a restart file: the record sequence of RT_Out_RestartSave as
 * Fortran unformatted sequential records (4-byte length markers, the files
 * IF_IO_FORMAT_BINARY opens, L1_IF/IO/IF_IO_File.f90:219-221):
 *   [version=1, 20260204, n_dof] (3 x int32), u, v, a (n_dof doubles each),
 *   [time, increment, lambda] (f64, i32, f64), [residual_norm, iterations,
 *   converged] (f64, i32, i32), [n_dof] (the reference's checksum),
 * so RT_Out_RestartRestore reads it; v and a are written as zeros (the
 * reference writes them only when allocated but always reads them back).
 * with_state = 1 appends the committed J2 state of ctx after those records:
 *   [n_elems, 8, 6, 7, n_chunks] (int32), then per chunk of at most 2^21
 *   elements one record of 13 doubles per Gauss point (stress(6), then the
 *   UMAT statev: eq. plastic strain, plastic strain(6),
 *   PH_Mat_Plast_J2_UMAT_Core.f90:521-531), elements in order.
 * The reference's reader stops before these records.  Errors as
 * RT_Out_RestartSave: empty path or no u -> UFC_STATUS_INVALID, the file
 * cannot be written -> UFC_STATUS_ERROR. */
int ufc_rt_restart_save(ufc_ctx* ctx, const char* path, const ufc_rt_restart_data* data,
                        int32_t with_state, ufc_status* st);
/* Reads a restart file written as above (RT_Out_RestartRestore,
 * RT_Out_Restart.f90:45-187): fills data (u must hold ctx's n_dof) and, with
 * with_state = 1, loads the J2 state into ctx (committed and trial copies;
 * ufc_ph_j2_init must have been called on the same mesh).  Missing file,
 * version != 1, n_dof <= 0 or != ctx's -> UFC_STATUS_INVALID; checksum
 * mismatch, truncated records or a missing state block ->
 * UFC_STATUS_ERROR. */
int ufc_rt_restart_load(ufc_ctx* ctx, const char* path, ufc_rt_restart_data* data,
                        int32_t with_state, ufc_status* st);

/* ====================================================================== */
/* L6_AP: input deck                                                       */
/* ====================================================================== */

/* A static C3D8 step read from an Abaqus-style input deck: replaces the
 * keyword reading of AP_Input_Domain_ParseKeyword (L6_AP/Input/Parser/
 * AP_Inp_Mgr.f90:99-125, keyword dispatch AP_Inp_KWBrg.f90:249) for the
 * keywords of this path.  Read: *NODE, *ELEMENT (TYPE=C3D8, C3D8R, C3D8H,
 * C3D8I, C3D8M, C3D8S: the PH_ELEM_C3D8_VARIANT_* of PH_Elem_C3D8.f90:110-116,
 * one per model, returned in variant), *NSET, *ELSET
 * (GENERATE), *SURFACE (TYPE=ELEMENT, S1..S6), *MATERIAL + *ELASTIC,
 * *SOLID SECTION, *BOUNDARY (dof ranges, ENCASTRE, PINNED, X/Y/ZSYMM),
 * *STEP + *STATIC, *CLOAD, *DLOAD (P1..P6), *DSLOAD, one *PART with one
 * *INSTANCE (labels "INSTANCE.label"); output requests and other keywords are
 * skipped.  Other element types, more than one part or step, and several
 * Poisson ratios return UFC_STATUS_UNSUPPORTED.  Pressures keep the
 * reference's sign (positive along the outward normal, the opposite of
 * Abaqus).  All arrays belong to the model until ufc_ap_free; indices are
 * 0-based.  Face label Sk (Pk) is column k of PH_ELEM_C3D8_FACE_NODES
 * (PH_Elem_C3D8.f90:120-121), as the reference maps *SURFACE labels
 * (MD_KW_Mapper.f90:2137-2155), returned as k - 1.  Errors carry the deck
 * line in st->line_number. */
typedef struct ufc_ap_model {
  int32_t n_nodes;
  int32_t n_elems;
  const int32_t* node_label; /* (n_nodes) labels of the deck                 */
  const int32_t* elem_label; /* (n_elems)                                    */
  const double* coords;      /* (n_nodes, 3)                                 */
  const int32_t* conn;       /* (n_elems, 8) node indices                    */
  const double* young;       /* (n_elems) E of each element's section        */
  double nu;
  int32_t n_bc;              /* prescribed equations (last value wins)       */
  int32_t n_cload;
  const int32_t* bc_eq;      /* (n_bc) eq = 3 node + dof                     */
  const double* bc_val;      /* (n_bc)                                       */
  const int32_t* cload_eq;   /* (n_cload)                                    */
  const double* cload_val;   /* (n_cload)                                    */
  int32_t n_face;            /* pressure-loaded element faces                */
  int32_t variant;           /* UFC_VARIANT_* of the element type            */
  const int32_t* face_elem;  /* (n_face)                                     */
  const int32_t* face_id;    /* (n_face) 0..5                                */
  const double* face_p;      /* (n_face) pressure                            */
  void* impl;                /* storage owned by the library                 */
} ufc_ap_model;
int ufc_ap_read_inp(const char* path, ufc_ap_model* model, ufc_status* st);
int ufc_ap_free(ufc_ap_model* model, ufc_status* st);

#ifdef __cplusplus
}
#endif
#endif /* UFC_HIP_H */
