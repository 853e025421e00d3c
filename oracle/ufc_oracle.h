/*
 * ufc_oracle.h -- FP64 CPU restatement of the UFC static-stiffness path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * library and the timed CPU baseline; it is never linked into, loaded by or
 * called from the product path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Conventions (mirroring the Fortran reference memory layout):
 *   coords   : (3, n_nodes) column-major  == [n_nodes][3] row-major, FP64
 *   conn     : (8, n_elems) column-major  == [n_elems][8], 0-based node ids,
 *              Abaqus C3D8 order (PH_Elem_C3D8.f90:1930-1933)
 *   Ke       : (24, 24) column-major, local dof j -> node j/3, comp j%3
 *   equation : eq = 3*node + d (0-based; RT_Asm_DofMap.f90:199-205 is the
 *              1-based form eq = 3(node-1)+d)
 *   CSR      : 0-based row_ptr (n+1) / col_idx (nnz), columns ascending
 *              (RT_CSR_FromTripletMerged, RT_Solv_Sparse.f90:336-447)
 */
#ifndef UFC_ORACLE_H
#define UFC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded counter-based fields ---------------------------------------- */
uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, uint64_t counter);
double   orc_rng_u01(uint64_t seed, uint64_t stream, uint64_t counter);
void orc_field_lognormal(int64_t n, uint64_t seed, double mu, double sigma, double* out);
void orc_field_uniform(int64_t n, uint64_t seed, double a, double b, double* out);

/* ---- structured box mesh ------------------------------------------------- */
void orc_mesh_box(int nx, int ny, int nz, double lx, double ly, double lz,
                  double* coords, int32_t* conn);
void orc_mesh_jitter(int nx, int ny, int nz, double lx, double ly, double lz,
                     double frac, uint64_t seed, double* coords);

/* ---- C3D8 element -------------------------------------------------------- */
void orc_c3d8_shape(double xi, double eta, double zeta, double N[8], double dNdxi[24]);
int  orc_c3d8_dndx(const double* coords, double xi, double eta, double zeta,
                   double dNdx[24], double* detJ);
int  orc_c3d8_ke(const double* coords, double E, double nu, double* Ke);
int  orc_c3d8_ke_dgp(const double* coords, const double* Dgp, double* Ke);
/* variants 0,1,5 standard, 2 reduced, 3 selective, 4 incompatible, 6 27-point */
int  orc_c3d8_ke_variant(const double* coords, double E, double nu, int variant, double* Ke);
void orc_c3d8_strain(const double* coords, const double* ue, double* eps_gp);
void orc_c3d8_fint(const double* coords, const double* sig_gp, double* Fe);
void orc_c3d8_face_pressure(const double* coords, int face, double p, double* Fe);
double orc_face_area(const double* coords, int face);

/* ---- CSR pattern / assembly ------------------------------------------------ */
int64_t orc_csr_pattern(int32_t n_nodes, int32_t n_elems, const int32_t* conn,
                        int32_t* row_ptr, int32_t* col_idx);
int64_t orc_csr_from_triplets(int32_t n_rows, int64_t n_trip, const int32_t* ti,
                              const int32_t* tj, const double* tv,
                              int32_t* row_ptr, int32_t* col_idx, double* val);
/* returns the number of element entries with no slot in the pattern (an
 * error); *n_degenerate (if not NULL) receives the number of elements with a
 * Gauss point skipped for |det J| <= 1e-12, as PH_Elem_C3D8_StiffMatrix skips
 * it (PH_Elem_C3D8.f90:775) */
int  orc_assemble(int32_t n_elems, const int32_t* conn, const double* coords,
                  const double* E, double nu, int32_t n_rows, const int32_t* row_ptr,
                  const int32_t* col_idx, double* val, int nthreads, int32_t* n_degenerate);
int  orc_assemble_dgp(int32_t n_elems, const int32_t* conn, const double* coords,
                      const double* Dgp, int32_t n_rows, const int32_t* row_ptr,
                      const int32_t* col_idx, double* val, int nthreads);
int  orc_apply_penalty(int32_t n_bc, const int32_t* eq, const double* value, double penalty,
                       const int32_t* row_ptr, const int32_t* col_idx, double* val, double* F);
void orc_face_loads_traction(int32_t n_faces, const int32_t* elem, const int32_t* face,
                             const int32_t* conn, const double* coords,
                             const double* traction, double* F);
void orc_face_loads_pressure(int32_t n_faces, const int32_t* elem, const int32_t* face,
                             const int32_t* conn, const double* coords, double p, double* F);

/* ---- solvers --------------------------------------------------------------- */
void orc_spmv(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
              const double* x, double* y, int nthreads);
void orc_spmv_t(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
                const double* x, double* y);
int  orc_pcg(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
             const double* b, double* x, int32_t max_iter, double rtol,
             int32_t* niter, double* rnorm, int nthreads);
int  orc_pcg_mark(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
                  const double* b, double* x, int32_t max_iter, double rtol, double rtol_mark,
                  int32_t* niter, double* rnorm, int32_t* iter_mark, double* sec_mark,
                  int nthreads, int32_t report_every);
int  orc_cg_transpose(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                      const double* val, const double* rhs, double* lam, double tol,
                      int32_t max_iter, int32_t* niter, double* rel);

/* ---- adjoint sensitivity ----------------------------------------------------- */
void orc_sensitivity(int32_t n_elems, const int32_t* conn, const double* coords, double nu,
                     const double* dEdrho, const double* u, const double* lam, double* g,
                     int nthreads);

/* ---- J2 plasticity -------------------------------------------------------- */
int  orc_j2_update(const double* props, const double* deps, double* sigma, double* peeq,
                   double* eps_p, double* Dt, int32_t* yielded);
/* props (11): E, nu, sigma_y0, H, hardening type (1 linear, 2 Swift, 3 Voce),
 * max local iterations (0: 25), K_swift, n_swift, eps0_swift, sigma_inf,
 * delta_voce; local_iters receives the local Newton iterations used */
int  orc_j2_update_ex(const double* props, const double* deps, double* sigma, double* peeq,
                      double* eps_p, double* Dt, int32_t* yielded, int32_t* local_iters);
int64_t orc_j2_internal_force(int32_t n_nodes, int32_t n_elems, const int32_t* conn,
                              const double* coords, const double* props, const double* u,
                              const double* u_n, const double* state_n, double* state,
                              double* Dgp, double* Fi, int nthreads);
int  orc_j2_newton_step(int32_t n_nodes, int32_t n_elems, const int32_t* conn,
                        const double* coords, const double* props,
                        int32_t n_bc, const int32_t* bc_eq, const double* bc_val, double penalty,
                        const double* state_n, double* state, double* u,
                        int32_t max_newton, double tol_r, double pcg_rtol,
                        int32_t* n_newton, double* rnorm_out);

#ifdef __cplusplus
}
#endif
#endif
