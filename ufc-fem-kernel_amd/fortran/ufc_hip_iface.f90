! ufc_hip_iface.f90 -- ISO_C_BINDING interface to libufc_hip.so (include/ufc_hip.h).
!
! The Fortran side of the drop-in boundary: BIND(C) mirrors of the C call
! bundles and one interface per C entry point, so UFC's step driver calls the
! MI355X path with its own (desc, state, algo, ctx, args) shape.  Paths in the
! citations are relative to the reference root:
!   ufc_status            <-> ErrorStatusType      (L1_IF/Error/IF_Err_Def.f90:121-136)
!   UFC_STATUS_*          <-> IF_STATUS_*          (L1_IF/Error/IF_Err_Brg.f90:95-109)
!   ufc_nm_solver_algo    <-> NM_Solver_Algo       (L2_NM/Solver/NM_Solv_Def.f90:86-100)
!   ufc_nm_solver_state   <-> NM_Solver_State      (L2_NM/Solver/NM_Solv_Def.f90:102-120)
! Arrays pass as C_LOC of contiguous Fortran arrays; index arrays are 1-based
! (index_base = 1).
MODULE ufc_hip_iface
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE
  PRIVATE

  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: UFC_STATUS_OK = 0, UFC_STATUS_INVALID = 1, &
       UFC_STATUS_ERROR = 2, UFC_STATUS_CONVERGED = 3, UFC_STATUS_NOT_CONVERGED = 4, &
       UFC_STATUS_WARN = 6, UFC_STATUS_MEM_ERROR = 7, UFC_STATUS_FATAL = 10, &
       UFC_STATUS_UNSUPPORTED = 11
  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: UFC_PREC_FP64 = 0, UFC_PREC_FP32 = 1, &
       UFC_PREC_MIXED = 2
  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: UFC_MAT_ELASTIC = 0, UFC_MAT_J2 = 1
  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: UFC_J2_HARD_LINEAR = 1, UFC_J2_HARD_SWIFT = 2, &
                                           UFC_J2_HARD_VOCE = 3, UFC_J2_HARD_AF = 4
  ! C3D8 variants (PH_ELEM_C3D8_VARIANT_*, PH_Elem_C3D8.f90:111-116)
  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: UFC_VARIANT_STANDARD = 1, UFC_VARIANT_REDUCED = 2, &
    UFC_VARIANT_SELECTIVE = 3, UFC_VARIANT_INCOMPAT = 4, UFC_VARIANT_27PT = 6
  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: UFC_LOAD_TRACTION = 0, UFC_LOAD_PRESSURE = 1
  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: UFC_KRYLOV_PCG = 0, UFC_KRYLOV_GMRES = 1
  ! FP32 operator views (ufc_rt_asm_info%fp32_view, ufc_rt_asm_set_fp32_view)
  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: UFC_FP32_VIEW_ID32 = 0, UFC_FP32_VIEW_OFF16 = 1, &
      UFC_FP32_VIEW_PATTERN = 2, UFC_FP32_VIEW_SYMDIAG = 3
  ! FP64 SpMV layouts (ufc_rt_asm_set_fp64_view)
  INTEGER(C_INT32_T), PARAMETER, PUBLIC :: UFC_FP64_VIEW_CSR = 0, UFC_FP64_VIEW_BLOCK = 1

  TYPE, BIND(C), PUBLIC :: ufc_status
    INTEGER(C_INT32_T) :: status_code = 0
    INTEGER(C_INT32_T) :: severity = 0
    INTEGER(C_INT32_T) :: category = 0
    INTEGER(C_INT32_T) :: line_number = 0
    CHARACTER(KIND=C_CHAR) :: message(512) = C_NULL_CHAR
    CHARACTER(KIND=C_CHAR) :: source(64) = C_NULL_CHAR
  END TYPE ufc_status

  TYPE, BIND(C), PUBLIC :: ufc_mesh_desc
    INTEGER(C_INT32_T) :: n_nodes = 0
    INTEGER(C_INT32_T) :: n_elems = 0
    INTEGER(C_INT32_T) :: index_base = 1
    INTEGER(C_INT32_T) :: reserved = 0
    TYPE(C_PTR) :: coords = C_NULL_PTR
    TYPE(C_PTR) :: conn = C_NULL_PTR
  END TYPE ufc_mesh_desc

  TYPE, BIND(C), PUBLIC :: ufc_rt_asm_info
    INTEGER(C_INT64_T) :: n_dof = 0
    INTEGER(C_INT64_T) :: nnz = 0
    INTEGER(C_INT32_T) :: n_nodes = 0
    INTEGER(C_INT32_T) :: n_elems = 0
    INTEGER(C_INT32_T) :: max_row_nnz = 0
    INTEGER(C_INT32_T) :: nb_index_bytes = 0
    INTEGER(C_INT32_T) :: fp32_view = 0
    INTEGER(C_INT32_T) :: sym_tile = 0
  END TYPE ufc_rt_asm_info

  TYPE, BIND(C), PUBLIC :: ufc_rt_asm_algo
    INTEGER(C_INT32_T) :: precision = UFC_PREC_FP64
    INTEGER(C_INT32_T) :: material = UFC_MAT_ELASTIC
    INTEGER(C_INT32_T) :: variant = UFC_VARIANT_STANDARD  ! PH_ELEM_C3D8_VARIANT_*
    INTEGER(C_INT32_T) :: reserved = 0
  END TYPE ufc_rt_asm_algo

  TYPE, BIND(C), PUBLIC :: ufc_rt_asm_args
    TYPE(C_PTR) :: young = C_NULL_PTR
    REAL(C_DOUBLE) :: nu = 0.3_C_DOUBLE
  END TYPE ufc_rt_asm_args

  TYPE, BIND(C), PUBLIC :: ufc_rt_bc_args
    INTEGER(C_INT32_T) :: n_bc = 0
    INTEGER(C_INT32_T) :: index_base = 1
    TYPE(C_PTR) :: eq = C_NULL_PTR
    TYPE(C_PTR) :: value = C_NULL_PTR
    REAL(C_DOUBLE) :: penalty = 1.0E12_C_DOUBLE
  END TYPE ufc_rt_bc_args

  TYPE, BIND(C), PUBLIC :: ufc_rt_load_args
    INTEGER(C_INT32_T) :: n_faces = 0
    INTEGER(C_INT32_T) :: kind = UFC_LOAD_TRACTION
    INTEGER(C_INT32_T) :: index_base = 1
    INTEGER(C_INT32_T) :: accumulate = 0
    TYPE(C_PTR) :: elem = C_NULL_PTR
    TYPE(C_PTR) :: face = C_NULL_PTR
    REAL(C_DOUBLE) :: magnitude = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: traction(3) = 0.0_C_DOUBLE
  END TYPE ufc_rt_load_args

  TYPE, BIND(C), PUBLIC :: ufc_nm_solver_algo
    INTEGER(C_INT32_T) :: precision = UFC_PREC_FP64
    INTEGER(C_INT32_T) :: max_iter = 100000
    REAL(C_DOUBLE) :: rtol = 1.0E-10_C_DOUBLE
    REAL(C_DOUBLE) :: inner_rtol = 1.0E-5_C_DOUBLE
    INTEGER(C_INT32_T) :: max_outer = 20
    INTEGER(C_INT32_T) :: check_every = 0
    REAL(C_DOUBLE) :: replace_rtol = 0.0_C_DOUBLE
    INTEGER(C_INT32_T) :: krylov = UFC_KRYLOV_PCG   ! adjoint: PCG (K symmetric) or GMRES
    INTEGER(C_INT32_T) :: restart = 30              ! GMRES restart length
  END TYPE ufc_nm_solver_algo

  TYPE, BIND(C), PUBLIC :: ufc_nm_solver_state
    TYPE(C_PTR) :: x = C_NULL_PTR
    TYPE(C_PTR) :: b = C_NULL_PTR
    INTEGER(C_INT32_T) :: niter = 0
    INTEGER(C_INT32_T) :: n_outer = 0
    REAL(C_DOUBLE) :: rnorm = 0.0_C_DOUBLE
    INTEGER(C_INT32_T) :: converged = 0
    INTEGER(C_INT32_T) :: x0_zero = 1
  END TYPE ufc_nm_solver_state

  ! PH_J2_Props (PH_Mat_Plast_J2_Iso_Core.f90:52-91); hardening_type is one
  ! of the PH_MAT_J2_HARD_* values (:33-36), 0 meaning linear
  TYPE, BIND(C), PUBLIC :: ufc_ph_j2_props
    REAL(C_DOUBLE) :: young = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: nu = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: sigma_y0 = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: hardening = 0.0_C_DOUBLE
    INTEGER(C_INT32_T) :: hardening_type = UFC_J2_HARD_LINEAR
    INTEGER(C_INT32_T) :: max_local_iter = 0       ! <= 0: 25
    REAL(C_DOUBLE) :: swift_k = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: swift_n = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: swift_eps0 = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: voce_q = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: voce_delta = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: af_c = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: af_gamma = 0.0_C_DOUBLE
  END TYPE ufc_ph_j2_props

  TYPE, BIND(C), PUBLIC :: ufc_timing
    REAL(C_DOUBLE) :: assemble_ms = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: solve_ms = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: spmv_ms_avg = 0.0_C_DOUBLE
    INTEGER(C_INT32_T) :: spmv_calls = 0
    INTEGER(C_INT32_T) :: n_replace = 0
    REAL(C_DOUBLE) :: j2_update_ms = 0.0_C_DOUBLE
  END TYPE ufc_timing

  ! PH_Elem_Eval_Ke_Arg (PH_Elem_Def.f90:218-222) for a batch of C3D8
  TYPE, BIND(C), PUBLIC :: ufc_ph_ke_algo
    INTEGER(C_INT32_T) :: precision = UFC_PREC_FP64
    INTEGER(C_INT32_T) :: variant = UFC_VARIANT_STANDARD
  END TYPE ufc_ph_ke_algo

  TYPE, BIND(C), PUBLIC :: ufc_ph_ke_args
    INTEGER(C_INT32_T) :: n_elems = 0
    INTEGER(C_INT32_T) :: reserved = 0
    TYPE(C_PTR) :: coords = C_NULL_PTR     ! (3, 8, n_elems)
    TYPE(C_PTR) :: young = C_NULL_PTR      ! (n_elems)
    REAL(C_DOUBLE) :: nu = 0.3_C_DOUBLE
    TYPE(C_PTR) :: ke = C_NULL_PTR         ! (24, 24, n_elems)
  END TYPE ufc_ph_ke_args

  TYPE, BIND(C), PUBLIC :: ufc_ph_sens_args
    TYPE(C_PTR) :: u = C_NULL_PTR          ! (n_dof)
    TYPE(C_PTR) :: lambda = C_NULL_PTR     ! (n_dof)
    TYPE(C_PTR) :: dedrho = C_NULL_PTR     ! (n_elems)
    REAL(C_DOUBLE) :: nu = 0.3_C_DOUBLE
    TYPE(C_PTR) :: grad = C_NULL_PTR       ! (n_elems) out
  END TYPE ufc_ph_sens_args

  ! ufc_rt_restart_data: the step data of MD_RestartData (RT_Out_Restart.f90)
  TYPE, BIND(C), PUBLIC :: ufc_rt_restart_data
    REAL(C_DOUBLE) :: time = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: lambda = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: residual_norm = 0.0_C_DOUBLE
    INTEGER(C_INT32_T) :: increment = 0
    INTEGER(C_INT32_T) :: iterations = 0
    INTEGER(C_INT32_T) :: converged = 1
    INTEGER(C_INT32_T) :: n_dof = 0
    TYPE(C_PTR) :: u = C_NULL_PTR          ! (n_dof)
  END TYPE ufc_rt_restart_data

  ! ufc_ap_model: a static C3D8 step read from an input deck (0-based arrays
  ! owned by the library until ufc_ap_free)
  TYPE, BIND(C), PUBLIC :: ufc_ap_model
    INTEGER(C_INT32_T) :: n_nodes = 0
    INTEGER(C_INT32_T) :: n_elems = 0
    TYPE(C_PTR) :: node_label = C_NULL_PTR ! (n_nodes)
    TYPE(C_PTR) :: elem_label = C_NULL_PTR ! (n_elems)
    TYPE(C_PTR) :: coords = C_NULL_PTR     ! (3, n_nodes)
    TYPE(C_PTR) :: conn = C_NULL_PTR       ! (8, n_elems), 0-based nodes
    TYPE(C_PTR) :: young = C_NULL_PTR      ! (n_elems)
    REAL(C_DOUBLE) :: nu = 0.0_C_DOUBLE
    INTEGER(C_INT32_T) :: n_bc = 0
    INTEGER(C_INT32_T) :: n_cload = 0
    TYPE(C_PTR) :: bc_eq = C_NULL_PTR      ! (n_bc), eq = 3 node + dof, 0-based
    TYPE(C_PTR) :: bc_val = C_NULL_PTR
    TYPE(C_PTR) :: cload_eq = C_NULL_PTR   ! (n_cload)
    TYPE(C_PTR) :: cload_val = C_NULL_PTR
    INTEGER(C_INT32_T) :: n_face = 0
    INTEGER(C_INT32_T) :: variant = 0      ! UFC_VARIANT_* of the element type
    TYPE(C_PTR) :: face_elem = C_NULL_PTR  ! (n_face), 0-based
    TYPE(C_PTR) :: face_id = C_NULL_PTR    ! (n_face), 0..5
    TYPE(C_PTR) :: face_p = C_NULL_PTR     ! (n_face)
    TYPE(C_PTR) :: impl = C_NULL_PTR
  END TYPE ufc_ap_model

  PUBLIC :: ufc_ctx_create, ufc_ctx_destroy, ufc_ctx_info, ufc_ctx_timing, &
            ufc_ctx_enable_timing, ufc_ph_c3d8_eval_ke, ufc_ph_c3d8_sensitivity, &
            ufc_rt_asm_build_pattern, ufc_rt_asm_get_pattern, ufc_rt_asm_assemble_k, &
            ufc_rt_asm_get_values, ufc_rt_asm_apply_bc, ufc_rt_asm_load, ufc_rt_set_rhs, &
            ufc_rt_get_rhs, ufc_rt_residual, ufc_nm_spmv, ufc_nm_pcg, ufc_nm_adjoint_solve, &
            ufc_ph_j2_init, ufc_ph_j2_update, ufc_ph_j2_commit, ufc_ph_j2_get_state, &
            ufc_ph_j2_set_state, ufc_mesh_box, ufc_field, ufc_ap_read_inp, ufc_ap_free, &
            ufc_rt_restart_save, ufc_rt_restart_load, ufc_nm_spmv_t, ufc_rt_asm_set_values, &
            ufc_rt_asm_set_fp32_view, ufc_rt_asm_set_fp64_view, ufc_status_message, &
            ufc_ph_c3d8_eval_ke_noalloc, ufc_mesh_box_numbering

  INTERFACE
    INTEGER(C_INT) FUNCTION ufc_ctx_create(device, ctx, st) BIND(C, NAME='ufc_ctx_create')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      INTEGER(C_INT32_T), VALUE :: device
      TYPE(C_PTR), INTENT(OUT) :: ctx
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ctx_destroy(ctx, st) BIND(C, NAME='ufc_ctx_destroy')
      IMPORT :: C_INT, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ctx_timing(ctx, t, st) BIND(C, NAME='ufc_ctx_timing')
      IMPORT :: C_INT, C_PTR, ufc_timing, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_timing), INTENT(INOUT) :: t
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_asm_build_pattern(ctx, mesh, info, st) &
        BIND(C, NAME='ufc_rt_asm_build_pattern')
      IMPORT :: C_INT, C_PTR, ufc_mesh_desc, ufc_rt_asm_info, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_mesh_desc), INTENT(IN) :: mesh
      TYPE(ufc_rt_asm_info), INTENT(INOUT) :: info
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_asm_get_pattern(ctx, index_base, row_ptr, col_idx, st) &
        BIND(C, NAME='ufc_rt_asm_get_pattern')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: index_base
      TYPE(C_PTR), VALUE :: row_ptr, col_idx
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_asm_assemble_k(ctx, algo, args, st) &
        BIND(C, NAME='ufc_rt_asm_assemble_k')
      IMPORT :: C_INT, C_PTR, ufc_rt_asm_algo, ufc_rt_asm_args, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_rt_asm_algo), INTENT(IN) :: algo
      TYPE(ufc_rt_asm_args), INTENT(IN) :: args
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_asm_apply_bc(ctx, args, st) BIND(C, NAME='ufc_rt_asm_apply_bc')
      IMPORT :: C_INT, C_PTR, ufc_rt_bc_args, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_rt_bc_args), INTENT(IN) :: args
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_asm_load(ctx, args, st) BIND(C, NAME='ufc_rt_asm_load')
      IMPORT :: C_INT, C_PTR, ufc_rt_load_args, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_rt_load_args), INTENT(IN) :: args
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_get_rhs(ctx, f, st) BIND(C, NAME='ufc_rt_get_rhs')
      IMPORT :: C_INT, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx, f
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_residual(ctx, u, r, norms, st) BIND(C, NAME='ufc_rt_residual')
      IMPORT :: C_INT, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx, u, r, norms
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_nm_spmv(ctx, precision, x, y, st) BIND(C, NAME='ufc_nm_spmv')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: precision
      TYPE(C_PTR), VALUE :: x, y
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_nm_spmv_t(ctx, precision, x, y, st) BIND(C, NAME='ufc_nm_spmv_t')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: precision
      TYPE(C_PTR), VALUE :: x, y
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_asm_set_values(ctx, precision, values, st) &
        BIND(C, NAME='ufc_rt_asm_set_values')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: precision
      TYPE(C_PTR), VALUE :: values
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_asm_set_fp32_view(ctx, view, st) &
        BIND(C, NAME='ufc_rt_asm_set_fp32_view')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: view
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_asm_set_fp64_view(ctx, view, st) &
        BIND(C, NAME='ufc_rt_asm_set_fp64_view')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: view
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_nm_pcg(ctx, algo, state, st) BIND(C, NAME='ufc_nm_pcg')
      IMPORT :: C_INT, C_PTR, ufc_nm_solver_algo, ufc_nm_solver_state, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_nm_solver_algo), INTENT(IN) :: algo
      TYPE(ufc_nm_solver_state), INTENT(INOUT) :: state
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ph_j2_init(ctx, props, st) BIND(C, NAME='ufc_ph_j2_init')
      IMPORT :: C_INT, C_PTR, ufc_ph_j2_props, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_ph_j2_props), INTENT(IN) :: props
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ph_j2_update(ctx, u, u_n, n_yield, st) &
        BIND(C, NAME='ufc_ph_j2_update')
      IMPORT :: C_INT, C_INT64_T, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx, u, u_n
      INTEGER(C_INT64_T), INTENT(OUT) :: n_yield
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ph_j2_commit(ctx, st) BIND(C, NAME='ufc_ph_j2_commit')
      IMPORT :: C_INT, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ph_j2_get_state(ctx, state, st) BIND(C, NAME='ufc_ph_j2_get_state')
      IMPORT :: C_INT, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx, state
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ph_j2_set_state(ctx, state, st) BIND(C, NAME='ufc_ph_j2_set_state')
      IMPORT :: C_INT, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx, state
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ctx_info(ctx, info, st) BIND(C, NAME='ufc_ctx_info')
      IMPORT :: C_INT, C_PTR, ufc_rt_asm_info, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_rt_asm_info), INTENT(INOUT) :: info
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ctx_enable_timing(ctx, per_kernel, st) &
        BIND(C, NAME='ufc_ctx_enable_timing')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: per_kernel
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ph_c3d8_eval_ke(ctx, algo, args, st) &
        BIND(C, NAME='ufc_ph_c3d8_eval_ke')
      IMPORT :: C_INT, C_PTR, ufc_ph_ke_algo, ufc_ph_ke_args, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_ph_ke_algo), INTENT(IN) :: algo
      TYPE(ufc_ph_ke_args), INTENT(INOUT) :: args
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    ! ufc_ph_c3d8_eval_ke without device allocation (no reference counterpart):
    ! coords, young and ke in args must be device memory of the context's GPU
    INTEGER(C_INT) FUNCTION ufc_ph_c3d8_eval_ke_noalloc(ctx, algo, args, st) &
        BIND(C, NAME='ufc_ph_c3d8_eval_ke_noalloc')
      IMPORT :: C_INT, C_PTR, ufc_ph_ke_algo, ufc_ph_ke_args, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_ph_ke_algo), INTENT(IN) :: algo
      TYPE(ufc_ph_ke_args), INTENT(INOUT) :: args
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ph_c3d8_sensitivity(ctx, args, st) &
        BIND(C, NAME='ufc_ph_c3d8_sensitivity')
      IMPORT :: C_INT, C_PTR, ufc_ph_sens_args, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_ph_sens_args), INTENT(IN) :: args
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_asm_get_values(ctx, precision, values, st) &
        BIND(C, NAME='ufc_rt_asm_get_values')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: precision
      TYPE(C_PTR), VALUE :: values
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_set_rhs(ctx, f, st) BIND(C, NAME='ufc_rt_set_rhs')
      IMPORT :: C_INT, C_PTR, ufc_status
      TYPE(C_PTR), VALUE :: ctx, f
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_nm_adjoint_solve(ctx, algo, dJdu, state, st) &
        BIND(C, NAME='ufc_nm_adjoint_solve')
      IMPORT :: C_INT, C_PTR, ufc_nm_solver_algo, ufc_nm_solver_state, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufc_nm_solver_algo), INTENT(IN) :: algo
      TYPE(C_PTR), VALUE :: dJdu
      TYPE(ufc_nm_solver_state), INTENT(INOUT) :: state
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_mesh_box(nx, ny, nz, lx, ly, lz, jitter, seed, index_base, &
                                         coords, conn, st) BIND(C, NAME='ufc_mesh_box')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR, ufc_status
      INTEGER(C_INT32_T), VALUE :: nx, ny, nz
      REAL(C_DOUBLE), VALUE :: lx, ly, lz, jitter
      INTEGER(C_INT64_T), VALUE :: seed
      INTEGER(C_INT32_T), VALUE :: index_base
      TYPE(C_PTR), VALUE :: coords, conn
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    ! lexicographic box numbering of a structured hex mesh in any node
    ! numbering (host only): perm(n_nodes), 0-based lattice numbers; conn in
    ! base index_base (UFC_STATUS_UNSUPPORTED when the mesh is not such a box)
    INTEGER(C_INT) FUNCTION ufc_mesh_box_numbering(n_nodes, n_elems, index_base, conn, perm, st) &
        BIND(C, NAME='ufc_mesh_box_numbering')
      IMPORT :: C_INT, C_INT32_T, C_PTR, ufc_status
      INTEGER(C_INT32_T), VALUE :: n_nodes, n_elems, index_base
      TYPE(C_PTR), VALUE :: conn, perm
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_field(kind, n, seed, a, b, out, st) BIND(C, NAME='ufc_field')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR, ufc_status
      INTEGER(C_INT32_T), VALUE :: kind
      INTEGER(C_INT64_T), VALUE :: n, seed
      REAL(C_DOUBLE), VALUE :: a, b
      TYPE(C_PTR), VALUE :: out
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    ! path: NUL-terminated (TRIM(name) // C_NULL_CHAR)
    INTEGER(C_INT) FUNCTION ufc_rt_restart_save(ctx, path, data, with_state, st) &
        BIND(C, NAME='ufc_rt_restart_save')
      IMPORT :: C_INT, C_INT32_T, C_PTR, C_CHAR, ufc_rt_restart_data, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: path
      TYPE(ufc_rt_restart_data), INTENT(IN) :: data
      INTEGER(C_INT32_T), VALUE :: with_state
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_rt_restart_load(ctx, path, data, with_state, st) &
        BIND(C, NAME='ufc_rt_restart_load')
      IMPORT :: C_INT, C_INT32_T, C_PTR, C_CHAR, ufc_rt_restart_data, ufc_status
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: path
      TYPE(ufc_rt_restart_data), INTENT(INOUT) :: data
      INTEGER(C_INT32_T), VALUE :: with_state
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ap_read_inp(path, model, st) BIND(C, NAME='ufc_ap_read_inp')
      IMPORT :: C_INT, C_CHAR, ufc_ap_model, ufc_status
      CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: path
      TYPE(ufc_ap_model), INTENT(INOUT) :: model
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION

    INTEGER(C_INT) FUNCTION ufc_ap_free(model, st) BIND(C, NAME='ufc_ap_free')
      IMPORT :: C_INT, ufc_ap_model, ufc_status
      TYPE(ufc_ap_model), INTENT(INOUT) :: model
      TYPE(ufc_status), INTENT(INOUT) :: st
    END FUNCTION
  END INTERFACE

CONTAINS

  ! message(512) of a ufc_status as a Fortran string (ErrorStatusType%message)
  FUNCTION ufc_status_message(st) RESULT(msg)
    TYPE(ufc_status), INTENT(IN) :: st
    CHARACTER(LEN=512) :: msg
    INTEGER :: i
    msg = ' '
    DO i = 1, 512
      IF (st%message(i) == C_NULL_CHAR) EXIT
      msg(i:i) = st%message(i)
    END DO
  END FUNCTION ufc_status_message

END MODULE ufc_hip_iface
