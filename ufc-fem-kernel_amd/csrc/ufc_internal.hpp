// ufc_internal.hpp -- context layout and kernel-launch interfaces shared by
// the HIP translation units of libufc_hip.so (not installed, not part of the
// C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ufc_hip.h"

namespace ufc {

// ---------------------------------------------------------------- errors
struct Error {
  int code;
  std::string msg;
};

#define UFC_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::ufc::Error{e_ == hipErrorOutOfMemory ? UFC_STATUS_MEM_ERROR : UFC_STATUS_ERROR, \
                         std::string(#expr) + ": " + hipGetErrorString(e_)};            \
  } while (0)

#define UFC_REQUIRE(cond, code, text)                    \
  do {                                                   \
    if (!(cond)) throw ::ufc::Error{(code), (text)};     \
  } while (0)

// ---------------------------------------------------------------- device buffer
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void resize(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    UFC_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  T* get() const { return p; }
  size_t bytes() const { return n * sizeof(T); }
};

// Layout of the scalars of one PCG run, kept on the device so that the
// iteration never waits for the host.
struct PcgScalars {
  double rz;      // r.z of the current iterate (after convergence: of the one before)
  double rzs[2];  // r.z of the iterate of iteration it in slot it & 1
  double pap;     // p.Ap
  double alpha;
  double beta;
  double rr;      // r.r
  double bnorm;   // ||b||
  double rnorm;   // ||r|| / ||b||
  double rtol;
  double gamma;   // (r.p)/(r.z) after a residual replacement
  int32_t iter;   // completed iterations
  int32_t done;   // 0 running, 1 converged, 2 breakdown, 3 max_iter
  int32_t max_iter;
  int32_t gate;   // 1: the current iteration's x/p pass does not run
};

constexpr int kSolveBlocks = 2048;  // vector kernels: 8 workgroups per CU on 256 CUs
constexpr int kSolveThreads = 256;  // vector kernels: 4 wave64 per workgroup
constexpr int kSpmvThreads = 512;   // SpMV: 8 wave64 = 16 node groups per workgroup
constexpr int kSpmvBlocks64 = 32768;  // FP64 CSR SpMV grid (measured best on gfx950)
constexpr int kSpmvBlocks32 = 32768;  // FP32 SpMV grid (measured best on gfx950: 0.415 ms at 16384, 0.405 at 32768, 0.415 at 65536 for 128^3)
constexpr int kNbPtrPad = 3;  // device nb_ptr entries past n_nodes, equal to nb_ptr[n_nodes]
constexpr int kMaxPartials = kSolveBlocks;  // per-workgroup partial sums (vector kernels)
constexpr int kMaxNodeElems = 8;    // elements per node of the slot map
constexpr int kMaxNodeNbrs = 32;    // neighbours per node of the slot map (5-bit codes) and the 32-lane views
// nodes beyond the slot map (general nodes): neighbours per node (the LDS
// row segment of the general assembly); their element count is unbounded
constexpr int kMaxGenNbrs = 256;
constexpr int kMaxPatterns = 256;   // stencil patterns of the FP32 view (one-byte ids)
constexpr int kMaxDiags = 13;       // upper block diagonals of the symmetric FP32 view
// doubles per Gauss point of the stored J2 tangent (mu', lam', c, n(6)):
// Dt = lam' 1x1 + 2 mu' I_s - c n x n
constexpr int kJ2Tangent = 9;
constexpr int kJ2Vals = 13;  // J2 state doubles per Gauss point: stress (6), eqps, plastic strain (6)
// nodes per tile of the symmetric view's block-diagonal layout, one per
// thread of the SpMV's workgroup: kDiagTile, or kDiagTileWide on meshes of at
// least kDiagWideMinNodes nodes (larger tiles keep more of the transposed
// re-reads one node row back inside the workgroup and stage less x window
// per node, but halve the workgroups: per PCG iteration 0.5% slower at 88^3,
// 2% faster at 96^3, 1.2% at 128^3, 19% slower at 64^3)
constexpr int kDiagTile = 256;
constexpr int kDiagTileWide = 512;
constexpr int64_t kDiagWideMinNodes = 800000;
// tiles one "plane" (ceil(o_last / kDiagTile) tiles) apart that run in a row
// (128^3 FP32 SpMV: 8 0.276 ms, 16 0.263, 32 0.270)
constexpr int kDiagPlanes = 16;
// FP32 SpMV views (ufc_rt_asm_info.fp32_view)
constexpr int kFp32ViewId32 = 0, kFp32ViewOff16 = 1, kFp32ViewPattern = 2, kFp32ViewSymDiag = 3;
// FP64 SpMV layouts: the CSR with a column index per value (the reference's
// NM_CSR_MatVec_Optimized), or the same values with one neighbour id per block
constexpr int kFp64ViewCsr = 0, kFp64ViewBlock = 1;

// offsets o_s of the symmetric FP32 view's block diagonals (kernel argument)
struct DiagOffsets {
  int32_t o[kMaxDiags];
};

// x window of a symmetric-SpMV tile in LDS: the nodes n + o for the tile's
// nodes n and every offset o in {0, +-o_s}, as at most kMaxXClusters runs of
// consecutive nodes (offsets closer than a tile share a run).  Run c starts
// lo[c] nodes after the tile's first node, holds len[c] nodes and sits at
// LDS node slot base[c]; up[s], dn[s] and self are the slots of n + o_s,
// n - o_s and n for the tile's first node (lane l adds l).
constexpr int kMaxXClusters = 4;
// window nodes (16 B each) per tile size: 26 KB for 256-node tiles (a
// structured box needs 3 x 516), 38 KB for 512-node tiles (3 x 772)
constexpr int kMaxXWin = 1664;
constexpr int kMaxXWinWide = 2432;
__host__ __device__ constexpr int max_xwin(int tile) {
  return tile == kDiagTileWide ? kMaxXWinWide : kMaxXWin;
}
inline int32_t tile_shift(int32_t tile) { return tile == kDiagTileWide ? 9 : 8; }
static_assert(kDiagTile == 256 && kDiagTileWide == 512, "tile_shift");
struct XWindow {
  int32_t nc;
  int32_t lo[kMaxXClusters], base[kMaxXClusters], len[kMaxXClusters];
  int32_t up[kMaxDiags], dn[kMaxDiags], self;
};

// C3D8 variants, PH_ELEM_C3D8_VARIANT_* (PH_Elem_C3D8.f90:111-116)
constexpr int kVarFull = 1, kVarReduced = 2, kVarSelective = 3, kVarIncompat = 4,
              kVarGauss27 = 6;
// the element each variant value computes (0 and 5 are the standard element,
// as PH_Elem_C3D8_StiffMatrixByVariant routes them), -1 if unknown
inline int canonical_variant(int v) {
  switch (v) {
    case 0:
    case 1:
    case 5:
      return kVarFull;
    case 2:
    case 3:
    case 4:
    case 6:
      return v;
    default:
      return -1;
  }
}
constexpr int kIncompatC = 24 * 9;  // condensation factor per element (doubles)

// ---------------------------------------------------------------- C-ABI status
inline void set_status(ufc_status* st, int code, const std::string& msg, const char* src) {
  if (!st) return;
  st->status_code = code;
  st->severity = code == UFC_STATUS_OK ? 0 : (code == UFC_STATUS_WARN ? 1 : 3);
  st->category = code;
  st->line_number = 0;
  std::snprintf(st->message, sizeof st->message, "%s", msg.c_str());
  std::snprintf(st->source, sizeof st->source, "%s", src);
}

template <typename F>
int guarded(ufc_status* st, const char* src, F&& body) {
  try {
    std::string msg;
    int code = body(msg);
    set_status(st, code, msg, src);
    return code;
  } catch (const Error& e) {
    set_status(st, e.code, e.msg, src);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_status(st, UFC_STATUS_MEM_ERROR, "host allocation failed", src);
    return UFC_STATUS_MEM_ERROR;
  } catch (...) {
    set_status(st, UFC_STATUS_FATAL, "unexpected exception", src);
    return UFC_STATUS_FATAL;
  }
}

}  // namespace ufc

// ---------------------------------------------------------------- context
struct ufc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;

  // mesh
  int32_t n_nodes = 0, n_elems = 0;
  int64_t n_dof = 0;
  std::vector<double> h_coords;  // (n_nodes, 3)
  std::vector<int32_t> h_conn;   // (n_elems, 8) 0-based
  ufc::DBuf<double> coords;
  ufc::DBuf<int32_t> conn;

  // pattern: node adjacency (the node-block index view of the CSR) + CSR
  int64_t nnz = 0;
  int32_t max_row_nnz = 0;
  std::vector<int32_t> h_nb_ptr;  // (n_nodes + 1)
  ufc::DBuf<int32_t> nb_ptr;      // (n_nodes + 1 + kNbPtrPad), tail = nb_ptr[n_nodes]
  ufc::DBuf<int32_t> nb_idx;      // (nnz / 9)
  // (nnz / 9) neighbour as m - n, when every |m - n| fits in 16 bits: the FP32
  // SpMV then reads 2 instead of 4 bytes per neighbour
  ufc::DBuf<int16_t> nb_off16;
  bool have_off16 = false;
  // Stencil patterns of the 16-bit view: when at most kMaxPatterns distinct
  // neighbour-offset lists (degree and offsets) occur -- any structured hex
  // box has at most 27 -- node n's list is pat_off[nb_pat[n]] and the FP32
  // SpMV reads one byte per node instead of 2 per neighbour.  n_pat = 0: no
  // pattern view.
  ufc::DBuf<uint8_t> nb_pat;      // (n_nodes rounded up to 4, + 4), pattern id
  // (n_pat, kMaxNodeNbrs): offsets m - n of the neighbours other than the
  // node itself (ascending), then 0 at position deg - 1, zero padded
  ufc::DBuf<int16_t> pat_off;
  int32_t n_pat = 0;
  // Symmetric block-diagonal view: when the offsets m - n > 0 of every node's
  // upper neighbours take at most kMaxDiags distinct values o_0 < ... (a
  // lexicographically numbered structured box has 13), each off-diagonal
  // block is stored once, for its upper node: K(n, n + o_s) is entry (s, n),
  // zero where n + o_s is no neighbour.  Node n's lower block
  // K(n, n - o_s) = K(n - o_s, n)^T is entry (s, n - o_s).  Tile-major: the
  // entries of the diag_tile nodes of tile T = n / diag_tile + diag_pad_tiles
  // for diagonal s are the diag_tile consecutive entries from
  // (T n_diag + s) diag_tile; diag_pad_tiles = ceil(o_last / diag_tile) zero
  // tiles in front keep the lower reads of the first nodes in bounds.
  // Entries: diag_a float4 (0,0) (0,1) (0,2) (1,0), diag_b float4 (1,1)
  // (1,2) (2,0) (2,1), diag_c float (2,2) -- three loads per block.
  int32_t n_diag = 0;
  int32_t h_diag_off[ufc::kMaxDiags] = {};
  int32_t diag_pad_tiles = 0;
  int32_t diag_tile = ufc::kDiagTile;  // kDiagTile or kDiagTileWide (a power of two)
  int64_t diag_entries = 0;     // (n_nodes / diag_tile + 1 + diag_pad_tiles) n_diag diag_tile
  ufc::DBuf<float> diag_a;      // (diag_entries, 4)
  ufc::DBuf<float> diag_b;      // (diag_entries, 4)
  ufc::DBuf<float> diag_c;      // (diag_entries)
  // Box numbering of the symmetric view: when the mesh's own numbering has
  // more than kMaxDiags offsets but the mesh is a structured box of hexes
  // (any node numbering), the view runs in the box's lexicographic numbering
  // instead.  sym_perm[n] = that number of node n, sym_iperm its inverse; the
  // entries, pen32, pen_flag, diag32p and the FP32 PCG's vectors are then in
  // the box numbering and converted at the solve's boundary.
  bool sym_permuted = false;
  // the block diagonals hold the current FP32 values (written by the FP32
  // assembly, or filled from val32 when the view was selected afterwards)
  bool diag_from_assembly = false;
  std::vector<int32_t> h_sym_perm;
  ufc::DBuf<int32_t> sym_perm;   // (n_nodes)
  ufc::DBuf<int32_t> sym_iperm;  // (n_nodes)
  ufc::DBuf<float> diag32p;      // diag32 in the box numbering
  ufc::DBuf<double> w_perm;      // FP64 right-hand side in the box numbering
  ufc::DBuf<double> w_perm2;     // FP64 iterate in the box numbering (restarted form)
  int32_t fp32_view = -1;       // view of the FP32 SpMV (kFp32View*); -1 before the pattern
  int32_t fp32_view_request = -1;  // -1: the fastest available
  int32_t fp64_view = 0;           // kFp64View*: layout the FP64 SpMV reads
  bool have_xwin = false;          // the symmetric SpMV stages its x window in LDS
  ufc::XWindow xwin{};
  ufc::DBuf<int32_t> row_ptr;     // (n_dof + 1) when nnz < 2^31 (the exported CSR)
  ufc::DBuf<int32_t> col_idx;     // (nnz)  built on demand (FP64 CSR SpMV, export)
  ufc::DBuf<int64_t> diag_pos;    // (n_dof) position of K(i,i) in values
  // element-to-slot map, node major: for node n and slot s < 8, the element
  // (ascending, -1 if none) and a code: bits 5b..5b+4 = position in n's
  // neighbour list of the element's local node b, bits 40..42 = n's local index
  ufc::DBuf<int32_t> slot_elem;   // (n_nodes * 8)
  ufc::DBuf<uint64_t> slot_code;  // (n_nodes * 8)
  // n's local index in the element of slot s, bits 4s..4s+2 (slot_code bits
  // 40..42 of all eight slots in one word, for the node gather)
  ufc::DBuf<uint32_t> slot_loc;   // (n_nodes)
  // General nodes: a node in more than kMaxNodeElems elements or with more
  // than kMaxNodeNbrs neighbours does not fit the slot map; every node of its
  // 8-node assembly block (nodes 8b .. 8b + 7) is then assembled and gathered
  // by the general kernels (one workgroup per node over its element list),
  // and the slot-map kernels skip the block.
  int32_t n_gen = 0;               // general nodes (whole blocks)
  ufc::DBuf<uint8_t> gen_flag;     // (ceil(n_nodes / 8)) 1: block of general nodes; empty if none
  ufc::DBuf<int32_t> gen_node;     // (n_gen) ascending
  ufc::DBuf<int32_t> gen_ne_ptr;   // (n_gen + 1)
  ufc::DBuf<int32_t> gen_ne;       // 8 e + a per element e of the node (local index a), e ascending
  bool wide = false;               // some node has more than kMaxNodeNbrs neighbours

  // values
  ufc::DBuf<double> val64;
  ufc::DBuf<float> val32;
  ufc::DBuf<double> diag64;
  ufc::DBuf<float> diag32;   // K(i,i) + penalty (Jacobi)
  ufc::DBuf<float> kdiag32;  // K(i,i) of the FP32 assembly, no penalty
  // Off-diagonal copy of the FP32 difference-form values for the stencil-
  // pattern SpMV (n_pat > 0): node n's three rows without the self block,
  // 9 (deg - 1) values at 9 (nb_ptr[n] - n), row d holding the neighbours
  // other than n in ascending order; the self block, which in difference form
  // holds only the penalty diagonal, becomes pen32 plus one bit per node
  // (pen_flag) marking nodes with a nonzero penalty.
  ufc::DBuf<float> val32m;     // (9 (nnz / 9 - n_nodes))
  ufc::DBuf<float> pen32;      // (n_dof)
  ufc::DBuf<uint32_t> pen_flag;  // (n_nodes / 32 + 1)
  bool have64 = false, have32 = false;
  int last_material = UFC_MAT_ELASTIC;
  int variant = 1;                 // element variant of the last elastic assembly
  bool user_values = false;        // FP64 values from ufc_rt_asm_set_values
  ufc::DBuf<int32_t> nb_mirror;    // per block (n, k): slot of n in node nb_idx's list
  bool have_mirror = false;
  ufc::DBuf<double> inc_c;         // (24, 9, n_elems) incompatible-mode factor

  // material of the last elastic assembly
  ufc::DBuf<double> young;
  double nu = 0.3;
  bool have_young = false;

  // boundary conditions (unique equations, summed penalties)
  std::vector<int32_t> h_bc_eq;
  std::vector<double> h_bc_pen;
  std::vector<double> h_bc_rhs;  // penalty * value
  ufc::DBuf<int32_t> bc_eq;
  ufc::DBuf<double> bc_pen;

  // load vector (FP64): F without penalty terms, and the penalty part
  std::vector<double> h_f;
  ufc::DBuf<double> f;  // F + P ubar

  // J2 state
  bool have_j2 = false;
  ufc_ph_j2_props j2{};
  ufc::DBuf<double> j2_commit;   // (13, 8, n_elems)
  ufc::DBuf<double> j2_trial;    // (13, 8, n_elems)
  ufc::DBuf<double> j2_tangent;  // (kJ2Tangent, 8, n_elems): factored consistent tangent
  ufc::DBuf<int32_t> j2_flags;   // (8, n_elems) 1 = plastic branch, 2 = failure

  // solver work (FP64)
  ufc::DBuf<double> w_r, w_p, w_ap, w_x, w_b;
  // solver work (FP32)
  ufc::DBuf<float> s_r, s_p, s_ap, s_x;
  // element force scratch (24 per element) and a spare FP64 vector
  ufc::DBuf<double> elem_w;
  ufc::DBuf<double> w_t;
  // reductions
  ufc::DBuf<double> partials;            // kSolveBlocks * 4
  ufc::DBuf<double> spmv_partials;       // kSpmvBlocks32: p.Ap per SpMV (FP32) or k_dot_pap (FP64) workgroup
  ufc::DBuf<double> pap_piece;           // 8: p.Ap in eight ordered pieces
  ufc::DBuf<ufc::PcgScalars> scalars;    // 1
  ufc::PcgScalars* h_scalars = nullptr;  // pinned mirror
  // Progress of the running PCG, written by the GPU into pinned, mapped,
  // coherent host memory with system-scope stores: [0] iterations completed,
  // [1] stop code (0 while running).  The host reads it without draining the
  // stream.
  int32_t* h_progress = nullptr;  // host address
  int32_t* d_progress = nullptr;  // device address of the same memory

  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  int32_t per_kernel = 0;           // bracket every per_kernel-th SpMV launch (0: none)
  std::vector<hipEvent_t> spmv_ev;  // pairs
  size_t spmv_ev_used = 0;
  std::vector<char> spmv_active;    // per pair: 1 = the launch did work
  std::vector<int64_t> spmv_pair_launch;  // per pair: index of the bracketed launch
  int64_t spmv_launches = 0;        // SpMV launches since the last reset
  bool spmv_sampled = false;        // the current launch is bracketed
  double assemble_ms = 0.0, solve_ms = 0.0, spmv_ms_avg = 0.0, j2_update_ms = 0.0;
  int32_t spmv_calls = 0;
  int32_t n_replace = 0;  // FP64 residual replacements of the last MIXED solve
};

namespace ufc {

// the FP32 view runs in the box numbering (sym_perm)
inline bool fp32_permuted(const ufc_ctx* c) {
  return c->sym_permuted && c->fp32_view == kFp32ViewSymDiag;
}

// ---------------------------------------------------------------- launchers
// device address of the allocation-free entry point's scratch array on the
// current device, resolved once per device (ufc_noalloc.hip)
void noalloc_setup(int device);
// element kernels (ufc_elem.hip)
void launch_eval_ke(int32_t n_elems, const double* coords, const double* young, double nu,
                    void* ke, bool fp32, int variant, double* inc_c, hipStream_t s);
// incompatible-mode condensation factor C (Ke = Kuu - C C^T) of every element;
// conn == nullptr: coords is (8, 3, n_elems) element coordinates
void launch_incompat_c(int32_t n_elems, const double* coords, const int32_t* conn,
                       const double* young, double nu, double* inc_c, hipStream_t s);
void launch_assemble(ufc_ctx* c, int precision, int material);
void launch_diag(ufc_ctx* c, bool fp32);
void launch_penalty(ufc_ctx* c, bool fp32, double sign);
// FP32 values -> difference form: K(i,i) saved to kdiag32, self blocks replaced
// by the penalty diagonal, diag32 = kdiag32 + penalty
void launch_fp32_difference_form(ufc_ctx* c);
void launch_diag32_penalty(ufc_ctx* c);
// the view-specific FP32 values from val32: val32m (stencil-pattern view) or
// the block diagonals (symmetric view), with pen32 and pen_flag
void launch_fp32_offdiag(ufc_ctx* c);
// the FP32 view the SpMV uses: the requested one, else the fastest the mesh allows
void select_fp32_view(ufc_ctx* c);
// y = K u in FP64 without penalties: element forces then node gather.
// mode 0: elastic with c->young; mode 1: J2 tangent D_gp; mode 2: J2 trial stress (u unused)
void launch_apply_elem64(ufc_ctx* c, int mode, const double* u, double* y);
void launch_sensitivity(ufc_ctx* c, const double* u, const double* lam, const double* dedrho,
                        double nu, double* grad);
void launch_j2_update(ufc_ctx* c, const double* u, const double* u_n);

// solver kernels (ufc_solve.hip)
void spmv64(ufc_ctx* c, const double* x, double* y);
void spmv32(ufc_ctx* c, const float* x, float* y);
// r = b - y - pen*u (penalty rows); returns nothing, norms via pcg scalars
void residual64(ufc_ctx* c, const double* b, const double* y, const double* u, double* r);
double norm2_64(ufc_ctx* c, const double* v);
int pcg64(ufc_ctx* c, const double* b, double* x, bool x0_zero, int max_iter, double rtol,
          int check_every, int* niter, double* rnorm);
// FP32 PCG into the FP64 iterate x; resume = continue the previous FP32 CG
// after a residual replacement (b is the new true residual)
int pcg32(ufc_ctx* c, const double* b64, double* x, bool zero_x, bool resume, int max_iter,
          double rtol, int check_every, int* niter, double* rnorm);
// FP32 PCG segment whose correction d (FP32, zeroed first) the caller adds to
// its FP64 iterate; resume as for pcg32
int pcg32_segment(ufc_ctx* c, const double* b64, float* d, bool resume, int max_iter,
                  double rtol, int check_every, int* niter, double* rnorm);
void axpy_f32_to_f64(ufc_ctx* c, const float* d, double* u);
void cast_f64_to_f32(ufc_ctx* c, const double* a, float* o);
void cast_f32_to_f64(ufc_ctx* c, const float* a, double* o);
// FP64 vector in node order <-> FP32 vector in the FP32 view's numbering
// (the same numbering unless fp32_permuted)
void to_view32(ufc_ctx* c, const double* a, float* o);
void from_view32(ufc_ctx* c, const float* a, double* o);
void build_col_idx(ufc_ctx* c);
double spmv_event_avg_ms(ufc_ctx* c, int* calls);
void reset_spmv_events(ufc_ctx* c);
void ensure_solver_buffers(ufc_ctx* c, bool need64, bool need32);
// transpose operator and GMRES on K^T (ufc_gmres.hip)
void ensure_mirror(ufc_ctx* c);
void spmv_t64(ufc_ctx* c, const double* x, double* y);
int gmres_t64(ufc_ctx* c, const double* g, double* lam, int restart, int max_iter, double rtol,
              int* niter, int* n_restart, double* rnorm, std::string& msg);

}  // namespace ufc
