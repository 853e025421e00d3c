// ufc_solve.hip -- gfx950 CSR SpMV and fused Jacobi-PCG.
//
// PCG follows NM_Solver_PCG (reference L2_NM/Solver/NM_Solv_Core.f90:363-423)
// step for step: r = b - K x0, z = r / diag (z = r where |diag| <= 1e-30,
// NM_Solver_Jacobi_Precond :131-147), p = z; per iteration Ap = K p,
// alpha = (r.z) / (p.Ap), x += alpha p, r -= alpha Ap, stop at the first k
// with ||r|| / ||b|| < rtol (||b|| < 1e-30 -> 1), z = r / diag,
// beta = (r.z)_new / (r.z)_old, p = z + beta p.
//
// One iteration is five launches (four for the FP32 view, whose SpMV forms
// the p.Ap partials itself); every scalar stays on the device:
//   spmv      : Ap = K p
//   dot_pap   : per-workgroup partials of p.Ap (FP64 CSR only)
//   fin_pap   : eight workgroups add the partials in index order (8 pieces)
//   update    : alpha from the pieces; r -= alpha Ap, partials of r.r and
//               r.(r/diag)
//   xp_update : every workgroup adds the r.r and r.z partials: ||r||/||b||,
//               convergence / breakdown / max_iter, beta; x += alpha p and
//               p = r/diag + beta p
// The ordered two-level sums make every dot product bit-reproducible.  Once
// `done` is set every later kernel returns at entry, so the host polls only
// every few iterations while the stopping iteration stays exact.
// (A last-workgroup-arrives reduction inside the producing kernel needs an
// agent-scope release per workgroup, which on gfx950 writes back the XCD's
// L2 and costs ~55 us per kernel; the separate one-workgroup finalize costs
// ~4 us.)
//
// SpMV layout: node n's three CSR rows are one contiguous segment of 9*deg
// values starting at 9*nb_ptr[n]; inside it, row d (0..2) holds the columns
// 3*nb_idx[nb_ptr[n]+k] + c for neighbours k < deg (ascending) and c = 0..2.
// A group of 32 lanes takes one node, lane k one neighbour k: it reads its
// three 3-value runs (one per row, contiguous across the lanes, so every load
// instruction is coalesced), gathers the neighbour's 3 entries of x and
// keeps three row partials; the group reduces them with shuffles.
//   FP64: plain CSR -- every value's column comes from col_idx, as
//         NM_CSR_MatVec_Optimized (NM_Mtx_Sparse.f90:341-365) reads it.
//   FP32: node-block index view of the same CSR -- one 4-byte neighbour id per
//         9 values replaces col_idx, and x is gathered once per neighbour --
//         applied in difference form.  Every element matrix annihilates rigid
//         translations, so each block row of K sums to zero and
//             (K x)_n = sum_{m != n} K_nm (x_m - x_n)   (+ P_n x_n on penalty rows).
//         The self blocks of the FP32 values therefore hold only the penalty
//         diagonal.  FP32 rounding then acts on x_m - x_n instead of x_m, so a
//         smooth x no longer loses ~eps32 * cond(K) to cancellation: the
//         refinement residual after one inner solve falls to the inner
//         tolerance instead of stalling near 1e-3.
#include <chrono>
#include <thread>

#include "ufc_internal.hpp"

namespace ufc {
namespace {

constexpr int WAVE = 64;
constexpr int GROUP = 32;  // lanes per node
// neighbour index of the FP32 node-block view
constexpr int kViewId32 = 0;     // 4-byte neighbour ids
constexpr int kViewOff16 = 1;    // 16-bit offsets m - n per neighbour
constexpr int kViewPattern = 2;  // one stencil-pattern byte per node
// 4-byte ids, neighbour lists longer than a 32-lane group (FP32 view 0 and
// the FP64 CSR on meshes with nodes of more than kMaxNodeNbrs neighbours)
constexpr int kViewWide = 3;
// FP64 node-block view: one 4-byte neighbour id per 3x3 block instead of a
// column index per value (narrow: neighbour lists within a 32-lane group)
constexpr int kViewBlock64 = 4, kViewBlock64Wide = 5;
constexpr int WAVES_PER_BLOCK = kSolveThreads / WAVE;
constexpr int GROUPS_PER_BLOCK = kSpmvThreads / GROUP;

template <typename T>
__device__ __forceinline__ T sum_lanes(T v, int width) {
  for (int o = width / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Sum over a 32-lane group, valid in every lane of the group.  FP32: four
// DPP steps inside each 16-lane row (quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror, row_mirror) and one cross-row exchange; the
// DPP forms run on the VALU instead of the LDS crossbar (measured 0.481 ->
// 0.466 ms for the 128^3 FP32 SpMV).  FP64: xor shuffles.
template <typename T>
__device__ __forceinline__ T sum_group32(T v) {
  if constexpr (sizeof(T) == 4) {
    v += dpp_f32<0xB1>(v);
    v += dpp_f32<0x4E>(v);
    v += dpp_f32<0x141>(v);
    v += dpp_f32<0x140>(v);
    return v + __shfl_xor(v, 16, WAVE);
  } else {
    return sum_lanes(v, 32);
  }
}

// Sum of v[k] over a workgroup of BS threads, valid in thread 0 (fixed order).
template <int NV, int BS>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  constexpr int NW = BS / WAVE;
  __shared__ double sh[NV][NW];
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    v[k] = sum_lanes(v[k], WAVE);
    if (lane == 0) sh[k][wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) s += sh[k][w];
      v[k] = s;
    }
  }
}

// partial k of workgroup b lives at partials[k * kMaxPartials + b]
template <int NV, int BS>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* partials) {
  block_sum<NV, BS>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * kMaxPartials + blockIdx.x] = v[k];
  }
}

constexpr int kFinThreads = 1024;

// the most partials a finalize reads per thread (the FP32 SpMV's grid)
constexpr int kFinPerThread = kSpmvBlocks32 / kFinThreads;
static_assert(kSpmvBlocks32 % kFinThreads == 0 && kMaxPartials <= kSpmvBlocks32, "partials");

// one workgroup of kFinThreads: ordered sum of nblocks <= kSpmvBlocks32
// partials per value, valid in thread 0.  Thread t adds partials t,
// t + kFinThreads, ... in that order; all its loads are issued before the
// first add (a loop of dependent adds would wait for each load in turn).
// The same for a workgroup of BS threads reading at most PER partials each.
template <int NV, int BS, int PER>
__device__ __forceinline__ void sum_partials_bs(double (&out)[NV], const double* partials,
                                                int nblocks) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double v[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int b = int(threadIdx.x) + j * BS;
      v[j] = b < nblocks ? partials[k * kMaxPartials + b] : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (int(threadIdx.x) + j * BS < nblocks) s += v[j];
    out[k] = s;
  }
  block_sum<NV, BS>(out);
}

template <int NV>
__device__ __forceinline__ void sum_partials(double (&out)[NV], const double* partials,
                                             int nblocks) {
  sum_partials_bs<NV, kFinThreads, kFinPerThread>(out, partials, nblocks);
}

// ---------------------------------------------------------------- SpMV body
// Plain CSR rows of node n in a[0..2] (valid in every lane of the 32-lane
// group): every value's column comes from col_idx.  WIDE: lane k also takes
// the neighbours k + 32, k + 64, ... (lists longer than the group).
template <typename T, bool WIDE = false>
__device__ __forceinline__ void node_rows_csr(int64_t n, int k, const int32_t* __restrict__ nb_ptr,
                                              const int32_t* __restrict__ col_idx,
                                              const T* __restrict__ val, const T* __restrict__ x,
                                              T (&a)[3]) {
  const int32_t q0 = nb_ptr[n];
  const int deg = nb_ptr[n + 1] - q0;
  a[0] = a[1] = a[2] = T(0);
  for (int kk = k; kk < deg; kk += GROUP) {
    const int64_t base = 9 * int64_t(q0) + 3 * kk;
    const int d3 = 3 * deg;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T* v = val + base + d * d3;
      const int32_t* c = col_idx + base + d * d3;
      a[d] += v[0] * x[c[0]] + v[1] * x[c[1]] + v[2] * x[c[2]];
    }
    if constexpr (!WIDE) break;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) a[d] = sum_group32(a[d]);
}

// FP64 node-block rows of node n: lane k on neighbour k (and k + 32, ... when
// WIDE) reads its id once and its three 3-value runs; the products and their
// order are those of node_rows_csr, so both give the same bits.
template <typename T, bool WIDE>
__device__ __forceinline__ void node_rows_block(int64_t n, int k,
                                                const int32_t* __restrict__ nb_ptr,
                                                const int32_t* __restrict__ nb_idx,
                                                const T* __restrict__ val, const T* __restrict__ x,
                                                T (&a)[3]) {
  const int32_t q0 = nb_ptr[n];
  const int deg = nb_ptr[n + 1] - q0;
  a[0] = a[1] = a[2] = T(0);
  for (int kk = k; kk < deg; kk += GROUP) {
    const int64_t m = nb_idx[q0 + kk];
    const int64_t base = 9 * int64_t(q0) + 3 * kk;
    const int d3 = 3 * deg;
    const T x0 = x[3 * m], x1 = x[3 * m + 1], x2 = x[3 * m + 2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T* v = val + base + d * d3;
      a[d] += v[0] * x0 + v[1] * x1 + v[2] * x2;
    }
    if constexpr (!WIDE) break;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) a[d] = sum_group32(a[d]);
}

__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// FP32 node-block view, difference form, for the two consecutive nodes n0
// (lanes 0-31) and n0 + 1 (lanes 32-63) of one wave; n0 is wave-uniform.
// Three things keep the texture path to five vector loads per wave and node
// pair (three value runs, the neighbour ids, one x gather):
//   - nb_ptr[n0 .. n0+2] are scalar loads (the wave's two nodes are adjacent;
//     the device row pointers are padded past the last node);
//   - x_n is not gathered a second time: the lane whose neighbour is the node
//     itself (every node is its own neighbour) has it, and readlane
//     broadcasts it to its group;
//   - after the DPP sums inside each 16-lane row, v_permlane16_swap pairs
//     rows 0/1 and 2/3, so every lane of a group holds the same two row sums
//     the xor-16 shuffle gave (same operands, same order, same result).
// Row sums of the group's node in a[0..2] and its own x_n[k] in xk (lanes
// k < 3); the caller writes lanes k < 3.
template <int VIEW>
__device__ __forceinline__ void node_pair_rows32(int64_t n0, int32_t n_nodes, int lane,
                                                 const int32_t* __restrict__ nb_ptr,
                                                 const int32_t* __restrict__ nb_idx,
                                                 const int16_t* __restrict__ nb_off,
                                                 const float* __restrict__ val,
                                                 const float* __restrict__ x, float (&a)[3],
                                                 float& xk) {
  const bool hi = lane >= GROUP;
  const int k = lane & (GROUP - 1);
  // The row pointers through the constant address space: wave-uniform
  // addresses there are always scalar loads.  n0 + 2 <= n_nodes + 1 is inside
  // the padded row pointers; past the last node the padding repeats
  // nb_ptr[n_nodes], so the hi group of a lone last node has degree 0.
  static_assert(VIEW == kViewOff16 || VIEW == kViewId32, "CSR-layout views");
  const auto* kptr = (const __attribute__((address_space(4))) int32_t*)nb_ptr;
  const int32_t p0 = kptr[n0], p1 = kptr[n0 + 1], p2 = kptr[n0 + 2];
  const int64_t n = n0 + (hi ? 1 : 0);
  const int32_t q0 = hi ? p1 : p0;
  const int deg = (hi ? p2 : p1) - q0;
  float v[9], xm[3] = {0.f, 0.f, 0.f};
  bool self = false;
#pragma unroll
  for (int j = 0; j < 9; ++j) v[j] = 0.f;
  if (k < deg) {
    // issue order: neighbour id, then the 9 values (they depend only on
    // nb_ptr), then -- behind a scheduling barrier, so the compiler cannot
    // sink the value loads below the wait for the id -- the x gather
    const int64_t m = VIEW == kViewOff16 ? n + int64_t(nb_off[q0 + k]) : int64_t(nb_idx[q0 + k]);
    const int64_t base = 9 * int64_t(q0) + 3 * k;
    const int d3 = 3 * deg;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c = 0; c < 3; ++c) v[3 * d + c] = val[base + d * d3 + c];
    __builtin_amdgcn_sched_barrier(0);
    self = m == n;
#pragma unroll
    for (int c = 0; c < 3; ++c) xm[c] = x[3 * m + c];
  }
  const uint64_t smask = __ballot(self);
  const uint32_t m_lo = uint32_t(smask), m_hi = uint32_t(smask >> 32);
  const int l0 = m_lo ? __builtin_ctz(m_lo) : 0;
  const int l1 = m_hi ? __builtin_ctz(m_hi) + GROUP : GROUP;
  const float s = self ? 0.f : 1.f;
  float xd[3], xn[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    // both broadcasts in every lane, then a select (a divergent readlane
    // would become two exec-masked branches)
    const float xn0 = lane_value(xm[c], l0);
    const float xn1 = lane_value(xm[c], l1);
    xn[c] = hi ? xn1 : xn0;
    xd[c] = xm[c] - s * xn[c];
  }
  xk = (k == 0) ? xn[0] : ((k == 1) ? xn[1] : xn[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float t = v[3 * d] * xd[0] + v[3 * d + 1] * xd[1] + v[3 * d + 2] * xd[2];
    t += dpp_f32<0xB1>(t);
    t += dpp_f32<0x4E>(t);
    t += dpp_f32<0x141>(t);
    t += dpp_f32<0x140>(t);
    const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(t), __float_as_uint(t),
                                                     false, false);
    a[d] = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
  }
}

// FP32 node-block view with 4-byte ids for neighbour lists of any length:
// the node pair n0, n0 + 1 of node_pair_rows32, lane k of a group on the
// neighbours k, k + 32, ...; x_n is loaded directly, the self block (the
// penalty diagonal in difference form) multiplies x_n itself, and the group
// sums its row partials with xor shuffles inside the 32 lanes.
__device__ __forceinline__ void node_pair_rows32_wide(int64_t n0, int32_t n_nodes, int lane,
                                                      const int32_t* __restrict__ nb_ptr,
                                                      const int32_t* __restrict__ nb_idx,
                                                      const float* __restrict__ val,
                                                      const float* __restrict__ x, float (&a)[3],
                                                      float& xk) {
  const bool hi = lane >= GROUP;
  const int k = lane & (GROUP - 1);
  const int64_t n = n0 + (hi ? 1 : 0);
  float xn[3] = {0.f, 0.f, 0.f}, t[3] = {0.f, 0.f, 0.f};
  if (n < n_nodes) {
    const int32_t q0 = nb_ptr[n];
    const int deg = nb_ptr[n + 1] - q0;
    const int d3 = 3 * deg;
#pragma unroll
    for (int c = 0; c < 3; ++c) xn[c] = x[3 * n + c];
    for (int kk = k; kk < deg; kk += GROUP) {
      const int64_t m = nb_idx[q0 + kk];
      const int64_t base = 9 * int64_t(q0) + 3 * kk;
      const float s = m == n ? 0.f : 1.f;
      float xd[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) xd[c] = x[3 * m + c] - s * xn[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float* v = val + base + d * d3;
        t[d] += v[0] * xd[0] + v[1] * xd[1] + v[2] * xd[2];
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int o = GROUP / 2; o > 0; o >>= 1) t[d] += __shfl_xor(t[d], o, 64);
    a[d] = t[d];
  }
  xk = (k == 0) ? xn[0] : ((k == 1) ? xn[1] : xn[2]);
}

// Stencil-pattern view over the off-diagonal values (val32m): the same pair
// of nodes n0, n0 + 1 as node_pair_rows32, with
//   - the pattern ids of both nodes as bytes of one scalar dword (n0 is even;
//     the id array is padded past the last node), issued with the row
//     pointers (the empty asm consumes the id there, so the compiler cannot
//     sink its load into the branch below, behind a wait of its own);
//   - lane k < deg - 1 on neighbour k other than the node itself (offset from
//     the pattern table, which stays in cache) with its 9 values, lane deg - 1
//     on the node itself (offset 0: x_n, no values), so readlane takes x_n
//     from a lane the row pointers name and no self test is needed;
//   - the penalty diagonal P_n x_n added after the row sums, from pen32, in
//     the pairs whose pen_flag bit is set (the clamped face only).
__device__ __forceinline__ void node_pair_rows32_offdiag(
    int64_t n0, int32_t n_nodes, int lane, const int32_t* __restrict__ nb_ptr,
    const uint8_t* __restrict__ nb_pat, const int16_t* __restrict__ pat_off,
    const float* __restrict__ valm, const float* __restrict__ pen,
    const uint32_t* __restrict__ pen_flag, const float* __restrict__ x, float (&a)[3],
    float& xk) {
  const bool hi = lane >= GROUP;
  const int k = lane & (GROUP - 1);
  const uint32_t word = ((const __attribute__((address_space(4))) uint32_t*)nb_pat)[n0 >> 2];
  const uint32_t flags =
      ((const __attribute__((address_space(4))) uint32_t*)pen_flag)[n0 >> 5];
  const auto* kptr = (const __attribute__((address_space(4))) int32_t*)nb_ptr;
  const int32_t p0 = kptr[n0], p1 = kptr[n0 + 1], p2 = kptr[n0 + 2];
  __asm__ volatile("" : : "s"(word));
  const uint32_t pid = (word >> (8 * int(n0 & 3) + (hi ? 8 : 0))) & 0xffu;
  // degrees of both groups (wave-uniform); past the last node deg_hi = 0
  const int deg_lo = p1 - p0, deg_hi = p2 - p1;
  const int64_t n = n0 + (hi ? 1 : 0);
  const int deg = hi ? deg_hi : deg_lo;
  const int nv = deg - 1;
  // node n's off-diagonal segment: every node before it dropped one block
  const int64_t seg = 9 * (int64_t(hi ? p1 : p0) - n);
  float v[9], xm[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < 9; ++j) v[j] = 0.f;
  if (k < deg) {
    // issue order: table offset, values, then (behind a scheduling barrier,
    // so the value loads are not sunk below the wait for the offset) x
    // (the self lane re-reads the last block, in bounds and without a
    // branch of its own, and multiplies it by zero below)
    const int64_t m = n + int64_t(pat_off[pid * kMaxNodeNbrs + k]);
    const int64_t base = seg + 3 * (k < nv ? k : nv - 1);
    const int d3 = 3 * nv;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c = 0; c < 3; ++c) v[3 * d + c] = valm[base + d * d3 + c];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int c = 0; c < 3; ++c) xm[c] = x[3 * m + c];
  }
  const int l0 = deg_lo - 1, l1 = GROUP + deg_hi - 1;
  const bool off = k < nv;
  float xd[3], xn[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float xn0 = lane_value(xm[c], l0);
    const float xn1 = lane_value(xm[c], l1);
    xn[c] = hi ? xn1 : xn0;
    xd[c] = off ? xm[c] - xn[c] : 0.f;
  }
  xk = (k == 0) ? xn[0] : ((k == 1) ? xn[1] : xn[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float t = v[3 * d] * xd[0] + v[3 * d + 1] * xd[1] + v[3 * d + 2] * xd[2];
    t += dpp_f32<0xB1>(t);
    t += dpp_f32<0x4E>(t);
    t += dpp_f32<0x141>(t);
    t += dpp_f32<0x140>(t);
    const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(t), __float_as_uint(t),
                                                     false, false);
    a[d] = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
  }
  // penalty rows: bits n0 & 31 and (n0 & 31) + 1 of the same word (n0 even)
  const uint32_t pbits = (flags >> (n0 & 31)) & 3u;
  if (pbits && k < 3 && n < n_nodes) {
    const float pk = pen[3 * n + k];
    const float ak = (k == 0) ? a[0] : ((k == 1) ? a[1] : a[2]);
    const float r = ak + pk * xk;
    if (k == 0) a[0] = r;
    else if (k == 1) a[1] = r;
    else a[2] = r;
  }
}

__device__ __forceinline__ double lane_value(double v, int lane) {
  const uint64_t b = uint64_t(__double_as_longlong(v));
  const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(b), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(b >> 32), lane);
  return __longlong_as_double(int64_t((uint64_t(hi) << 32) | lo));
}

// (v[0] + v[1] + v[2]) + (v[32] + v[33] + v[34]): the sum over a wave of a
// value held only by lanes 0-2 of each 32-lane group (wave-uniform result)
__device__ __forceinline__ double group3_sum(double v) {
  return ((lane_value(v, 0) + lane_value(v, 1)) + lane_value(v, 2)) +
         ((lane_value(v, GROUP) + lane_value(v, GROUP + 1)) + lane_value(v, GROUP + 2));
}

// y = K x; with sc != null an early return once the solve is done.  VIEW, the
// FP32 view's neighbour index: kViewId32 (4-byte node ids), kViewOff16
// (16-bit offsets m - n, measured 0.456 -> 0.440 ms at 128^3 against 4-byte
// ids) or kViewPattern (one stencil-pattern byte per node; the offsets come
// from the small pattern table, which stays in cache; the values are the
// off-diagonal copy val32m, 9 values per node fewer than the CSR, plus the
// penalty diagonal on the clamped nodes).
// pap != null (FP32 view inside PCG): also the workgroup partials of p.Ap,
// pap[blockIdx.x] = sum over the workgroup's nodes of double(x_i) double(y_i)
// (x = p, y = Ap as stored).  Lanes k < 3 have x_n[k] and y_n[k] in registers,
// so the dot costs no memory traffic.  The workgroup's waves combine without
// a barrier: each wave sums its lanes, stores the sum in an LDS slot and
// increments an LDS counter; the wave that arrives last adds the slots in
// wave order.  (A barrier there holds every wave until the slowest has
// finished; that cost 0.04 ms per launch at 128^3.)  The order of every sum
// is fixed, so p.Ap is bit-reproducible.
template <typename T, bool BLOCK_INDEX, int VIEW>
__global__ __launch_bounds__(kSpmvThreads) void k_spmv(
    int32_t n_nodes, const int32_t* __restrict__ nb_ptr, const int32_t* __restrict__ nb_idx,
    const int16_t* __restrict__ nb_off, const uint8_t* __restrict__ nb_pat,
    const int16_t* __restrict__ pat_off, const int32_t* __restrict__ col_idx,
    const T* __restrict__ val, const float* __restrict__ pen,
    const uint32_t* __restrict__ pen_flag, const T* __restrict__ x, T* __restrict__ y,
    const PcgScalars* __restrict__ sc, double* __restrict__ pap) {
  if (sc && sc->done) return;
  // A workgroup takes 2 x GROUPS_PER_BLOCK consecutive nodes per pass (group
  // g: nodes g and g + GROUPS_PER_BLOCK; wave w: nodes 2w, 2w + 1 and
  // 2w + GROUPS_PER_BLOCK, 2w + GROUPS_PER_BLOCK + 1), so the y it writes is
  // whole 128-byte lines and no line is shared with a workgroup on another XCD
  // (measured 0.443 -> 0.433 ms for the FP32 SpMV at 128^3)
  constexpr int kNodesPerPass = 2 * GROUPS_PER_BLOCK;
  const int64_t stride = int64_t(gridDim.x) * kNodesPerPass;
  if constexpr (BLOCK_INDEX) {
    static_assert(sizeof(T) == 4, "the node-block view is the FP32 operator");
    constexpr int kWaves = kSpmvThreads / WAVE;
    __shared__ double slot[kWaves];
    __shared__ int arrived;
    const int lane = threadIdx.x & (WAVE - 1);
    const int k = lane & (GROUP - 1);
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / WAVE);
    if (pap) {
      if (threadIdx.x == 0) arrived = 0;
      __syncthreads();
    }
    double dot = 0.0;
    for (int64_t b = int64_t(blockIdx.x) * kNodesPerPass; b < n_nodes; b += stride) {
#pragma unroll 1
      for (int h = 0; h < 2; ++h) {
        const int64_t n0 = b + 2 * w + h * GROUPS_PER_BLOCK;
        if (n0 >= n_nodes) break;
        float a[3], xk;
        if constexpr (VIEW == kViewPattern)
          node_pair_rows32_offdiag(n0, n_nodes, lane, nb_ptr, nb_pat, pat_off, val, pen,
                                   pen_flag, x, a, xk);
        else if constexpr (VIEW == kViewWide)
          node_pair_rows32_wide(n0, n_nodes, lane, nb_ptr, nb_idx, val, x, a, xk);
        else
          node_pair_rows32<VIEW>(n0, n_nodes, lane, nb_ptr, nb_idx, nb_off, val, x, a, xk);
        const int64_t n = n0 + (lane >= GROUP ? 1 : 0);
        if (k < 3 && n < n_nodes) {
          const float yk = (k == 0) ? a[0] : ((k == 1) ? a[1] : a[2]);
          y[3 * n + k] = yk;
          dot += double(xk) * double(yk);
        }
      }
    }
    if (pap) {
      dot = group3_sum(dot);
      if (lane == 0) {
        slot[w] = dot;
        const int before = __hip_atomic_fetch_add(&arrived, 1, __ATOMIC_ACQ_REL,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
        if (before == kWaves - 1) {
          double t = 0.0;
#pragma unroll
          for (int i = 0; i < kWaves; ++i)
            t += __hip_atomic_load(&slot[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          pap[blockIdx.x] = t;
        }
      }
    }
  } else {
    const int k = threadIdx.x & (GROUP - 1);
    const int gi = threadIdx.x / GROUP;
    for (int64_t b = int64_t(blockIdx.x) * kNodesPerPass; b < n_nodes; b += stride) {
#pragma unroll 1
      for (int h = 0; h < 2; ++h) {
        const int64_t n = b + gi + h * GROUPS_PER_BLOCK;
        if (n >= n_nodes) break;
        T a[3];
        if constexpr (VIEW == kViewBlock64 || VIEW == kViewBlock64Wide)
          node_rows_block<T, VIEW == kViewBlock64Wide>(n, k, nb_ptr, nb_idx, val, x, a);
        else
          node_rows_csr<T, VIEW == kViewWide>(n, k, nb_ptr, col_idx, val, x, a);
        if (k < 3) y[3 * n + k] = (k == 0) ? a[0] : ((k == 1) ? a[1] : a[2]);
      }
    }
  }
}

// the three values of a node as one 12-byte access (global_load_dwordx3 /
// global_store_dwordx3)
__device__ __forceinline__ float3 load3(const float* __restrict__ p) {
  float3 v;
  __builtin_memcpy(&v, p, 12);
  return v;
}
__device__ __forceinline__ void store3(float* __restrict__ p, float a, float b, float c) {
  const float3 v = make_float3(a, b, c);
  __builtin_memcpy(p, &v, 12);
}

// y = K x in the symmetric block-diagonal view, one node per thread:
//   y_n = sum_s K(n, n + o_s) (x_{n+o_s} - x_n) + K(n - o_s, n)^T (x_{n-o_s} - x_n)
//         + P_n x_n,
// the difference form of the assembled K, whose blocks are symmetric
// (K(n, m) = K(m, n)^T), so each off-diagonal block is read from HBM once,
// for its upper node, and its second use (as the lower block of the other
// node) is mostly a cache re-read.  A neighbour index outside [0, n_nodes)
// is replaced by n: its block is zero, and so is x_n - x_n.  Each block is
// three loads (float4, float4, float); measured 0.26 ms at 128^3 against
// 0.38 ms for the same one-node-per-thread kernel over all 26 stored blocks.
// Tile order: a workgroup takes one TILE-node tile t at a time.  Tiles
// are visited in blocks of kDiagPlanes "planes" of F = diag_pad_tiles tiles
// (F tiles span the largest offset, one node plane of a structured box):
// within a block, the tiles p, p + F, p + 2F, ... run one after another, so
// the lower blocks one plane back were fetched one tile earlier and those one
// node row back kDiagPlanes tiles earlier -- both still in L2 -- and each
// tile's entries are one contiguous range of the tile-major layout.  XCD
// order on top: workgroup b runs on XCD b % 8 and takes the positions of that
// XCD's contiguous share in steps of gridDim / 8 (gridDim a multiple of 8),
// so the blocks an XCD re-reads are in its own L2 (in a probe: 0.301 ms plain
// tile order, 0.279 XCD order, 0.250 with the blocked tile-major order).
// pap != null: the workgroup partials of p.Ap as in k_spmv (fixed order,
// barrier-free LDS combine).
// XW: the tile's x window (XWindow: the 27 neighbour offsets' nodes as a few
// runs of consecutive nodes, 3 (TILE + 260) nodes on a 128^3 box) is first loaded
// into LDS with coalesced loads, and every x_{n +- o_s} and x_n is then an LDS
// read: 27 global gathers per node become 6 streaming loads per thread.
// Window nodes outside the mesh hold 0; their blocks are zero.
// Tile order of the symmetric view, shared by k_spmv_sym and its grid: the
// P = ceil(ntiles / F) planes of F tiles (F = diag_pad_tiles) are visited in
// blocks of kDiagPlanes planes, the last block holding the P mod kDiagPlanes
// that remain; within a block of nb planes, position r is tile
// (first plane + r mod nb) F + r / nb.  P F positions, of which fewer than F
// (in the last plane) hold no tile.  Rounding the last block up to
// kDiagPlanes planes instead left up to 15 planes of empty positions at the
// end, i.e. whole XCDs idle on a box a few blocks thick (a 128 x 128 x 16
// slab: half of its positions empty, the FP32 SpMV 1.5x slower than on a
// 64^3 cube of the same size).
struct SymOrder {
  int32_t F, nfull, rem, npos;
};
__host__ __device__ inline SymOrder sym_order(int32_t n_nodes, int32_t F, int32_t tile) {
  const int32_t ntiles = (n_nodes + tile - 1) / tile;
  const int32_t P = (ntiles + F - 1) / F;
  return SymOrder{F, P / kDiagPlanes, P % kDiagPlanes, P * F};
}
__device__ __forceinline__ int32_t sym_tile(const SymOrder& o, int32_t w) {
  const int32_t blk = o.F * kDiagPlanes;
  const int32_t zb = w / blk;
  const int32_t nb = zb < o.nfull ? kDiagPlanes : o.rem;
  const int32_t r = w - zb * blk;
  const int32_t p = r / nb, zi = r - p * nb;
  return (zb * kDiagPlanes + zi) * o.F + p;
}

// diagonals whose loads are issued together (128^3: 4 and 7 0.266 ms, 13
// 0.272 ms; one at a time, as the compiler schedules it unstaged, drains the
// loads after every diagonal)
constexpr int kDiagBatch = 4;
// TILE: the layout's tile size, one node per thread (ufc_ctx::diag_tile)
template <bool XW, int TILE>
__global__ __launch_bounds__(TILE) void k_spmv_sym(
    int32_t n_nodes, DiagOffsets off, int nd, int32_t pad_tiles, XWindow win,
    const float4* __restrict__ da, const float4* __restrict__ db, const float* __restrict__ dc,
    const float* __restrict__ pen, const uint32_t* __restrict__ pen_flag,
    const float* __restrict__ x, float* __restrict__ y, const PcgScalars* __restrict__ sc,
    double* __restrict__ pap) {
  if (sc && sc->done) return;
  constexpr int kWaves = TILE / WAVE;
  __shared__ double slot[kWaves];
  __shared__ int arrived;
  __shared__ float4 xw[XW ? max_xwin(TILE) : 1];
  if (pap) {
    if (threadIdx.x == 0) arrived = 0;
    __syncthreads();
  }
  const int32_t ntiles = (n_nodes + TILE - 1) / TILE;
  const SymOrder ord = sym_order(n_nodes, pad_tiles, TILE);
  const int32_t npos = ord.npos;
  const int32_t ppx = (npos + 7) / 8;
  const int32_t lo = int32_t(blockIdx.x % 8) * ppx;
  const int32_t hi = lo + ppx < npos ? lo + ppx : npos;
  const int32_t step = int32_t(gridDim.x / 8);
  double dot = 0.0;
  for (int32_t w = lo + int32_t(blockIdx.x / 8); w < hi; w += step) {
    const int32_t t = sym_tile(ord, w);
    // 32-bit indices: n < n_nodes < 2^31 and the entries < 2^31 (the host
    // checks), which keeps the address arithmetic off 64-bit VALU pairs
    const int32_t n = t * TILE + int32_t(threadIdx.x);
    if (t >= ntiles) continue;  // uniform over the workgroup
    const int tid = int(threadIdx.x);
    if constexpr (XW) {
      // all of a thread's window loads are issued before the LDS writes (in
      // a loop of load, store per slot each load waited for the previous
      // one: 244.4 -> 243.1 us at 128^3)
      const int32_t t0 = t * TILE;
      const int32_t total = win.base[win.nc - 1] + win.len[win.nc - 1];
      constexpr int kPer = (max_xwin(TILE) + TILE - 1) / TILE;
      float3 q[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int32_t j = tid + k * TILE;
        int32_t m = -1;
#pragma unroll
        for (int c = 0; c < kMaxXClusters; ++c)
          if (c < win.nc && j >= win.base[c] && j < win.base[c] + win.len[c])
            m = t0 + win.lo[c] + (j - win.base[c]);
        q[k] = make_float3(0.f, 0.f, 0.f);
        if (j < total && m >= 0 && m < n_nodes) q[k] = load3(x + 3 * m);
      }
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int32_t j = tid + k * TILE;
        if (j < total) xw[j] = make_float4(q[k].x, q[k].y, q[k].z, 0.f);
      }
      __syncthreads();
    }
    if (n < n_nodes) {
    const int32_t tu = (t + pad_tiles) * nd;  // tile row of the upper entries
    float3 xn;
    if constexpr (XW) {
      const float4 q = xw[win.self + tid];
      xn = make_float3(q.x, q.y, q.z);
    } else {
      xn = load3(x + 3 * n);
    }
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    // Branch-free over kMaxDiags diagonals, in batches of kDiagBatch whose
    // loads are all issued before the batch's arithmetic (without the
    // staging the compiler waits for each diagonal's loads in turn): a
    // diagonal s >= nd re-reads diagonal nd - 1 at offset 0, whose
    // differences x_n - x_n make its terms exactly zero.
#pragma unroll
    for (int s0 = 0; s0 < kMaxDiags; s0 += kDiagBatch) {
      constexpr int G = kDiagBatch;
      float3 xu[G], xl[G];
      float4 ua[G], ub[G], va[G], vb[G];
      float u8[G], v8[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int s = s0 + g < kMaxDiags ? s0 + g : kMaxDiags - 1;
        const int32_t se = s < nd ? s : nd - 1;
        const int32_t o = (s0 + g < kMaxDiags && s < nd) ? off.o[s] : 0;
        if constexpr (XW) {
          // a padding diagonal (o = 0) reads x_n: its differences are zero
          const bool real = s0 + g < kMaxDiags && s < nd;
          const float4 qu = xw[(real ? win.up[s] : win.self) + tid];
          const float4 ql = xw[(real ? win.dn[s] : win.self) + tid];
          xu[g] = make_float3(qu.x, qu.y, qu.z);
          xl[g] = make_float3(ql.x, ql.y, ql.z);
        } else {
          const int32_t mu = n < n_nodes - o ? n + o : n;
          const int32_t ml = n >= o ? n - o : n;
          xu[g] = load3(x + 3 * mu);
          xl[g] = load3(x + 3 * ml);
        }
        const int32_t e = (tu + se) * TILE + int32_t(threadIdx.x);
        const int32_t m = n - o + pad_tiles * TILE;  // >= 0: o <= pad_tiles TILE
        const int32_t el = ((m / TILE) * nd + se) * TILE + m % TILE;
        ua[g] = da[e];
        ub[g] = db[e];
        u8[g] = dc[e];
        va[g] = da[el];
        vb[g] = db[el];
        v8[g] = dc[el];
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const float d0 = xu[g].x - xn.x, d1 = xu[g].y - xn.y, d2 = xu[g].z - xn.z;
        const float l0 = xl[g].x - xn.x, l1 = xl[g].y - xn.y, l2 = xl[g].z - xn.z;
        // upper block rows (ua.xyz), (ua.w ub.xy), (ub.zw u8); the lower
        // block is the transpose of the entry o_s back: columns of (va vb v8)
        a0 += ua[g].x * d0 + ua[g].y * d1 + ua[g].z * d2 + va[g].x * l0 + va[g].w * l1 +
              vb[g].z * l2;
        a1 += ua[g].w * d0 + ub[g].x * d1 + ub[g].y * d2 + va[g].y * l0 + vb[g].x * l1 +
              vb[g].w * l2;
        a2 += ub[g].z * d0 + ub[g].w * d1 + u8[g] * d2 + va[g].z * l0 + vb[g].y * l1 +
              v8[g] * l2;
      }
    }
    if ((pen_flag[n >> 5] >> (n & 31)) & 1u) {
      const float3 pn = load3(pen + 3 * n);
      a0 += pn.x * xn.x;
      a1 += pn.y * xn.y;
      a2 += pn.z * xn.z;
    }
    store3(y + 3 * n, a0, a1, a2);
    dot += (double(xn.x) * double(a0) + double(xn.y) * double(a1)) + double(xn.z) * double(a2);
    }
    if constexpr (XW) __syncthreads();  // the next tile's window reuses the LDS
  }
  if (pap) {
    dot = sum_lanes(dot, WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = int(threadIdx.x) / WAVE;
    if (lane == 0) {
      slot[w] = dot;
      const int before = __hip_atomic_fetch_add(&arrived, 1, __ATOMIC_ACQ_REL,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
      if (before == kWaves - 1) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < kWaves; ++i)
          t += __hip_atomic_load(&slot[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        pap[blockIdx.x] = t;
      }
    }
  }
}

// 16-byte accesses to V consecutive elements starting at a multiple of V
// (every solver vector is a context buffer, 256-byte aligned); a 4-byte load
// per lane streams ~40% slower than a 16-byte one on gfx950
template <typename T, int V>
__device__ __forceinline__ void load_vec(const T* __restrict__ a, int64_t i0, T (&v)[V]) {
  static_assert((sizeof(T) * V) % 16 == 0, "whole 16-byte words");
  constexpr int E = 16 / sizeof(T);
#pragma unroll
  for (int h = 0; h < V / E; ++h) {
    const uint4 q = *reinterpret_cast<const uint4*>(a + i0 + h * E);
    __builtin_memcpy(&v[h * E], &q, 16);
  }
}
template <typename T, int V>
__device__ __forceinline__ void store_vec(T* __restrict__ a, int64_t i0, const T (&v)[V]) {
  constexpr int E = 16 / sizeof(T);
#pragma unroll
  for (int h = 0; h < V / E; ++h) {
    uint4 q;
    __builtin_memcpy(&q, &v[h * E], 16);
    *reinterpret_cast<uint4*>(a + i0 + h * E) = q;
  }
}

// partials of p.Ap.  A pass of its own: inside the SpMV the per-workgroup
// reduction's barrier holds every group of a workgroup until the slowest has
// finished, which cost 0.04 ms per FP32 SpMV at 128^3; re-reading p and Ap
// (2 x 26 MB) costs 0.01 ms.
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void k_dot_pap(int64_t n, const T* __restrict__ p,
                                                          const T* __restrict__ ap,
                                                          double* __restrict__ partials,
                                                          const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  double s = 0.0;
  constexpr int V = 16 / sizeof(T);
  const int64_t tid = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x;
  const int64_t nth = int64_t(gridDim.x) * kSolveThreads;
  const int64_t nvec = n / V;
  for (int64_t j = tid; j < nvec; j += nth) {
    T pv[V], av[V];
    load_vec(p, j * V, pv);
    load_vec(ap, j * V, av);
#pragma unroll
    for (int e = 0; e < V; ++e) s += double(pv[e]) * double(av[e]);
  }
  for (int64_t i = nvec * V + tid; i < n; i += nth) s += double(p[i]) * double(ap[i]);
  double v[1] = {s};
  write_partials<1, kSolveThreads>(v, partials);
}

// p.Ap in kFinPapBlocks pieces: workgroup g adds partials
// [g chunk, (g + 1) chunk) in index order into pap_piece[g]; the update kernel
// adds the pieces (in order) and forms alpha.  Eight workgroups read the
// up-to-32768 SpMV partials in parallel (one workgroup took 8 us at 128^3).
constexpr int kFinPapBlocks = 8;
__global__ __launch_bounds__(kFinThreads) void k_fin_pap(const double* __restrict__ partials,
                                                        int nblocks,
                                                        double* __restrict__ pap_piece,
                                                        const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  const int chunk = (nblocks + kFinPapBlocks - 1) / kFinPapBlocks;
  const int lo = min(nblocks, int(blockIdx.x) * chunk);
  const int hi = min(nblocks, lo + chunk);
  double s[1];
  sum_partials<1>(s, partials + lo, hi - lo);
  if (threadIdx.x == 0) pap_piece[blockIdx.x] = s[0];
}

// x += alpha p, r -= alpha Ap; partials of r.r and r.z with z = r/diag.
// The iterate X: FP64 for the FP64 solver and for a restarted FP32 inner
// solve -- an FP32 iterate carries rounding noise of eps32 |x|, which K
// amplifies to eps32 * lambda_max |x| in the true residual, and for the
// smooth fields of an elastic solve that stalls a refinement step near
// eps32 * cond(K) (~1e-3) instead of its 1e-5 tolerance.  FP32 for an inner
// segment between residual replacements: the segment's correction is flushed
// into the FP64 u when its 0.1 reduction is reached, so the same noise is
// ~1e-3 of a residual the segment only has to cut by 10.
// the host-visible progress record (pinned host memory; null: none)
__device__ __forceinline__ void report(int32_t* prog, int slot, int32_t v) {
  if (prog) __hip_atomic_store(&prog[slot], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// p.Ap partials the update pass adds itself (per thread kFusedPapPer of them,
// one launch less per iteration than k_fin_pap's pieces; the 128^3
// symmetric SpMV's 4224: 290.2 -> 289.6 us per iteration, 34 KB from L2 per
// workgroup)
constexpr int kFusedPapPer = 18;
constexpr int kMaxFusedPap = kFusedPapPer * kSolveThreads;

// alpha of iteration it = (r.z of its iterate) / p.Ap, with p.Ap the ordered
// sum of the kFinPapBlocks pieces (the same in every workgroup), or, nspart >
// 0, the ordered sum of the nspart <= kMaxFusedPap workgroup partials the
// p.Ap pass wrote, which every workgroup adds itself (64^3: 1088 partials,
// 8.7 KB from L2 per workgroup); p.Ap <= 0 or
// not finite is a breakdown (done = 2, set by workgroup 0).  r -= alpha Ap
// and the partials of r.r and r.z with z = r/diag; the iterate is updated by
// the next pass (k_xpupdate), which reads p anyway.  Workgroup 0 leaves
// sc->gate = 1 when the iteration does not run (stopped earlier, or this
// breakdown), so that pass skips it: a flag written in the kernel before it,
// never by one of its own workgroups.
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void k_update(int64_t n, T* __restrict__ r,
                                                         const T* __restrict__ ap,
                                                         const T* __restrict__ diag,
                                                         double* __restrict__ partials,
                                                         const double* __restrict__ pap_piece,
                                                         const double* __restrict__ spart,
                                                         int nspart, PcgScalars* __restrict__ sc,
                                                         int it, int32_t* prog) {
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (sc->done) {
    if (lead) sc->gate = 1;
    return;
  }
  constexpr int V = 16 / sizeof(T);
  const int64_t tid = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x;
  const int64_t nth = int64_t(gridDim.x) * kSolveThreads;
  const int64_t nvec = n / V;
  // the thread's first chunk does not depend on alpha: its loads are in
  // flight while the p.Ap pieces are read
  T rv0[V], av0[V], dv0[V];
  const bool first = tid < nvec;
  if (first) {
    load_vec(r, tid * V, rv0);
    load_vec(ap, tid * V, av0);
    load_vec(diag, tid * V, dv0);
  }
  double pap = 0.0;
  if (nspart > 0) {
    __shared__ double sh_pap;
    double s[1];
    sum_partials_bs<1, kSolveThreads, kFusedPapPer>(s, spart, nspart);
    if (threadIdx.x == 0) sh_pap = s[0];
    __syncthreads();
    pap = sh_pap;
  } else {
#pragma unroll
    for (int g = 0; g < kFinPapBlocks; ++g) pap += pap_piece[g];
  }
  if (!(pap > 0.0) || !isfinite(pap)) {
    if (lead) {
      sc->pap = pap;
      sc->done = 2;
      sc->gate = 1;
      report(prog, 1, 2);
    }
    return;
  }
  const double alpha = sc->rzs[it & 1] / pap;
  if (lead) {
    sc->pap = pap;
    sc->alpha = alpha;
    sc->gate = 0;
  }
  const T a = T(alpha);
  double rr = 0.0, rz = 0.0;
  auto step = [&](T& ri, T api, T d) {
    ri = ri - a * api;
    const T zi = (d > T(1e-30) || d < T(-1e-30)) ? ri / d : ri;
    rr += double(ri) * double(ri);
    rz += double(ri) * double(zi);
  };
  // 16-byte chunks of V elements, then the n % V tail element by element
  if (first) {
#pragma unroll
    for (int e = 0; e < V; ++e) step(rv0[e], av0[e], dv0[e]);
    store_vec(r, tid * V, rv0);
  }
  for (int64_t j = tid + nth; j < nvec; j += nth) {
    const int64_t i0 = j * V;
    T rv[V], av[V], dv[V];
    load_vec(r, i0, rv);
    load_vec(ap, i0, av);
    load_vec(diag, i0, dv);
#pragma unroll
    for (int e = 0; e < V; ++e) step(rv[e], av[e], dv[e]);
    store_vec(r, i0, rv);
  }
  for (int64_t i = nvec * V + tid; i < n; i += nth) {
    T ri = r[i];
    step(ri, ap[i], diag[i]);
    r[i] = ri;
  }
  double v[2] = {rr, rz};
  write_partials<2, kSolveThreads>(v, partials);
}

// x += alpha p (alpha of iteration it), then, with the finalize of iteration
// it folded in, p = r/diag + beta p: every workgroup adds the update's r.r
// and r.z partials itself (2 x gridDim doubles from L2, at most 2 x 2048; the
// update pass runs on the same grid; the same ordered sum in every workgroup) and forms ||r||/||b||, the stop decision and
// beta = (r.z)_new / (r.z)_it.  The iterate is updated in every iteration
// that ran, the stopping one included; p only while the solve goes on.
// Workgroup 0 publishes iter, rnorm, rr, done and, while running, the new r.z
// into slot (it + 1) & 1 (read by the next iteration) and rz -- never the
// slot it & 1 the other workgroups read.  The entry test reads sc->gate, set
// by the update kernel before this one, not sc->done, which workgroup 0 may
// set while the others still have to update x.
// The iterate X: FP64 for the FP64 solver and for a restarted FP32 inner
// solve -- an FP32 iterate carries rounding noise of eps32 |x|, which K
// amplifies to eps32 * lambda_max |x| in the true residual, and for the
// smooth fields of an elastic solve that stalls a refinement step near
// eps32 * cond(K) (~1e-3) instead of its 1e-5 tolerance.  FP32 for an inner
// segment between residual replacements: the segment's correction is flushed
// into the FP64 u when its 0.1 reduction is reached, so the same noise is
// ~1e-3 of a residual the segment only has to cut by 10.
template <typename T, typename X>
__global__ __launch_bounds__(kSolveThreads) void k_xpupdate(int64_t n, X* __restrict__ x,
                                                           T* __restrict__ p,
                                                           const T* __restrict__ r,
                                                           const T* __restrict__ diag,
                                                           const double* __restrict__ partials,
                                                           PcgScalars* __restrict__ sc, int it,
                                                           int32_t* prog) {
  if (sc->gate) return;
  constexpr int V = 16 / sizeof(T);
  const int64_t tid = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x;
  const int64_t nth = int64_t(gridDim.x) * kSolveThreads;
  const int64_t nvec = n / V;
  // the thread's first chunk does not depend on beta: its loads are in
  // flight while the partials are added
  X xv0[V];
  T pv0[V], rv0[V], dv0[V];
  const bool first = tid < nvec;
  if (first) {
    load_vec(x, tid * V, xv0);
    load_vec(p, tid * V, pv0);
    load_vec(r, tid * V, rv0);
    load_vec(diag, tid * V, dv0);
  }
  double s[2];
  sum_partials_bs<2, kSolveThreads, kSolveBlocks / kSolveThreads>(s, partials, int(gridDim.x));
  __shared__ double sh_beta;
  __shared__ int sh_stop;
  if (threadIdx.x == 0) {
    const double rn = sqrt(s[0]) / sc->bnorm;
    int stop = 0;
    if (!isfinite(rn))
      stop = 2;
    else if (rn < sc->rtol)
      stop = 1;
    else if (it + 1 >= sc->max_iter)
      stop = 3;
    const double beta = s[1] / sc->rzs[it & 1];
    sh_beta = beta;
    sh_stop = stop;
    if (blockIdx.x == 0) {
      sc->rr = s[0];
      sc->iter = it + 1;
      sc->rnorm = rn;
      if (stop) {
        sc->done = stop;
        report(prog, 1, stop);
      } else {
        sc->beta = beta;
        sc->rz = s[1];
        sc->rzs[(it + 1) & 1] = s[1];
      }
      report(prog, 0, it + 1);
    }
  }
  __syncthreads();
  const bool stop = sh_stop != 0;
  const double alpha = sc->alpha;
  const T a = T(alpha);
  const T beta = T(sh_beta);
  auto step = [&](X& xi, T& pi, T ri, T d) {
    if constexpr (sizeof(X) == 8)
      xi = xi + alpha * double(pi);
    else
      xi = xi + a * pi;
    const T zi = (d > T(1e-30) || d < T(-1e-30)) ? ri / d : ri;
    pi = zi + beta * pi;
  };
  if (first) {
#pragma unroll
    for (int e = 0; e < V; ++e) step(xv0[e], pv0[e], rv0[e], dv0[e]);
    store_vec(x, tid * V, xv0);
    if (!stop) store_vec(p, tid * V, pv0);
  }
  for (int64_t j = tid + nth; j < nvec; j += nth) {
    const int64_t i0 = j * V;
    X xv[V];
    T pv[V], rv[V], dv[V];
    load_vec(x, i0, xv);
    load_vec(p, i0, pv);
    load_vec(r, i0, rv);
    load_vec(diag, i0, dv);
#pragma unroll
    for (int e = 0; e < V; ++e) step(xv[e], pv[e], rv[e], dv[e]);
    store_vec(x, i0, xv);
    if (!stop) store_vec(p, i0, pv);
  }
  for (int64_t i = nvec * V + tid; i < n; i += nth) {
    X xi = x[i];
    T pi = p[i];
    step(xi, pi, r[i], diag[i]);
    x[i] = xi;
    if (!stop) p[i] = pi;
  }
}

// r = b - Ax (ax null: x0 = 0), p = z = r/diag (p kept when keep_p);
// partials of r.z, b.b, r.p (with keep_p) and r.r
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void k_setup(int64_t n, const double* __restrict__ b,
                                                        const T* __restrict__ ax,
                                                        T* __restrict__ r, T* __restrict__ p,
                                                        const T* __restrict__ diag,
                                                        double* __restrict__ partials,
                                                        int keep_p) {
  double rz = 0.0, bb = 0.0, rp = 0.0, rr = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kSolveThreads) {
    const double bi = b[i];
    const T ri = ax ? T(bi - double(ax[i])) : T(bi);
    rr += double(ri) * double(ri);
    r[i] = ri;
    const T d = diag[i];
    const T zi = (d > T(1e-30) || d < T(-1e-30)) ? ri / d : ri;
    if (keep_p)
      rp += double(ri) * double(p[i]);
    else
      p[i] = zi;
    rz += double(ri) * double(zi);
    bb += bi * bi;
  }
  double v[4] = {rz, bb, rp, rr};
  write_partials<4, kSolveThreads>(v, partials);
}

// replace != 0: residual replacement in a running CG.  The new r.z enters
// beta = (r.z)_new / (r.z)_old, and gamma = (r.p)/(r.z) restores r.p = 0
// (the Galerkin condition the replaced residual has lost) in
// p = z + beta (p_old - gamma z), k_preplace.
__global__ __launch_bounds__(kFinThreads) void k_fin_setup(const double* __restrict__ partials,
                                                            PcgScalars* __restrict__ sc,
                                                            double rtol, int32_t max_iter,
                                                            int replace, int32_t* prog) {
  double s[4];
  sum_partials<4>(s, partials, kSolveBlocks);
  if (threadIdx.x == 0) {
    const double rz_old = sc->rz;
    sc->rz = s[0];
    sc->rzs[0] = s[0];
    double bn = sqrt(s[1]);
    if (bn < 1.0e-30) bn = 1.0;
    sc->bnorm = bn;
    sc->rtol = rtol;
    sc->max_iter = max_iter;
    sc->iter = 0;
    sc->done = max_iter > 0 ? 0 : 3;
    sc->rnorm = 0.0;
    sc->alpha = sc->pap = sc->rr = 0.0;
    const bool ok = replace && rz_old > 0.0 && isfinite(rz_old) && s[0] > 0.0;
    sc->beta = ok ? s[0] / rz_old : 0.0;
    sc->gamma = ok ? s[2] / s[0] : 0.0;
    // an initial residual already below rtol (b = 0, or x0 the solution)
    // converges with 0 iterations; NM_Solver_PCG would form alpha = 0/0 here
    const double rn0 = sqrt(s[3]) / bn;
    if (sc->done == 0 && rn0 < rtol) {
      sc->done = 1;
      sc->rnorm = rn0;
    }
    if (sc->done) report(prog, 1, sc->done);
  }
}

// p = z + beta (p - gamma z) after a residual replacement
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void k_preplace(int64_t n, T* __restrict__ p,
                                                           const T* __restrict__ r,
                                                           const T* __restrict__ diag,
                                                           const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  const double beta = sc->beta, gamma = sc->gamma;
  for (int64_t i = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kSolveThreads) {
    const T d = diag[i];
    const T ri = r[i];
    const T zi = (d > T(1e-30) || d < T(-1e-30)) ? ri / d : ri;
    p[i] = T(double(zi) + beta * (double(p[i]) - gamma * double(zi)));
  }
}

// v.v, ordered two-level sum; result in sc->rr
__global__ __launch_bounds__(kSolveThreads) void k_sumsq(int64_t n, const double* __restrict__ v,
                                                        double* __restrict__ partials) {
  double s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kSolveThreads)
    s += v[i] * v[i];
  double a[1] = {s};
  write_partials<1, kSolveThreads>(a, partials);
}

__global__ __launch_bounds__(kFinThreads) void k_fin_sumsq(const double* __restrict__ partials,
                                                            PcgScalars* __restrict__ sc) {
  double s[1];
  sum_partials<1>(s, partials, kSolveBlocks);
  if (threadIdx.x == 0) sc->rr = s[0];
}

__global__ void k_residual(int64_t n, const double* __restrict__ b, const double* __restrict__ y,
                           double* __restrict__ r) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
}

__global__ void k_residual_pen(int32_t nb, const int32_t* __restrict__ eq,
                               const double* __restrict__ pen, const double* __restrict__ u,
                               double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) r[eq[i]] -= pen[i] * u[eq[i]];
}

__global__ void k_axpy_f32(int64_t n, const float* __restrict__ d, double* __restrict__ u) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) u[i] += double(d[i]);
}

__global__ void k_cast_f32(int64_t n, const double* __restrict__ a, float* __restrict__ o) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) o[i] = float(a[i]);
}

__global__ void k_cast_f64(int64_t n, const float* __restrict__ a, double* __restrict__ o) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) o[i] = double(a[i]);
}

// Conversions between node order and the box numbering of the symmetric
// FP32 view (perm[n] = box number of node n, iperm its inverse); one thread
// per box number i, which writes its 3 contiguous entries and reads node
// iperm[i]'s 3 contiguous ones.
template <typename A, typename B>
__global__ void k_to_box(int32_t n_nodes, const int32_t* __restrict__ iperm,
                         const A* __restrict__ a, B* __restrict__ o) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const int64_t n = iperm[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) o[3 * i + d] = B(a[3 * n + d]);
}
template <typename A, typename B>
__global__ void k_from_box(int32_t n_nodes, const int32_t* __restrict__ iperm,
                           const A* __restrict__ a, B* __restrict__ o) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const int64_t n = iperm[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) o[3 * n + d] = B(a[3 * i + d]);
}
// u (node order) += d (box numbering)
__global__ void k_axpy_f32_box(int32_t n_nodes, const int32_t* __restrict__ iperm,
                               const float* __restrict__ d, double* __restrict__ u) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const int64_t n = iperm[i];
#pragma unroll
  for (int k = 0; k < 3; ++k) u[3 * n + k] += double(d[3 * i + k]);
}

// col_idx of node n's three rows from its neighbour list (one wave per node)
__global__ __launch_bounds__(kSolveThreads) void k_expand_cols(int32_t n_nodes,
                                                              const int32_t* __restrict__ nb_ptr,
                                                              const int32_t* __restrict__ nb_idx,
                                                              int32_t* __restrict__ col_idx) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = int64_t(blockIdx.x) * WAVES_PER_BLOCK + threadIdx.x / WAVE;
  const int64_t nw = int64_t(gridDim.x) * WAVES_PER_BLOCK;
  for (int64_t n = wave; n < n_nodes; n += nw) {
    const int32_t q0 = nb_ptr[n];
    const int deg = nb_ptr[n + 1] - q0;
    const int d3 = 3 * deg;
    for (int o = lane; o < 9 * deg; o += WAVE) {
      const int d = (o >= d3) + (o >= 2 * d3);
      const int rem = o - d * d3;
      const int k = rem / 3;
      col_idx[9 * int64_t(q0) + o] = 3 * nb_idx[q0 + k] + (rem - 3 * k);
    }
  }
}

inline unsigned blocks_for(int64_t n, int t) { return unsigned((n + t - 1) / t); }

// SpMV launches are bracketed by event pairs when per-kernel timing is on:
// every per_kernel-th launch (an event record costs ~5.6 us of queue time, so
// sampling keeps the timed solve close to an untimed one).  Each pair keeps
// the index of the launch it brackets, so pairs of launches that returned at
// entry (after convergence) can be left out of the average.
void spmv_event_begin(ufc_ctx* c) {
  c->spmv_sampled = c->per_kernel > 0 && c->spmv_launches % c->per_kernel == 0;
  if (!c->spmv_sampled) return;
  if (c->spmv_ev_used + 2 > c->spmv_ev.size()) {
    for (int k = 0; k < 256; ++k) {
      hipEvent_t e;
      UFC_HIP_CHECK(hipEventCreate(&e));
      c->spmv_ev.push_back(e);
    }
  }
  UFC_HIP_CHECK(hipEventRecord(c->spmv_ev[c->spmv_ev_used], c->stream));
}
void spmv_event_end(ufc_ctx* c) {
  if (c->spmv_sampled) {
    UFC_HIP_CHECK(hipEventRecord(c->spmv_ev[c->spmv_ev_used + 1], c->stream));
    c->spmv_ev_used += 2;
    c->spmv_active.push_back(1);
    c->spmv_pair_launch.push_back(c->spmv_launches);
  }
  ++c->spmv_launches;
}

void poll(ufc_ctx* c) {
  UFC_HIP_CHECK(hipMemcpyAsync(c->h_scalars, c->scalars.get(), sizeof(PcgScalars),
                               hipMemcpyDeviceToHost, c->stream));
  UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
}

// SpMV grid: up to kSpmvBlocks{64,32} workgroups, never more than the nodes need
template <typename T>
int spmv_blocks(const ufc_ctx* c) {
  const int cap = sizeof(T) == 8 ? kSpmvBlocks64 : kSpmvBlocks32;
  const int64_t need = (int64_t(c->n_nodes) + 2 * GROUPS_PER_BLOCK - 1) / (2 * GROUPS_PER_BLOCK);
  return int(need < cap ? need : cap);
}

// grid of the per-iteration vector passes (p.Ap, update, x/p): kSolveBlocks
// workgroups, fewer when the vector has fewer 16-byte chunks than that many
// threads (64^3 FP32: 805), so that no workgroup only adds partials; a
// function of n alone, so the partial sums keep their order run to run
template <typename T>
int vec_blocks(int64_t n) {
  constexpr int V = 16 / sizeof(T);
  const int64_t need = (n / V + kSolveThreads - 1) / kSolveThreads;
  return int(std::max<int64_t>(1, std::min<int64_t>(need, kSolveBlocks)));
}

// symmetric block-diagonal SpMV grid: 8 x the tile positions per XCD, capped (a
// multiple of 8 either way)
int sym_blocks(const ufc_ctx* c) {
  const int64_t npos = sym_order(c->n_nodes, c->diag_pad_tiles, c->diag_tile).npos;
  const int64_t ppx = (npos + 7) / 8;
  return int(8 * std::min<int64_t>(ppx, kSpmvBlocks32 / 8));
}

// workgroup partials the FP32 SpMV writes (its grid)
int fp32_spmv_blocks(const ufc_ctx* c) {
  return c->fp32_view == kFp32ViewSymDiag ? sym_blocks(c) : spmv_blocks<float>(c);
}

// pap: workgroup partials of x.y (FP32 view only), or null
template <typename T, bool BLOCK_INDEX>
void launch_spmv(ufc_ctx* c, const T* val, const T* x, T* y, bool in_pcg,
                 double* pap = nullptr) {
  const PcgScalars* sc = in_pcg ? c->scalars.get() : nullptr;
  const dim3 grid(spmv_blocks<T>(c)), block(kSpmvThreads);
  if constexpr (BLOCK_INDEX) {
    if (c->fp32_view == kFp32ViewSymDiag) {
      DiagOffsets off;
      for (int s = 0; s < kMaxDiags; ++s) off.o[s] = c->h_diag_off[s];
      auto launch = [&](auto kernel, int threads) {
        hipLaunchKernelGGL(kernel, dim3(sym_blocks(c)), dim3(threads), 0, c->stream, c->n_nodes,
                           off, int(c->n_diag), c->diag_pad_tiles, c->xwin,
                           reinterpret_cast<const float4*>(c->diag_a.get()),
                           reinterpret_cast<const float4*>(c->diag_b.get()), c->diag_c.get(),
                           c->pen32.get(), c->pen_flag.get(), x, y, sc, pap);
      };
      if (c->diag_tile == kDiagTileWide) {
        if (c->have_xwin)
          launch(k_spmv_sym<true, kDiagTileWide>, kDiagTileWide);
        else
          launch(k_spmv_sym<false, kDiagTileWide>, kDiagTileWide);
      } else if (c->have_xwin) {
        launch(k_spmv_sym<true, kDiagTile>, kDiagTile);
      } else {
        launch(k_spmv_sym<false, kDiagTile>, kDiagTile);
      }
      return;
    }
    // the stencil-pattern view reads the off-diagonal copy of val (val32m)
    if (c->fp32_view == kFp32ViewPattern)
      hipLaunchKernelGGL((k_spmv<T, true, kViewPattern>), grid, block, 0, c->stream, c->n_nodes,
                         c->nb_ptr.get(), c->nb_idx.get(), c->nb_off16.get(), c->nb_pat.get(),
                         c->pat_off.get(), c->col_idx.get(), c->val32m.get(), c->pen32.get(),
                         c->pen_flag.get(), x, y, sc, pap);
    else if (c->fp32_view == kFp32ViewOff16)
      hipLaunchKernelGGL((k_spmv<T, true, kViewOff16>), grid, block, 0, c->stream, c->n_nodes,
                         c->nb_ptr.get(), c->nb_idx.get(), c->nb_off16.get(), c->nb_pat.get(),
                         c->pat_off.get(), c->col_idx.get(), val, nullptr, nullptr, x, y, sc,
                         pap);
    else if (c->wide)
      hipLaunchKernelGGL((k_spmv<T, true, kViewWide>), grid, block, 0, c->stream, c->n_nodes,
                         c->nb_ptr.get(), c->nb_idx.get(), c->nb_off16.get(), c->nb_pat.get(),
                         c->pat_off.get(), c->col_idx.get(), val, nullptr, nullptr, x, y, sc,
                         pap);
    else
      hipLaunchKernelGGL((k_spmv<T, true, kViewId32>), grid, block, 0, c->stream, c->n_nodes,
                         c->nb_ptr.get(), c->nb_idx.get(), c->nb_off16.get(), c->nb_pat.get(),
                         c->pat_off.get(), c->col_idx.get(), val, nullptr, nullptr, x, y, sc,
                         pap);
  } else if (c->fp64_view == kFp64ViewBlock) {
    if (c->wide)
      hipLaunchKernelGGL((k_spmv<T, false, kViewBlock64Wide>), grid, block, 0, c->stream,
                         c->n_nodes, c->nb_ptr.get(), c->nb_idx.get(), c->nb_off16.get(),
                         c->nb_pat.get(), c->pat_off.get(), c->col_idx.get(), val, nullptr,
                         nullptr, x, y, sc, nullptr);
    else
      hipLaunchKernelGGL((k_spmv<T, false, kViewBlock64>), grid, block, 0, c->stream,
                         c->n_nodes, c->nb_ptr.get(), c->nb_idx.get(), c->nb_off16.get(),
                         c->nb_pat.get(), c->pat_off.get(), c->col_idx.get(), val, nullptr,
                         nullptr, x, y, sc, nullptr);
  } else if (c->wide) {
    hipLaunchKernelGGL((k_spmv<T, false, kViewWide>), grid, block, 0, c->stream, c->n_nodes,
                       c->nb_ptr.get(), c->nb_idx.get(), c->nb_off16.get(), c->nb_pat.get(),
                       c->pat_off.get(), c->col_idx.get(), val, nullptr, nullptr, x, y, sc,
                       nullptr);
  } else {
    hipLaunchKernelGGL((k_spmv<T, false, kViewId32>), grid, block, 0, c->stream, c->n_nodes,
                       c->nb_ptr.get(), c->nb_idx.get(), c->nb_off16.get(), c->nb_pat.get(),
                       c->pat_off.get(), c->col_idx.get(), val, nullptr, nullptr, x, y, sc,
                       nullptr);
  }
}

// resume: continue a CG after a residual replacement -- r := b (the new true
// residual), p := z + beta (p - gamma z) keeps the search direction (and with
// it the Krylov information) instead of restarting from p = z
template <typename T, bool BLOCK_INDEX, typename X>
int pcg_run(ufc_ctx* c, const double* b, X* x, const T* ax_or_null, T* r, T* p, T* ap,
            const T* diag, const T* val, int max_iter, double rtol, int check_every,
            int* niter, double* rnorm, bool resume = false) {
  const int64_t n = c->n_dof;
  double* part = c->partials.get();
  double* pap_piece = c->pap_piece.get();
  PcgScalars* sc = c->scalars.get();
  // Host loop: with check_every > 0, batches of check_every iterations, each
  // followed by a poll that drains the stream.  By default the GPU reports
  // every finished iteration (and a stop) into pinned host memory and the
  // host keeps at most kAhead iterations queued beyond the last one reported:
  // no poll bubbles (~40 us each) and at most kAhead iterations launched
  // after convergence (each returns at entry but costs ~55 us of launches).
  // A report that does not arrive within 50 ms falls back to a poll.  The
  // stopping iteration is exact either way.
  constexpr int kAhead = 3;
  const bool flagged = check_every <= 0;
  volatile int32_t* prog = c->h_progress;
  int32_t* dprog = flagged ? c->d_progress : nullptr;
  prog[0] = 0;
  prog[1] = 0;
  hipLaunchKernelGGL(k_setup<T>, dim3(kSolveBlocks), dim3(kSolveThreads), 0, c->stream, n, b,
                     ax_or_null, r, p, diag, part, int(resume));
  hipLaunchKernelGGL(k_fin_setup, dim3(1), dim3(kFinThreads), 0, c->stream, part, sc, rtol,
                     int32_t(max_iter), int(resume), dprog);
  if (resume)
    hipLaunchKernelGGL(k_preplace<T>, dim3(kSolveBlocks), dim3(kSolveThreads), 0, c->stream, n,
                       p, r, diag, sc);
  UFC_HIP_CHECK(hipGetLastError());
  const int64_t first_launch = c->spmv_launches;
  // p.Ap partials: written by the FP32 SpMV itself (its grid), or by
  // k_dot_pap (FP64, the vector grid); at most kMaxFusedPap of them are added
  // by the update pass directly, more go through k_fin_pap's eight pieces
  const int nvb = vec_blocks<T>(n);
  double* spart = c->spmv_partials.get();
  const int npap = BLOCK_INDEX ? fp32_spmv_blocks(c) : nvb;
  const int fused = npap <= kMaxFusedPap ? npap : 0;
  auto launch_iteration = [&](int it) {
    spmv_event_begin(c);
    if constexpr (BLOCK_INDEX) {
      // the FP32 SpMV forms the p.Ap partials itself
      launch_spmv<T, BLOCK_INDEX>(c, val, p, ap, true, spart);
      spmv_event_end(c);
    } else {
      launch_spmv<T, BLOCK_INDEX>(c, val, p, ap, true);
      spmv_event_end(c);
      hipLaunchKernelGGL(k_dot_pap<T>, dim3(nvb), dim3(kSolveThreads), 0, c->stream, n, p, ap,
                         spart, sc);
    }
    if (!fused)
      hipLaunchKernelGGL(k_fin_pap, dim3(kFinPapBlocks), dim3(kFinThreads), 0, c->stream, spart,
                         npap, pap_piece, sc);
    hipLaunchKernelGGL(k_update<T>, dim3(nvb), dim3(kSolveThreads), 0, c->stream, n, r, ap, diag,
                       part, pap_piece, spart, fused, sc, it, dprog);
    hipLaunchKernelGGL((k_xpupdate<T, X>), dim3(nvb), dim3(kSolveThreads), 0, c->stream, n, x, p,
                       r, diag, part, sc, it, dprog);
  };
  if (flagged) {
    for (int launched = 0; launched < max_iter; ++launched) {
      if (prog[1] != 0) break;
      if (launched - prog[0] >= kAhead) {
        const auto t0 = std::chrono::steady_clock::now();
        bool drained = false;
        while (prog[1] == 0 && launched - prog[0] >= kAhead) {
          std::this_thread::yield();
          if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
            poll(c);  // no report seen: drain the stream and read the scalars
            drained = true;
            break;
          }
        }
        if (prog[1] != 0 || (drained && c->h_scalars->done)) break;
      }
      launch_iteration(launched);
    }
    UFC_HIP_CHECK(hipGetLastError());
    poll(c);
  } else {
    int launched = 0;
    for (;;) {
      for (int k = 0; k < check_every && launched < max_iter; ++k, ++launched)
        launch_iteration(launched);
      UFC_HIP_CHECK(hipGetLastError());
      poll(c);
      if (c->h_scalars->done || launched >= max_iter) break;
    }
  }
  *niter = c->h_scalars->iter;
  *rnorm = c->h_scalars->rnorm;
  // SpMV launches after the converging iteration returned at entry
  for (size_t i = 0; i < c->spmv_active.size(); ++i)
    if (c->spmv_pair_launch[i] >= first_launch + *niter) c->spmv_active[i] = 0;
  const int done = c->h_scalars->done;
  if (done == 1) return UFC_STATUS_OK;
  if (done == 2) return UFC_STATUS_ERROR;
  return UFC_STATUS_INVALID;
}

}  // namespace

// ---------------------------------------------------------------- host entry points
void ensure_solver_buffers(ufc_ctx* c, bool need64, bool need32) {
  const size_t n = size_t(c->n_dof);
  c->w_r.resize(n);
  c->w_x.resize(n);
  c->w_b.resize(n);
  c->w_t.resize(n);
  if (need64) {
    c->w_p.resize(n);
    c->w_ap.resize(n);
  }
  if (need32) {
    c->s_r.resize(n);
    c->s_p.resize(n);
    c->s_ap.resize(n);
    c->s_x.resize(n);
  }
  c->spmv_partials.resize(size_t(kSpmvBlocks32));
  c->partials.resize(size_t(kMaxPartials) * 4);
  c->pap_piece.resize(size_t(kFinPapBlocks));
  if (!c->scalars.get()) {
    c->scalars.resize(1);
    UFC_HIP_CHECK(hipMemsetAsync(c->scalars.get(), 0, sizeof(PcgScalars), c->stream));
  }
}

void build_col_idx(ufc_ctx* c) {
  c->col_idx.resize(size_t(c->nnz));
  hipLaunchKernelGGL(k_expand_cols, dim3(kSolveBlocks), dim3(kSolveThreads), 0, c->stream,
                     c->n_nodes, c->nb_ptr.get(), c->nb_idx.get(), c->col_idx.get());
  UFC_HIP_CHECK(hipGetLastError());
}

void spmv64(ufc_ctx* c, const double* x, double* y) {
  spmv_event_begin(c);
  launch_spmv<double, false>(c, c->val64.get(), x, y, false);
  spmv_event_end(c);
  UFC_HIP_CHECK(hipGetLastError());
}

void spmv32(ufc_ctx* c, const float* x, float* y) {
  spmv_event_begin(c);
  launch_spmv<float, true>(c, c->val32.get(), x, y, false);
  spmv_event_end(c);
  UFC_HIP_CHECK(hipGetLastError());
}

void residual64(ufc_ctx* c, const double* b, const double* y, const double* u, double* r) {
  hipLaunchKernelGGL(k_residual, dim3(blocks_for(c->n_dof, 256)), dim3(256), 0, c->stream,
                     c->n_dof, b, y, r);
  const int32_t nb = int32_t(c->h_bc_eq.size());
  if (nb)
    hipLaunchKernelGGL(k_residual_pen, dim3(blocks_for(nb, 256)), dim3(256), 0, c->stream, nb,
                       c->bc_eq.get(), c->bc_pen.get(), u, r);
  UFC_HIP_CHECK(hipGetLastError());
}

double norm2_64(ufc_ctx* c, const double* v) {
  hipLaunchKernelGGL(k_sumsq, dim3(kSolveBlocks), dim3(kSolveThreads), 0, c->stream, c->n_dof, v,
                     c->partials.get());
  hipLaunchKernelGGL(k_fin_sumsq, dim3(1), dim3(kFinThreads), 0, c->stream, c->partials.get(),
                     c->scalars.get());
  UFC_HIP_CHECK(hipGetLastError());
  poll(c);
  return std::sqrt(c->h_scalars->rr);
}

int pcg64(ufc_ctx* c, const double* b, double* x, bool x0_zero, int max_iter, double rtol,
          int check_every, int* niter, double* rnorm) {
  const double* ax = nullptr;
  if (x0_zero) {
    UFC_HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * size_t(c->n_dof), c->stream));
  } else {
    spmv64(c, x, c->w_ap.get());
    ax = c->w_ap.get();
  }
  return pcg_run<double, false>(c, b, x, ax, c->w_r.get(), c->w_p.get(), c->w_ap.get(),
                                c->diag64.get(), c->val64.get(), max_iter, rtol, check_every,
                                niter, rnorm);
}

// In the box numbering (fp32_permuted) the right-hand side is carried over
// into w_perm and the Jacobi diagonal is diag32p; the FP32 vectors never leave
// that numbering between segments, so a resumed CG keeps its direction.
static const double* rhs_view(ufc_ctx* c, const double* b64) {
  if (!fp32_permuted(c)) return b64;
  c->w_perm.resize(size_t(c->n_dof));
  hipLaunchKernelGGL((k_to_box<double, double>), dim3(blocks_for(c->n_nodes, 256)), dim3(256),
                     0, c->stream, c->n_nodes, c->sym_iperm.get(), b64, c->w_perm.get());
  UFC_HIP_CHECK(hipGetLastError());
  return c->w_perm.get();
}

static const float* diag_view(ufc_ctx* c) {
  return fp32_permuted(c) ? c->diag32p.get() : c->diag32.get();
}

int pcg32(ufc_ctx* c, const double* b64, double* x, bool zero_x, bool resume, int max_iter,
          double rtol, int check_every, int* niter, double* rnorm) {
  // the iteration starts from 0; without zero_x the correction accumulates
  // into the caller's x (the refinement's u += d)
  if (zero_x) UFC_HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * size_t(c->n_dof), c->stream));
  const double* b = rhs_view(c, b64);
  double* xv = x;
  const unsigned g = blocks_for(c->n_nodes, 256);
  if (fp32_permuted(c)) {
    c->w_perm2.resize(size_t(c->n_dof));
    xv = c->w_perm2.get();
    hipLaunchKernelGGL((k_to_box<double, double>), dim3(g), dim3(256), 0, c->stream, c->n_nodes,
                       c->sym_iperm.get(), x, xv);
  }
  const int code = pcg_run<float, true, double>(c, b, xv, nullptr, c->s_r.get(), c->s_p.get(),
                                                c->s_ap.get(), diag_view(c), c->val32.get(),
                                                max_iter, rtol, check_every, niter, rnorm,
                                                resume);
  if (xv != x)
    hipLaunchKernelGGL((k_from_box<double, double>), dim3(g), dim3(256), 0, c->stream,
                       c->n_nodes, c->sym_iperm.get(), xv, x);
  UFC_HIP_CHECK(hipGetLastError());
  return code;
}

int pcg32_segment(ufc_ctx* c, const double* b64, float* d, bool resume, int max_iter,
                  double rtol, int check_every, int* niter, double* rnorm) {
  UFC_HIP_CHECK(hipMemsetAsync(d, 0, sizeof(float) * size_t(c->n_dof), c->stream));
  const double* b = rhs_view(c, b64);
  return pcg_run<float, true, float>(c, b, d, nullptr, c->s_r.get(), c->s_p.get(),
                                     c->s_ap.get(), diag_view(c), c->val32.get(), max_iter,
                                     rtol, check_every, niter, rnorm, resume);
}

void to_view32(ufc_ctx* c, const double* a, float* o) {
  if (!fp32_permuted(c)) return cast_f64_to_f32(c, a, o);
  hipLaunchKernelGGL((k_to_box<double, float>), dim3(blocks_for(c->n_nodes, 256)), dim3(256), 0,
                     c->stream, c->n_nodes, c->sym_iperm.get(), a, o);
  UFC_HIP_CHECK(hipGetLastError());
}

void from_view32(ufc_ctx* c, const float* a, double* o) {
  if (!fp32_permuted(c)) return cast_f32_to_f64(c, a, o);
  hipLaunchKernelGGL((k_from_box<float, double>), dim3(blocks_for(c->n_nodes, 256)), dim3(256),
                     0, c->stream, c->n_nodes, c->sym_iperm.get(), a, o);
  UFC_HIP_CHECK(hipGetLastError());
}

void cast_f64_to_f32(ufc_ctx* c, const double* a, float* o) {
  hipLaunchKernelGGL(k_cast_f32, dim3(blocks_for(c->n_dof, 256)), dim3(256), 0, c->stream,
                     c->n_dof, a, o);
  UFC_HIP_CHECK(hipGetLastError());
}

void cast_f32_to_f64(ufc_ctx* c, const float* a, double* o) {
  hipLaunchKernelGGL(k_cast_f64, dim3(blocks_for(c->n_dof, 256)), dim3(256), 0, c->stream,
                     c->n_dof, a, o);
  UFC_HIP_CHECK(hipGetLastError());
}

double spmv_event_avg_ms(ufc_ctx* c, int* calls) {
  double tot = 0.0;
  int k = 0;
  for (size_t i = 0; i + 1 < c->spmv_ev_used; i += 2) {
    if (!c->spmv_active[i / 2]) continue;
    float ms = 0.f;
    UFC_HIP_CHECK(hipEventElapsedTime(&ms, c->spmv_ev[i], c->spmv_ev[i + 1]));
    tot += ms;
    ++k;
  }
  *calls = k;
  return k ? tot / k : 0.0;
}

void reset_spmv_events(ufc_ctx* c) {
  c->spmv_ev_used = 0;
  c->spmv_active.clear();
  c->spmv_pair_launch.clear();
  c->spmv_launches = 0;
}

void axpy_f32_to_f64(ufc_ctx* c, const float* d, double* u) {
  if (fp32_permuted(c))
    hipLaunchKernelGGL(k_axpy_f32_box, dim3(blocks_for(c->n_nodes, 256)), dim3(256), 0,
                       c->stream, c->n_nodes, c->sym_iperm.get(), d, u);
  else
    hipLaunchKernelGGL(k_axpy_f32, dim3(blocks_for(c->n_dof, 256)), dim3(256), 0, c->stream,
                       c->n_dof, d, u);
  UFC_HIP_CHECK(hipGetLastError());
}

}  // namespace ufc
