// ufc_elem.hip -- gfx950 element kernels: C3D8 stiffness (batch and direct
// into the CSR values through the element-to-slot map), element forces and
// their atomics-free node gather, adjoint sensitivities, J2 material update.
//
// Element math follows PH_Elem_C3D8.f90 (reference paths relative to its
// root) with two documented corrections: trilinear shape functions carry the
// factor 1/8 (the reference's QUARTER at :1936-1939 doubles sum N), and
// dN/dx = J^-T dN/dxi for J(j,k) = dx_j/dxi_k (the reference applies J^-1 at
// :936-943; identical for the symmetric J of axis-aligned elements).
//
// Instead of 24x24 dense B^T D B products (PH_Elem_C3D8.f90:776-778) the 3x3
// block of nodes (a, b) is formed directly:
//   K_ab(i,j) = sum_gp w detJ [lambda g_a,i g_b,j + mu g_a,j g_b,i + mu d_ij g_a.g_b]
// which is the same sum of the same products with B's zeros skipped.
#include "ufc_internal.hpp"

namespace ufc {
namespace {

// PH_ELEM_GAUSS_PT (PH_Elem_C3D8.f90:118), the reference's 15-digit literal
constexpr double kGaussPt = 0.577350269189626;

template <typename T>
__device__ __forceinline__ T tabs(T v) {
  return v < T(0) ? -v : v;
}

// Shape-function derivatives w.r.t. x at Gauss point gp (order of
// PH_Elem_C3D8_GaussPoints, :694-715: xi outer, zeta inner; weights 1).
// Returns false when |detJ| <= 1e-12 (the point is skipped, :775).
template <typename T>
__device__ __forceinline__ bool c3d8_dndx(const T (&x)[8][3], int gp, T (&g)[8][3], T& detJ) {
  constexpr int XS[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
  constexpr int YS[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
  constexpr int ZS[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  const T xi = (gp & 4) ? T(kGaussPt) : T(-kGaussPt);
  const T eta = (gp & 2) ? T(kGaussPt) : T(-kGaussPt);
  const T zeta = (gp & 1) ? T(kGaussPt) : T(-kGaussPt);
  T dn[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const T fx = T(1) + T(XS[a]) * xi;
    const T fy = T(1) + T(YS[a]) * eta;
    const T fz = T(1) + T(ZS[a]) * zeta;
    dn[a][0] = T(0.125) * T(XS[a]) * fy * fz;
    dn[a][1] = T(0.125) * fx * T(YS[a]) * fz;
    dn[a][2] = T(0.125) * fx * fy * T(ZS[a]);
  }
  T J[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T s = T(0);
#pragma unroll
      for (int a = 0; a < 8; ++a) s += dn[a][k] * x[a][j];
      J[j][k] = s;
    }
  const T c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const T c01 = J[1][0] * J[2][2] - J[1][2] * J[2][0];
  const T c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const T d = J[0][0] * c00 - J[0][1] * c01 + J[0][2] * c02;
  detJ = d;
  if (tabs(d) <= T(1.0e-12)) return false;
  T Ji[3][3];
  Ji[0][0] = c00 / d;
  Ji[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / d;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
  Ji[1][0] = -c01 / d;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
  Ji[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / d;
  Ji[2][0] = c02 / d;
  Ji[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / d;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int j = 0; j < 3; ++j) g[a][j] = Ji[0][j] * dn[a][0] + Ji[1][j] * dn[a][1] + Ji[2][j] * dn[a][2];
  return true;
}

template <typename T>
__device__ __forceinline__ void gather_coords(const double* __restrict__ coords,
                                              const int32_t* __restrict__ conn, int64_t e,
                                              T (&x)[8][3]) {
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t n = conn[8 * e + a];
    x[a][0] = T(coords[3 * n + 0]);
    x[a][1] = T(coords[3 * n + 1]);
    x[a][2] = T(coords[3 * n + 2]);
  }
}

// Block (a, b) of B_a^T D B_b for a general (6,6) column-major D.
template <typename T>
__device__ __forceinline__ void btdb_general(const T (&bd)[3][6], const T* gb, T w,
                                             T (&k)[3][3]) {
  const T bx = gb[0], by = gb[1], bz = gb[2];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    k[i][0] += w * (bd[i][0] * bx + bd[i][3] * by + bd[i][5] * bz);
    k[i][1] += w * (bd[i][1] * by + bd[i][3] * bx + bd[i][4] * bz);
    k[i][2] += w * (bd[i][2] * bz + bd[i][4] * by + bd[i][5] * bx);
  }
}

// B_a^T D for column-major D (6,6): rows of B_a^T are
// [gx,0,0,gy,0,gz], [0,gy,0,gx,gz,0], [0,0,gz,0,gy,gx].
template <typename T>
__device__ __forceinline__ void bt_d(const T* ga, const double* __restrict__ D, T (&bd)[3][6]) {
  const T gx = ga[0], gy = ga[1], gz = ga[2];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const T d0 = T(D[0 + 6 * m]), d1 = T(D[1 + 6 * m]), d2 = T(D[2 + 6 * m]);
    const T d3 = T(D[3 + 6 * m]), d4 = T(D[4 + 6 * m]), d5 = T(D[5 + 6 * m]);
    bd[0][m] = gx * d0 + gy * d3 + gz * d5;
    bd[1][m] = gy * d1 + gx * d3 + gz * d4;
    bd[2][m] = gz * d2 + gy * d4 + gx * d5;
  }
}

// ---------------------------------------------------------------- batch Ke
// One wave64 per element, lane = (a, b): the 3x3 block K_ab over 8 Gauss
// points, written column-major into Ke(24,24,e).
template <typename T>
__global__ __launch_bounds__(64) void k_eval_ke(int32_t n_elems, const double* __restrict__ xe,
                                                const double* __restrict__ young, double nu,
                                                T* __restrict__ ke) {
  const int64_t e = blockIdx.x;
  if (e >= n_elems) return;
  const int a = threadIdx.x >> 3, b = threadIdx.x & 7;
  T x[8][3];
#pragma unroll
  for (int q = 0; q < 8; ++q)
#pragma unroll
    for (int d = 0; d < 3; ++d) x[q][d] = T(xe[24 * e + 3 * q + d]);
  const double E = young[e];
  const T lam = T(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
  const T mu = T(E / (2.0 * (1.0 + nu)));
  T k[3][3] = {};
#pragma unroll
  for (int gp = 0; gp < 8; ++gp) {
    T g[8][3], dj;
    if (!c3d8_dndx(x, gp, g, dj)) continue;
    const T ga0 = g[a][0], ga1 = g[a][1], ga2 = g[a][2];
    const T gb0 = g[b][0], gb1 = g[b][1], gb2 = g[b][2];
    const T ga[3] = {ga0, ga1, ga2}, gb[3] = {gb0, gb1, gb2};
    const T dotab = ga0 * gb0 + ga1 * gb1 + ga2 * gb2;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        k[i][j] += dj * (lam * ga[i] * gb[j] + mu * ga[j] * gb[i] + (i == j ? mu * dotab : T(0)));
  }
  T* out = ke + 576 * e;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) out[(3 * a + i) + 24 * (3 * b + j)] = k[i][j];
}

// ---------------------------------------------------------------- assembly
// One wave64 per 8 consecutive nodes; lane = (node, slot).  Each lane forms
// the 3x24 rows of K_e for its node's local index in its element (slot s of
// the node's ascending element list) in registers, then the lanes add them
// into the nodes' CSR segments staged in LDS in slot order -- every CSR value
// is the sum over its elements in ascending element order, written to HBM
// exactly once, without atomics.  The 8 nodes' segments are contiguous in the
// values array (node n's 3 rows start at 9*nb_ptr[n]), so the write-out is a
// single coalesced stream.
template <typename T, int MAT>
__global__ __launch_bounds__(64) void k_assemble(
    int32_t n_nodes, const double* __restrict__ coords, const int32_t* __restrict__ conn,
    const int32_t* __restrict__ slot_elem, const uint64_t* __restrict__ slot_code,
    const int32_t* __restrict__ nb_ptr, const double* __restrict__ young, double nu,
    const double* __restrict__ dgp, T* __restrict__ val) {
  __shared__ T acc[8 * 9 * kMaxNodeNbrs];
  const int lane = threadIdx.x;
  const int nl = lane >> 3, s = lane & 7;
  const int64_t node0 = int64_t(blockIdx.x) * 8;
  const int64_t nodeN = min(node0 + 8, int64_t(n_nodes));
  const int64_t q0 = nb_ptr[node0];
  const int total = int(9 * (int64_t(nb_ptr[nodeN]) - q0));
  for (int i = lane; i < total; i += 64) acc[i] = T(0);

  const int64_t node = node0 + nl;
  int32_t elem = -1;
  uint64_t code = 0;
  int deg = 0, off = 0;
  if (node < n_nodes) {
    elem = slot_elem[8 * node + s];
    code = slot_code[8 * node + s];
    deg = nb_ptr[node + 1] - nb_ptr[node];
    off = int(9 * (int64_t(nb_ptr[node]) - q0));
  }
  T kb[8][3][3];
#pragma unroll
  for (int b = 0; b < 8; ++b)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) kb[b][i][j] = T(0);

  if (elem >= 0) {
    const int a = int((code >> 40) & 7);
    T x[8][3];
    gather_coords(coords, conn, elem, x);
    T lam = T(0), mu = T(0);
    if (MAT == UFC_MAT_ELASTIC) {
      const double E = young[elem];
      lam = T(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
      mu = T(E / (2.0 * (1.0 + nu)));
    }
#pragma unroll
    for (int gp = 0; gp < 8; ++gp) {
      T g[8][3], dj;
      if (!c3d8_dndx(x, gp, g, dj)) continue;
      T ga[3];
      // select row a of g without dynamic register indexing
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        T v = g[0][q];
#pragma unroll
        for (int r = 1; r < 8; ++r) v = (a == r) ? g[r][q] : v;
        ga[q] = v;
      }
      if (MAT == UFC_MAT_ELASTIC) {
        const T la = lam * dj, mw = mu * dj;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const T gb0 = g[b][0], gb1 = g[b][1], gb2 = g[b][2];
          const T dd = mw * (ga[0] * gb0 + ga[1] * gb1 + ga[2] * gb2);
          const T gb[3] = {gb0, gb1, gb2};
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
              kb[b][i][j] += la * ga[i] * gb[j] + mw * ga[j] * gb[i] + (i == j ? dd : T(0));
        }
      } else {
        T bd[3][6];
        bt_d(ga, dgp + 36 * (8 * int64_t(elem) + gp), bd);
#pragma unroll
        for (int b = 0; b < 8; ++b) btdb_general(bd, g[b], dj, kb[b]);
      }
    }
  }
  __syncthreads();
  const int d3 = 3 * deg;
  for (int t = 0; t < 8; ++t) {
    if (s == t && elem >= 0) {
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int pos = int((code >> (5 * b)) & 31);
        T* dst = acc + off + 3 * pos;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) dst[i * d3 + j] += kb[b][i][j];
      }
    }
    __syncthreads();
  }
  T* out = val + 9 * q0;
  for (int i = lane; i < total; i += 64) out[i] = acc[i];
}

template <typename T>
__global__ void k_diag(int64_t n, const int32_t* __restrict__ pos, const T* __restrict__ val,
                       T* __restrict__ diag) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) diag[i] = val[pos[i]];
}

// FP32 difference form: zero the 3x3 self block of every node (the SpMV then
// applies sum_{m != n} K_nm (x_m - x_n) plus the penalty diagonal stored in
// the self block)
__global__ void k_zero_self(int32_t n_nodes, const int32_t* __restrict__ nb_ptr,
                            const int32_t* __restrict__ diag_pos, float* __restrict__ val) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  const int d3 = 3 * (nb_ptr[n + 1] - nb_ptr[n]);
  const int64_t base = diag_pos[3 * n];  // self block entry (0, 0)
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int c = 0; c < 3; ++c) val[base + d * d3 + c] = 0.f;
}

// diag[eq[i]] += pen[i]
__global__ void k_penalty_vec(int32_t n_bc, const int32_t* __restrict__ eq,
                              const double* __restrict__ pen, float* __restrict__ diag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_bc) diag[eq[i]] = float(double(diag[eq[i]]) + pen[i]);
}

template <typename T>
__global__ void k_penalty(int32_t n_bc, const int32_t* __restrict__ eq,
                          const double* __restrict__ pen, const int32_t* __restrict__ pos,
                          T* __restrict__ val, double sign) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_bc) {
    const int64_t p = pos[eq[i]];
    val[p] = T(double(val[p]) + sign * pen[i]);
  }
}

// ---------------------------------------------------------------- element forces
// Fe = sum_gp w detJ B^T sigma_gp with sigma from
//   mode 0: D(E_e, nu) B u_e            (elastic K_e u_e)
//   mode 1: D_gp B u_e                  (J2 tangent K_e u_e)
//   mode 2: the J2 trial stress         (internal force, RT_Asm_ComputeResidual)
// one thread per element, result in w[24 e + j].
template <int MODE>
__global__ __launch_bounds__(128) void k_elem_force(int32_t n_elems,
                                                   const double* __restrict__ coords,
                                                   const int32_t* __restrict__ conn,
                                                   const double* __restrict__ young, double nu,
                                                   const double* __restrict__ dgp,
                                                   const double* __restrict__ stress,
                                                   const double* __restrict__ u,
                                                   double* __restrict__ w) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_elems) return;
  double x[8][3], ue[8][3], fe[8][3];
  gather_coords(coords, conn, e, x);
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t n = conn[8 * e + a];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      ue[a][d] = (MODE == 2) ? 0.0 : u[3 * n + d];
      fe[a][d] = 0.0;
    }
  }
  double lam = 0.0, mu = 0.0;
  if (MODE == 0) {
    const double E = young[e];
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    mu = E / (2.0 * (1.0 + nu));
  }
  for (int gp = 0; gp < 8; ++gp) {
    double g[8][3], dj;
    if (!c3d8_dndx(x, gp, g, dj)) continue;
    double sig[6];
    if (MODE == 2) {
      const double* sp = stress + 13 * (8 * e + gp);
#pragma unroll
      for (int m = 0; m < 6; ++m) sig[m] = sp[m];
    } else {
      double eps[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        eps[0] += g[a][0] * ue[a][0];
        eps[1] += g[a][1] * ue[a][1];
        eps[2] += g[a][2] * ue[a][2];
        eps[3] += g[a][1] * ue[a][0] + g[a][0] * ue[a][1];
        eps[4] += g[a][2] * ue[a][1] + g[a][1] * ue[a][2];
        eps[5] += g[a][2] * ue[a][0] + g[a][0] * ue[a][2];
      }
      if (MODE == 0) {
        const double tr = lam * (eps[0] + eps[1] + eps[2]);
        sig[0] = tr + 2.0 * mu * eps[0];
        sig[1] = tr + 2.0 * mu * eps[1];
        sig[2] = tr + 2.0 * mu * eps[2];
        sig[3] = mu * eps[3];
        sig[4] = mu * eps[4];
        sig[5] = mu * eps[5];
      } else {
        const double* D = dgp + 36 * (8 * e + gp);
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          double s = 0.0;
#pragma unroll
          for (int m = 0; m < 6; ++m) s += D[i + 6 * m] * eps[m];
          sig[i] = s;
        }
      }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      fe[a][0] += dj * (g[a][0] * sig[0] + g[a][1] * sig[3] + g[a][2] * sig[5]);
      fe[a][1] += dj * (g[a][1] * sig[1] + g[a][0] * sig[3] + g[a][2] * sig[4]);
      fe[a][2] += dj * (g[a][2] * sig[2] + g[a][1] * sig[4] + g[a][0] * sig[5]);
    }
  }
  double* out = w + 24 * e;
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) out[3 * a + d] = fe[a][d];
}

// y[3n+d] = sum over n's elements (ascending) of w[24 e + 3 a + d]
__global__ void k_node_gather(int32_t n_nodes, const int32_t* __restrict__ slot_elem,
                              const uint64_t* __restrict__ slot_code,
                              const double* __restrict__ w, double* __restrict__ y) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int32_t e = slot_elem[8 * n + s];
    if (e < 0) continue;
    const int a = int((slot_code[8 * n + s] >> 40) & 7);
    const double* p = w + 24 * int64_t(e) + 3 * a;
    s0 += p[0];
    s1 += p[1];
    s2 += p[2];
  }
  y[3 * n + 0] = s0;
  y[3 * n + 1] = s1;
  y[3 * n + 2] = s2;
}

// ---------------------------------------------------------------- sensitivity
// grad_e = -dE/drho_e * sum_gp w detJ eps(lam_e)^T D(E=1) eps(u_e)
__global__ __launch_bounds__(128) void k_sensitivity(int32_t n_elems,
                                                    const double* __restrict__ coords,
                                                    const int32_t* __restrict__ conn, double nu,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ lam,
                                                    const double* __restrict__ dedrho,
                                                    double* __restrict__ grad) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_elems) return;
  double x[8][3], ue[8][3], le[8][3];
  gather_coords(coords, conn, e, x);
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t n = conn[8 * e + a];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      ue[a][d] = u[3 * n + d];
      le[a][d] = lam[3 * n + d];
    }
  }
  const double l0 = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double m0 = 1.0 / (2.0 * (1.0 + nu));
  double s = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    double g[8][3], dj;
    if (!c3d8_dndx(x, gp, g, dj)) continue;
    double eu[6] = {0, 0, 0, 0, 0, 0}, el[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      eu[0] += g[a][0] * ue[a][0];
      eu[1] += g[a][1] * ue[a][1];
      eu[2] += g[a][2] * ue[a][2];
      eu[3] += g[a][1] * ue[a][0] + g[a][0] * ue[a][1];
      eu[4] += g[a][2] * ue[a][1] + g[a][1] * ue[a][2];
      eu[5] += g[a][2] * ue[a][0] + g[a][0] * ue[a][2];
      el[0] += g[a][0] * le[a][0];
      el[1] += g[a][1] * le[a][1];
      el[2] += g[a][2] * le[a][2];
      el[3] += g[a][1] * le[a][0] + g[a][0] * le[a][1];
      el[4] += g[a][2] * le[a][1] + g[a][1] * le[a][2];
      el[5] += g[a][2] * le[a][0] + g[a][0] * le[a][2];
    }
    const double e_ = l0 * (eu[0] + eu[1] + eu[2]) * (el[0] + el[1] + el[2]) +
                      m0 * (2.0 * (eu[0] * el[0] + eu[1] * el[1] + eu[2] * el[2]) +
                            eu[3] * el[3] + eu[4] * el[4] + eu[5] * el[5]);
    s += dj * e_;
  }
  grad[e] = -dedrho[e] * s;
}

// ---------------------------------------------------------------- J2
// PH_J2_ComputeStress_Core (PH_Mat_Plast_J2_Iso_Core.f90:146-207) for linear
// isotropic hardening: trial stress (:254-317), yield check, radial return
// with the local Newton of :355-456; the tangent is the consistent tangent
//   Dt = K 1x1 + 2G th1 (I_s - 1x1/3) - 2G (th1 + th2 - 1) n x n
// (th1 = 1 - 3G dg/q, th2 = 3G/(3G+H), I_s = diag(1,1,1,1/2,1/2,1/2)),
// which replaces the reference's :509-537.  st: sigma(6), peeq, eps_p(6).
// Returns 0 elastic, 1 plastic, 2 failure.
__device__ int j2_point(double E, double nu, double sy0, double H, const double* deps,
                        double* st, double* Dt) {
  const double mu = E / (2.0 * (1.0 + nu));
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double strial[6];
  const double tr = lam * (deps[0] + deps[1] + deps[2]);
  strial[0] = st[0] + tr + 2.0 * mu * deps[0];
  strial[1] = st[1] + tr + 2.0 * mu * deps[1];
  strial[2] = st[2] + tr + 2.0 * mu * deps[2];
  strial[3] = st[3] + mu * deps[3];
  strial[4] = st[4] + mu * deps[4];
  strial[5] = st[5] + mu * deps[5];
  const double pm = (strial[0] + strial[1] + strial[2]) / 3.0;
  double s[6] = {strial[0] - pm, strial[1] - pm, strial[2] - pm, strial[3], strial[4], strial[5]};
  const double q = sqrt(1.5) * sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] +
                                    2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
  const double f = q - (sy0 + H * st[6]);
  if (f <= 0.0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) st[i] = strial[i];
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
      for (int i = 0; i < 6; ++i)
        Dt[i + 6 * j] = (i < 3 && j < 3 ? lam : 0.0) + (i == j ? (i < 3 ? 2.0 * mu : mu) : 0.0);
    return 0;
  }
  if (q < 1.0e-30) return 2;
  const double G = mu;
  double n[6];
  const double snorm = q / sqrt(1.5);
#pragma unroll
  for (int i = 0; i < 6; ++i) n[i] = s[i] / snorm;
  double dg = f / (3.0 * G + H);
  int it;
  for (it = 1; it <= 25; ++it) {
    const double R = q - 3.0 * G * dg - (sy0 + H * (st[6] + dg));
    if (fabs(R) < 1.0e-10 * sy0) break;
    dg = dg - R / (-(3.0 * G + H));
    if (dg < 0.0) dg = 0.0;
  }
  if (it > 25) return 2;
  const double beta = 1.0 - 3.0 * G * dg / q;
  st[0] = beta * s[0] + pm;
  st[1] = beta * s[1] + pm;
  st[2] = beta * s[2] + pm;
  st[3] = beta * s[3];
  st[4] = beta * s[4];
  st[5] = beta * s[5];
  st[6] += dg;
#pragma unroll
  for (int i = 0; i < 6; ++i) st[7 + i] += dg * n[i];
  const double K = E / (3.0 * (1.0 - 2.0 * nu));
  const double th1 = beta, th2 = 3.0 * G / (3.0 * G + H);
  const double cn = 2.0 * G * (th1 + th2 - 1.0);
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double v = (i < 3 && j < 3) ? K - 2.0 * G * th1 / 3.0 : 0.0;
      if (i == j) v += 2.0 * G * th1 * (i < 3 ? 1.0 : 0.5);
      Dt[i + 6 * j] = v - cn * n[i] * n[j];
    }
  return 1;
}

// one thread per Gauss point: strain increment B(u - u_n), J2 update from the
// committed state into the trial state and tangent
__global__ __launch_bounds__(128) void k_j2_update(int32_t n_elems,
                                                  const double* __restrict__ coords,
                                                  const int32_t* __restrict__ conn,
                                                  ufc_ph_j2_props pr, const double* __restrict__ u,
                                                  const double* __restrict__ u_n,
                                                  const double* __restrict__ commit,
                                                  double* __restrict__ trial,
                                                  double* __restrict__ tangent,
                                                  int32_t* __restrict__ flags) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= int64_t(n_elems) * 8) return;
  const int64_t e = idx >> 3;
  const int gp = int(idx & 7);
  double x[8][3];
  gather_coords(coords, conn, e, x);
  double g[8][3], dj;
  double deps[6] = {0, 0, 0, 0, 0, 0};
  if (c3d8_dndx(x, gp, g, dj)) {
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int64_t n = conn[8 * e + a];
      const double d0 = u[3 * n] - u_n[3 * n];
      const double d1 = u[3 * n + 1] - u_n[3 * n + 1];
      const double d2 = u[3 * n + 2] - u_n[3 * n + 2];
      deps[0] += g[a][0] * d0;
      deps[1] += g[a][1] * d1;
      deps[2] += g[a][2] * d2;
      deps[3] += g[a][1] * d0 + g[a][0] * d1;
      deps[4] += g[a][2] * d1 + g[a][1] * d2;
      deps[5] += g[a][2] * d0 + g[a][0] * d2;
    }
  }
  double st[13];
#pragma unroll
  for (int k = 0; k < 13; ++k) st[k] = commit[13 * idx + k];
  double Dt[36];
  const int r = j2_point(pr.young, pr.nu, pr.sigma_y0, pr.hardening, deps, st, Dt);
#pragma unroll
  for (int k = 0; k < 13; ++k) trial[13 * idx + k] = st[k];
#pragma unroll
  for (int k = 0; k < 36; ++k) tangent[36 * idx + k] = Dt[k];
  flags[idx] = r;
}

inline unsigned blocks_for(int64_t n, int t) { return unsigned((n + t - 1) / t); }

}  // namespace

// ---------------------------------------------------------------- launchers
void launch_eval_ke(int32_t n_elems, const double* coords, const double* young, double nu,
                    void* ke, bool fp32, hipStream_t s) {
  if (n_elems <= 0) return;
  if (fp32)
    hipLaunchKernelGGL(k_eval_ke<float>, dim3(n_elems), dim3(64), 0, s, n_elems, coords, young,
                       nu, static_cast<float*>(ke));
  else
    hipLaunchKernelGGL(k_eval_ke<double>, dim3(n_elems), dim3(64), 0, s, n_elems, coords, young,
                       nu, static_cast<double*>(ke));
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_assemble(ufc_ctx* c, int precision, int material) {
  const unsigned grid = blocks_for(c->n_nodes, 8);
  const double* young = c->young.get();
  const double* dgp = c->j2_tangent.get();
  if (precision == UFC_PREC_FP64) {
    if (material == UFC_MAT_ELASTIC)
      hipLaunchKernelGGL((k_assemble<double, UFC_MAT_ELASTIC>), dim3(grid), dim3(64), 0, c->stream,
                         c->n_nodes, c->coords.get(), c->conn.get(), c->slot_elem.get(),
                         c->slot_code.get(), c->nb_ptr.get(), young, c->nu, dgp, c->val64.get());
    else
      hipLaunchKernelGGL((k_assemble<double, UFC_MAT_J2>), dim3(grid), dim3(64), 0, c->stream,
                         c->n_nodes, c->coords.get(), c->conn.get(), c->slot_elem.get(),
                         c->slot_code.get(), c->nb_ptr.get(), young, c->nu, dgp, c->val64.get());
  } else {
    if (material == UFC_MAT_ELASTIC)
      hipLaunchKernelGGL((k_assemble<float, UFC_MAT_ELASTIC>), dim3(grid), dim3(64), 0, c->stream,
                         c->n_nodes, c->coords.get(), c->conn.get(), c->slot_elem.get(),
                         c->slot_code.get(), c->nb_ptr.get(), young, c->nu, dgp, c->val32.get());
    else
      hipLaunchKernelGGL((k_assemble<float, UFC_MAT_J2>), dim3(grid), dim3(64), 0, c->stream,
                         c->n_nodes, c->coords.get(), c->conn.get(), c->slot_elem.get(),
                         c->slot_code.get(), c->nb_ptr.get(), young, c->nu, dgp, c->val32.get());
  }
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_diag(ufc_ctx* c, bool fp32) {
  const unsigned grid = blocks_for(c->n_dof, 256);
  if (fp32)
    hipLaunchKernelGGL(k_diag<float>, dim3(grid), dim3(256), 0, c->stream, c->n_dof,
                       c->diag_pos.get(), c->val32.get(), c->diag32.get());
  else
    hipLaunchKernelGGL(k_diag<double>, dim3(grid), dim3(256), 0, c->stream, c->n_dof,
                       c->diag_pos.get(), c->val64.get(), c->diag64.get());
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_fp32_difference_form(ufc_ctx* c) {
  // kdiag32 = diagonal of the assembled FP32 K (no penalties)
  hipLaunchKernelGGL(k_diag<float>, dim3(blocks_for(c->n_dof, 256)), dim3(256), 0, c->stream,
                     c->n_dof, c->diag_pos.get(), c->val32.get(), c->kdiag32.get());
  hipLaunchKernelGGL(k_zero_self, dim3(blocks_for(c->n_nodes, 256)), dim3(256), 0, c->stream,
                     c->n_nodes, c->nb_ptr.get(), c->diag_pos.get(), c->val32.get());
  UFC_HIP_CHECK(hipGetLastError());
  launch_penalty(c, true, +1.0);
  launch_diag32_penalty(c);
}

void launch_diag32_penalty(ufc_ctx* c) {
  UFC_HIP_CHECK(hipMemcpyAsync(c->diag32.get(), c->kdiag32.get(), c->kdiag32.bytes(),
                               hipMemcpyDeviceToDevice, c->stream));
  const int32_t nb = int32_t(c->h_bc_eq.size());
  if (nb)
    hipLaunchKernelGGL(k_penalty_vec, dim3(blocks_for(nb, 256)), dim3(256), 0, c->stream, nb,
                       c->bc_eq.get(), c->bc_pen.get(), c->diag32.get());
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_penalty(ufc_ctx* c, bool fp32, double sign) {
  const int32_t nb = int32_t(c->h_bc_eq.size());
  if (nb == 0) return;
  const unsigned grid = blocks_for(nb, 256);
  if (fp32)
    hipLaunchKernelGGL(k_penalty<float>, dim3(grid), dim3(256), 0, c->stream, nb, c->bc_eq.get(),
                       c->bc_pen.get(), c->diag_pos.get(), c->val32.get(), sign);
  else
    hipLaunchKernelGGL(k_penalty<double>, dim3(grid), dim3(256), 0, c->stream, nb, c->bc_eq.get(),
                       c->bc_pen.get(), c->diag_pos.get(), c->val64.get(), sign);
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_apply_elem64(ufc_ctx* c, int mode, const double* u, double* y) {
  c->elem_w.resize(size_t(c->n_elems) * 24);
  const unsigned grid = blocks_for(c->n_elems, 128);
  if (mode == 0)
    hipLaunchKernelGGL(k_elem_force<0>, dim3(grid), dim3(128), 0, c->stream, c->n_elems,
                       c->coords.get(), c->conn.get(), c->young.get(), c->nu, nullptr, nullptr, u,
                       c->elem_w.get());
  else if (mode == 1)
    hipLaunchKernelGGL(k_elem_force<1>, dim3(grid), dim3(128), 0, c->stream, c->n_elems,
                       c->coords.get(), c->conn.get(), nullptr, 0.0, c->j2_tangent.get(), nullptr,
                       u, c->elem_w.get());
  else
    hipLaunchKernelGGL(k_elem_force<2>, dim3(grid), dim3(128), 0, c->stream, c->n_elems,
                       c->coords.get(), c->conn.get(), nullptr, 0.0, nullptr, c->j2_trial.get(),
                       nullptr, c->elem_w.get());
  UFC_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_node_gather, dim3(blocks_for(c->n_nodes, 256)), dim3(256), 0, c->stream,
                     c->n_nodes, c->slot_elem.get(), c->slot_code.get(), c->elem_w.get(), y);
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_sensitivity(ufc_ctx* c, const double* u, const double* lam, const double* dedrho,
                        double nu, double* grad) {
  hipLaunchKernelGGL(k_sensitivity, dim3(blocks_for(c->n_elems, 128)), dim3(128), 0, c->stream,
                     c->n_elems, c->coords.get(), c->conn.get(), nu, u, lam, dedrho, grad);
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_j2_update(ufc_ctx* c, const double* u, const double* u_n) {
  const int64_t ngp = int64_t(c->n_elems) * 8;
  hipLaunchKernelGGL(k_j2_update, dim3(blocks_for(ngp, 128)), dim3(128), 0, c->stream,
                     c->n_elems, c->coords.get(), c->conn.get(), c->j2, u, u_n,
                     c->j2_commit.get(), c->j2_trial.get(), c->j2_tangent.get(),
                     c->j2_flags.get());
  UFC_HIP_CHECK(hipGetLastError());
}

}  // namespace ufc
