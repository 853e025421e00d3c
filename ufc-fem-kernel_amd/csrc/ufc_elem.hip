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
constexpr int XS[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
constexpr int YS[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
constexpr int ZS[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// dN/dx at (xi, eta, zeta); also J^-1 (Ji[k][j] = dxi_k/dx_j) when Ji != null
template <typename T>
__device__ __forceinline__ bool c3d8_dndx_at(const T (&x)[8][3], T xi, T eta, T zeta,
                                             T (&g)[8][3], T& detJ, T (*Jinv)[3] = nullptr) {
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
  if (Jinv) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j) Jinv[k][j] = Ji[k][j];
  }
  return true;
}

// at Gauss point gp of PH_Elem_C3D8_GaussPoints (xi outer, zeta inner)
template <typename T>
__device__ __forceinline__ bool c3d8_dndx(const T (&x)[8][3], int gp, T (&g)[8][3], T& detJ) {
  const T xi = (gp & 4) ? T(kGaussPt) : T(-kGaussPt);
  const T eta = (gp & 2) ? T(kGaussPt) : T(-kGaussPt);
  const T zeta = (gp & 1) ? T(kGaussPt) : T(-kGaussPt);
  return c3d8_dndx_at(x, xi, eta, zeta, g, detJ);
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

// The J2 consistent tangent of a Gauss point, stored factored as kJ2Tangent
// doubles tc = (mu', lam', c, n(6)):
//   Dt = lam' 1x1 + 2 mu' I_s - c n x n
// (Voigt [xx,yy,zz,xy,yz,xz], engineering shear, I_s = diag(1,1,1,1/2,1/2,1/2)),
// which is the form the radial return produces it in (elastic points: mu,
// lam, c = 0): 72 bytes per point instead of the 288 of the 6x6 matrix, read
// by the tangent assembly and the FP64 tangent residual, written by the update.
// v = B(h)^T n for the B-matrix rows [hx,0,0], [0,hy,0], [0,0,hz], [hy,hx,0],
// [0,hz,hy], [hz,0,hx].
template <typename T>
__device__ __forceinline__ void bt_n(const T* h, const T (&n)[6], T (&v)[3]) {
  v[0] = h[0] * n[0] + h[1] * n[3] + h[2] * n[5];
  v[1] = h[1] * n[1] + h[0] * n[3] + h[2] * n[4];
  v[2] = h[2] * n[2] + h[1] * n[4] + h[0] * n[5];
}

// sigma = Dt eps for the factored tangent tc
__device__ __forceinline__ void j2_tangent_apply(const double* __restrict__ tc,
                                                 const double (&eps)[6], double (&sig)[6]) {
  const double mu = tc[0], lam = tc[1], c = tc[2];
  double n[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) n[i] = tc[3 + i];
  double ne = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) ne += n[i] * eps[i];
  ne *= c;
  const double tr = lam * (eps[0] + eps[1] + eps[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i) sig[i] = tr + 2.0 * mu * eps[i] - ne * n[i];
#pragma unroll
  for (int i = 3; i < 6; ++i) sig[i] = mu * eps[i] - ne * n[i];
}

// ga = row a of the (8, 3) register table g without dynamic register
// indexing.  A macro, not a function: with g behind a reference the select
// chain is folded into an indexed load before inlining, and g then lives in
// scratch memory.
#define UFC_SELECT_ROW(g, a, ga)                                     \
  do {                                                               \
    _Pragma("unroll") for (int q_ = 0; q_ < 3; ++q_) {               \
      auto v_ = (g)[0][q_];                                          \
      _Pragma("unroll") for (int r_ = 1; r_ < 8; ++r_) v_ = ((a) == r_) ? (g)[r_][q_] : v_; \
      (ga)[q_] = v_;                                                 \
    }                                                                \
  } while (0)

// kb[b] += la ga gb^T + mw gb ga^T + mw (ga . gb) I, the isotropic B_a^T D B_b
template <typename T>
__device__ __forceinline__ void add_iso_blocks(const T (&g)[8][3], const T (&ga)[3], T la, T mw,
                                               T (&kb)[8][3][3]) {
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
}

// The trilinear map x(xi, eta, zeta) = sum_a N_a x_a in monomial form: its
// derivatives are dx/dxi = e1 + e12 eta + e13 zeta + e123 eta zeta and the
// two cyclic analogues, with e_S = (1/8) sum_a (prod_{s in S} s_a) x_a over
// the node signs s_a = (XS, YS, ZS)[a].  21 values per element replace the
// 8 x 3 coordinates in the Jacobian at each point.
template <typename T>
struct TriMap {
  T e1[3], e2[3], e3[3], e12[3], e13[3], e23[3], e123[3];
};

template <typename T>
__device__ __forceinline__ TriMap<T> tri_map(const T (&x)[8][3]) {
  TriMap<T> m;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    T s1 = T(0), s2 = T(0), s3 = T(0), s12 = T(0), s13 = T(0), s23 = T(0), s123 = T(0);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const T v = x[a][c];
      s1 += T(XS[a]) * v;
      s2 += T(YS[a]) * v;
      s3 += T(ZS[a]) * v;
      s12 += T(XS[a] * YS[a]) * v;
      s13 += T(XS[a] * ZS[a]) * v;
      s23 += T(YS[a] * ZS[a]) * v;
      s123 += T(XS[a] * YS[a] * ZS[a]) * v;
    }
    m.e1[c] = T(0.125) * s1;
    m.e2[c] = T(0.125) * s2;
    m.e3[c] = T(0.125) * s3;
    m.e12[c] = T(0.125) * s12;
    m.e13[c] = T(0.125) * s13;
    m.e23[c] = T(0.125) * s23;
    m.e123[c] = T(0.125) * s123;
  }
  return m;
}

// dN_b/dxi at (xi, eta, zeta) for node signs (sx, sy, sz)
template <typename T>
__device__ __forceinline__ void node_dn(T sx, T sy, T sz, T xi, T eta, T zeta, T (&dn)[3]) {
  const T fx = T(1) + sx * xi, fy = T(1) + sy * eta, fz = T(1) + sz * zeta;
  dn[0] = T(0.125) * sx * fy * fz;
  dn[1] = T(0.125) * fx * sy * fz;
  dn[2] = T(0.125) * fx * fy * sz;
}

// At one point: the Jacobian J[j][k] = dx_j/dxi_k, d = det J and its
// adjugate A (J^-1 = A / d, so dN_b/dx = A^T dn_b / d).  False when
// |d| <= 1e-12 (the point is skipped, PH_Elem_C3D8.f90:775).
template <typename T>
__device__ __forceinline__ bool tri_point(const TriMap<T>& m, T xi, T eta, T zeta, T (&A)[3][3],
                                          T& d) {
  T J[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    J[j][0] = m.e1[j] + m.e12[j] * eta + m.e13[j] * zeta + m.e123[j] * (eta * zeta);
    J[j][1] = m.e2[j] + m.e12[j] * xi + m.e23[j] * zeta + m.e123[j] * (xi * zeta);
    J[j][2] = m.e3[j] + m.e13[j] * xi + m.e23[j] * eta + m.e123[j] * (xi * eta);
  }
  A[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  A[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]);
  A[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  d = J[0][0] * A[0][0] + J[0][1] * A[1][0] + J[0][2] * A[2][0];
  if (tabs(d) <= T(1.0e-12)) return false;
  A[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]);
  A[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  A[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  A[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]);
  A[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]);
  A[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  return true;
}

// h = A^T dn (= d dN/dx)
template <typename T>
__device__ __forceinline__ void adj_t(const T (&A)[3][3], const T (&dn)[3], T (&h)[3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) h[j] = A[0][j] * dn[0] + A[1][j] * dn[1] + A[2][j] * dn[2];
}

// One integration point of weight w for the isotropic blocks: with
// h_b = d dN_b/dx, w detJ g_a g_b^T = (w / d) h_a h_b^T, so
// P_b += (w / d) h_a h_b^T needs one division per point and no table of
// gradients; sg_b += w h_b (= w detJ g_b) when SG.  Node a's signs are
// (sx, sy, sz).  Returns w d (the weighted volume), 0 for a skipped point.
template <typename T, bool SG>
__device__ __forceinline__ T iso_point(const TriMap<T>& m, T xi, T eta, T zeta, T w, T sx, T sy,
                                       T sz, T (&P)[8][3][3], T (&sg)[8][3]) {
  T A[3][3], d;
  if (!tri_point(m, xi, eta, zeta, A, d)) return T(0);
  T dna[3], ha[3];
  node_dn(sx, sy, sz, xi, eta, zeta, dna);
  adj_t(A, dna, ha);
  const T r = w / d;
#pragma unroll
  for (int i = 0; i < 3; ++i) ha[i] *= r;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    T dnb[3], hb[3];
    node_dn(T(XS[b]), T(YS[b]), T(ZS[b]), xi, eta, zeta, dnb);
    adj_t(A, dnb, hb);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) P[b][i][j] += ha[i] * hb[j];
    if constexpr (SG) {
#pragma unroll
      for (int j = 0; j < 3; ++j) sg[b][j] += w * hb[j];
    }
  }
  return w * d;
}

// iso_point over the 2x2x2 Gauss points of PH_Elem_C3D8_GaussPoints (xi
// outer, zeta inner); returns the volume.  The FP64 loop stays rolled:
// unrolled, the overlap of consecutive points takes it past 256 registers
// (one wave per SIMD), rolled it fits in 254 (two); FP32 fits either way and
// keeps the folded constants.
template <typename T, bool SG>
__device__ __forceinline__ T iso_gauss8(const TriMap<T>& m, T sx, T sy, T sz, T (&P)[8][3][3],
                                        T (&sg)[8][3]) {
  T V = T(0);
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int gp = 0; gp < 8; ++gp) {
      const T xi = (gp & 4) ? T(kGaussPt) : T(-kGaussPt);
      const T eta = (gp & 2) ? T(kGaussPt) : T(-kGaussPt);
      const T zeta = (gp & 1) ? T(kGaussPt) : T(-kGaussPt);
      V += iso_point<T, SG>(m, xi, eta, zeta, T(1), sx, sy, sz, P, sg);
    }
  } else {
#pragma unroll 1
    for (int gp = 0; gp < 8; ++gp) {
      const T xi = (gp & 4) ? T(kGaussPt) : T(-kGaussPt);
      const T eta = (gp & 2) ? T(kGaussPt) : T(-kGaussPt);
      const T zeta = (gp & 1) ? T(kGaussPt) : T(-kGaussPt);
      V += iso_point<T, SG>(m, xi, eta, zeta, T(1), sx, sy, sz, P, sg);
    }
  }
  return V;
}

// kb[b] += la P_b + mw P_b^T + mw tr(P_b) I: the isotropic B_a^T D B_b
// summed over the points, P_b = sum w detJ g_a g_b^T
template <typename T>
__device__ __forceinline__ void iso_blocks_from_p(const T (&P)[8][3][3], T la, T mw,
                                                  T (&kb)[8][3][3]) {
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const T tr = mw * (P[b][0][0] + P[b][1][1] + P[b][2][2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        kb[b][i][j] += la * P[b][i][j] + mw * P[b][j][i] + (i == j ? tr : T(0));
  }
}

template <typename T>
__device__ __forceinline__ void node_signs(int a, T& sx, T& sy, T& sz) {
  sx = ((a ^ (a >> 1)) & 1) ? T(1) : T(-1);
  sy = ((a >> 1) & 1) ? T(1) : T(-1);
  sz = ((a >> 2) & 1) ? T(1) : T(-1);
}

// Rows of node a of the elastic K_e, kb[b] = K_ab (added to kb), for the
// variant VAR of PH_Elem_C3D8_StiffMatrixByVariant (PH_Elem_C3D8.f90:400-417):
//   kVarFull      2x2x2 Gauss (PH_Elem_C3D8_StiffMatrix, :759-780)
//   kVarGauss27   3x3x3 Gauss, points 0, -/+0.774596669241483, weights 5/9,
//                 8/9 (PH_Elem_C3D8_StiffMatrix27, :1976-1995)
//   kVarSelective B-bar (PH_Elem_C3D8_StiffMatrixSelective, :500-540): with
//                 the volumetric rows replaced by their element average the
//                 block is sum_gp dV [(lam - kap) ga gb^T + mu (gb ga^T +
//                 ga.gb I)] + kap s_a s_b^T / V, kap = lam + 2 mu / 3,
//                 s_a = sum_gp dV g_a, V = sum_gp dV
//   kVarReduced   one point at the centre, weight 8, plus the hourglass
//                 stiffness k sum_m gamma_m,a gamma_m,b I, k = 0.1 mu V^(1/3)
//                 (PH_Elem_C3D8_StiffMatrixReduced, :466-498) with base
//                 vectors xi.eta.zeta, eta.zeta, xi.zeta, xi.eta projected
//                 off the linear fields (Flanagan & Belytschko 1981; the
//                 reference's GAMMA8, :123-127, holds -zeta in place of xi.eta)
//   kVarIncompat  the full block minus the incompatible-mode condensation
//                 C_a C_b^T, C (24, 9) per element from k_incompat_c
template <typename T, int VAR>
__device__ __forceinline__ void elem_rows(const T (&x)[8][3], int a, T lam, T mu,
                                          const double* __restrict__ cinc, T (&kb)[8][3][3]) {
  T sx, sy, sz;
  node_signs(a, sx, sy, sz);
  if constexpr (VAR == kVarFull || VAR == kVarIncompat) {
    const TriMap<T> m = tri_map(x);
    T P[8][3][3] = {}, sg[8][3];
    iso_gauss8<T, false>(m, sx, sy, sz, P, sg);
    iso_blocks_from_p(P, lam, mu, kb);
    if constexpr (VAR == kVarIncompat) {
      T ca[3][9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 9; ++k) ca[i][k] = T(cinc[9 * (3 * a + i) + k]);
#pragma unroll
      for (int b = 0; b < 8; ++b)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          T cb[9];
#pragma unroll
          for (int k = 0; k < 9; ++k) cb[k] = T(cinc[9 * (3 * b + j) + k]);
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            T t = T(0);
#pragma unroll
            for (int k = 0; k < 9; ++k) t += ca[i][k] * cb[k];
            kb[b][i][j] -= t;
          }
        }
    }
  } else if constexpr (VAR == kVarSelective) {
    const T kap = lam + T(2) * mu / T(3);
    const TriMap<T> m = tri_map(x);
    T P[8][3][3] = {}, sg[8][3] = {};
    const T V = iso_gauss8<T, true>(m, sx, sy, sz, P, sg);
    iso_blocks_from_p(P, lam - kap, mu, kb);
    if (V <= T(1.0e-20)) {  // the reference returns Ke = 0
#pragma unroll
      for (int b = 0; b < 8; ++b)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) kb[b][i][j] = T(0);
      return;
    }
    T sa[3];
    UFC_SELECT_ROW(sg, a, sa);
    const T f = kap / V;
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) kb[b][i][j] += f * sa[i] * sg[b][j];
  } else if constexpr (VAR == kVarGauss27) {
    const TriMap<T> m = tri_map(x);
    T P[8][3][3] = {}, sg[8][3];
#pragma unroll 1
    for (int p = 0; p < 27; ++p) {
      const int i = p / 9, j = (p / 3) % 3, k = p % 3;
      const T c = T(0.774596669241483);
      const T xi = i == 0 ? -c : (i == 1 ? T(0) : c);
      const T eta = j == 0 ? -c : (j == 1 ? T(0) : c);
      const T zeta = k == 0 ? -c : (k == 1 ? T(0) : c);
      const T wi = i == 1 ? T(8.0 / 9.0) : T(5.0 / 9.0);
      const T wj = j == 1 ? T(8.0 / 9.0) : T(5.0 / 9.0);
      const T wk = k == 1 ? T(8.0 / 9.0) : T(5.0 / 9.0);
      iso_point<T, false>(m, xi, eta, zeta, wi * wj * wk, sx, sy, sz, P, sg);
    }
    iso_blocks_from_p(P, lam, mu, kb);
  } else if constexpr (VAR == kVarReduced) {
    // everything from the trilinear map in the adjugate form of the standard
    // element (no gradient tables: with them the FP64 kernel ran at one wave
    // per SIMD)
    const TriMap<T> m = tri_map(x);
    // x only enters the hourglass terms through hx = sum_n h_n x_n of the four
    // base vectors
    T hx[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        T t = T(0);
#pragma unroll
        for (int n = 0; n < 8; ++n) {
          const int hv = q == 0 ? -XS[n] * YS[n] * ZS[n]
                                : (q == 1 ? YS[n] * ZS[n] : (q == 2 ? XS[n] * ZS[n] : XS[n] * YS[n]));
          t += T(hv) * x[n][i];
        }
        hx[q][i] = t;
      }
    // centre point and element volume first: the trilinear map is then no
    // longer needed while the 8 row blocks accumulate
    T A0[3][3], d0;
    const bool ok0 = tri_point(m, T(0), T(0), T(0), A0, d0);
    // element volume from the 2x2x2 points (the reference's hourglass scale)
    T V = T(0);
#pragma unroll 1
    for (int gp = 0; gp < 8; ++gp) {
      T A[3][3], d;
      tri_point(m, (gp & 4) ? T(kGaussPt) : T(-kGaussPt), (gp & 2) ? T(kGaussPt) : T(-kGaussPt),
                (gp & 1) ? T(kGaussPt) : T(-kGaussPt), A, d);
      V += d;
    }
    // the one point at the centre, weight 8: kb[b] += 8 d0 (lam ga gb^T +
    // mu gb ga^T + mu (ga . gb) I) with the centre gradients g = A0^T s / (8 d0)
    if (ok0) {
      const T sa[3] = {sx, sy, sz};
      T ga[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) ga[j] = A0[0][j] * sa[0] + A0[1][j] * sa[1] + A0[2][j] * sa[2];
      // with h = A0^T s = 8 d0 g: 8 d0 g_a g_b^T = h_a h_b^T / (8 d0)
      const T w = T(0.125) / d0;
      const T la = w * lam, mw = w * mu;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const T sb[3] = {T(XS[b]), T(YS[b]), T(ZS[b])};
        T gb[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) gb[j] = A0[0][j] * sb[0] + A0[1][j] * sb[1] + A0[2][j] * sb[2];
        const T dd = mw * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            kb[b][i][j] += la * ga[i] * gb[j] + mw * ga[j] * gb[i] + (i == j ? dd : T(0));
      }
    }
    if (V <= T(1.0e-20)) return;
    const T L = T(cbrt(double(V)));
    const T khg = T(0.1) * mu * V / (L * L);
    // centre gradients g0_n = (1/8) A0^T s_n / d0 (s_n the node signs; zero when
    // the centre point is degenerate, where A0 is only partly formed)
    const T r0 = ok0 ? T(0.125) / d0 : T(0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // hx . g0_n = (A0 hx) . s_n r0
      T ah[3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
        ah[k] = ok0 ? r0 * (A0[k][0] * hx[q][0] + A0[k][1] * hx[q][1] + A0[k][2] * hx[q][2])
                    : T(0);
      T gam[8], gma = T(0);
#pragma unroll
      for (int n = 0; n < 8; ++n) {
        const int hv = q == 0 ? -XS[n] * YS[n] * ZS[n]
                              : (q == 1 ? YS[n] * ZS[n] : (q == 2 ? XS[n] * ZS[n] : XS[n] * YS[n]));
        gam[n] = T(hv) - (ah[0] * T(XS[n]) + ah[1] * T(YS[n]) + ah[2] * T(ZS[n]));
        gma = (a == n) ? gam[n] : gma;
      }
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const T t = khg * gma * gam[b];
        kb[b][0][0] += t;
        kb[b][1][1] += t;
        kb[b][2][2] += t;
      }
    }
  }
}

// element coordinates: conn == nullptr reads (8, 3, n) element coordinates
template <typename T>
__device__ __forceinline__ void load_x(const double* __restrict__ coords,
                                       const int32_t* __restrict__ conn, int64_t e,
                                       T (&x)[8][3]) {
  if (conn) {
    gather_coords(coords, conn, e, x);
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[q][d] = T(coords[24 * e + 3 * q + d]);
  }
}

// ---------------------------------------------------------------- incompatible modes
// C (24, 9) with K_cond = Kuu - C C^T: the modes P_k = 1 - xi_k^2 of each
// displacement component (Wilson et al. 1973) with gradients
// J0^-T dP/dxi detJ0/detJ (Taylor, Beresford & Wilson 1976) in place of the
// reference's three scalar modes (PH_Elem_C3D8.f90:240-259, which fail the
// patch test); Kaa = L L^T, C = Kua L^-T.  One wave per element: lanes 0-23
// form the rows of Kua, lanes 24-32 the rows of Kaa, every lane factors Kaa
// and lanes 0-23 solve their row of C.  A Kaa that is not positive definite
// (degenerate element) gives C = 0, i.e. Kuu, as the reference's
// det(Kaa) ~ 0 branch (:457-460).
__global__ __launch_bounds__(64) void k_incompat_c(int32_t n_elems,
                                                   const double* __restrict__ coords,
                                                   const int32_t* __restrict__ conn,
                                                   const double* __restrict__ young, double nu,
                                                   double* __restrict__ cout) {
  __shared__ double kua[24][9];
  __shared__ double kaa[9][9];
  __shared__ double pg[8][11][3];  // per Gauss point: 8 node + 3 mode gradients
  __shared__ double pdj[8];        // detJ, 0 for a skipped point
  const int64_t e = blockIdx.x;
  const int r = threadIdx.x;
  const double E = young[e];
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  // lanes 0-7: one Gauss point each (gradients of the nodes and of the modes)
  bool c_ok = true;
  {
    double x[8][3];
    load_x(coords, conn, e, x);
    double g0[8][3], dj0, J0i[3][3];
    c_ok = c3d8_dndx_at(x, 0.0, 0.0, 0.0, g0, dj0, J0i);
    if (r < 8) {
      const int gp = r;
      const double xc[3] = {(gp & 4) ? kGaussPt : -kGaussPt, (gp & 2) ? kGaussPt : -kGaussPt,
                            (gp & 1) ? kGaussPt : -kGaussPt};
      double g[8][3], dj;
      const bool ok = c3d8_dndx(x, gp, g, dj) && c_ok;
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) pg[gp][a][j] = ok ? g[a][j] : 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          pg[gp][8 + k][j] = ok ? J0i[k][j] * (-2.0 * xc[k]) * dj0 / dj : 0.0;
      pdj[gp] = ok ? dj : 0.0;
    }
  }
  __syncthreads();
  // lanes 0-23: row (node r/3, component r%3) of Kua; lanes 24-32: row
  // (mode (r-24)/3, component (r-24)%3) of Kaa; columns (mode q, component j)
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (r < 33) {
    const int pn = r < 24 ? r / 3 : 8 + (r - 24) / 3;
    const int i = r < 24 ? r % 3 : (r - 24) % 3;
    for (int gp = 0; gp < 8; ++gp) {
      const double dj = pdj[gp];
      const double gr0 = pg[gp][pn][0], gr1 = pg[gp][pn][1], gr2 = pg[gp][pn][2];
      const double gri = i == 0 ? gr0 : (i == 1 ? gr1 : gr2);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double gq0 = pg[gp][8 + q][0], gq1 = pg[gp][8 + q][1], gq2 = pg[gp][8 + q][2];
        const double gqi = i == 0 ? gq0 : (i == 1 ? gq1 : gq2);
        const double dd = mu * (gr0 * gq0 + gr1 * gq1 + gr2 * gq2);
        acc[3 * q + 0] += dj * (lam * gri * gq0 + mu * gr0 * gqi + (i == 0 ? dd : 0.0));
        acc[3 * q + 1] += dj * (lam * gri * gq1 + mu * gr1 * gqi + (i == 1 ? dd : 0.0));
        acc[3 * q + 2] += dj * (lam * gri * gq2 + mu * gr2 * gqi + (i == 2 ? dd : 0.0));
      }
    }
  }
  if (r < 24) {
#pragma unroll
    for (int k = 0; k < 9; ++k) kua[r][k] = acc[k];
  } else if (r < 33) {
#pragma unroll
    for (int k = 0; k < 9; ++k) kaa[r - 24][k] = acc[k];
  }
  __syncthreads();
  // every lane factors Kaa = L L^T in registers (the same arithmetic in each
  // lane, no LDS round trips in the dependent chain), lanes 0-23 then solve
  // L w = Kua row r
  double l[45];  // lower triangle, row-major: l[i (i + 1) / 2 + j]
  bool good = c_ok;
#pragma unroll
  for (int i2 = 0; i2 < 9; ++i2)
#pragma unroll
    for (int j = 0; j <= i2; ++j) l[i2 * (i2 + 1) / 2 + j] = kaa[i2][j];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    double d = l[j * (j + 1) / 2 + j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= l[j * (j + 1) / 2 + k] * l[j * (j + 1) / 2 + k];
    good = good && d > 0.0;
    d = good ? sqrt(d) : 1.0;
    l[j * (j + 1) / 2 + j] = d;
    const double inv = 1.0 / d;
#pragma unroll
    for (int i2 = j + 1; i2 < 9; ++i2) {
      double t = l[i2 * (i2 + 1) / 2 + j];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= l[i2 * (i2 + 1) / 2 + k] * l[j * (j + 1) / 2 + k];
      l[i2 * (i2 + 1) / 2 + j] = t * inv;
    }
  }
  if (r < 24) {
    double w[9];
#pragma unroll
    for (int i2 = 0; i2 < 9; ++i2) {
      double t = kua[r][i2];
#pragma unroll
      for (int k = 0; k < i2; ++k) t -= l[i2 * (i2 + 1) / 2 + k] * w[k];
      w[i2] = good ? t / l[i2 * (i2 + 1) / 2 + i2] : 0.0;
    }
    double* out = cout + int64_t(kIncompatC) * e + 9 * r;
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = w[k];
  }
}

// ---------------------------------------------------------------- batch Ke
// Eight lanes per element, lane = node a: rows 3a..3a+2 of Ke, written
// column-major into Ke(24,24,e).
template <typename T, int VAR>
__global__ __launch_bounds__(64) void k_eval_ke(int32_t n_elems, const double* __restrict__ xe,
                                                const double* __restrict__ young, double nu,
                                                const double* __restrict__ inc_c,
                                                T* __restrict__ ke) {
  const int64_t e = int64_t(blockIdx.x) * 8 + (threadIdx.x >> 3);
  const int a = threadIdx.x & 7;
  if (e >= n_elems) return;
  T x[8][3];
  load_x(xe, nullptr, e, x);
  const double E = young[e];
  const T lam = T(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
  const T mu = T(E / (2.0 * (1.0 + nu)));
  T kb[8][3][3] = {};
  elem_rows<T, VAR>(x, a, lam, mu, VAR == kVarIncompat ? inc_c + int64_t(kIncompatC) * e : nullptr,
                    kb);
  T* out = ke + 576 * e;
#pragma unroll
  for (int b = 0; b < 8; ++b)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) out[(3 * a + i) + 24 * (3 * b + j)] = kb[b][i][j];
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
// What the FP32 assembly writes besides the CSR values.
//   kdiag != null: the diagonal of K (kdiag[3n + d], from the self blocks) and
//     the self blocks zeroed in the values: the FP32 values are kept in the
//     difference form (K x)_n = sum_{m != n} K_nm (x_m - x_n), whose self
//     blocks hold only the penalty diagonal.
//   nd > 0: block diagonals of the symmetric FP32 view: for node n and
//     diagonal s, the block K(n, m) with m the node at box number
//     perm[n] + o_s, found in n's sorted neighbour list, goes to entry
//     (s, perm[n]) (perm = null: the node numbering is the box's).
struct DiagWrite {
  DiagOffsets off;
  int nd;
  int32_t pad_tiles;
  int32_t tile_shift;  // log2 of the layout's tile size
  const int32_t* perm;
  const int32_t* iperm;
  const int32_t* nb_idx;
  float4* da;
  float4* db;
  float* dc;
  float* kdiag;
  const int64_t* diag_pos;
};

// The 8 row blocks K_e(a, b) of element elem for its local node a: elastic
// (variant VAR, E_e = young[elem]) or the J2 tangent (MAT = UFC_MAT_J2, the
// Gauss-point tangents dgp).
template <typename T, int MAT, int VAR>
__device__ __forceinline__ void elem_node_rows(const double* __restrict__ coords,
                                               const int32_t* __restrict__ conn, int32_t elem,
                                               int a, const double* __restrict__ young,
                                               double nu, const double* __restrict__ dgp,
                                               const double* __restrict__ inc_c,
                                               T (&kb)[8][3][3]) {
  T x[8][3];
  gather_coords(coords, conn, elem, x);
  if constexpr (MAT == UFC_MAT_ELASTIC) {
    const double E = young[elem];
    const T lam = T(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
    const T mu = T(E / (2.0 * (1.0 + nu)));
    elem_rows<T, VAR>(x, a, lam, mu,
                      VAR == kVarIncompat ? inc_c + int64_t(kIncompatC) * elem : nullptr, kb);
  } else {
    // sum_gp detJ B_a^T D_gp B_b = sum_gp (1 / d) H_a^T D_gp H_b with the
    // B-matrices H of h = d dN/dx (one division per point, no gradient
    // table); with the factored tangent each block is the isotropic one of
    // (lam', mu') less the rank-one c v_a v_b^T, v = H^T n
    const TriMap<T> m = tri_map(x);
    T sx, sy, sz;
    node_signs(a, sx, sy, sz);
#pragma unroll 1
    for (int gp = 0; gp < 8; ++gp) {
      const T xi = (gp & 4) ? T(kGaussPt) : T(-kGaussPt);
      const T eta = (gp & 2) ? T(kGaussPt) : T(-kGaussPt);
      const T zeta = (gp & 1) ? T(kGaussPt) : T(-kGaussPt);
      T A[3][3], d;
      if (!tri_point(m, xi, eta, zeta, A, d)) continue;
      T dna[3], ha[3];
      node_dn(sx, sy, sz, xi, eta, zeta, dna);
      adj_t(A, dna, ha);
      const T r = T(1) / d;
#pragma unroll
      for (int i = 0; i < 3; ++i) ha[i] *= r;
      const double* tc = dgp + kJ2Tangent * (8 * int64_t(elem) + gp);
      const T mu = T(tc[0]), lam = T(tc[1]), cn = T(tc[2]);
      T n[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) n[i] = T(tc[3 + i]);
      T va[3];
      bt_n(ha, n, va);
#pragma unroll
      for (int i = 0; i < 3; ++i) va[i] *= cn;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        T dnb[3], hb[3], vb[3];
        node_dn(T(XS[b]), T(YS[b]), T(ZS[b]), xi, eta, zeta, dnb);
        adj_t(A, dnb, hb);
        bt_n(hb, n, vb);
        const T dot = mu * (ha[0] * hb[0] + ha[1] * hb[1] + ha[2] * hb[2]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            kb[b][i][j] += lam * ha[i] * hb[j] + mu * hb[i] * ha[j] - va[i] * vb[j] +
                           (i == j ? dot : T(0));
      }
    }
  }
}

// gen_flag != null: blocks of general nodes (k_assemble_gen) are skipped
template <typename T, int MAT, int VAR>
__global__ __launch_bounds__(64) void k_assemble(
    int32_t n_nodes, const double* __restrict__ coords, const int32_t* __restrict__ conn,
    const int32_t* __restrict__ slot_elem, const uint64_t* __restrict__ slot_code,
    const int32_t* __restrict__ nb_ptr, const double* __restrict__ young, double nu,
    const double* __restrict__ dgp, const double* __restrict__ inc_c, T* __restrict__ val,
    DiagWrite dw, const uint8_t* __restrict__ gen_flag) {
  __shared__ T acc[8 * 9 * kMaxNodeNbrs];
  __shared__ int32_t nbl[8 * kMaxNodeNbrs];  // the 8 nodes' neighbour lists (diagonal write)
  if (gen_flag && gen_flag[blockIdx.x]) return;
  const int lane = threadIdx.x;
  const int nl = lane >> 3, s = lane & 7;
  const int64_t node0 = int64_t(blockIdx.x) * 8;
  const int64_t nodeN = min(node0 + 8, int64_t(n_nodes));
  const int64_t q0 = nb_ptr[node0];
  const int total = int(9 * (int64_t(nb_ptr[nodeN]) - q0));
  for (int i = lane; i < total; i += 64) acc[i] = T(0);
  if constexpr (sizeof(T) == 4) {
    if (dw.nd > 0)
      for (int i = lane; i < total / 9; i += 64) nbl[i] = dw.nb_idx[q0 + i];
  }

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

  if (elem >= 0)
    elem_node_rows<T, MAT, VAR>(coords, conn, elem, int((code >> 40) & 7), young, nu, dgp, inc_c,
                                kb);
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
  if constexpr (sizeof(T) == 4) {
    if (dw.kdiag) {
      for (int i = lane; i < 24; i += 64) {
        const int64_t nn0 = node0 + i / 3;
        const int d = i % 3;
        if (nn0 < nodeN) dw.kdiag[3 * nn0 + d] = acc[dw.diag_pos[3 * nn0 + d] - 9 * q0];
      }
      __syncthreads();
      for (int i = lane; i < 72; i += 64) {
        const int64_t nn0 = node0 + i / 9;
        const int r = i % 9;
        if (nn0 < nodeN) {
          const int dg = int(nb_ptr[nn0 + 1] - nb_ptr[nn0]);
          acc[dw.diag_pos[3 * nn0] - 9 * q0 + (r / 3) * 3 * dg + r % 3] = T(0);
        }
      }
      __syncthreads();
    }
  }
  T* out = val + 9 * q0;
  for (int i = lane; i < total; i += 64) out[i] = acc[i];
  if constexpr (sizeof(T) == 4) {
    if (dw.nd > 0) {
      // lane = (diagonal s, node): eight consecutive nodes of one diagonal
      // write consecutive entries
      for (int pr = lane; pr < 8 * kMaxDiags; pr += 64) {
        const int ln = pr & 7, sd = pr >> 3;
        const int64_t nn0 = node0 + ln;
        if (sd >= dw.nd || nn0 >= nodeN) continue;
        const int64_t bn = dw.perm ? dw.perm[nn0] : nn0;
        const int64_t bt = bn + dw.off.o[sd];
        if (bt >= n_nodes) continue;
        const int32_t m = dw.iperm ? dw.iperm[bt] : int32_t(bt);
        const int64_t p0 = nb_ptr[nn0];
        const int dg = int(nb_ptr[nn0 + 1] - p0);
        int lo = 0, hi = dg - 1, q = -1;
        while (lo <= hi) {
          const int mid = (lo + hi) >> 1;
          const int32_t v = nbl[p0 - q0 + mid];
          if (v == m) {
            q = mid;
            break;
          }
          if (v < m)
            lo = mid + 1;
          else
            hi = mid - 1;
        }
        if (q < 0) continue;  // no neighbour at this offset: the entry stays zero
        const T* src = acc + 9 * (p0 - q0) + 3 * q;
        const int d3 = 3 * dg;
        const int64_t e =
            ((((bn >> dw.tile_shift) + dw.pad_tiles) * dw.nd + sd) << dw.tile_shift) +
            (bn & ((int64_t(1) << dw.tile_shift) - 1));
        dw.da[e] = make_float4(src[0], src[1], src[2], src[d3]);
        dw.db[e] = make_float4(src[d3 + 1], src[d3 + 2], src[2 * d3], src[2 * d3 + 1]);
        dw.dc[e] = src[2 * d3 + 2];
      }
    }
  }
}

// position of node m in the sorted list L[0 .. deg) (m is in it)
__device__ __forceinline__ int list_pos(const int32_t* __restrict__ L, int deg, int32_t m) {
  int lo = 0, hi = deg - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (L[mid] < m)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The node rows of the general nodes (outside the slot map): one workgroup
// per node gen_node[blockIdx.x]; lane i takes the node's element i of a
// chunk of 64 (gen_ne = 8 e + a, elements ascending), computes its 8 row
// blocks and finds the columns of the element's nodes in the node's
// neighbour list; the lanes then add their blocks to the LDS row segment one
// after another, in element order (the order of the slot-map kernel), and the
// segment is written once.
template <typename T, int MAT, int VAR>
__global__ __launch_bounds__(64) void k_assemble_gen(
    const int32_t* __restrict__ gen_node, const int32_t* __restrict__ gen_ne_ptr,
    const int32_t* __restrict__ gen_ne, const double* __restrict__ coords,
    const int32_t* __restrict__ conn, const int32_t* __restrict__ nb_ptr,
    const int32_t* __restrict__ nb_idx, const double* __restrict__ young, double nu,
    const double* __restrict__ dgp, const double* __restrict__ inc_c, T* __restrict__ val,
    float* __restrict__ kdiag, const int64_t* __restrict__ diag_pos) {
  __shared__ T acc[9 * kMaxGenNbrs];
  const int lane = threadIdx.x;
  const int64_t n = gen_node[blockIdx.x];
  const int64_t q0 = nb_ptr[n];
  const int deg = int(nb_ptr[n + 1] - q0);
  const int total = 9 * deg, d3 = 3 * deg;
  for (int i = lane; i < total; i += 64) acc[i] = T(0);
  const int32_t p0 = gen_ne_ptr[blockIdx.x];
  const int ne = gen_ne_ptr[blockIdx.x + 1] - p0;
  __syncthreads();
  for (int c0 = 0; c0 < ne; c0 += 64) {
    const int i = c0 + lane;
    T kb[8][3][3];
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int j = 0; j < 3; ++j) kb[b][r][j] = T(0);
    int pos[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // incompatible modes: the standard rows first, then -C_a C_b^T in a second
    // pass, so that the two are never live together (one pass spilled here)
    constexpr int RV = VAR == kVarIncompat ? kVarFull : VAR;
    int32_t pk = 0;
    if (i < ne) {
      pk = gen_ne[p0 + i];
      const int32_t elem = pk >> 3;
      elem_node_rows<T, MAT, RV>(coords, conn, elem, pk & 7, young, nu, dgp, inc_c, kb);
#pragma unroll
      for (int b = 0; b < 8; ++b) pos[b] = list_pos(nb_idx + q0, deg, conn[8 * int64_t(elem) + b]);
    }
    const int cnt = ne - c0 < 64 ? ne - c0 : 64;
    auto add_rows = [&]() {
      for (int t = 0; t < cnt; ++t) {
        if (lane == t) {
#pragma unroll
          for (int b = 0; b < 8; ++b) {
            T* dst = acc + 3 * pos[b];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
              for (int j = 0; j < 3; ++j) dst[r * d3 + j] += kb[b][r][j];
          }
        }
        __syncthreads();
      }
    };
    add_rows();
    if constexpr (VAR == kVarIncompat) {
      if (i < ne) {
        const double* cinc = inc_c + int64_t(kIncompatC) * (pk >> 3);
        const int a = pk & 7;
        T ca[3][9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int k = 0; k < 9; ++k) ca[r][k] = T(cinc[9 * (3 * a + r) + k]);
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            T cb[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) cb[k] = T(cinc[9 * (3 * b + j) + k]);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              T t = T(0);
#pragma unroll
              for (int k = 0; k < 9; ++k) t += ca[r][k] * cb[k];
              kb[b][r][j] = -t;
            }
          }
      }
      add_rows();
    }
  }
  if constexpr (sizeof(T) == 4) {
    if (kdiag) {  // difference form, as in k_assemble
      if (lane < 3) kdiag[3 * n + lane] = acc[diag_pos[3 * n + lane] - 9 * q0];
      __syncthreads();
      if (lane < 9) acc[diag_pos[3 * n] - 9 * q0 + (lane / 3) * d3 + lane % 3] = T(0);
      __syncthreads();
    }
  }
  T* out = val + 9 * q0;
  for (int i = lane; i < total; i += 64) out[i] = acc[i];
}

template <typename T>
__global__ void k_diag(int64_t n, const int64_t* __restrict__ pos, const T* __restrict__ val,
                       T* __restrict__ diag) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) diag[i] = val[pos[i]];
}

// Off-diagonal FP32 values: a 32-lane group per node, lane k on neighbour k
// other than the node itself (column position kself follows from diag_pos),
// copies its three 3-value runs from the CSR rows into the compact layout at
// 9 (nb_ptr[n] - n), where it is neighbour k - (k > kself) of 3 (deg - 1)
// per row.  Both sides are contiguous across the lanes.
__global__ __launch_bounds__(256) void k_offdiag_compact(int32_t n_nodes,
                                                         const int32_t* __restrict__ nb_ptr,
                                                         const int64_t* __restrict__ diag_pos,
                                                         const float* __restrict__ val,
                                                         float* __restrict__ valm) {
  const int k = threadIdx.x & 31;
  const int64_t ng = int64_t(gridDim.x) * (blockDim.x / 32);
  for (int64_t n = int64_t(blockIdx.x) * (blockDim.x / 32) + threadIdx.x / 32; n < n_nodes;
       n += ng) {
    const int64_t q0 = nb_ptr[n];
    const int deg = int(nb_ptr[n + 1] - q0);
    const int kself = int((int64_t(diag_pos[3 * n]) - 9 * q0) / 3);
    if (k >= deg || k == kself) continue;
    const int kk = k - (k > kself ? 1 : 0);
    const int64_t src = 9 * q0 + 3 * k, dst = 9 * (q0 - n) + 3 * kk;
    const int s3 = 3 * deg, d3 = 3 * (deg - 1);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c = 0; c < 3; ++c) valm[dst + d * d3 + c] = val[src + d * s3 + c];
  }
}

// Symmetric block-diagonal values from the difference-form val32: a 32-lane
// group per node, lane k on neighbour m = nb_idx[q0 + k] > n copies block
// K(n, m) (node n's rows d, columns 3 m + c) to entry (s, n), o_s = m - n:
// tile-major, ((n / tile + pad_tiles) nd + s) tile + n % tile (tile =
// 2^tile_shift nodes),
// as float4 (0,0) (0,1) (0,2) (1,0), float4 (1,1) (1,2) (2,0) (2,1) and
// float (2,2).  The pairs of the pattern are fixed, so the entries never
// written stay zero from the allocation.
// perm != null: the entries are in the box numbering (node n's rows go to
// entry (s, perm[n]) for the neighbours m with perm[m] - perm[n] = o_s > 0).
__global__ __launch_bounds__(256) void k_diag_fill(int32_t n_nodes,
                                                   const int32_t* __restrict__ nb_ptr,
                                                   const int32_t* __restrict__ nb_idx,
                                                   const float* __restrict__ val, DiagOffsets off,
                                                   int nd, int32_t pad_tiles,
                                                   int32_t tile_shift,
                                                   const int32_t* __restrict__ perm,
                                                   float4* __restrict__ da,
                                                   float4* __restrict__ db,
                                                   float* __restrict__ dc) {
  const int k = threadIdx.x & 31;
  const int64_t ng = int64_t(gridDim.x) * (blockDim.x / 32);
  for (int64_t n0 = int64_t(blockIdx.x) * (blockDim.x / 32) + threadIdx.x / 32; n0 < n_nodes;
       n0 += ng) {
    const int64_t q0 = nb_ptr[n0];
    const int deg = int(nb_ptr[n0 + 1] - q0);
    if (k >= deg) continue;
    const int64_t m0 = nb_idx[q0 + k];
    const int64_t n = perm ? perm[n0] : n0;
    const int64_t o = (perm ? perm[m0] : m0) - n;
    if (o <= 0) continue;
    int s = -1;
#pragma unroll
    for (int j = 0; j < kMaxDiags; ++j)
      if (j < nd && off.o[j] == o) s = j;
    if (s < 0) continue;  // not one of the pattern's offsets (cannot occur)
    const int64_t src = 9 * q0 + 3 * k, s3 = 3 * deg;
    float v[9];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) v[3 * d + cc] = val[src + d * s3 + cc];
    const int64_t e = ((((n >> tile_shift) + pad_tiles) * nd + s) << tile_shift) +
                      (n & ((int64_t(1) << tile_shift) - 1));
    da[e] = make_float4(v[0], v[1], v[2], v[3]);
    db[e] = make_float4(v[4], v[5], v[6], v[7]);
    dc[e] = v[8];
  }
}

// pen32 and pen_flag from the penalty list (both zeroed before): the penalty
// diagonal of the FP32 difference form (pen32[3 b + d] = the FP32 self
// diagonal, float(P)) and bit b of the flags for box node b = perm[node]
// (perm = null: b = node); the equations of the list are distinct.
__global__ void k_pen_scatter(int32_t n_bc, const int32_t* __restrict__ eq,
                              const double* __restrict__ pen, const int32_t* __restrict__ perm,
                              float* __restrict__ pen32, uint32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_bc) return;
  const int32_t q = eq[i];
  const int32_t node = q / 3;
  const int64_t b = perm ? perm[node] : node;
  pen32[3 * b + (q - 3 * node)] = float(pen[i]);
  atomicOr(&flag[b >> 5], 1u << uint32_t(b & 31));
}

// o[3 i + d] = a[3 iperm[i] + d]: a node-ordered FP32 vector in the box numbering
__global__ void k_permute_f32(int32_t n_nodes, const int32_t* __restrict__ iperm,
                              const float* __restrict__ a, float* __restrict__ o) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const int64_t n = iperm[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) o[3 * i + d] = a[3 * n + d];
}

// diag[eq[i]] += pen[i]
__global__ void k_penalty_vec(int32_t n_bc, const int32_t* __restrict__ eq,
                              const double* __restrict__ pen, float* __restrict__ diag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_bc) diag[eq[i]] = float(double(diag[eq[i]]) + pen[i]);
}

template <typename T>
__global__ void k_penalty(int32_t n_bc, const int32_t* __restrict__ eq,
                          const double* __restrict__ pen, const int64_t* __restrict__ pos,
                          T* __restrict__ val, double sign) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_bc) {
    const int64_t p = pos[eq[i]];
    if constexpr (sizeof(T) == 4)
      val[p] = sign > 0.0 ? T(pen[i]) : T(0);  // the difference form's self diagonal is the penalty
    else
      val[p] = T(double(val[p]) + sign * pen[i]);
  }
}

// ---------------------------------------------------------------- element forces
// Fe = sum_gp w detJ B^T sigma_gp, result in w[24 e + j], with sigma from
//   mode 0: D(E_e, nu) B u_e            (elastic K_e u_e: k_elem_force_elastic)
//   mode 1: D_gp B u_e                  (J2 tangent K_e u_e: k_elem_force<1>)
//   mode 2: the J2 trial stress         (internal force, RT_Asm_ComputeResidual:
//                                        k_elem_force<2>)
// The J2 modes: eight lanes per element, lane = Gauss point.  A workgroup's
// 16 elements first stage their nodes' coordinates and u in LDS, one node slot
// per lane (slot 8e + a = node a of element e, the same index range as the
// points), each lane forms its point's 24 force terms, and the eight lanes
// sum them in a fixed order of halving exchanges (nodes 4-7 <-> 0-3 with the
// lane 4 apart, then 2 apart, then 1 apart), after which lane a holds node
// a's three sums and stores them.  64^3: mode 1 57 us, mode 2 60 us against
// 71 and 75 us for one thread per element.
constexpr int kForceThreads = 128;
template <int MODE>
__global__ __launch_bounds__(kForceThreads) void k_elem_force(int32_t n_elems,
                                                             const double* __restrict__ coords,
                                                             const int32_t* __restrict__ conn,
                                                             const double* __restrict__ dgp,
                                                             const double* __restrict__ stress,
                                                             const double* __restrict__ u,
                                                             double* __restrict__ w) {
  static_assert(kForceThreads % 8 == 0, "whole elements per workgroup");
  static_assert(MODE == 1 || MODE == 2, "the J2 modes");
  __shared__ double sx[kForceThreads][3];
  __shared__ double su[MODE == 2 ? 1 : kForceThreads][3];
  const int64_t npts = int64_t(n_elems) * 8;
  const int64_t idx = int64_t(blockIdx.x) * kForceThreads + threadIdx.x;
  const int t = int(threadIdx.x);
  if (idx < npts) {
    const int64_t n = conn[idx];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      sx[t][k] = coords[3 * n + k];
      if constexpr (MODE != 2) su[t][k] = u[3 * n + k];
    }
  }
  __syncthreads();
  if (idx >= npts) return;  // whole elements: npts and the workgroup are multiples of 8
  const int64_t e = idx >> 3;
  const int gp = t & 7, e0 = t & ~7;
  double x[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) x[a][k] = sx[e0 + a][k];
  double fe[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a) fe[a][0] = fe[a][1] = fe[a][2] = 0.0;
  double g[8][3], dj;
  if (c3d8_dndx(x, gp, g, dj)) {
    double sig[6];
    if constexpr (MODE == 2) {
      const double* sp = stress + 13 * idx;
#pragma unroll
      for (int m = 0; m < 6; ++m) sig[m] = sp[m];
    } else {
      double eps[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const double u0 = su[e0 + a][0], u1 = su[e0 + a][1], u2 = su[e0 + a][2];
        eps[0] += g[a][0] * u0;
        eps[1] += g[a][1] * u1;
        eps[2] += g[a][2] * u2;
        eps[3] += g[a][1] * u0 + g[a][0] * u1;
        eps[4] += g[a][2] * u1 + g[a][1] * u2;
        eps[5] += g[a][2] * u0 + g[a][0] * u2;
      }
      j2_tangent_apply(dgp + kJ2Tangent * idx, eps, sig);
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      fe[a][0] = dj * (g[a][0] * sig[0] + g[a][1] * sig[3] + g[a][2] * sig[5]);
      fe[a][1] = dj * (g[a][1] * sig[1] + g[a][0] * sig[3] + g[a][2] * sig[4]);
      fe[a][2] = dj * (g[a][2] * sig[2] + g[a][1] * sig[4] + g[a][0] * sig[5]);
    }
  }
  // halving exchanges: keep the half of the nodes whose index bit matches the
  // lane's, send the other half to the partner lane and add what it sends
  const bool b2 = gp & 4, b1 = gp & 2, b0 = gp & 1;
  double h1[4][3], h2[2][3], h3[3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double keep = b2 ? fe[a + 4][d] : fe[a][d];
      const double send = b2 ? fe[a][d] : fe[a + 4][d];
      h1[a][d] = keep + __shfl_xor(send, 4, 8);
    }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double keep = b1 ? h1[a + 2][d] : h1[a][d];
      const double send = b1 ? h1[a][d] : h1[a + 2][d];
      h2[a][d] = keep + __shfl_xor(send, 2, 8);
    }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double keep = b0 ? h2[1][d] : h2[0][d];
    const double send = b0 ? h2[0][d] : h2[1][d];
    h3[d] = keep + __shfl_xor(send, 1, 8);
  }
  double* out = w + 24 * e + 3 * gp;
  out[0] = h3[0];
  out[1] = h3[1];
  out[2] = h3[2];
}

// Mode 0 (elastic K_e u_e): one thread per element looping over its 8
// points.  The eight-lane form above re-reads the element's 48 node values
// from LDS in every lane and is slower here at every size (48^3 to 128^3:
// 0.029 / 0.056 / 0.177 / 0.410 ms against 0.028 / 0.048 / 0.138 / 0.306).
__global__ __launch_bounds__(128) void k_elem_force_elastic(int32_t n_elems,
                                                           const double* __restrict__ coords,
                                                           const int32_t* __restrict__ conn,
                                                           const double* __restrict__ young,
                                                           double nu,
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
      ue[a][d] = u[3 * n + d];
      fe[a][d] = 0.0;
    }
  }
  const double E = young[e];
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  for (int gp = 0; gp < 8; ++gp) {
    double g[8][3], dj;
    if (!c3d8_dndx(x, gp, g, dj)) continue;
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
    const double tr = lam * (eps[0] + eps[1] + eps[2]);
    const double sig[6] = {tr + 2.0 * mu * eps[0], tr + 2.0 * mu * eps[1],
                           tr + 2.0 * mu * eps[2], mu * eps[3], mu * eps[4], mu * eps[5]};
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

// y[3n+d] = sum over n's elements of w[24 e + 3 a + d]: eight lanes per node,
// lane s on slot s (its element id and n's local index a from the packed
// slot_loc word), so the slot reads of a wave are one contiguous 256-byte run;
// the eight contributions are added by a fixed xor tree ((s0+s1)+(s2+s3)) +
// ((s4+s5)+(s6+s7)) and lanes s = 0..2 write component s.
// gen_flag != null: the nodes of general blocks are left to k_node_gather_gen.
__global__ __launch_bounds__(256) void k_node_gather(int32_t n_nodes,
                                                     const int32_t* __restrict__ slot_elem,
                                                     const uint32_t* __restrict__ slot_loc,
                                                     const uint8_t* __restrict__ gen_flag,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ y) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = t >> 3;
  const int s = int(t & 7);
  const bool live = n < n_nodes && !(gen_flag && gen_flag[n >> 3]);
  const int64_t nn = live ? n : 0;
  const int32_t e = slot_elem[8 * nn + s];
  const int a = int((slot_loc[nn] >> (4 * s)) & 7);
  double f[3] = {0.0, 0.0, 0.0};
  if (live && e >= 0) {
    const double* p = w + 24 * int64_t(e) + 3 * a;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] = p[d];
  }
#pragma unroll
  for (int o = 1; o < 8; o <<= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] += __shfl_xor(f[d], o, 64);
  if (live && s < 3) y[3 * n + s] = (s == 0) ? f[0] : ((s == 1) ? f[1] : f[2]);
}

// y[3n+d] for the general nodes: one thread per node over its elements in
// ascending order
__global__ __launch_bounds__(256) void k_node_gather_gen(int32_t n_gen,
                                                         const int32_t* __restrict__ gen_node,
                                                         const int32_t* __restrict__ gen_ne_ptr,
                                                         const int32_t* __restrict__ gen_ne,
                                                         const double* __restrict__ w,
                                                         double* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_gen) return;
  double f[3] = {0.0, 0.0, 0.0};
  for (int32_t q = gen_ne_ptr[i]; q < gen_ne_ptr[i + 1]; ++q) {
    const int32_t pk = gen_ne[q];
    const double* p = w + 24 * int64_t(pk >> 3) + 3 * (pk & 7);
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] += p[d];
  }
  const int64_t n = gen_node[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) y[3 * n + d] = f[d];
}

// ---------------------------------------------------------------- variant apply
// Eight lanes per element, lane = node a: f_a = sum_b K_ab u_b with the
// elastic K_e of variant VAR.  SENS = false: w[24 e + 3 a + d] = f_a (element
// forces for the node gather); SENS = true: grad_e = -dE/drho_e lam_e^T K_e u_e / E_e
// (K_e is linear in E_e for every variant).
template <int VAR, bool SENS>
__global__ __launch_bounds__(64) void k_elem_apply(int32_t n_elems,
                                                   const double* __restrict__ coords,
                                                   const int32_t* __restrict__ conn,
                                                   const double* __restrict__ young, double nu,
                                                   const double* __restrict__ inc_c,
                                                   const double* __restrict__ u,
                                                   const double* __restrict__ lamv,
                                                   const double* __restrict__ dedrho,
                                                   double* __restrict__ out) {
  const int64_t e0 = int64_t(blockIdx.x) * 8 + (threadIdx.x >> 3);
  const int a = threadIdx.x & 7;
  const bool live = e0 < n_elems;
  const int64_t e = live ? e0 : 0;
  double x[8][3];
  gather_coords(coords, conn, e, x);
  const double E = young[e];
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  // incompatible modes: K_e = K_std - C C^T (C = K_ua L^-T, 24 x 9), applied
  // as K_std u - C (C^T u) without forming C C^T: each lane adds its node's
  // rows of C^T u and the eight lanes sum them (three exchanges)
  constexpr int RV = VAR == kVarIncompat ? kVarFull : VAR;
  double kb[8][3][3] = {};
  elem_rows<double, RV>(x, a, lam, mu, nullptr, kb);
  double f[3] = {0.0, 0.0, 0.0};
  double ua[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int64_t n = conn[8 * e + b];
    const double u0 = u[3 * n], u1 = u[3 * n + 1], u2 = u[3 * n + 2];
    if (b == a) {
      ua[0] = u0;
      ua[1] = u1;
      ua[2] = u2;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) f[i] += kb[b][i][0] * u0 + kb[b][i][1] * u1 + kb[b][i][2] * u2;
  }
  if constexpr (VAR == kVarIncompat) {
    const double* ca = inc_c + int64_t(kIncompatC) * e + 27 * a;  // rows 3a..3a+2 of C
    double w[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      double t = ca[k] * ua[0] + ca[9 + k] * ua[1] + ca[18 + k] * ua[2];
      t += __shfl_xor(t, 1, 8);
      t += __shfl_xor(t, 2, 8);
      t += __shfl_xor(t, 4, 8);
      w[k] = t;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) t += ca[9 * i + k] * w[k];
      f[i] -= t;
    }
  }
  if constexpr (!SENS) {
    if (live) {
#pragma unroll
      for (int i = 0; i < 3; ++i) out[24 * e + 3 * a + i] = f[i];
    }
  } else {
    const int64_t na = conn[8 * e + a];
    double t = lamv[3 * na] * f[0] + lamv[3 * na + 1] * f[1] + lamv[3 * na + 2] * f[2];
    t += __shfl_xor(t, 1, 8);
    t += __shfl_xor(t, 2, 8);
    t += __shfl_xor(t, 4, 8);
    if (live && a == 0) out[e] = -dedrho[e] * t / E;
  }
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
// sigma_y(peeq) and H'(peeq) of PH_J2_ComputeHardening /
// PH_J2_ComputeHardeningTangent (PH_Mat_Plast_J2_Iso_Core.f90:549-605); the
// law was validated and normalised (0 -> linear) by ufc_ph_j2_init
__device__ __forceinline__ double j2_yield(const ufc_ph_j2_props& p, double e) {
  if (p.hardening_type == UFC_J2_HARD_SWIFT) return p.swift_k * pow(p.swift_eps0 + e, p.swift_n);
  if (p.hardening_type == UFC_J2_HARD_VOCE)
    return p.sigma_y0 + p.voce_q * (1.0 - exp(-p.voce_delta * e));
  return p.sigma_y0 + p.hardening * e;
}
__device__ __forceinline__ double j2_hard_tan(const ufc_ph_j2_props& p, double e) {
  if (p.hardening_type == UFC_J2_HARD_SWIFT)
    return p.swift_k * p.swift_n * pow(p.swift_eps0 + e, p.swift_n - 1.0);
  if (p.hardening_type == UFC_J2_HARD_VOCE)
    return p.voce_q * p.voce_delta * exp(-p.voce_delta * e);
  return p.hardening;
}

// PH_J2_ComputeStress_Core (PH_Mat_Plast_J2_Iso_Core.f90:146-207): trial
// stress (:254-317), yield check against sigma_y(peeq_n), radial return with
// the local Newton of :355-456 -- dg from f / (3G + H'(peeq_n)), then
// R = q - 3G dg - sigma_y(peeq_n + dg), dg -= R / -(3G + H'), dg >= 0, until
// |R| < 1e-10 sigma_y0 within max_local_iter iterations (25 by default);
// the tangent is the consistent tangent
//   Dt = K 1x1 + 2G th1 (I_s - 1x1/3) - 2G (th1 + th2 - 1) n x n
// (th1 = 1 - 3G dg/q, th2 = 3G/(3G + H'(peeq_n + dg)),
// I_s = diag(1,1,1,1/2,1/2,1/2)), which replaces the reference's :509-537,
// kept factored as tc = (G th1, K - 2G th1/3, 2G (th1 + th2 - 1), n).
// A point that has yielded and whose trial state lies on the yield surface
// (f > -1e-8 sigma_y0: the zero strain increment of an increment's first
// Newton iteration) gets the plastic branch's tangent at dg = 0 (th1 = 1,
// th2 = 3G/(3G + H'(peeq_n))), the direction a loading step takes, instead of
// D_el, which would make that iteration an elastic predictor; the stress is
// the trial stress either way.
// st: sigma(6), peeq, eps_p(6).  Returns 0 elastic, 1 plastic, 2 failure
// (degenerate trial state, zero Jacobian or no local convergence: st and tc
// are then not meaningful).
__device__ __forceinline__ void j2_tangent(double K, double G, double th1, double th2,
                                           const double (&n)[6], double* tc) {
  tc[0] = G * th1;
  tc[1] = K - 2.0 * G * th1 / 3.0;
  tc[2] = 2.0 * G * (th1 + th2 - 1.0);
#pragma unroll
  for (int i = 0; i < 6; ++i) tc[3 + i] = n[i];
}

__device__ __forceinline__ int j2_point(const ufc_ph_j2_props& pr, const double* deps, double* st,
                                         double* tc) {
  const double E = pr.young, nu = pr.nu;
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
  const double peeq_n = st[6];
  const double f = q - j2_yield(pr, peeq_n);
  const double G = mu, K = E / (3.0 * (1.0 - 2.0 * nu));
  if (f <= 0.0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) st[i] = strial[i];
    if (peeq_n > 0.0 && f > -1.0e-8 * pr.sigma_y0 && q >= 1.0e-30) {
      double n0[6];
      const double sn0 = q / sqrt(1.5);
#pragma unroll
      for (int i = 0; i < 6; ++i) n0[i] = s[i] / sn0;
      j2_tangent(K, G, 1.0, 3.0 * G / (3.0 * G + j2_hard_tan(pr, peeq_n)), n0, tc);
    } else {
      tc[0] = mu;
      tc[1] = lam;
#pragma unroll
      for (int i = 2; i < kJ2Tangent; ++i) tc[i] = 0.0;
    }
    return 0;
  }
  if (q < 1.0e-30) return 2;
  double n[6];
  const double snorm = q / sqrt(1.5);
#pragma unroll
  for (int i = 0; i < 6; ++i) n[i] = s[i] / snorm;
  double dg = f / (3.0 * G + j2_hard_tan(pr, peeq_n));
  const int max_it = pr.max_local_iter > 0 ? pr.max_local_iter : 25;
  int it;
  for (it = 1; it <= max_it; ++it) {
    const double R = q - 3.0 * G * dg - j2_yield(pr, peeq_n + dg);
    if (fabs(R) < 1.0e-10 * pr.sigma_y0) break;
    const double dR = -(3.0 * G + j2_hard_tan(pr, peeq_n + dg));
    if (fabs(dR) < 1.0e-30) return 2;
    dg = dg - R / dR;
    if (dg < 0.0) dg = 0.0;
  }
  if (it > max_it) return 2;
  const double beta = 1.0 - 3.0 * G * dg / q;
  st[0] = beta * s[0] + pm;
  st[1] = beta * s[1] + pm;
  st[2] = beta * s[2] + pm;
  st[3] = beta * s[3];
  st[4] = beta * s[4];
  st[5] = beta * s[5];
  st[6] = peeq_n + dg;
#pragma unroll
  for (int i = 0; i < 6; ++i) st[7 + i] += dg * n[i];
  j2_tangent(K, G, beta, 3.0 * G / (3.0 * G + j2_hard_tan(pr, st[6])), n, tc);
  return 1;
}

// one thread per Gauss point: strain increment B(u - u_n), J2 update from the
// committed state into the trial state and tangent.  A workgroup holds 16
// whole elements (8 points each); their nodes' coordinates and u - u_n are
// staged in LDS first, one node slot per lane (slot 8e + a is node a of
// element e, the same index range as the points), so each lane gathers one
// node instead of all eight of its element (0.327 -> 0.300 ms per update at
// 64^3).  The workgroup's points hold one contiguous range of the state
// (13 doubles each) and of the tangent (kJ2Tangent each): it is read and
// written with 16-byte accesses that are contiguous across the lanes and
// passed through LDS, instead of 13 + 13 + 9 loads and stores per lane 104 or
// 72 bytes apart.
constexpr int kJ2Threads = 128;
constexpr int kJ2StateV = kJ2Vals * kJ2Threads / 2;     // 16-byte words of state per workgroup
constexpr int kJ2TanV = kJ2Tangent * kJ2Threads / 2;   // and of tangent
static_assert(kJ2Threads % 8 == 0 && (kJ2Vals * 8) % 2 == 0 && (kJ2Tangent * 8) % 2 == 0,
              "whole elements per workgroup, whole 16-byte words per element");
__global__ __launch_bounds__(kJ2Threads) void k_j2_update(int32_t n_elems,
                                                         const double* __restrict__ coords,
                                                         const int32_t* __restrict__ conn,
                                                         ufc_ph_j2_props pr,
                                                         const double* __restrict__ u,
                                                         const double* __restrict__ u_n,
                                                         const double* __restrict__ commit,
                                                         double* __restrict__ trial,
                                                         double* __restrict__ tangent,
                                                         int32_t* __restrict__ flags) {
  // LDS: the state records (committed in, trial out), and the nodes'
  // coordinates and u - u_n, whose space the tangent records take over once
  // the points are computed (22.5 KB: six workgroups, three waves per SIMD)
  constexpr int kBufV = kJ2TanV > 3 * kJ2Threads ? kJ2TanV : 3 * kJ2Threads;
  __shared__ double2 sst[kJ2StateV];
  __shared__ double2 stn[kBufV];
  double (*const sx)[3] = reinterpret_cast<double (*)[3]>(stn);
  double (*const sd)[3] = reinterpret_cast<double (*)[3]>(stn) + kJ2Threads;
  const int64_t npts = int64_t(n_elems) * 8;
  const int64_t first = int64_t(blockIdx.x) * kJ2Threads;
  const int t = int(threadIdx.x);
  const int64_t idx = first + t;
  const bool active = idx < npts;
  // the workgroup's points: a multiple of 8, so whole 16-byte words
  const int np = int(npts - first < kJ2Threads ? npts - first : kJ2Threads);
  const int nsv = kJ2Vals * np / 2, ntv = kJ2Tangent * np / 2;
  constexpr int kPer = (kJ2StateV + kJ2Threads - 1) / kJ2Threads;
  const double2* c2 = reinterpret_cast<const double2*>(commit + kJ2Vals * first);
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int j = t + k * kJ2Threads;
    if (j < nsv) sst[j] = c2[j];
  }
  if (active) {
    const int64_t n = conn[idx];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      sx[t][k] = coords[3 * n + k];
      sd[t][k] = u[3 * n + k] - u_n[3 * n + k];
    }
  }
  __syncthreads();
  double* const ss = reinterpret_cast<double*>(sst);
  double st[kJ2Vals];
  double tc[kJ2Tangent];
  if (active) {
    const int e0 = t & ~7;  // the element's first slot
    const int gp = t & 7;
    double x[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) x[a][k] = sx[e0 + a][k];
    double g[8][3], dj;
    double deps[6] = {0, 0, 0, 0, 0, 0};
    if (c3d8_dndx(x, gp, g, dj)) {
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const double d0 = sd[e0 + a][0], d1 = sd[e0 + a][1], d2 = sd[e0 + a][2];
        deps[0] += g[a][0] * d0;
        deps[1] += g[a][1] * d1;
        deps[2] += g[a][2] * d2;
        deps[3] += g[a][1] * d0 + g[a][0] * d1;
        deps[4] += g[a][2] * d1 + g[a][1] * d2;
        deps[5] += g[a][2] * d0 + g[a][0] * d2;
      }
    }
#pragma unroll
    for (int k = 0; k < kJ2Vals; ++k) st[k] = ss[kJ2Vals * t + k];
    flags[idx] = j2_point(pr, deps, st, tc);
  }
  __syncthreads();  // every committed record and node read before their space is reused
  if (active) {
#pragma unroll
    for (int k = 0; k < kJ2Vals; ++k) ss[kJ2Vals * t + k] = st[k];
    double* const sn = reinterpret_cast<double*>(stn);
#pragma unroll
    for (int k = 0; k < kJ2Tangent; ++k) sn[kJ2Tangent * t + k] = tc[k];
  }
  __syncthreads();
  double2* const t2 = reinterpret_cast<double2*>(trial + kJ2Vals * first);
  double2* const g2 = reinterpret_cast<double2*>(tangent + kJ2Tangent * first);
  for (int j = t; j < nsv; j += kJ2Threads) t2[j] = sst[j];
  for (int j = t; j < ntv; j += kJ2Threads) g2[j] = stn[j];
}

inline unsigned blocks_for(int64_t n, int t) { return unsigned((n + t - 1) / t); }

template <int VAR>
void eval_ke_var(int32_t n_elems, const double* coords, const double* young, double nu, void* ke,
                 bool fp32, const double* inc_c, hipStream_t s) {
  const unsigned grid = blocks_for(n_elems, 8);
  if (fp32)
    hipLaunchKernelGGL((k_eval_ke<float, VAR>), dim3(grid), dim3(64), 0, s, n_elems, coords, young,
                       nu, inc_c, static_cast<float*>(ke));
  else
    hipLaunchKernelGGL((k_eval_ke<double, VAR>), dim3(grid), dim3(64), 0, s, n_elems, coords,
                       young, nu, inc_c, static_cast<double*>(ke));
}

// the general nodes' rows, after the slot-map kernel
template <typename T, int MAT, int VAR>
void assemble_general(ufc_ctx* c, const double* dgp, T* val) {
  if (c->n_gen == 0) return;
  hipLaunchKernelGGL((k_assemble_gen<T, MAT, VAR>), dim3(unsigned(c->n_gen)), dim3(64), 0,
                     c->stream, c->gen_node.get(), c->gen_ne_ptr.get(), c->gen_ne.get(),
                     c->coords.get(), c->conn.get(), c->nb_ptr.get(), c->nb_idx.get(),
                     MAT == UFC_MAT_ELASTIC ? c->young.get() : nullptr, c->nu, dgp,
                     c->inc_c.get(), val, sizeof(T) == 4 ? c->kdiag32.get() : nullptr,
                     c->diag_pos.get());
}

template <typename T, int VAR>
void assemble_elastic(ufc_ctx* c, T* val, const DiagWrite& dw) {
  hipLaunchKernelGGL((k_assemble<T, UFC_MAT_ELASTIC, VAR>), dim3(blocks_for(c->n_nodes, 8)),
                     dim3(64), 0, c->stream, c->n_nodes, c->coords.get(), c->conn.get(),
                     c->slot_elem.get(), c->slot_code.get(), c->nb_ptr.get(), c->young.get(),
                     c->nu, nullptr, c->inc_c.get(), val, dw, c->gen_flag.get());
  assemble_general<T, UFC_MAT_ELASTIC, VAR>(c, nullptr, val);
}

template <typename T>
void assemble_variant(ufc_ctx* c, T* val, const DiagWrite& dw) {
  switch (c->variant) {
    case kVarReduced:
      assemble_elastic<T, kVarReduced>(c, val, dw);
      break;
    case kVarSelective:
      assemble_elastic<T, kVarSelective>(c, val, dw);
      break;
    case kVarIncompat:
      assemble_elastic<T, kVarIncompat>(c, val, dw);
      break;
    case kVarGauss27:
      assemble_elastic<T, kVarGauss27>(c, val, dw);
      break;
    default:
      assemble_elastic<T, kVarFull>(c, val, dw);
  }
}

// the symmetric view's arrays, zeroed once: the pattern fixes which entries
// are ever written, so the others stay zero
void ensure_diag_arrays(ufc_ctx* c) {
  if (c->diag_c.get()) return;
  const size_t entries = size_t(c->diag_entries);
  c->diag_a.resize(4 * entries);
  c->diag_b.resize(4 * entries);
  c->diag_c.resize(entries);
  UFC_HIP_CHECK(hipMemsetAsync(c->diag_a.get(), 0, c->diag_a.bytes(), c->stream));
  UFC_HIP_CHECK(hipMemsetAsync(c->diag_b.get(), 0, c->diag_b.bytes(), c->stream));
  UFC_HIP_CHECK(hipMemsetAsync(c->diag_c.get(), 0, c->diag_c.bytes(), c->stream));
}

DiagWrite diag_write(ufc_ctx* c, bool fp32) {
  DiagWrite dw{};
  dw.nd = 0;
  if (!fp32) return dw;
  dw.kdiag = c->kdiag32.get();
  dw.diag_pos = c->diag_pos.get();
  if (c->fp32_view != kFp32ViewSymDiag || c->n_diag == 0) return dw;
  ensure_diag_arrays(c);
  for (int s = 0; s < kMaxDiags; ++s) dw.off.o[s] = c->h_diag_off[s];
  dw.nd = c->n_diag;
  dw.pad_tiles = c->diag_pad_tiles;
  dw.tile_shift = tile_shift(c->diag_tile);
  dw.perm = fp32_permuted(c) ? c->sym_perm.get() : nullptr;
  dw.iperm = fp32_permuted(c) ? c->sym_iperm.get() : nullptr;
  dw.nb_idx = c->nb_idx.get();
  dw.da = reinterpret_cast<float4*>(c->diag_a.get());
  dw.db = reinterpret_cast<float4*>(c->diag_b.get());
  dw.dc = c->diag_c.get();
  return dw;
}

template <int VAR, bool SENS>
void apply_var(ufc_ctx* c, const double* u, const double* lamv, const double* dedrho,
               double* out) {
  hipLaunchKernelGGL((k_elem_apply<VAR, SENS>), dim3(blocks_for(c->n_elems, 8)), dim3(64), 0,
                     c->stream, c->n_elems, c->coords.get(), c->conn.get(), c->young.get(), c->nu,
                     c->inc_c.get(), u, lamv, dedrho, out);
}

template <bool SENS>
void apply_variant(ufc_ctx* c, const double* u, const double* lamv, const double* dedrho,
                   double* out) {
  switch (c->variant) {
    case kVarReduced:
      apply_var<kVarReduced, SENS>(c, u, lamv, dedrho, out);
      break;
    case kVarSelective:
      apply_var<kVarSelective, SENS>(c, u, lamv, dedrho, out);
      break;
    case kVarIncompat:
      apply_var<kVarIncompat, SENS>(c, u, lamv, dedrho, out);
      break;
    case kVarGauss27:
      apply_var<kVarGauss27, SENS>(c, u, lamv, dedrho, out);
      break;
    default:
      apply_var<kVarFull, SENS>(c, u, lamv, dedrho, out);
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers
void launch_incompat_c(int32_t n_elems, const double* coords, const int32_t* conn,
                       const double* young, double nu, double* inc_c, hipStream_t s) {
  if (n_elems <= 0) return;
  hipLaunchKernelGGL(k_incompat_c, dim3(n_elems), dim3(64), 0, s, n_elems, coords, conn, young,
                     nu, inc_c);
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_eval_ke(int32_t n_elems, const double* coords, const double* young, double nu,
                    void* ke, bool fp32, int variant, double* inc_c, hipStream_t s) {
  if (n_elems <= 0) return;
  switch (variant) {
    case kVarReduced:
      eval_ke_var<kVarReduced>(n_elems, coords, young, nu, ke, fp32, nullptr, s);
      break;
    case kVarSelective:
      eval_ke_var<kVarSelective>(n_elems, coords, young, nu, ke, fp32, nullptr, s);
      break;
    case kVarIncompat:
      launch_incompat_c(n_elems, coords, nullptr, young, nu, inc_c, s);
      eval_ke_var<kVarIncompat>(n_elems, coords, young, nu, ke, fp32, inc_c, s);
      break;
    case kVarGauss27:
      eval_ke_var<kVarGauss27>(n_elems, coords, young, nu, ke, fp32, nullptr, s);
      break;
    default:
      eval_ke_var<kVarFull>(n_elems, coords, young, nu, ke, fp32, nullptr, s);
  }
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_assemble(ufc_ctx* c, int precision, int material) {
  const unsigned grid = blocks_for(c->n_nodes, 8);
  const double* dgp = c->j2_tangent.get();
  if (material == UFC_MAT_ELASTIC && c->variant == kVarIncompat)
    launch_incompat_c(c->n_elems, c->coords.get(), c->conn.get(), c->young.get(), c->nu,
                      c->inc_c.get(), c->stream);
  const DiagWrite dw = diag_write(c, precision != UFC_PREC_FP64);
  if (precision != UFC_PREC_FP64) c->diag_from_assembly = dw.nd > 0;
  if (precision == UFC_PREC_FP64) {
    if (material == UFC_MAT_ELASTIC) {
      assemble_variant<double>(c, c->val64.get(), dw);
    } else {
      hipLaunchKernelGGL((k_assemble<double, UFC_MAT_J2, kVarFull>), dim3(grid), dim3(64), 0,
                         c->stream, c->n_nodes, c->coords.get(), c->conn.get(),
                         c->slot_elem.get(), c->slot_code.get(), c->nb_ptr.get(), nullptr, c->nu,
                         dgp, nullptr, c->val64.get(), dw, c->gen_flag.get());
      assemble_general<double, UFC_MAT_J2, kVarFull>(c, dgp, c->val64.get());
    }
  } else {
    if (material == UFC_MAT_ELASTIC) {
      assemble_variant<float>(c, c->val32.get(), dw);
    } else {
      hipLaunchKernelGGL((k_assemble<float, UFC_MAT_J2, kVarFull>), dim3(grid), dim3(64), 0,
                         c->stream, c->n_nodes, c->coords.get(), c->conn.get(),
                         c->slot_elem.get(), c->slot_code.get(), c->nb_ptr.get(), nullptr, c->nu,
                         dgp, nullptr, c->val32.get(), dw, c->gen_flag.get());
      assemble_general<float, UFC_MAT_J2, kVarFull>(c, dgp, c->val32.get());
    }
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
  // kdiag32 and the zeroed self blocks come from the FP32 assembly itself
  launch_penalty(c, true, +1.0);
  launch_diag32_penalty(c);
  launch_fp32_offdiag(c);
}

void launch_fp32_offdiag(ufc_ctx* c) {
  const bool pattern = c->fp32_view == kFp32ViewPattern, sym = c->fp32_view == kFp32ViewSymDiag;
  if (!pattern && !sym) return;
  c->pen32.resize(size_t(c->n_dof));
  const int64_t nwords = int64_t(c->n_nodes) / 32 + 1;
  c->pen_flag.resize(size_t(nwords));
  const unsigned cgrid = unsigned(std::min<int64_t>(blocks_for(c->n_nodes, 8), 32768));
  if (pattern) {
    const int64_t nv = c->nnz - 9 * int64_t(c->n_nodes);
    c->val32m.resize(size_t(nv));
    hipLaunchKernelGGL(k_offdiag_compact, dim3(cgrid), dim3(256), 0, c->stream, c->n_nodes,
                       c->nb_ptr.get(), c->diag_pos.get(), c->val32.get(), c->val32m.get());
  } else {
    const bool perm = fp32_permuted(c);
    if (!c->diag_from_assembly) {
      // the view was chosen after the FP32 assembly: fill it from val32
      ensure_diag_arrays(c);
      DiagOffsets off;
      for (int s = 0; s < kMaxDiags; ++s) off.o[s] = c->h_diag_off[s];
      hipLaunchKernelGGL(k_diag_fill, dim3(cgrid), dim3(256), 0, c->stream, c->n_nodes,
                         c->nb_ptr.get(), c->nb_idx.get(), c->val32.get(), off, int(c->n_diag),
                         c->diag_pad_tiles, tile_shift(c->diag_tile),
                         perm ? c->sym_perm.get() : nullptr,
                         reinterpret_cast<float4*>(c->diag_a.get()),
                         reinterpret_cast<float4*>(c->diag_b.get()), c->diag_c.get());
      c->diag_from_assembly = true;  // filled now: a later penalty change keeps them
    }
    if (perm) {
      // the Jacobi diagonal of the FP32 PCG in the box numbering
      c->diag32p.resize(size_t(c->n_dof));
      hipLaunchKernelGGL(k_permute_f32, dim3(blocks_for(c->n_nodes, 256)), dim3(256), 0,
                         c->stream, c->n_nodes, c->sym_iperm.get(), c->diag32.get(),
                         c->diag32p.get());
    }
  }
  UFC_HIP_CHECK(hipMemsetAsync(c->pen32.get(), 0, c->pen32.bytes(), c->stream));
  UFC_HIP_CHECK(hipMemsetAsync(c->pen_flag.get(), 0, c->pen_flag.bytes(), c->stream));
  const int32_t nb = int32_t(c->h_bc_eq.size());
  if (nb)
    hipLaunchKernelGGL(k_pen_scatter, dim3(blocks_for(nb, 256)), dim3(256), 0, c->stream, nb,
                       c->bc_eq.get(), c->bc_pen.get(),
                       fp32_permuted(c) ? c->sym_perm.get() : nullptr, c->pen32.get(),
                       c->pen_flag.get());
  UFC_HIP_CHECK(hipGetLastError());
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
  const unsigned gpt = blocks_for(int64_t(c->n_elems) * 8, kForceThreads);
  if (mode == 0 && c->variant != kVarFull)
    apply_variant<false>(c, u, nullptr, nullptr, c->elem_w.get());
  else if (mode == 0)
    hipLaunchKernelGGL(k_elem_force_elastic, dim3(blocks_for(c->n_elems, 128)), dim3(128), 0,
                       c->stream, c->n_elems, c->coords.get(), c->conn.get(), c->young.get(),
                       c->nu, u, c->elem_w.get());
  else if (mode == 1)
    hipLaunchKernelGGL(k_elem_force<1>, dim3(gpt), dim3(kForceThreads), 0, c->stream,
                       c->n_elems, c->coords.get(), c->conn.get(), c->j2_tangent.get(), nullptr,
                       u, c->elem_w.get());
  else
    hipLaunchKernelGGL(k_elem_force<2>, dim3(gpt), dim3(kForceThreads), 0, c->stream,
                       c->n_elems, c->coords.get(), c->conn.get(), nullptr, c->j2_trial.get(),
                       nullptr, c->elem_w.get());
  UFC_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_node_gather, dim3(blocks_for(int64_t(c->n_nodes) * 8, 256)), dim3(256), 0,
                     c->stream, c->n_nodes, c->slot_elem.get(), c->slot_loc.get(),
                     c->gen_flag.get(), c->elem_w.get(), y);
  if (c->n_gen)
    hipLaunchKernelGGL(k_node_gather_gen, dim3(blocks_for(c->n_gen, 256)), dim3(256), 0,
                       c->stream, c->n_gen, c->gen_node.get(), c->gen_ne_ptr.get(),
                       c->gen_ne.get(), c->elem_w.get(), y);
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_sensitivity(ufc_ctx* c, const double* u, const double* lam, const double* dedrho,
                        double nu, double* grad) {
  if (c->variant != kVarFull) {  // the variant and E_e of the last elastic assembly
    apply_variant<true>(c, u, lam, dedrho, grad);
    UFC_HIP_CHECK(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_sensitivity, dim3(blocks_for(c->n_elems, 128)), dim3(128), 0, c->stream,
                     c->n_elems, c->coords.get(), c->conn.get(), nu, u, lam, dedrho, grad);
  UFC_HIP_CHECK(hipGetLastError());
}

void launch_j2_update(ufc_ctx* c, const double* u, const double* u_n) {
  const int64_t ngp = int64_t(c->n_elems) * 8;
  hipLaunchKernelGGL(k_j2_update, dim3(blocks_for(ngp, kJ2Threads)), dim3(kJ2Threads), 0,
                     c->stream, c->n_elems, c->coords.get(), c->conn.get(), c->j2, u, u_n,
                     c->j2_commit.get(), c->j2_trial.get(), c->j2_tangent.get(),
                     c->j2_flags.get());
  UFC_HIP_CHECK(hipGetLastError());
}

}  // namespace ufc
