/*
 * ufc_oracle.c -- FP64 CPU restatement of the UFC static-stiffness hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
 * timed CPU baseline.  Nothing in the product path links or calls it.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the reference root).  Where the reference is defective the restatement
 * follows the documented resolution and says so at the function:
 *   - shape functions use 1/8, not QUARTER            (PH_Elem_C3D8.f90:1936-1939)
 *   - dN/dx = J^-T dN/dxi, not J^-1 dN/dxi            (PH_Elem_C3D8.f90:936-943)
 *   - CSR is merged and column-sorted                  (RT_Solv_Sparse.f90:336-447)
 *   - J2 tangent is the textbook consistent tangent    (PH_Mat_Plast_J2_Iso_Core.f90:467-539)
 */
#define _POSIX_C_SOURCE 199309L  /* clock_gettime */
#include "ufc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* PH_ELEM_GAUSS_PT, PH_Elem_C3D8.f90:118 (the reference's 15-digit literal) */
static const double GP = 0.577350269189626;

/* Abaqus hex node natural coordinates, PH_Elem_C3D8.f90:1930-1933 */
static const double XI_N[8]   = {-1, 1, 1, -1, -1, 1, 1, -1};
static const double ETA_N[8]  = {-1, -1, 1, 1, -1, -1, 1, 1};
static const double ZETA_N[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

/* PH_ELEM_C3D8_FACE_NODES, PH_Elem_C3D8.f90:120-121 (1-based there) */
static const int FACE_NODES[6][4] = {
    {0, 1, 2, 3}, {4, 7, 6, 5}, {0, 1, 5, 4}, {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 5, 6, 2}};

/* ========================================================================= */
/* Seeded counter-based generator (SplitMix64 finaliser, Steele et al. 2014). */
/* ========================================================================= */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, uint64_t counter) {
  const uint64_t G = 0x9E3779B97F4A7C15ULL;
  return mix64(mix64(seed + G * (stream + 1)) + G * (counter + 1));
}

/* uniform on (0, 1] with 53 random bits */
double orc_rng_u01(uint64_t seed, uint64_t stream, uint64_t counter) {
  return (double)((orc_rng_u64(seed, stream, counter) >> 11) + 1) * 0x1.0p-53;
}

/* E_e = exp(mu + sigma z), z by Box-Muller from counters 2e and 2e+1, stream 0 */
void orc_field_lognormal(int64_t n, uint64_t seed, double mu, double sigma, double* out) {
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t e = 0; e < n; ++e) {
    double u1 = orc_rng_u01(seed, 0, 2 * (uint64_t)e);
    double u2 = orc_rng_u01(seed, 0, 2 * (uint64_t)e + 1);
    double z = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
    out[e] = exp(mu + sigma * z);
  }
}

/* a + (b - a) u, stream 1 */
void orc_field_uniform(int64_t n, uint64_t seed, double a, double b, double* out) {
  for (int64_t e = 0; e < n; ++e) out[e] = a + (b - a) * orc_rng_u01(seed, 1, (uint64_t)e);
}

/* ========================================================================= */
/* Structured box mesh.  The reference has no generator on the path; node id */
/* = i + (nx+1)(j + (ny+1)k), elements in the same lexicographic order, each */
/* element's nodes in the Abaqus order of PH_Elem_C3D8.f90:1930-1933.        */
/* ========================================================================= */
void orc_mesh_box(int nx, int ny, int nz, double lx, double ly, double lz,
                  double* coords, int32_t* conn) {
  const int64_t sx = nx + 1, sy = ny + 1;
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        int64_t n = i + sx * (j + sy * k);
        coords[3 * n + 0] = lx * (double)i / (double)nx;
        coords[3 * n + 1] = ly * (double)j / (double)ny;
        coords[3 * n + 2] = lz * (double)k / (double)nz;
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int64_t e = i + (int64_t)nx * (j + (int64_t)ny * k);
        int64_t n0 = i + sx * (j + sy * k);
        int32_t* c = conn + 8 * e;
        c[0] = (int32_t)n0;
        c[1] = (int32_t)(n0 + 1);
        c[2] = (int32_t)(n0 + 1 + sx);
        c[3] = (int32_t)(n0 + sx);
        c[4] = (int32_t)(n0 + sx * sy);
        c[5] = (int32_t)(n0 + 1 + sx * sy);
        c[6] = (int32_t)(n0 + 1 + sx + sx * sy);
        c[7] = (int32_t)(n0 + sx + sx * sy);
      }
}

/* interior nodes moved by U[-frac, frac] * h per coordinate, stream 2 */
void orc_mesh_jitter(int nx, int ny, int nz, double lx, double ly, double lz,
                     double frac, uint64_t seed, double* coords) {
  const int64_t sx = nx + 1, sy = ny + 1;
  const double h[3] = {lx / nx, ly / ny, lz / nz};
  for (int k = 1; k < nz; ++k)
    for (int j = 1; j < ny; ++j)
      for (int i = 1; i < nx; ++i) {
        int64_t n = i + sx * (j + sy * k);
        for (int d = 0; d < 3; ++d) {
          double u = orc_rng_u01(seed, 2, 3 * (uint64_t)n + d);
          coords[3 * n + d] += (2.0 * u - 1.0) * frac * h[d];
        }
      }
}

/* ========================================================================= */
/* C3D8 element.                                                             */
/* ========================================================================= */

/* PH_Elem_C3D8_ShapeFunc, PH_Elem_C3D8.f90:1922-1944, with the trilinear
 * factor 1/8 (the reference's QUARTER makes sum N = 2).  dNdxi is (3,8)
 * column-major: dNdxi[3*a + k] = dN_a/dxi_k. */
void orc_c3d8_shape(double xi, double eta, double zeta, double N[8], double dNdxi[24]) {
  for (int a = 0; a < 8; ++a) {
    double fx = 1.0 + XI_N[a] * xi, fy = 1.0 + ETA_N[a] * eta, fz = 1.0 + ZETA_N[a] * zeta;
    N[a] = 0.125 * fx * fy * fz;
    dNdxi[3 * a + 0] = 0.125 * XI_N[a] * fy * fz;
    dNdxi[3 * a + 1] = 0.125 * fx * ETA_N[a] * fz;
    dNdxi[3 * a + 2] = 0.125 * fx * fy * ZETA_N[a];
  }
}

/* PH_Elem_C3D8_Jac (:863-885) + JacB (:887-963).  J(j,k) = sum_a dN_a/dxi_k x_ja,
 * detJ by cofactor expansion, J^-1 by cofactors / detJ.  The reference then
 * forms dNdx = J^-1 dNdxi (:936-943); the chain rule with this J needs
 * dNdx = J^-T dNdxi, which is what is computed here (identical whenever J is
 * symmetric, e.g. every axis-aligned box element).  Returns 0 when
 * |detJ| <= 1e-12 (the reference's skip threshold, :775). */
int orc_c3d8_dndx(const double* coords, double xi, double eta, double zeta,
                  double dNdx[24], double* detJ) {
  double N[8], dN[24], J[3][3] = {{0}};
  orc_c3d8_shape(xi, eta, zeta, N, dN);
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 8; ++a) J[j][k] += dN[3 * a + k] * coords[3 * a + j];
  double d = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  *detJ = d;
  if (fabs(d) <= 1.0e-12) {
    memset(dNdx, 0, 24 * sizeof(double));
    return 0;
  }
  double Ji[3][3];
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / d;
  Ji[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / d;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
  Ji[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) / d;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
  Ji[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / d;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / d;
  Ji[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / d;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
  for (int a = 0; a < 8; ++a)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += Ji[k][j] * dN[3 * a + k];
      dNdx[3 * a + j] = s;
    }
  return 1;
}

/* PH_Elem_C3D8_BMatrix, PH_Elem_C3D8.f90:550-573: rows [xx,yy,zz,gxy,gyz,gxz],
 * B is (6,24) column-major. */
static void bmatrix_n(const double* dNdx, int n, double* B) {
  memset(B, 0, 18 * n * sizeof(double));
  for (int a = 0; a < n; ++a) {
    const double gx = dNdx[3 * a], gy = dNdx[3 * a + 1], gz = dNdx[3 * a + 2];
    const int c = 3 * a;
    B[0 + 6 * c] = gx;
    B[1 + 6 * (c + 1)] = gy;
    B[2 + 6 * (c + 2)] = gz;
    B[3 + 6 * c] = gy;
    B[3 + 6 * (c + 1)] = gx;
    B[4 + 6 * (c + 1)] = gz;
    B[4 + 6 * (c + 2)] = gy;
    B[5 + 6 * c] = gz;
    B[5 + 6 * (c + 2)] = gx;
  }
}
static void bmatrix(const double dNdx[24], double B[144]) { bmatrix_n(dNdx, 8, B); }

/* PH_Elem_C3D8_ConstMatrix, PH_Elem_C3D8.f90:603-625 ((6,6) column-major) */
static void const_matrix(double E, double nu, double D[36]) {
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  memset(D, 0, 36 * sizeof(double));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D[i + 6 * j] = lam;
  for (int i = 0; i < 3; ++i) D[i + 6 * i] = lam + 2.0 * mu;
  for (int i = 3; i < 6; ++i) D[i + 6 * i] = mu;
}

/* Gauss point order of PH_Elem_C3D8_GaussPoints (:694-715): i (xi) outer,
 * j (eta), k (zeta) inner; all weights 1. */
static void gauss_point(int g, double* xi, double* eta, double* zeta) {
  *xi = (g & 4) ? GP : -GP;
  *eta = (g & 2) ? GP : -GP;
  *zeta = (g & 1) ? GP : -GP;
}

/* Ke = sum_gp detJ w (B^T D_gp) B, the dense form of PH_Elem_C3D8_StiffMatrix
 * (:759-780) with one (6,6) D per Gauss point. */
int orc_c3d8_ke_dgp(const double* coords, const double* Dgp, double* Ke) {
  double dNdx[24], B[144], BTD[144];
  int ok = 1;
  memset(Ke, 0, 576 * sizeof(double));
  for (int g = 0; g < 8; ++g) {
    double xi, eta, zeta, detJ;
    gauss_point(g, &xi, &eta, &zeta);
    if (!orc_c3d8_dndx(coords, xi, eta, zeta, dNdx, &detJ)) {
      ok = 0;
      continue;
    }
    const double dV = detJ * 1.0;
    const double* D = Dgp + 36 * g;
    bmatrix(dNdx, B);
    /* BTD(24,6) = B^T D */
    for (int i = 0; i < 24; ++i)
      for (int k = 0; k < 6; ++k) {
        double s = 0.0;
        for (int m = 0; m < 6; ++m) s += B[m + 6 * i] * D[m + 6 * k];
        BTD[i + 24 * k] = s;
      }
    for (int j = 0; j < 24; ++j)
      for (int i = 0; i < 24; ++i) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += BTD[i + 24 * k] * B[k + 6 * j];
        Ke[i + 24 * j] += dV * s;
      }
  }
  return ok;
}

/* PH_Elem_C3D8_StiffMatrix, PH_Elem_C3D8.f90:759-780 (isotropic elastic) */
int orc_c3d8_ke(const double* coords, double E, double nu, double* Ke) {
  double D[36], Dgp[288];
  const_matrix(E, nu, D);
  for (int g = 0; g < 8; ++g) memcpy(Dgp + 36 * g, D, sizeof D);
  return orc_c3d8_ke_dgp(coords, Dgp, Ke);
}

/* ---------------------------------------------------------------- variants
 * PH_Elem_C3D8_StiffMatrixByVariant (PH_Elem_C3D8.f90:400-417): 1 standard
 * (0 and 5, which the reference routes to the standard element, too),
 * 2 reduced (C3D8R), 3 selective (the reference's VARIANT_HYBRID routes to
 * PH_Elem_C3D8_StiffMatrixSelective), 4 incompatible modes, 6 27-point. */

/* Ke += w B^T D B for one (6,ncol) B column-major into Ke (ncol,ncol) */
static void add_btdb(const double* Ba, int na, const double* Bb, int nb, const double* D,
                     double w, double* K, int ld) {
  for (int j = 0; j < nb; ++j) {
    double db[6];
    for (int m = 0; m < 6; ++m) {
      double s = 0.0;
      for (int q = 0; q < 6; ++q) s += D[m + 6 * q] * Bb[q + 6 * j];
      db[m] = s;
    }
    for (int i = 0; i < na; ++i) {
      double s = 0.0;
      for (int m = 0; m < 6; ++m) s += Ba[m + 6 * i] * db[m];
      K[i + ld * j] += w * s;
    }
  }
}

/* J (j,k) = dx_j/dxi_k at (xi,eta,zeta), its determinant and inverse */
static double jacobian(const double* coords, double xi, double eta, double zeta,
                       double J[3][3], double Ji[3][3]) {
  double N[8], dN[24];
  orc_c3d8_shape(xi, eta, zeta, N, dN);
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) {
      J[j][k] = 0.0;
      for (int a = 0; a < 8; ++a) J[j][k] += dN[3 * a + k] * coords[3 * a + j];
    }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][0] * J[2][2] - J[1][2] * J[2][0];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double d = J[0][0] * c00 - J[0][1] * c01 + J[0][2] * c02;
  if (fabs(d) > 1.0e-20) {
    Ji[0][0] = c00 / d;
    Ji[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / d;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
    Ji[1][0] = -c01 / d;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
    Ji[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / d;
    Ji[2][0] = c02 / d;
    Ji[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / d;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
  } else {
    memset(Ji, 0, 9 * sizeof(double));
  }
  return d;
}

/* PH_Elem_C3D8_StiffMatrixReduced (:466-498): one point at the centre with
 * weight 8, plus k_hg sum_m gamma_m gamma_m^T (x) I3, k_hg = 0.1 mu V / V^(2/3),
 * V the 8-point volume (PH_ELEM_C3D8_Volume_8pt, :291-304).  The reference's
 * GAMMA8 (:123-127) holds xi.eta.zeta, eta.zeta, xi.zeta and, as its 4th
 * column, the linear mode -zeta in place of xi.eta; here the 4th vector is
 * xi.eta, and each base vector h is projected off the linear fields,
 * gamma = h - sum_i (h . x_i) b_i with b_i the centre gradients
 * (Flanagan & Belytschko 1981), which leaves h unchanged on parallelepipeds. */
static int ke_reduced(const double* coords, double E, double nu, double* Ke) {
  double D[36], g0[24], B[144], dj0;
  const_matrix(E, nu, D);
  memset(Ke, 0, 576 * sizeof(double));
  if (orc_c3d8_dndx(coords, 0.0, 0.0, 0.0, g0, &dj0)) {
    bmatrix(g0, B);
    add_btdb(B, 24, B, 24, D, 8.0 * dj0, Ke, 24);
  }
  double V = 0.0;
  for (int g = 0; g < 8; ++g) {
    double xi, eta, zeta, J[3][3], Ji[3][3];
    gauss_point(g, &xi, &eta, &zeta);
    V += jacobian(coords, xi, eta, zeta, J, Ji);
  }
  if (V <= 1.0e-20) return 1;
  const double mu = E / (2.0 * (1.0 + nu));
  const double L = cbrt(V);
  const double k = 0.1 * mu * V / (L * L);
  for (int m = 0; m < 4; ++m) {
    double h[8], gam[8], hx[3] = {0, 0, 0};
    for (int a = 0; a < 8; ++a) {
      const double x = XI_N[a], y = ETA_N[a], z = ZETA_N[a];
      h[a] = m == 0 ? -x * y * z : m == 1 ? y * z : m == 2 ? x * z : x * y;
    }
    for (int i = 0; i < 3; ++i)
      for (int a = 0; a < 8; ++a) hx[i] += h[a] * coords[3 * a + i];
    for (int a = 0; a < 8; ++a)
      gam[a] = h[a] - (hx[0] * g0[3 * a] + hx[1] * g0[3 * a + 1] + hx[2] * g0[3 * a + 2]);
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b)
        for (int d = 0; d < 3; ++d) Ke[(3 * a + d) + 24 * (3 * b + d)] += k * gam[a] * gam[b];
  }
  return 1;
}

/* PH_Elem_C3D8_StiffMatrixSelective (:500-540): B-bar, the volumetric rows
 * (B1+B2+B3)/3 replaced by their volume average over the element. */
static int ke_selective(const double* coords, double E, double nu, double* Ke) {
  double D[36], g[24], B[144], bvol[24], bavg[24] = {0}, V = 0.0;
  const_matrix(E, nu, D);
  memset(Ke, 0, 576 * sizeof(double));
  for (int p = 0; p < 8; ++p) {
    double xi, eta, zeta, dj;
    gauss_point(p, &xi, &eta, &zeta);
    if (!orc_c3d8_dndx(coords, xi, eta, zeta, g, &dj)) continue;
    bmatrix(g, B);
    V += dj;
    for (int c = 0; c < 24; ++c) bavg[c] += dj * (B[0 + 6 * c] + B[1 + 6 * c] + B[2 + 6 * c]) / 3.0;
  }
  if (V <= 1.0e-20) return 1;
  for (int c = 0; c < 24; ++c) bavg[c] /= V;
  for (int p = 0; p < 8; ++p) {
    double xi, eta, zeta, dj;
    gauss_point(p, &xi, &eta, &zeta);
    if (!orc_c3d8_dndx(coords, xi, eta, zeta, g, &dj)) continue;
    bmatrix(g, B);
    for (int c = 0; c < 24; ++c) {
      bvol[c] = (B[0 + 6 * c] + B[1 + 6 * c] + B[2 + 6 * c]) / 3.0;
      for (int r = 0; r < 3; ++r) B[r + 6 * c] += bavg[c] - bvol[c];
    }
    add_btdb(B, 24, B, 24, D, dj, Ke, 24);
  }
  return 1;
}

/* Incompatible modes.  The reference's PH_Elem_C3D8_StiffMatrixIncompatible
 * (:419-464) enriches u by three scalar modes along (1,1,1), two of which
 * (xi(1-eta^2)(1-zeta^2), eta(1-xi^2)(1-zeta^2)) have a non-zero mean
 * gradient, so the condensed element fails the patch test (2% soft in
 * uniform strain on the unit cube).  Here: the incompatible modes
 * P_k = 1 - xi_k^2 for each displacement component (9 modes, Wilson et al.
 * 1973) with gradients J0^-T dP/dxi detJ0/detJ (J0 at the centre; Taylor,
 * Beresford & Wilson 1976), whose integral over the element vanishes, and
 * Ke = Kuu - Kua Kaa^-1 Kua^T. */
static int ke_incompat(const double* coords, double E, double nu, double* Ke) {
  double D[36], g[24], B[144], Be[54], J0[3][3], J0i[3][3];
  double Kua[24 * 9] = {0}, Kaa[81] = {0};
  const_matrix(E, nu, D);
  memset(Ke, 0, 576 * sizeof(double));
  const double dj0 = jacobian(coords, 0.0, 0.0, 0.0, J0, J0i);
  for (int p = 0; p < 8; ++p) {
    double xc[3], dj;
    gauss_point(p, &xc[0], &xc[1], &xc[2]);
    if (!orc_c3d8_dndx(coords, xc[0], xc[1], xc[2], g, &dj)) continue;
    bmatrix(g, B);
    double ge[9];  /* gradient of mode k: ge[3k + j] */
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j) ge[3 * k + j] = J0i[k][j] * (-2.0 * xc[k]) * dj0 / dj;
    bmatrix_n(ge, 3, Be);
    add_btdb(B, 24, B, 24, D, dj, Ke, 24);
    add_btdb(B, 24, Be, 9, D, dj, Kua, 24);
    add_btdb(Be, 9, Be, 9, D, dj, Kaa, 9);
  }
  /* Kaa = L L^T, W = L^-1 Kua^T (9,24); Ke -= W^T W */
  double L[81] = {0};
  for (int j = 0; j < 9; ++j) {
    double s = Kaa[j + 9 * j];
    for (int k = 0; k < j; ++k) s -= L[j + 9 * k] * L[j + 9 * k];
    if (s <= 0.0) return 0;
    L[j + 9 * j] = sqrt(s);
    for (int i = j + 1; i < 9; ++i) {
      double t = Kaa[i + 9 * j];
      for (int k = 0; k < j; ++k) t -= L[i + 9 * k] * L[j + 9 * k];
      L[i + 9 * j] = t / L[j + 9 * j];
    }
  }
  double W[9 * 24];
  for (int c = 0; c < 24; ++c)
    for (int i = 0; i < 9; ++i) {
      double t = Kua[c + 24 * i];
      for (int k = 0; k < i; ++k) t -= L[i + 9 * k] * W[k + 9 * c];
      W[i + 9 * c] = t / L[i + 9 * i];
    }
  for (int b = 0; b < 24; ++b)
    for (int a = 0; a < 24; ++a) {
      double s = 0.0;
      for (int k = 0; k < 9; ++k) s += W[k + 9 * a] * W[k + 9 * b];
      Ke[a + 24 * b] -= s;
    }
  return 1;
}

/* PH_Elem_C3D8_StiffMatrix27 (:1976-1995): 3x3x3 Gauss, points GP3 =
 * -/+0.774596669241483 and 0 (the reference's literal, :130), weights 5/9,
 * 8/9, 5/9, order of PH_Elem_C3D8_GaussPoints27 (:717-732). */
static int ke_gauss27(const double* coords, double E, double nu, double* Ke) {
  static const double P3[3] = {-0.774596669241483, 0.0, 0.774596669241483};
  const double W3[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  double D[36], g[24], B[144];
  const_matrix(E, nu, D);
  memset(Ke, 0, 576 * sizeof(double));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        double dj;
        if (!orc_c3d8_dndx(coords, P3[i], P3[j], P3[k], g, &dj)) continue;
        bmatrix(g, B);
        add_btdb(B, 24, B, 24, D, dj * W3[i] * W3[j] * W3[k], Ke, 24);
      }
  return 1;
}

int orc_c3d8_ke_variant(const double* coords, double E, double nu, int variant, double* Ke) {
  switch (variant) {
    case 0:
    case 1:
    case 5:
      return orc_c3d8_ke(coords, E, nu, Ke);
    case 2:
      return ke_reduced(coords, E, nu, Ke);
    case 3:
      return ke_selective(coords, E, nu, Ke);
    case 4:
      return ke_incompat(coords, E, nu, Ke);
    case 6:
      return ke_gauss27(coords, E, nu, Ke);
    default:
      return -1;
  }
}

/* strain at each Gauss point, eps = B u_e (PH_Elem_C3D8_Strain as used by
 * PH_Elem_C3D8_IntForce, :734-757); eps_gp is (6,8) column-major */
void orc_c3d8_strain(const double* coords, const double* ue, double* eps_gp) {
  double dNdx[24], B[144];
  for (int g = 0; g < 8; ++g) {
    double xi, eta, zeta, detJ;
    gauss_point(g, &xi, &eta, &zeta);
    orc_c3d8_dndx(coords, xi, eta, zeta, dNdx, &detJ);
    bmatrix(dNdx, B);
    for (int m = 0; m < 6; ++m) {
      double s = 0.0;
      for (int j = 0; j < 24; ++j) s += B[m + 6 * j] * ue[j];
      eps_gp[6 * g + m] = s;
    }
  }
}

/* Fe = sum_gp detJ w B^T sigma_gp, PH_Elem_C3D8_FormIntForceFromStress (:782-798)
 * with one stress per Gauss point */
void orc_c3d8_fint(const double* coords, const double* sig_gp, double* Fe) {
  double dNdx[24], B[144];
  memset(Fe, 0, 24 * sizeof(double));
  for (int g = 0; g < 8; ++g) {
    double xi, eta, zeta, detJ;
    gauss_point(g, &xi, &eta, &zeta);
    if (!orc_c3d8_dndx(coords, xi, eta, zeta, dNdx, &detJ)) continue;
    bmatrix(dNdx, B);
    for (int j = 0; j < 24; ++j) {
      double s = 0.0;
      for (int m = 0; m < 6; ++m) s += B[m + 6 * j] * sig_gp[6 * g + m];
      Fe[j] += detJ * s;
    }
  }
}

/* Consistent face load F = int_face N p n_out dA, 2x2 Gauss on the face,
 * PH_Elem_C3D8_FormFacePressure (:2360-2541) with the 1/8 shape functions.
 * Positive p acts along the outward normal (the reference's sign). */
void orc_c3d8_face_pressure(const double* coords, int face, double p, double* Fe) {
  /* which natural coordinate is fixed on each face, its value, the two
   * in-face coordinates (u,v) and the orientation sign of d/du x d/dv */
  static const int fixed[6] = {2, 2, 1, 1, 0, 0};
  static const double fval[6] = {-1, 1, -1, 1, -1, 1};
  static const int cu[6] = {0, 0, 0, 0, 1, 1};
  static const int cv[6] = {1, 1, 2, 2, 2, 2};
  static const double sgn[6] = {-1, 1, -1, 1, -1, 1};
  static const double fu[4] = {-GP, GP, -GP, GP};
  static const double fv[4] = {-GP, -GP, GP, GP};
  memset(Fe, 0, 24 * sizeof(double));
  for (int ip = 0; ip < 4; ++ip) {
    double nat[3];
    nat[fixed[face]] = fval[face];
    nat[cu[face]] = fu[ip];
    nat[cv[face]] = fv[ip];
    double N[8], dN[24], du[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
    orc_c3d8_shape(nat[0], nat[1], nat[2], N, dN);
    for (int q = 0; q < 4; ++q) {
      int a = FACE_NODES[face][q];
      for (int d = 0; d < 3; ++d) {
        du[d] += dN[3 * a + cu[face]] * coords[3 * a + d];
        dv[d] += dN[3 * a + cv[face]] * coords[3 * a + d];
      }
    }
    double n[3] = {du[1] * dv[2] - du[2] * dv[1], du[2] * dv[0] - du[0] * dv[2],
                   du[0] * dv[1] - du[1] * dv[0]};
    double dA = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (dA < 1.0e-15) continue;
    for (int d = 0; d < 3; ++d) n[d] = sgn[face] * n[d] / dA;
    for (int q = 0; q < 4; ++q) {
      int a = FACE_NODES[face][q];
      for (int d = 0; d < 3; ++d) Fe[3 * a + d] += N[a] * p * n[d] * dA;
    }
  }
}

/* area of a (flat or warped) face by the same 2x2 rule */
double orc_face_area(const double* coords, int face) {
  double A = 0.0;
  static const int fixed[6] = {2, 2, 1, 1, 0, 0};
  static const double fval[6] = {-1, 1, -1, 1, -1, 1};
  static const int cu[6] = {0, 0, 0, 0, 1, 1};
  static const int cv[6] = {1, 1, 2, 2, 2, 2};
  static const double fu[4] = {-GP, GP, -GP, GP};
  static const double fv[4] = {-GP, -GP, GP, GP};
  for (int ip = 0; ip < 4; ++ip) {
    double nat[3];
    nat[fixed[face]] = fval[face];
    nat[cu[face]] = fu[ip];
    nat[cv[face]] = fv[ip];
    double N[8], dN[24], du[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
    orc_c3d8_shape(nat[0], nat[1], nat[2], N, dN);
    for (int q = 0; q < 4; ++q) {
      int a = FACE_NODES[face][q];
      for (int d = 0; d < 3; ++d) {
        du[d] += dN[3 * a + cu[face]] * coords[3 * a + d];
        dv[d] += dN[3 * a + cv[face]] * coords[3 * a + d];
      }
    }
    double n0 = du[1] * dv[2] - du[2] * dv[1], n1 = du[2] * dv[0] - du[0] * dv[2],
           n2 = du[0] * dv[1] - du[1] * dv[0];
    A += sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  }
  return A;
}

/* ========================================================================= */
/* CSR pattern and assembly.                                                 */
/* ========================================================================= */
static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Structural pattern of the merged, column-sorted CSR of RT_CSR_FromTripletMerged
 * (RT_Solv_Sparse.f90:336-447) for C3D8 connectivity: row 3n+d holds columns
 * 3m+e for every node m sharing an element with n (itself included), m and e
 * ascending.  Built from node adjacency rather than a sort of all 576 triplets
 * per element; orc_csr_from_triplets is the literal algorithm and the tests
 * check that both agree.  With col_idx == NULL only row_ptr and nnz are made. */
int64_t orc_csr_pattern(int32_t n_nodes, int32_t n_elems, const int32_t* conn,
                        int32_t* row_ptr, int32_t* col_idx) {
  /* node -> element incidence */
  int64_t* ne_ptr = (int64_t*)calloc((size_t)n_nodes + 1, sizeof(int64_t));
  int32_t* ne = (int32_t*)malloc(sizeof(int32_t) * (size_t)n_elems * 8);
  for (int64_t e = 0; e < n_elems; ++e)
    for (int a = 0; a < 8; ++a) ne_ptr[conn[8 * e + a] + 1]++;
  for (int64_t n = 0; n < n_nodes; ++n) ne_ptr[n + 1] += ne_ptr[n];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)n_nodes);
  for (int64_t n = 0; n < n_nodes; ++n) fill[n] = ne_ptr[n];
  for (int64_t e = 0; e < n_elems; ++e)
    for (int a = 0; a < 8; ++a) {
      int32_t n = conn[8 * e + a];
      ne[fill[n]++] = (int32_t)e;
    }
  free(fill);
  int64_t nnz = 0;
  row_ptr[0] = 0;
  int32_t nb[512];
  for (int64_t n = 0; n < n_nodes; ++n) {
    int cnt = 0;
    for (int64_t k = ne_ptr[n]; k < ne_ptr[n + 1]; ++k) {
      int64_t e = ne[k];
      for (int a = 0; a < 8; ++a) {
        int32_t m = conn[8 * e + a];
        int dup = 0;
        for (int q = 0; q < cnt; ++q)
          if (nb[q] == m) { dup = 1; break; }
        if (!dup && cnt < 512) nb[cnt++] = m;
      }
    }
    qsort(nb, (size_t)cnt, sizeof(int32_t), cmp_i32);
    for (int d = 0; d < 3; ++d) {
      if (col_idx)
        for (int q = 0; q < cnt; ++q)
          for (int c = 0; c < 3; ++c) col_idx[nnz + 3 * q + c] = 3 * nb[q] + c;
      nnz += 3 * cnt;
      row_ptr[3 * n + d + 1] = (int32_t)nnz;
    }
  }
  free(ne_ptr);
  free(ne);
  return nnz;
}

typedef struct { int32_t i, j; int64_t k; } trip_t;
static int cmp_trip(const void* a, const void* b) {
  const trip_t* x = (const trip_t*)a;
  const trip_t* y = (const trip_t*)b;
  if (x->i != y->i) return (x->i > y->i) - (x->i < y->i);
  if (x->j != y->j) return (x->j > y->j) - (x->j < y->j);
  return (x->k > y->k) - (x->k < y->k);
}

/* RT_CSR_FromTripletMerged (RT_Solv_Sparse.f90:336-447): order triplets by
 * (row, col), sum duplicates (in insertion order here), 0-based output.
 * Returns nnz; with col_idx == NULL only counts. */
int64_t orc_csr_from_triplets(int32_t n_rows, int64_t n_trip, const int32_t* ti,
                              const int32_t* tj, const double* tv,
                              int32_t* row_ptr, int32_t* col_idx, double* val) {
  trip_t* t = (trip_t*)malloc(sizeof(trip_t) * (size_t)(n_trip > 0 ? n_trip : 1));
  for (int64_t k = 0; k < n_trip; ++k) {
    t[k].i = ti[k];
    t[k].j = tj[k];
    t[k].k = k;
  }
  qsort(t, (size_t)n_trip, sizeof(trip_t), cmp_trip);
  int64_t nnz = 0;
  for (int32_t r = 0; r <= n_rows; ++r) row_ptr[r] = 0;
  for (int64_t k = 0; k < n_trip;) {
    int64_t k2 = k;
    double s = 0.0;
    while (k2 < n_trip && t[k2].i == t[k].i && t[k2].j == t[k].j) {
      if (tv) s += tv[t[k2].k];
      ++k2;
    }
    if (col_idx) col_idx[nnz] = t[k].j;
    if (val) val[nnz] = s;
    row_ptr[t[k].i + 1]++;
    ++nnz;
    k = k2;
  }
  for (int32_t r = 0; r < n_rows; ++r) row_ptr[r + 1] += row_ptr[r];
  free(t);
  return nnz;
}

static int64_t find_slot(const int32_t* row_ptr, const int32_t* col_idx, int32_t r, int32_t c) {
  int64_t lo = row_ptr[r], hi = row_ptr[r + 1] - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    if (col_idx[mid] == c) return mid;
    if (col_idx[mid] < c) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

static void gather_coords(const int32_t* conn, const double* coords, int64_t e, double* xe) {
  for (int a = 0; a < 8; ++a) {
    int64_t n = conn[8 * e + a];
    xe[3 * a] = coords[3 * n];
    xe[3 * a + 1] = coords[3 * n + 1];
    xe[3 * a + 2] = coords[3 * n + 2];
  }
}

/* RT_Asm_GlobalStiffness reuse path (RT_Asm_Solv.f90:1840-1900) with
 * RT_Asm_ScatterKe_CSR_Atomic (:3467-3493): K%values = 0, element loop,
 * column search in the row, add.  nthreads == 1 is the serial, element-ordered
 * sum; nthreads > 1 uses OpenMP dynamic scheduling with atomic adds like the
 * reference (summation order then varies run to run). */
int orc_assemble(int32_t n_elems, const int32_t* conn, const double* coords,
                 const double* E, double nu, int32_t n_rows, const int32_t* row_ptr,
                 const int32_t* col_idx, double* val, int nthreads, int32_t* n_degenerate) {
  int64_t nnz = row_ptr[n_rows];
  memset(val, 0, sizeof(double) * (size_t)nnz);
  int bad = 0, degen = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : bad, degen)
#endif
  for (int64_t e = 0; e < n_elems; ++e) {
    double xe[24], Ke[576];
    gather_coords(conn, coords, e, xe);
    if (!orc_c3d8_ke(xe, E[e], nu, Ke)) degen++;
    for (int j = 0; j < 24; ++j) {
      int32_t cj = 3 * conn[8 * e + j / 3] + j % 3;
      for (int i = 0; i < 24; ++i) {
        int32_t ri = 3 * conn[8 * e + i / 3] + i % 3;
        int64_t s = find_slot(row_ptr, col_idx, ri, cj);
        if (s < 0) { bad++; continue; }
        if (nthreads > 1) {
#ifdef _OPENMP
#pragma omp atomic
#endif
          val[s] += Ke[i + 24 * j];
        } else {
          val[s] += Ke[i + 24 * j];
        }
      }
    }
  }
  if (n_degenerate) *n_degenerate = degen;
  return bad;
}

/* the same with one D per Gauss point: Dgp is (6,6,8,n_elems); nthreads as
 * orc_assemble */
int orc_assemble_dgp(int32_t n_elems, const int32_t* conn, const double* coords,
                     const double* Dgp, int32_t n_rows, const int32_t* row_ptr,
                     const int32_t* col_idx, double* val, int nthreads) {
  int64_t nnz = row_ptr[n_rows];
  memset(val, 0, sizeof(double) * (size_t)nnz);
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : bad)
#endif
  for (int64_t e = 0; e < n_elems; ++e) {
    double xe[24], Ke[576];
    gather_coords(conn, coords, e, xe);
    if (!orc_c3d8_ke_dgp(xe, Dgp + 288 * e, Ke)) bad++;
    for (int j = 0; j < 24; ++j) {
      int32_t cj = 3 * conn[8 * e + j / 3] + j % 3;
      for (int i = 0; i < 24; ++i) {
        int32_t ri = 3 * conn[8 * e + i / 3] + i % 3;
        int64_t s = find_slot(row_ptr, col_idx, ri, cj);
        if (s < 0) { bad++; continue; }
        if (nthreads > 1) {
#ifdef _OPENMP
#pragma omp atomic
#endif
          val[s] += Ke[i + 24 * j];
        } else {
          val[s] += Ke[i + 24 * j];
        }
      }
    }
  }
  return bad;
}

/* RT_Asm_ApplyBC _Idx path, RT_Asm_Solv.f90:318-344:
 * K(eq,eq) += penalty, F(eq) += penalty * value.  Returns #eq without a
 * diagonal slot (those only get the load, as in the reference). */
int orc_apply_penalty(int32_t n_bc, const int32_t* eq, const double* value, double penalty,
                      const int32_t* row_ptr, const int32_t* col_idx, double* val, double* F) {
  int missing = 0;
  for (int32_t i = 0; i < n_bc; ++i) {
    int64_t s = find_slot(row_ptr, col_idx, eq[i], eq[i]);
    if (s >= 0 && val) val[s] += penalty; else if (s < 0) missing++;
    if (F) F[eq[i]] += penalty * (value ? value[i] : 0.0);
  }
  return missing;
}

/* LoadBC_DistributeLoad_ToSurface, MD_LBC_Apply.f90:258-293: per facet the
 * force traction * area / n_face_nodes goes to each face node. */
void orc_face_loads_traction(int32_t n_faces, const int32_t* elem, const int32_t* face,
                             const int32_t* conn, const double* coords,
                             const double* traction, double* F) {
  for (int32_t f = 0; f < n_faces; ++f) {
    double xe[24];
    gather_coords(conn, coords, elem[f], xe);
    double area = orc_face_area(xe, face[f]);
    for (int q = 0; q < 4; ++q) {
      int32_t n = conn[8 * (int64_t)elem[f] + FACE_NODES[face[f]][q]];
      for (int d = 0; d < 3; ++d) F[3 * (int64_t)n + d] += traction[d] * (area / 4.0);
    }
  }
}

/* element-consistent face pressure (PH_Elem_C3D8_FormFacePressure) summed into F */
void orc_face_loads_pressure(int32_t n_faces, const int32_t* elem, const int32_t* face,
                             const int32_t* conn, const double* coords, double p, double* F) {
  for (int32_t f = 0; f < n_faces; ++f) {
    double xe[24], Fe[24];
    gather_coords(conn, coords, elem[f], xe);
    orc_c3d8_face_pressure(xe, face[f], p, Fe);
    for (int a = 0; a < 8; ++a) {
      int64_t n = conn[8 * (int64_t)elem[f] + a];
      for (int d = 0; d < 3; ++d) F[3 * n + d] += Fe[3 * a + d];
    }
  }
}

/* ========================================================================= */
/* SpMV and Krylov solvers.                                                   */
/* ========================================================================= */

/* NM_CSR_MatVec_Optimized, NM_Mtx_Sparse.f90:341-365: row loop, storage order */
void orc_spmv(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
              const double* x, double* y, int nthreads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int32_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) s += val[p] * x[col_idx[p]];
    y[i] = s;
  }
}

/* y = A^T x for a square CSR A: the product NM_CSR_Transpose + SpMV forms
 * (NM_Mtx_SpMVCSRTranspose.f90:46-145), as a scatter over the rows of A */
void orc_spmv_t(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
                const double* x, double* y) {
  for (int32_t i = 0; i < n; ++i) y[i] = 0.0;
  for (int32_t i = 0; i < n; ++i)
    for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) y[col_idx[p]] += val[p] * x[i];
}

static double dot(int32_t n, const double* a, const double* b, int nthreads) {
  double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : s)
#endif
  for (int32_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* NM_Solver_PCG, NM_Solv_Core.f90:363-423, with NM_Solver_Jacobi_Precond
 * (:131-147): diag taken from the CSR diagonal, z = r/d when |d| > 1e-30
 * else z = r; stop at the first k with ||r||/||b|| < rtol (||b|| < 1e-30 ->
 * 1).  Returns 0 (IF_STATUS_OK) when converged, 1 (IF_STATUS_INVALID, the
 * reference's non-convergence code) otherwise, 2 on p.Ap breakdown. */
static double wall_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1.0e-9 * (double)ts.tv_nsec;
}

/* The PCG loop of orc_pcg.  rtol_mark > 0: the first iteration whose
 * ||r||/||b|| falls below rtol_mark, and the wall time from entry to the end
 * of that iteration, go to *iter_mark / *sec_mark (the iteration continues to
 * rtol). */
static int pcg_core(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
                    const double* b, double* x, int32_t max_iter, double rtol,
                    int32_t* niter, double* rnorm, int nthreads, double rtol_mark,
                    int32_t* iter_mark, double* sec_mark, int32_t report_every) {
  const double t_start = wall_seconds();
  if (iter_mark) *iter_mark = -1;
  if (sec_mark) *sec_mark = -1.0;
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* Ap = (double*)malloc(sizeof(double) * (size_t)n);
  double* dg = (double*)malloc(sizeof(double) * (size_t)n);
  int nt = nthreads > 0 ? nthreads : 1;
  (void)nt;
  for (int32_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int32_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q)
      if (col_idx[q] == i) d = val[q];
    dg[i] = d;
  }
  orc_spmv(n, row_ptr, col_idx, val, x, Ap, nthreads);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (int32_t i = 0; i < n; ++i) {
    r[i] = b[i] - Ap[i];
    z[i] = fabs(dg[i]) > 1.0e-30 ? r[i] / dg[i] : r[i];
    p[i] = z[i];
  }
  double rz_old = dot(n, r, z, nthreads);
  double bn = sqrt(dot(n, b, b, nthreads));
  if (bn < 1.0e-30) bn = 1.0;
  int status = 1;
  int32_t k;
  double rn = 0.0;
  for (k = 1; k <= max_iter; ++k) {
    orc_spmv(n, row_ptr, col_idx, val, p, Ap, nthreads);
    double pAp = dot(n, p, Ap, nthreads);
    if (!(pAp > 0.0)) { status = 2; break; }
    double alpha = rz_old / pAp;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    rn = sqrt(dot(n, r, r, nthreads)) / bn;
    if (rtol_mark > 0.0 && rn < rtol_mark && iter_mark && *iter_mark < 0) {
      *iter_mark = k;
      if (sec_mark) *sec_mark = wall_seconds() - t_start;
    }
    if (report_every > 0 && k % report_every == 0) {
      fprintf(stderr, "[oracle pcg] iteration %d ||r||/||b|| %.3e %.1f s\n", (int)k, rn,
              wall_seconds() - t_start);
      fflush(stderr);
    }
    if (rn < rtol) { status = 0; break; }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int32_t i = 0; i < n; ++i) z[i] = fabs(dg[i]) > 1.0e-30 ? r[i] / dg[i] : r[i];
    double rz_new = dot(n, r, z, nthreads);
    double beta = rz_new / rz_old;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz_old = rz_new;
  }
  if (niter) *niter = (k > max_iter) ? max_iter : k;
  if (rnorm) *rnorm = rn;
  free(r); free(z); free(p); free(Ap); free(dg);
  return status;
}

int orc_pcg(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
            const double* b, double* x, int32_t max_iter, double rtol,
            int32_t* niter, double* rnorm, int nthreads) {
  return pcg_core(n, row_ptr, col_idx, val, b, x, max_iter, rtol, niter, rnorm, nthreads, 0.0,
                  NULL, NULL, 0);
}

/* orc_pcg that also reports when ||r||/||b|| first fell below rtol_mark
 * (iteration and wall seconds from entry): one run to a tight rtol gives the
 * solution for error checks and the time to the benchmark's tolerance.
 * report_every > 0 prints a progress line to stderr every that many
 * iterations (long runs). */
int orc_pcg_mark(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
                 const double* b, double* x, int32_t max_iter, double rtol, double rtol_mark,
                 int32_t* niter, double* rnorm, int32_t* iter_mark, double* sec_mark,
                 int nthreads, int32_t report_every) {
  return pcg_core(n, row_ptr, col_idx, val, b, x, max_iter, rtol, niter, rnorm, nthreads,
                  rtol_mark, iter_mark, sec_mark, report_every);
}

/* NM_CG_Solve_Transpose, NM_Solv_GMRESTranspose.f90:118-224: unpreconditioned
 * CG for symmetric K, lambda_0 taken as 0 (r = rhs), stop when
 * ||r||/||r_0|| < tol, beta = (r.r)/((r+alpha Ap).(r+alpha Ap)).  lam is
 * overwritten (the reference adds to the incoming lambda while assuming it is
 * zero; callers pass zero).  Returns 0 converged, 6 (IF_STATUS_WARN) not. */
int orc_cg_transpose(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                     const double* val, const double* rhs, double* lam, double tol,
                     int32_t max_iter, int32_t* niter, double* rel) {
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* Ap = (double*)malloc(sizeof(double) * (size_t)n);
  for (int32_t i = 0; i < n; ++i) { r[i] = rhs[i]; p[i] = r[i]; lam[i] = 0.0; }
  double r0 = sqrt(dot(n, r, r, 1));
  int status = 6;
  if (niter) *niter = 0;
  if (rel) *rel = 0.0;
  if (r0 < 2.220446049250313e-16) {
    free(r); free(p); free(Ap);
    return 0;
  }
  for (int32_t it = 1; it <= max_iter; ++it) {
    orc_spmv(n, row_ptr, col_idx, val, p, Ap, 1);
    double rr = dot(n, r, r, 1);
    double alpha = rr / dot(n, p, Ap, 1);
    for (int32_t i = 0; i < n; ++i) {
      lam[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    double rn = sqrt(dot(n, r, r, 1));
    if (niter) *niter = it;
    if (rel) *rel = rn / r0;
    if (rn / r0 < tol) { status = 0; break; }
    double num = dot(n, r, r, 1), den = 0.0;
    for (int32_t i = 0; i < n; ++i) {
      double t = r[i] + alpha * Ap[i];
      den += t * t;
    }
    double beta = num / den;
    for (int32_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
  }
  free(r); free(p); free(Ap);
  return status;
}

/* ========================================================================= */
/* Adjoint sensitivity: g_e = -dE_e/drho_e * lam_e^T K_e(E=1) u_e.  No such   */
/* kernel exists in the reference; this is its definition.                  */
/* ========================================================================= */
void orc_sensitivity(int32_t n_elems, const int32_t* conn, const double* coords, double nu,
                     const double* dEdrho, const double* u, const double* lam, double* g,
                     int nthreads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t e = 0; e < n_elems; ++e) {
    double xe[24], Ke[576], ue[24], le[24];
    gather_coords(conn, coords, e, xe);
    orc_c3d8_ke(xe, 1.0, nu, Ke);
    for (int j = 0; j < 24; ++j) {
      int64_t q = 3 * (int64_t)conn[8 * e + j / 3] + j % 3;
      ue[j] = u[q];
      le[j] = lam[q];
    }
    double s = 0.0;
    /* s = lam^T (Ke u) */
    for (int i = 0; i < 24; ++i) {
      double t = 0.0;
      for (int j = 0; j < 24; ++j) t += Ke[i + 24 * j] * ue[j];
      s += le[i] * t;
    }
    g[e] = -dEdrho[e] * s;
  }
}

/* ========================================================================= */
/* J2 plasticity with isotropic hardening.                                   */
/* props = [E, nu, sigma_y0, H, type, max_local_iter, K_swift, n_swift,      */
/*          eps0_swift, sigma_inf, delta_voce]; a shorter array from the     */
/* Python front end is zero-padded (type 0 = linear, max_local_iter 0 = 25). */
/* Hardening PH_J2_ComputeHardening / _Tangent (PH_Mat_Plast_J2_Iso_Core.f90 */
/* :549-605): 1 linear sigma_y0 + H e, 2 Swift K (eps0 + e)^n, 3 Voce        */
/* sigma_y0 + Q (1 - exp(-delta e)).  Stress update per                      */
/* PH_J2_ComputeStress_Core (:146-207): trial stress (:254-317), yield check */
/* against sigma_y(peeq_n), radial return with the local Newton of :355-456  */
/* (dg0 = f / (3G + H'(peeq_n)); R = q - 3G dg - sigma_y(peeq_n + dg);       */
/* dg -= R / -(3G + H'(peeq_n + dg)), clipped at 0; |R| < 1e-10 sigma_y0;     */
/* at most max_local_iter iterations).  The tangent is the textbook          */
/* consistent tangent (Simo & Hughes 1998 sec. 3.3; de Souza Neto et al.     */
/* 2008 eq. 7.120) in engineering-shear Voigt form with H' at the updated    */
/* peeq:                                                                     */
/*   Dt = K 1x1 + 2G th1 (I_s - 1x1/3) - 2G (th1 + th2 - 1) n x n,           */
/*   th1 = 1 - 3G dg/q_tr, th2 = 3G/(3G+H'), I_s = diag(1,1,1,1/2,1/2,1/2),  */
/* replacing the reference's :509-537, which is not the consistent tangent.  */
/* Returns 0 OK, 2 (IF_STATUS_ERROR) for a degenerate trial state, a zero    */
/* Jacobian or a local Newton that did not converge (*local_iters, when not  */
/* NULL, gets the iterations the local Newton used).                         */
/* ========================================================================= */
static double j2_yield(const double* pr, double e) {
  const int type = (int)pr[4];
  if (type == 2) return pr[6] * pow(pr[8] + e, pr[7]);
  if (type == 3) return pr[2] + pr[9] * (1.0 - exp(-pr[10] * e));
  return pr[2] + pr[3] * e;
}

static double j2_hard_tan(const double* pr, double e) {
  const int type = (int)pr[4];
  if (type == 2) return pr[6] * pr[7] * pow(pr[8] + e, pr[7] - 1.0);
  if (type == 3) return pr[9] * pr[10] * exp(-pr[10] * e);
  return pr[3];
}

/* Dt = K 1x1 + 2G th1 (I_s - 1x1/3) - 2G (th1 + th2 - 1) n x n (column-major) */
static void j2_tangent(double K, double G, double th1, double th2, const double* n, double* Dt) {
  const double cn = 2.0 * G * (th1 + th2 - 1.0);
  for (int j = 0; j < 6; ++j)
    for (int i = 0; i < 6; ++i) {
      double v = 0.0;
      if (i < 3 && j < 3) v += K - 2.0 * G * th1 / 3.0;
      if (i == j) v += 2.0 * G * th1 * (i < 3 ? 1.0 : 0.5);
      v -= cn * n[i] * n[j];
      Dt[i + 6 * j] = v;
    }
}

int orc_j2_update_ex(const double* props, const double* deps, double* sigma, double* peeq,
                     double* eps_p, double* Dt, int32_t* yielded, int32_t* local_iters) {
  const double E = props[0], nu = props[1], sy0 = props[2];
  const double mu = E / (2.0 * (1.0 + nu));
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  if (local_iters) *local_iters = 0;
  double De[36];
  memset(De, 0, sizeof De);
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) De[i + 6 * j] = lam;
    De[i + 6 * i] = lam + 2.0 * mu;
  }
  for (int i = 3; i < 6; ++i) De[i + 6 * i] = mu;
  double st[6];
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
    for (int j = 0; j < 6; ++j) s += De[i + 6 * j] * deps[j];
    st[i] = sigma[i] + s;
  }
  const double pm = (st[0] + st[1] + st[2]) / 3.0;
  double s[6];
  for (int i = 0; i < 6; ++i) s[i] = st[i];
  s[0] -= pm; s[1] -= pm; s[2] -= pm;
  double q = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] +
                  2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
  q = sqrt(1.5) * q;
  const double peeq_n = *peeq;
  const double f = q - j2_yield(props, peeq_n);
  const double G = mu, K = E / (3.0 * (1.0 - 2.0 * nu));
  if (f <= 0.0) {
    for (int i = 0; i < 6; ++i) sigma[i] = st[i];
    if (yielded) *yielded = 0;
    if (!Dt) return 0;
    if (peeq_n > 0.0 && f > -1.0e-8 * sy0 && q >= 1.0e-30) {
      /* on the yield surface of a point that has yielded (the strain
       * increment of the first Newton iteration of an increment is zero):
       * the plastic branch's tangent at dg = 0, th1 = 1, th2 = 3G/(3G+H'),
       * the direction a loading step takes (D_el would make that iteration
       * an elastic predictor) */
      double n0[6];
      for (int i = 0; i < 6; ++i) n0[i] = s[i] / (q / sqrt(1.5));
      j2_tangent(K, G, 1.0, 3.0 * G / (3.0 * G + j2_hard_tan(props, peeq_n)), n0, Dt);
    } else {
      memcpy(Dt, De, sizeof De);
    }
    return 0;
  }
  if (q < 1.0e-30) return 2;
  double n[6];
  for (int i = 0; i < 6; ++i) n[i] = s[i] / (q / sqrt(1.5));
  double dg = f / (3.0 * G + j2_hard_tan(props, peeq_n));
  const int max_it = props[5] > 0.0 ? (int)props[5] : 25;
  int it;
  for (it = 1; it <= max_it; ++it) {
    double R = q - 3.0 * G * dg - j2_yield(props, peeq_n + dg);
    if (fabs(R) < 1.0e-10 * sy0) break;
    double dR = -(3.0 * G + j2_hard_tan(props, peeq_n + dg));
    if (fabs(dR) < 1.0e-30) return 2;
    dg = dg - R / dR;
    if (dg < 0.0) dg = 0.0;
  }
  if (local_iters) *local_iters = it;
  if (it > max_it) return 2;
  const double beta = 1.0 - 3.0 * G * dg / q;
  sigma[0] = beta * s[0] + pm;
  sigma[1] = beta * s[1] + pm;
  sigma[2] = beta * s[2] + pm;
  sigma[3] = beta * s[3];
  sigma[4] = beta * s[4];
  sigma[5] = beta * s[5];
  *peeq = peeq_n + dg;
  for (int i = 0; i < 6; ++i) eps_p[i] += dg * n[i];
  if (yielded) *yielded = 1;
  if (Dt) j2_tangent(K, G, beta, 3.0 * G / (3.0 * G + j2_hard_tan(props, *peeq)), n, Dt);
  return 0;
}

int orc_j2_update(const double* props, const double* deps, double* sigma, double* peeq,
                  double* eps_p, double* Dt, int32_t* yielded) {
  return orc_j2_update_ex(props, deps, sigma, peeq, eps_p, Dt, yielded, NULL);
}

/* Strain increments B(u - u_n) at every Gauss point, the J2 update from the
 * committed state state_n into state (with the tangent into Dgp,
 * (6,6,8,n_elems)), and the internal force F_int = sum_e B^T sigma into Fi
 * (RT_Asm_ComputeResidual, RT_Asm_Solv.f90:1311-1426): elements in parallel
 * (nthreads), element forces added in element order.  Returns the number of
 * Gauss points whose local Newton failed. */
int64_t orc_j2_internal_force(int32_t n_nodes, int32_t n_elems, const int32_t* conn,
                              const double* coords, const double* props, const double* u,
                              const double* u_n, const double* state_n, double* state,
                              double* Dgp, double* Fi, int nthreads) {
  double* Fe_all = (double*)malloc(sizeof(double) * 24 * (size_t)n_elems);
  int64_t local_fail = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : local_fail)
#endif
  for (int64_t e = 0; e < n_elems; ++e) {
    double xe[24], due[24], deps[48], sig[48];
    gather_coords(conn, coords, e, xe);
    for (int j = 0; j < 24; ++j) {
      int64_t q = 3 * (int64_t)conn[8 * e + j / 3] + j % 3;
      due[j] = u[q] - u_n[q];
    }
    orc_c3d8_strain(xe, due, deps);
    for (int g = 0; g < 8; ++g) {
      const double* sn = state_n + 13 * (8 * e + g);
      double* sc = state + 13 * (8 * e + g);
      memcpy(sc, sn, 13 * sizeof(double));
      if (orc_j2_update(props, deps + 6 * g, sc, sc + 6, sc + 7, Dgp + 288 * e + 36 * g,
                        NULL) != 0)
        local_fail++;
      memcpy(sig + 6 * g, sc, 6 * sizeof(double));
    }
    orc_c3d8_fint(xe, sig, Fe_all + 24 * e);
  }
  memset(Fi, 0, sizeof(double) * 3 * (size_t)n_nodes);
  for (int64_t e = 0; e < n_elems; ++e)
    for (int j = 0; j < 24; ++j) Fi[3 * (int64_t)conn[8 * e + j / 3] + j % 3] += Fe_all[24 * e + j];
  free(Fe_all);
  return local_fail;
}

/* One load increment of the small-strain static Newton loop
 * (RT_NLSolver_NewtonRaph, RT_Solv_Nonlin.f90:807-1000): per iteration
 * strain increments from u - u_n, J2 update from the committed state,
 * internal force (RT_Asm_ComputeResidual, RT_Asm_Solv.f90:1311-1426),
 * tangent stiffness with per-Gauss-point D, penalty Dirichlet BC
 * (RT_Asm_Solv.f90:318-344), R = P ubar - (F_int + P u) on the constrained
 * rows, Jacobi-PCG for du, u += du.  Converged when ||R|| <= tol_r ||F_int||
 * (the reference compares absolute norms with 1e-6 / 1e-5,
 * RT_Step_Exec.f90:380-382, which are meaningless in SI units).
 * state (13 doubles per Gauss point: sigma(6), peeq, eps_p(6)) holds the
 * trial state on return; state_n is the committed state; u holds u_n on
 * entry.  Returns 0 converged, 4 (IF_STATUS_NOT_CONVERGED) otherwise, 2 when
 * a Gauss point's local Newton failed (the reference's pnewdt cut-back
 * signal, PH_Mat_Plast_J2_Iso_Core.f90:419-427). */
int orc_j2_newton_step(int32_t n_nodes, int32_t n_elems, const int32_t* conn,
                       const double* coords, const double* props,
                       int32_t n_bc, const int32_t* bc_eq, const double* bc_val, double penalty,
                       const double* state_n, double* state, double* u,
                       int32_t max_newton, double tol_r, double pcg_rtol,
                       int32_t* n_newton, double* rnorm_out) {
  const int32_t n = 3 * n_nodes;
  int32_t* row_ptr = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  int64_t nnz = orc_csr_pattern(n_nodes, n_elems, conn, row_ptr, NULL);
  int32_t* col_idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)nnz);
  orc_csr_pattern(n_nodes, n_elems, conn, row_ptr, col_idx);
  double* val = (double*)malloc(sizeof(double) * (size_t)nnz);
  double* Dgp = (double*)malloc(sizeof(double) * 288 * (size_t)n_elems);
  double* u_n = (double*)malloc(sizeof(double) * (size_t)n);
  double* R = (double*)malloc(sizeof(double) * (size_t)n);
  double* Fi = (double*)malloc(sizeof(double) * (size_t)n);
  double* du = (double*)malloc(sizeof(double) * (size_t)n);
  memcpy(u_n, u, sizeof(double) * (size_t)n);
  int status = 4;
  int32_t it;
  double rn = 0.0;
  for (it = 1; it <= max_newton; ++it) {
    if (orc_j2_internal_force(n_nodes, n_elems, conn, coords, props, u, u_n, state_n, state,
                              Dgp, Fi, 1)) {
      status = 2;
      break;
    }
    for (int32_t i = 0; i < n; ++i) R[i] = -Fi[i];
    for (int32_t b = 0; b < n_bc; ++b) R[bc_eq[b]] += penalty * (bc_val[b] - u[bc_eq[b]]);
    double fn = sqrt(dot(n, Fi, Fi, 1));
    rn = sqrt(dot(n, R, R, 1));
    if (it > 1 && rn <= tol_r * (fn > 0.0 ? fn : 1.0)) { status = 0; break; }
    orc_assemble_dgp(n_elems, conn, coords, Dgp, n, row_ptr, col_idx, val, 1);
    orc_apply_penalty(n_bc, bc_eq, NULL, penalty, row_ptr, col_idx, val, NULL);
    memset(du, 0, sizeof(double) * (size_t)n);
    int32_t ni;
    double rr;
    orc_pcg(n, row_ptr, col_idx, val, R, du, 100000, pcg_rtol, &ni, &rr, 1);
    for (int32_t i = 0; i < n; ++i) u[i] += du[i];
  }
  if (n_newton) *n_newton = it > max_newton ? max_newton : it;
  if (rnorm_out) *rnorm_out = rn;
  free(row_ptr); free(col_idx); free(val); free(Dgp); free(u_n); free(R); free(Fi); free(du);
  return status;
}
