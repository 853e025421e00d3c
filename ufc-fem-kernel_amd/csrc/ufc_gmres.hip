// ufc_gmres.hip -- y = K^T x on the node-block values and right-Jacobi
// preconditioned GMRES(m) on K^T: the non-symmetric adjoint path of
// NM_Adjoint_Solve (L2_NM/Solver/NM_Solv_GMRESTranspose.f90:229-291, is_symmetric
// = .FALSE.: GMRES, restart 30, tolerance 1e-8, 1000 iterations), whose body in
// the reference is a damped Richardson placeholder (:337-393).  The reference
// forms K^T explicitly (NM_CSR_Transpose, NM_Mtx_SpMVCSRTranspose.f90:46-145);
// here the transpose is read in place: block (m, n) of K is block mirror[q] of
// node m's row segment, for q the slot of m in node n's neighbour list.
#include <cmath>
#include <vector>

#include "ufc_internal.hpp"

namespace ufc {
namespace {

constexpr int kGroup = 32;        // lanes per node
constexpr int kTThreads = 512;
constexpr int kDotBlocks = 256;   // workgroups per dot product
constexpr int kDotThreads = 256;

// position of node n in the (sorted) neighbour list of each of its neighbours
__global__ void k_build_mirror(int32_t n_nodes, const int32_t* __restrict__ nb_ptr,
                               const int32_t* __restrict__ nb_idx, int32_t* __restrict__ mirror) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  for (int32_t q = nb_ptr[n]; q < nb_ptr[n + 1]; ++q) {
    const int32_t m = nb_idx[q];
    int32_t lo = nb_ptr[m], hi = nb_ptr[m + 1] - 1, pos = -1;
    while (lo <= hi) {
      const int32_t mid = (lo + hi) >> 1;
      const int32_t v = nb_idx[mid];
      if (v == n) {
        pos = mid - nb_ptr[m];
        break;
      }
      if (v < n)
        lo = mid + 1;
      else
        hi = mid - 1;
    }
    mirror[q] = pos;
  }
}

// y_n = sum_k K_{m_k n}^T x_{m_k}: 32-lane group per node, lane k reads the
// transposed block of neighbour m_k (3 runs of 3 values in m_k's rows)
__global__ __launch_bounds__(kTThreads) void k_spmv_t(int32_t n_nodes,
                                                     const int32_t* __restrict__ nb_ptr,
                                                     const int32_t* __restrict__ nb_idx,
                                                     const int32_t* __restrict__ mirror,
                                                     const double* __restrict__ val,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y) {
  const int k = threadIdx.x & (kGroup - 1);
  const int64_t g0 = int64_t(blockIdx.x) * (kTThreads / kGroup) + threadIdx.x / kGroup;
  const int64_t ng = int64_t(gridDim.x) * (kTThreads / kGroup);
  for (int64_t n = g0; n < n_nodes; n += ng) {
    const int32_t q0 = nb_ptr[n];
    const int deg = nb_ptr[n + 1] - q0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    // lane k on the neighbours k, k + 32, ... (one pass unless the list is
    // longer than the group)
    for (int kk = k; kk < deg; kk += kGroup) {
      const int32_t m = nb_idx[q0 + kk];
      const int32_t p = mirror[q0 + kk];
      if (p >= 0) {
        const int32_t mq = nb_ptr[m];
        const int64_t d3 = 3 * int64_t(nb_ptr[m + 1] - mq);
        const double* blk = val + 9 * int64_t(mq) + 3 * int64_t(p);
        const double x0 = x[3 * int64_t(m)], x1 = x[3 * int64_t(m) + 1], x2 = x[3 * int64_t(m) + 2];
        // row i of block (m, n) times x_m[i], accumulated into column j
        a0 += blk[0] * x0 + blk[d3] * x1 + blk[2 * d3] * x2;
        a1 += blk[1] * x0 + blk[d3 + 1] * x1 + blk[2 * d3 + 1] * x2;
        a2 += blk[2] * x0 + blk[d3 + 2] * x1 + blk[2 * d3 + 2] * x2;
      }
    }
#pragma unroll
    for (int o = kGroup / 2; o > 0; o >>= 1) {
      a0 += __shfl_xor(a0, o, 64);
      a1 += __shfl_xor(a1, o, 64);
      a2 += __shfl_xor(a2, o, 64);
    }
    if (k < 3) y[3 * n + k] = k == 0 ? a0 : (k == 1 ? a1 : a2);
  }
}

// partials[i * kDotBlocks + b] = sum over block b of w . V_i, i = blockIdx.y
__global__ __launch_bounds__(kDotThreads) void k_mdot(int64_t n, const double* __restrict__ V,
                                                      const double* __restrict__ w,
                                                      double* __restrict__ partials) {
  const double* v = V + int64_t(blockIdx.y) * n;
  double s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kDotThreads + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kDotThreads)
    s += v[i] * w[i];
  __shared__ double sh[kDotThreads / 64];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kDotThreads / 64; ++k) t += sh[k];
    partials[int64_t(blockIdx.y) * kDotBlocks + blockIdx.x] = t;
  }
}

// h[i] = ordered sum of the kDotBlocks partials of dot i (one workgroup each)
__global__ __launch_bounds__(kDotBlocks) void k_mdot_fin(const double* __restrict__ partials,
                                                         double* __restrict__ h) {
  __shared__ double sh[kDotBlocks];
  sh[threadIdx.x] = partials[int64_t(blockIdx.x) * kDotBlocks + threadIdx.x];
  __syncthreads();
  for (int s = kDotBlocks / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) h[blockIdx.x] = sh[0];
}

// w -= sum_{i < k} h[i] V_i
__global__ void k_vsub(int64_t n, int k, const double* __restrict__ V,
                       const double* __restrict__ h, double* __restrict__ w) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int j = 0; j < k; ++j) s -= h[j] * V[int64_t(j) * n + i];
  w[i] = s;
}

// out = a * v
__global__ void k_scale(int64_t n, double a, const double* v, double* out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * v[i];
}

// z = v / diag (the PCG's Jacobi guard: |d| <= 1e-30 leaves v)
__global__ void k_jacobi(int64_t n, const double* __restrict__ v, const double* __restrict__ d,
                         double* __restrict__ z) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double di = d[i];
  z[i] = (di > 1e-30 || di < -1e-30) ? v[i] / di : v[i];
}

// x += M^-1 sum_{j < k} y[j] V_j
__global__ void k_combine(int64_t n, int k, const double* __restrict__ V,
                          const double* __restrict__ yv, const double* __restrict__ d,
                          double* __restrict__ x) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += yv[j] * V[int64_t(j) * n + i];
  const double di = d[i];
  x[i] += (di > 1e-30 || di < -1e-30) ? s / di : s;
}

__global__ void k_sub(int64_t n, const double* __restrict__ b, const double* __restrict__ y,
                      double* __restrict__ r) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
}

inline unsigned blocks_for(int64_t n, int t) { return unsigned((n + t - 1) / t); }

// h_host[0..k) = V_i . w
void mdot(ufc_ctx* c, int k, const double* V, const double* w, double* partials, double* h_dev,
          double* h_host) {
  hipLaunchKernelGGL(k_mdot, dim3(kDotBlocks, k), dim3(kDotThreads), 0, c->stream, c->n_dof, V,
                     w, partials);
  hipLaunchKernelGGL(k_mdot_fin, dim3(k), dim3(kDotBlocks), 0, c->stream, partials, h_dev);
  UFC_HIP_CHECK(hipGetLastError());
  UFC_HIP_CHECK(hipMemcpyAsync(h_host, h_dev, sizeof(double) * size_t(k), hipMemcpyDeviceToHost,
                               c->stream));
  UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
}

}  // namespace

void ensure_mirror(ufc_ctx* c) {
  if (c->have_mirror) return;
  c->nb_mirror.resize(size_t(c->nnz / 9));
  hipLaunchKernelGGL(k_build_mirror, dim3(blocks_for(c->n_nodes, 256)), dim3(256), 0, c->stream,
                     c->n_nodes, c->nb_ptr.get(), c->nb_idx.get(), c->nb_mirror.get());
  UFC_HIP_CHECK(hipGetLastError());
  c->have_mirror = true;
}

void spmv_t64(ufc_ctx* c, const double* x, double* y) {
  ensure_mirror(c);
  const int64_t per_block = kTThreads / kGroup;
  const unsigned grid = unsigned(std::min<int64_t>(32768, (c->n_nodes + per_block - 1) / per_block));
  hipLaunchKernelGGL(k_spmv_t, dim3(grid), dim3(kTThreads), 0, c->stream, c->n_nodes,
                     c->nb_ptr.get(), c->nb_idx.get(), c->nb_mirror.get(), c->val64.get(), x, y);
  UFC_HIP_CHECK(hipGetLastError());
}

// GMRES(m) on K^T lambda = g with right Jacobi preconditioning
// (K^T M^-1 u = g, lambda = M^-1 u; diag(K^T) = diag(K)), classical
// Gram-Schmidt applied twice, Givens rotations on the host.  lambda_0 = 0.
// Stops when the (preconditioned-system) residual estimate drops below
// rtol ||g||, then reports the true residual ||g - K^T lambda|| / ||g||.
int gmres_t64(ufc_ctx* c, const double* g, double* lam, int restart, int max_iter, double rtol,
              int* niter, int* n_restart, double* rnorm, std::string& msg) {
  const int64_t n = c->n_dof;
  const int m = restart;
  DBuf<double> V, w, z, part, hd;
  V.resize(size_t(m + 1) * size_t(n));
  w.resize(size_t(n));
  z.resize(size_t(n));
  part.resize(size_t(m + 1) * kDotBlocks);
  hd.resize(size_t(m + 2));
  const unsigned gb = blocks_for(n, 256);
  UFC_HIP_CHECK(hipMemsetAsync(lam, 0, sizeof(double) * size_t(n), c->stream));
  const size_t mm = static_cast<size_t>(m);
  std::vector<double> H((mm + 1) * mm, 0.0), cs(mm, 0.0), sn(mm, 0.0), s(mm + 1, 0.0),
      h(mm + 1, 0.0), h2(mm + 1, 0.0), yv(mm, 0.0);
  double gnorm;
  mdot(c, 1, g, g, part.get(), hd.get(), &gnorm);
  gnorm = std::sqrt(gnorm);
  *niter = 0;
  *n_restart = 0;
  if (!(gnorm > 0.0) || !std::isfinite(gnorm)) {
    *rnorm = 0.0;
    return std::isfinite(gnorm) ? UFC_STATUS_OK : UFC_STATUS_ERROR;
  }
  int its = 0;
  double est = 1.0;
  bool conv = false;
  while (its < max_iter && !conv) {
    // r = g - K^T lambda into V_0
    if (its == 0) {
      UFC_HIP_CHECK(hipMemcpyAsync(V.get(), g, sizeof(double) * size_t(n),
                                   hipMemcpyDeviceToDevice, c->stream));
    } else {
      spmv_t64(c, lam, w.get());
      hipLaunchKernelGGL(k_sub, dim3(gb), dim3(256), 0, c->stream, n, g, w.get(), V.get());
    }
    double beta;
    mdot(c, 1, V.get(), V.get(), part.get(), hd.get(), &beta);
    beta = std::sqrt(beta);
    if (!std::isfinite(beta)) {
      msg = "GMRES breakdown: non-finite residual";
      return UFC_STATUS_ERROR;
    }
    if (beta <= rtol * gnorm) {
      conv = true;
      break;
    }
    hipLaunchKernelGGL(k_scale, dim3(gb), dim3(256), 0, c->stream, n, 1.0 / beta, V.get(),
                       V.get());
    std::fill(s.begin(), s.end(), 0.0);
    s[0] = beta;
    int k = 0;
    for (; k < m && its < max_iter; ++k) {
      double* vk = V.get() + int64_t(k) * n;
      hipLaunchKernelGGL(k_jacobi, dim3(gb), dim3(256), 0, c->stream, n, vk, c->diag64.get(),
                         z.get());
      spmv_t64(c, z.get(), w.get());
      // classical Gram-Schmidt, twice
      mdot(c, k + 1, V.get(), w.get(), part.get(), hd.get(), h.data());
      hipLaunchKernelGGL(k_vsub, dim3(gb), dim3(256), 0, c->stream, n, k + 1, V.get(), hd.get(),
                         w.get());
      mdot(c, k + 1, V.get(), w.get(), part.get(), hd.get(), h2.data());
      hipLaunchKernelGGL(k_vsub, dim3(gb), dim3(256), 0, c->stream, n, k + 1, V.get(), hd.get(),
                         w.get());
      for (int i = 0; i <= k; ++i) h[size_t(i)] += h2[size_t(i)];
      double hn;
      mdot(c, 1, w.get(), w.get(), part.get(), hd.get(), &hn);
      hn = std::sqrt(hn);
      ++its;
      if (!std::isfinite(hn)) {
        msg = "GMRES breakdown: non-finite Krylov vector";
        return UFC_STATUS_ERROR;
      }
      if (hn > 0.0)
        hipLaunchKernelGGL(k_scale, dim3(gb), dim3(256), 0, c->stream, n, 1.0 / hn, w.get(),
                           V.get() + int64_t(k + 1) * n);
      // apply the previous rotations to the new column, then its own
      for (int i = 0; i < k; ++i) {
        const double t = cs[size_t(i)] * h[size_t(i)] + sn[size_t(i)] * h[size_t(i + 1)];
        h[size_t(i + 1)] = -sn[size_t(i)] * h[size_t(i)] + cs[size_t(i)] * h[size_t(i + 1)];
        h[size_t(i)] = t;
      }
      const double den = std::hypot(h[size_t(k)], hn);
      cs[size_t(k)] = den > 0.0 ? h[size_t(k)] / den : 1.0;
      sn[size_t(k)] = den > 0.0 ? hn / den : 0.0;
      h[size_t(k)] = den;
      s[size_t(k + 1)] = -sn[size_t(k)] * s[size_t(k)];
      s[size_t(k)] = cs[size_t(k)] * s[size_t(k)];
      for (int i = 0; i <= k; ++i) H[size_t(i) + size_t(m + 1) * size_t(k)] = h[size_t(i)];
      est = std::fabs(s[size_t(k + 1)]) / gnorm;
      if (est <= rtol || hn == 0.0) {
        ++k;
        conv = true;
        break;
      }
    }
    // y = R^-1 s (upper triangular k x k), lambda += M^-1 V y
    for (int i = k - 1; i >= 0; --i) {
      double t = s[size_t(i)];
      for (int j = i + 1; j < k; ++j) t -= H[size_t(i) + size_t(m + 1) * size_t(j)] * yv[size_t(j)];
      const double d = H[size_t(i) + size_t(m + 1) * size_t(i)];
      yv[size_t(i)] = d != 0.0 ? t / d : 0.0;
    }
    UFC_HIP_CHECK(hipMemcpyAsync(hd.get(), yv.data(), sizeof(double) * size_t(k),
                                 hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_combine, dim3(gb), dim3(256), 0, c->stream, n, k, V.get(), hd.get(),
                       c->diag64.get(), lam);
    UFC_HIP_CHECK(hipGetLastError());
    if (!conv) ++*n_restart;
  }
  // true residual
  spmv_t64(c, lam, w.get());
  hipLaunchKernelGGL(k_sub, dim3(gb), dim3(256), 0, c->stream, n, g, w.get(), z.get());
  double rr;
  mdot(c, 1, z.get(), z.get(), part.get(), hd.get(), &rr);
  *rnorm = std::sqrt(rr) / gnorm;
  *niter = its;
  if (!std::isfinite(*rnorm)) {
    msg = "GMRES breakdown: non-finite residual";
    return UFC_STATUS_ERROR;
  }
  if (!conv) {
    msg = "[NM_GMRES_Solve_Transpose]: did not converge";
    return UFC_STATUS_INVALID;
  }
  return UFC_STATUS_OK;
}

}  // namespace ufc
