// ufc_solve.hip -- gfx950 CSR SpMV and fused Jacobi-PCG.
//
// PCG follows NM_Solver_PCG (reference L2_NM/Solver/NM_Solv_Core.f90:363-423)
// step for step: r = b - K x0, z = r / diag (z = r where |diag| <= 1e-30,
// NM_Solver_Jacobi_Precond :131-147), p = z; per iteration Ap = K p,
// alpha = (r.z) / (p.Ap), x += alpha p, r -= alpha Ap, stop at the first k
// with ||r|| / ||b|| < rtol (||b|| < 1e-30 -> 1), z = r / diag,
// beta = (r.z)_new / (r.z)_old, p = z + beta p.
//
// Each iteration is three kernels, all scalars live on the device:
//   spmv_dot : Ap = K p and p.Ap; the last workgroup forms alpha
//   update   : x += alpha p, r -= alpha Ap, r.r and r.(r/diag); the last
//              workgroup forms ||r||/||b||, decides convergence, forms beta
//   p_update : p = r/diag + beta p
// Reductions are deterministic: every workgroup writes its partial sum, the
// last one to arrive (ticket counter) adds the partials in index order.
// Once `done` is set every later kernel returns at entry, so the host only
// polls every few iterations and the stopping iteration is still exact.
//
// SpMV: one wave64 per node, i.e. per group of 3 CSR rows whose values are
// one contiguous segment (node n's rows start at 9*nb_ptr[n]); lanes stride
// the segment (coalesced), keep one partial per row and reduce with shuffles.
//   FP64: scalar CSR (values + col_idx), the reference's storage.
//   FP32: node-block index view of the same CSR -- the column of value
//         3*deg*d + 3*k + c is 3*nb_idx[nb_ptr[n]+k] + c, so one 4-byte index
//         per 9 values replaces col_idx.
#include "ufc_internal.hpp"

namespace ufc {
namespace {

constexpr int WAVE = 64;
constexpr int WAVES_PER_BLOCK = kSolveThreads / WAVE;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// Sum of v over the workgroup (kSolveThreads threads), valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*sh)[WAVES_PER_BLOCK]) {
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    v[k] = wave_sum(v[k]);
    if (lane == 0) sh[k][wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < WAVES_PER_BLOCK; ++w) s += sh[k][w];
      v[k] = s;
    }
  }
}

// Writes this workgroup's partials, then returns true in every thread of the
// last workgroup to finish, with all partials of the grid visible to it.
template <int NV>
__device__ __forceinline__ bool publish_partials(const double (&v)[NV], double* partials,
                                                 uint32_t* ticket) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * kSolveBlocks + blockIdx.x] = v[k];
    const uint32_t t =
        __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return last;
}

// In the last workgroup: ordered sum of the grid's partials (same order on
// every run), valid in thread 0.  Resets the ticket.
template <int NV>
__device__ __forceinline__ void reduce_partials(double (&out)[NV], const double* partials,
                                                uint32_t* ticket,
                                                double (*sh)[WAVES_PER_BLOCK]) {
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < int(gridDim.x); b += kSolveThreads)
      s += partials[k * kSolveBlocks + b];
    v[k] = s;
  }
  __syncthreads();
  block_sum<NV>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = v[k];
    *ticket = 0u;
  }
}

// ---------------------------------------------------------------- SpMV bodies
// y(3n..3n+2) for node n, FP64 scalar CSR; returns the row sums in lane 0..2 order
__device__ __forceinline__ void node_rows64(int64_t n, const int32_t* __restrict__ row_ptr,
                                            const int32_t* __restrict__ col_idx,
                                            const double* __restrict__ val,
                                            const double* __restrict__ x, double (&acc)[3]) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int32_t s0 = row_ptr[3 * n], s1 = row_ptr[3 * n + 1], s2 = row_ptr[3 * n + 2],
                s3 = row_ptr[3 * n + 3];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int32_t o = s0 + lane; o < s3; o += WAVE) {
    const double v = val[o] * x[col_idx[o]];
    if (o < s1) a0 += v;
    else if (o < s2) a1 += v;
    else a2 += v;
  }
  acc[0] = wave_sum(a0);
  acc[1] = wave_sum(a1);
  acc[2] = wave_sum(a2);
}

// FP32 node-block view; products and row sums in FP32
__device__ __forceinline__ void node_rows32(int64_t n, const int32_t* __restrict__ nb_ptr,
                                            const int32_t* __restrict__ nb_idx,
                                            const float* __restrict__ val,
                                            const float* __restrict__ x, float (&acc)[3]) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int32_t q0 = nb_ptr[n];
  const int deg = nb_ptr[n + 1] - q0;
  const int d3 = 3 * deg;
  const int len = 9 * deg;
  const float* __restrict__ v = val + 9 * int64_t(q0);
  const int32_t* __restrict__ nb = nb_idx + q0;
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;
  for (int o = lane; o < len; o += WAVE) {
    const int d = (o >= d3) + (o >= 2 * d3);
    const int rem = o - d * d3;
    const int k = rem / 3;
    const int c = rem - 3 * k;
    const float t = v[o] * x[3 * nb[k] + c];
    if (d == 0) a0 += t;
    else if (d == 1) a1 += t;
    else a2 += t;
  }
  acc[0] = wave_sum(a0);
  acc[1] = wave_sum(a1);
  acc[2] = wave_sum(a2);
}

// ---------------------------------------------------------------- kernels
// plain y = K x
__global__ __launch_bounds__(kSolveThreads) void k_spmv64(int32_t n_nodes,
                                                         const int32_t* __restrict__ row_ptr,
                                                         const int32_t* __restrict__ col_idx,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ y) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = int64_t(blockIdx.x) * WAVES_PER_BLOCK + threadIdx.x / WAVE;
  const int64_t nw = int64_t(gridDim.x) * WAVES_PER_BLOCK;
  for (int64_t n = wave; n < n_nodes; n += nw) {
    double acc[3];
    node_rows64(n, row_ptr, col_idx, val, x, acc);
    if (lane < 3) y[3 * n + lane] = acc[lane == 0 ? 0 : (lane == 1 ? 1 : 2)];
  }
}

__global__ __launch_bounds__(kSolveThreads) void k_spmv32(int32_t n_nodes,
                                                         const int32_t* __restrict__ nb_ptr,
                                                         const int32_t* __restrict__ nb_idx,
                                                         const float* __restrict__ val,
                                                         const float* __restrict__ x,
                                                         float* __restrict__ y) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = int64_t(blockIdx.x) * WAVES_PER_BLOCK + threadIdx.x / WAVE;
  const int64_t nw = int64_t(gridDim.x) * WAVES_PER_BLOCK;
  for (int64_t n = wave; n < n_nodes; n += nw) {
    float acc[3];
    node_rows32(n, nb_ptr, nb_idx, val, x, acc);
    if (lane < 3) y[3 * n + lane] = acc[lane == 0 ? 0 : (lane == 1 ? 1 : 2)];
  }
}

// Ap = K p, p.Ap; last workgroup: alpha = rz / pAp (breakdown if pAp <= 0 or not finite)
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void k_spmv_dot(
    int32_t n_nodes, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col_idx,
    const int32_t* __restrict__ nb_ptr, const int32_t* __restrict__ nb_idx,
    const T* __restrict__ val, const T* __restrict__ p, T* __restrict__ ap,
    double* __restrict__ partials, PcgScalars* __restrict__ sc) {
  __shared__ double sh[1][WAVES_PER_BLOCK];
  if (sc->done) return;
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = int64_t(blockIdx.x) * WAVES_PER_BLOCK + threadIdx.x / WAVE;
  const int64_t nw = int64_t(gridDim.x) * WAVES_PER_BLOCK;
  double dot = 0.0;
  for (int64_t n = wave; n < n_nodes; n += nw) {
    T acc[3];
    if constexpr (sizeof(T) == 8) {
      double a[3];
      node_rows64(n, row_ptr, col_idx, reinterpret_cast<const double*>(val),
                  reinterpret_cast<const double*>(p), a);
      acc[0] = a[0]; acc[1] = a[1]; acc[2] = a[2];
    } else {
      float a[3];
      node_rows32(n, nb_ptr, nb_idx, reinterpret_cast<const float*>(val),
                  reinterpret_cast<const float*>(p), a);
      acc[0] = a[0]; acc[1] = a[1]; acc[2] = a[2];
    }
    if (lane < 3) {
      const T yv = acc[lane == 0 ? 0 : (lane == 1 ? 1 : 2)];
      ap[3 * n + lane] = yv;
      dot += double(yv) * double(p[3 * n + lane]);
    }
  }
  double v[1] = {dot};
  block_sum<1>(v, sh);
  if (publish_partials<1>(v, partials, &sc->ticket[0])) {
    double out[1];
    reduce_partials<1>(out, partials, &sc->ticket[0], sh);
    if (threadIdx.x == 0) {
      const double pap = out[0];
      sc->pap = pap;
      if (!(pap > 0.0) || !isfinite(pap)) {
        sc->done = 2;
      } else {
        sc->alpha = sc->rz / pap;
      }
    }
  }
}

// x += alpha p, r -= alpha Ap; r.r and r.z with z = r/diag
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void k_update(int64_t n, T* __restrict__ x,
                                                         T* __restrict__ r,
                                                         const T* __restrict__ p,
                                                         const T* __restrict__ ap,
                                                         const T* __restrict__ diag,
                                                         double* __restrict__ partials,
                                                         PcgScalars* __restrict__ sc) {
  __shared__ double sh[2][WAVES_PER_BLOCK];
  if (sc->done) return;
  const double alpha = sc->alpha;
  const T a = T(alpha);
  double rr = 0.0, rz = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kSolveThreads) {
    x[i] = x[i] + a * p[i];
    const T ri = r[i] - a * ap[i];
    r[i] = ri;
    const T d = diag[i];
    const T zi = (d > T(1e-30) || d < T(-1e-30)) ? ri / d : ri;
    rr += double(ri) * double(ri);
    rz += double(ri) * double(zi);
  }
  double v[2] = {rr, rz};
  block_sum<2>(v, sh);
  if (publish_partials<2>(v, partials, &sc->ticket[1])) {
    double out[2];
    reduce_partials<2>(out, partials, &sc->ticket[1], sh);
    if (threadIdx.x == 0) {
      sc->rr = out[0];
      sc->iter += 1;
      const double rn = sqrt(out[0]) / sc->bnorm;
      sc->rnorm = rn;
      if (!isfinite(rn)) {
        sc->done = 2;
      } else if (rn < sc->rtol) {
        sc->done = 1;
      } else if (sc->iter >= sc->max_iter) {
        sc->done = 3;
      } else {
        sc->beta = out[1] / sc->rz;
        sc->rz = out[1];
      }
    }
  }
}

// p = r/diag + beta p
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void k_pupdate(int64_t n, T* __restrict__ p,
                                                          const T* __restrict__ r,
                                                          const T* __restrict__ diag,
                                                          const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  const T beta = T(sc->beta);
  for (int64_t i = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kSolveThreads) {
    const T d = diag[i];
    const T ri = r[i];
    const T zi = (d > T(1e-30) || d < T(-1e-30)) ? ri / d : ri;
    p[i] = zi + beta * p[i];
  }
}

// Setup: r = b - Ax (Ax may be null: x0 = 0), z = r/diag, p = z, x cast in;
// r.z and b.b; last workgroup initialises the scalars.
template <typename T>
__global__ __launch_bounds__(kSolveThreads) void k_setup(int64_t n, const double* __restrict__ b,
                                                        const T* __restrict__ ax,
                                                        T* __restrict__ r, T* __restrict__ p,
                                                        const T* __restrict__ diag,
                                                        double* __restrict__ partials,
                                                        PcgScalars* __restrict__ sc, double rtol,
                                                        int32_t max_iter) {
  __shared__ double sh[2][WAVES_PER_BLOCK];
  double rz = 0.0, bb = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kSolveThreads) {
    const double bi = b[i];
    const T ri = ax ? T(bi - double(ax[i])) : T(bi);
    r[i] = ri;
    const T d = diag[i];
    const T zi = (d > T(1e-30) || d < T(-1e-30)) ? ri / d : ri;
    p[i] = zi;
    rz += double(ri) * double(zi);
    bb += bi * bi;
  }
  double v[2] = {rz, bb};
  block_sum<2>(v, sh);
  if (publish_partials<2>(v, partials, &sc->ticket[2])) {
    double out[2];
    reduce_partials<2>(out, partials, &sc->ticket[2], sh);
    if (threadIdx.x == 0) {
      sc->rz = out[0];
      double bn = sqrt(out[1]);
      if (bn < 1.0e-30) bn = 1.0;
      sc->bnorm = bn;
      sc->rtol = rtol;
      sc->max_iter = max_iter;
      sc->iter = 0;
      sc->done = max_iter > 0 ? 0 : 3;
      sc->rnorm = 0.0;
      sc->alpha = sc->beta = sc->pap = sc->rr = 0.0;
    }
  }
}

// v.v with the deterministic two-level reduction; result in sc->rr
__global__ __launch_bounds__(kSolveThreads) void k_norm2(int64_t n, const double* __restrict__ v,
                                                        double* __restrict__ partials,
                                                        PcgScalars* __restrict__ sc) {
  __shared__ double sh[1][WAVES_PER_BLOCK];
  double s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kSolveThreads + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kSolveThreads)
    s += v[i] * v[i];
  double a[1] = {s};
  block_sum<1>(a, sh);
  if (publish_partials<1>(a, partials, &sc->ticket[3])) {
    double out[1];
    reduce_partials<1>(out, partials, &sc->ticket[3], sh);
    if (threadIdx.x == 0) sc->rr = out[0];
  }
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

void spmv_event_begin(ufc_ctx* c) {
  if (!c->per_kernel) return;
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
  if (!c->per_kernel) return;
  UFC_HIP_CHECK(hipEventRecord(c->spmv_ev[c->spmv_ev_used + 1], c->stream));
  c->spmv_ev_used += 2;
}

void poll(ufc_ctx* c) {
  UFC_HIP_CHECK(hipMemcpyAsync(c->h_scalars, c->scalars.get(), sizeof(PcgScalars),
                               hipMemcpyDeviceToHost, c->stream));
  UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
}

template <typename T>
int pcg_run(ufc_ctx* c, const double* b, T* x, const T* ax_or_null, T* r, T* p, T* ap,
            const T* diag, const T* val, int max_iter, double rtol, int check_every,
            int* niter, double* rnorm) {
  const int64_t n = c->n_dof;
  const unsigned vgrid = kSolveBlocks;
  hipLaunchKernelGGL(k_setup<T>, dim3(vgrid), dim3(kSolveThreads), 0, c->stream, n, b,
                     ax_or_null, r, p, diag, c->partials.get(), c->scalars.get(), rtol,
                     int32_t(max_iter));
  UFC_HIP_CHECK(hipGetLastError());
  const int every = check_every > 0 ? check_every : 32;
  int launched = 0;
  for (;;) {
    for (int k = 0; k < every && launched < max_iter; ++k, ++launched) {
      spmv_event_begin(c);
      hipLaunchKernelGGL(k_spmv_dot<T>, dim3(kSolveBlocks), dim3(kSolveThreads), 0, c->stream,
                         c->n_nodes, c->row_ptr.get(), c->col_idx.get(), c->nb_ptr.get(),
                         c->nb_idx.get(), val, p, ap, c->partials.get(), c->scalars.get());
      spmv_event_end(c);
      hipLaunchKernelGGL(k_update<T>, dim3(vgrid), dim3(kSolveThreads), 0, c->stream, n, x, r, p,
                         ap, diag, c->partials.get(), c->scalars.get());
      hipLaunchKernelGGL(k_pupdate<T>, dim3(vgrid), dim3(kSolveThreads), 0, c->stream, n, p, r,
                         diag, c->scalars.get());
    }
    UFC_HIP_CHECK(hipGetLastError());
    poll(c);
    if (c->h_scalars->done || launched >= max_iter) break;
  }
  *niter = c->h_scalars->iter;
  *rnorm = c->h_scalars->rnorm;
  const int done = c->h_scalars->done;
  if (done == 1) return UFC_STATUS_OK;
  if (done == 2) return UFC_STATUS_ERROR;
  return UFC_STATUS_INVALID;
}

}  // namespace

// ---------------------------------------------------------------- host entry points
void ensure_solver_buffers(ufc_ctx* c, bool need64, bool need32) {
  const size_t n = size_t(c->n_dof);
  if (need64) {
    c->w_r.resize(n);
    c->w_p.resize(n);
    c->w_ap.resize(n);
    c->w_x.resize(n);
    c->w_b.resize(n);
    c->w_t.resize(n);
  }
  if (need32) {
    c->s_r.resize(n);
    c->s_p.resize(n);
    c->s_ap.resize(n);
    c->s_x.resize(n);
    c->w_b.resize(n);
    c->w_t.resize(n);
    c->w_r.resize(n);
    c->w_x.resize(n);
  }
  c->partials.resize(size_t(kSolveBlocks) * 4);
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
  hipLaunchKernelGGL(k_spmv64, dim3(kSolveBlocks), dim3(kSolveThreads), 0, c->stream,
                     c->n_nodes, c->row_ptr.get(), c->col_idx.get(), c->val64.get(), x, y);
  spmv_event_end(c);
  UFC_HIP_CHECK(hipGetLastError());
}

void spmv32(ufc_ctx* c, const float* x, float* y) {
  spmv_event_begin(c);
  hipLaunchKernelGGL(k_spmv32, dim3(kSolveBlocks), dim3(kSolveThreads), 0, c->stream,
                     c->n_nodes, c->nb_ptr.get(), c->nb_idx.get(), c->val32.get(), x, y);
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
  hipLaunchKernelGGL(k_norm2, dim3(kSolveBlocks), dim3(kSolveThreads), 0, c->stream, c->n_dof, v,
                     c->partials.get(), c->scalars.get());
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
  return pcg_run<double>(c, b, x, ax, c->w_r.get(), c->w_p.get(), c->w_ap.get(),
                         c->diag64.get(), c->val64.get(), max_iter, rtol, check_every, niter,
                         rnorm);
}

int pcg32(ufc_ctx* c, const double* b64, float* x, int max_iter, double rtol, int check_every,
          int* niter, double* rnorm) {
  UFC_HIP_CHECK(hipMemsetAsync(x, 0, sizeof(float) * size_t(c->n_dof), c->stream));
  return pcg_run<float>(c, b64, x, nullptr, c->s_r.get(), c->s_p.get(), c->s_ap.get(),
                        c->diag32.get(), c->val32.get(), max_iter, rtol, check_every, niter,
                        rnorm);
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
    float ms = 0.f;
    UFC_HIP_CHECK(hipEventElapsedTime(&ms, c->spmv_ev[i], c->spmv_ev[i + 1]));
    tot += ms;
    ++k;
  }
  *calls = k;
  return k ? tot / k : 0.0;
}

void axpy_f32_to_f64(ufc_ctx* c, const float* d, double* u) {
  hipLaunchKernelGGL(k_axpy_f32, dim3(blocks_for(c->n_dof, 256)), dim3(256), 0, c->stream,
                     c->n_dof, d, u);
  UFC_HIP_CHECK(hipGetLastError());
}

}  // namespace ufc
