// ufc_api.cpp -- the C-ABI of libufc_hip.so: context lifetime, pattern and
// element-to-slot map construction, load and boundary-condition bundles, and
// the solver orchestration (FP64 / FP32 PCG, FP32-in-FP64 iterative
// refinement, adjoint).  Compiled as HIP host code.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <thread>

#include "ufc_internal.hpp"

using ufc::Error;
using ufc::guarded;
using ufc::set_status;

namespace {

template <typename F>
void parallel_for(int64_t n, F&& f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 65536) nt = 1;
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t b = int64_t(t) * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([&, b, e] {
      for (int64_t i = b; i < e; ++i) f(i);
    });
  }
  for (auto& t : th) t.join();
}

// true when p is device memory of the current device
bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

// device view of a caller array of n elements: the pointer itself when it is
// device memory, else a copy in tmp
template <typename T>
const T* dev_in(ufc_ctx* c, const T* p, size_t n, ufc::DBuf<T>& tmp) {
  if (is_device_ptr(p)) return p;
  tmp.resize(n);
  UFC_HIP_CHECK(hipMemcpyAsync(tmp.get(), p, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return tmp.get();
}

template <typename T>
void copy_out(ufc_ctx* c, T* dst, const T* src, size_t n) {
  UFC_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDefault, c->stream));
  UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
}

void check_ctx(ufc_ctx* c) { UFC_REQUIRE(c != nullptr, UFC_STATUS_INVALID, "null context"); }
void check_mesh(ufc_ctx* c) {
  check_ctx(c);
  UFC_REQUIRE(c->n_nodes > 0 && c->nb_ptr.get(), UFC_STATUS_INVALID,
              "no mesh: call ufc_rt_asm_build_pattern first");
}

// face node table, PH_ELEM_C3D8_FACE_NODES (PH_Elem_C3D8.f90:120-121), 0-based
constexpr int kFaceNodes[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 1, 5, 4},
                                  {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 5, 6, 2}};
constexpr double kGaussPt = 0.577350269189626;

// 2x2 Gauss over face f of element coordinates x (8,3): per point the four
// face shape values (trilinear 1/8 form on the face), the surface Jacobian
// dA and the outward unit normal -- PH_Elem_C3D8_FormFacePressure,
// PH_Elem_C3D8.f90:2360-2541.
template <typename F>
void face_quadrature(const double (&x)[8][3], int f, F&& visit) {
  static const int fixed[6] = {2, 2, 1, 1, 0, 0};
  static const double fval[6] = {-1, 1, -1, 1, -1, 1};
  static const int cu[6] = {0, 0, 0, 0, 1, 1};
  static const int cv[6] = {1, 1, 2, 2, 2, 2};
  static const double sgn[6] = {-1, 1, -1, 1, -1, 1};
  static const double fu[4] = {-kGaussPt, kGaussPt, -kGaussPt, kGaussPt};
  static const double fv[4] = {-kGaussPt, -kGaussPt, kGaussPt, kGaussPt};
  static const double XS[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
  static const double YS[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
  static const double ZS[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  for (int ip = 0; ip < 4; ++ip) {
    double nat[3];
    nat[fixed[f]] = fval[f];
    nat[cu[f]] = fu[ip];
    nat[cv[f]] = fv[ip];
    double N[8], dn[8][3];
    for (int a = 0; a < 8; ++a) {
      const double fx = 1 + XS[a] * nat[0], fy = 1 + YS[a] * nat[1], fz = 1 + ZS[a] * nat[2];
      N[a] = 0.125 * fx * fy * fz;
      dn[a][0] = 0.125 * XS[a] * fy * fz;
      dn[a][1] = 0.125 * fx * YS[a] * fz;
      dn[a][2] = 0.125 * fx * fy * ZS[a];
    }
    double du[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
    for (int q = 0; q < 4; ++q) {
      const int a = kFaceNodes[f][q];
      for (int d = 0; d < 3; ++d) {
        du[d] += dn[a][cu[f]] * x[a][d];
        dv[d] += dn[a][cv[f]] * x[a][d];
      }
    }
    double n[3] = {du[1] * dv[2] - du[2] * dv[1], du[2] * dv[0] - du[0] * dv[2],
                   du[0] * dv[1] - du[1] * dv[0]};
    const double dA = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (dA < 1.0e-15) continue;
    for (double& v : n) v = v / dA;
    for (double& v : n) v *= sgn[f];
    double Nf[4];
    for (int q = 0; q < 4; ++q) Nf[q] = N[kFaceNodes[f][q]];
    visit(Nf, dA, n);
  }
}

// rebuild the device load vector: f = F + sum P*ubar
void upload_rhs(ufc_ctx* c) {
  std::vector<double> h(c->h_f);
  for (size_t i = 0; i < c->h_bc_eq.size(); ++i) h[size_t(c->h_bc_eq[i])] += c->h_bc_rhs[i];
  c->f.resize(size_t(c->n_dof));
  UFC_HIP_CHECK(hipMemcpyAsync(c->f.get(), h.data(), h.size() * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
  UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
}

bool verbose() {
  static const bool v = [] {
    const char* e = std::getenv("UFC_VERBOSE");
    return e && e[0] && e[0] != '0';
  }();
  return v;
}

// FP64 operator for refinement residuals: y = K u (no penalties), matrix-free
void apply_fp64_operator(ufc_ctx* c, const double* u, double* y) {
  if (c->last_material == UFC_MAT_J2)
    ufc::launch_apply_elem64(c, 1, u, y);
  else
    ufc::launch_apply_elem64(c, 0, u, y);
}

// Width of the FP32 view's neighbour index: 0 (symmetric block diagonals:
// the offsets are kernel arguments), 1 (one stencil-pattern byte per node),
// 2 (16-bit offset per neighbour) or 4 (node id per neighbour).
int32_t nb_index_bytes(const ufc_ctx* c) {
  switch (c->fp32_view) {
    case ufc::kFp32ViewSymDiag:
      return 0;
    case ufc::kFp32ViewPattern:
      return 1;
    case ufc::kFp32ViewOff16:
      return 2;
    default:
      return 4;
  }
}

bool fp32_view_available(const ufc_ctx* c, int view) {
  switch (view) {
    case ufc::kFp32ViewSymDiag:
      return c->n_diag > 0;
    case ufc::kFp32ViewPattern:
      return c->n_pat > 0;
    case ufc::kFp32ViewOff16:
      return c->have_off16;
    case ufc::kFp32ViewId32:
      return true;
    default:
      return false;
  }
}

// Lexicographic numbering of a mesh that is a structured box of hexes under
// some node numbering: perm[n] = i + NX (j + NY k) for the lattice position
// (i, j, k) of node n, or an empty vector when the mesh is not such a box.
// The lattice is recovered from the element edges (the 12 edges of the C3D8
// node order): a corner node (one element, three edge neighbours) is the
// origin and its edge neighbours the three axes; a lattice line continues
// from (prev, cur) to the edge neighbour of cur that shares no element with
// prev; every other position is the one common edge neighbour of two
// labelled positions besides the one they already share.  The result is
// accepted only if it labels every node once and every element is a unit
// cell of the lattice.  node_elem: (n_nodes, 8) element ids, -1 padded.
std::vector<int32_t> box_numbering(int32_t nn, const std::vector<int32_t>& conn,
                                   const std::vector<int32_t>& node_elem) {
  static const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                   {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  const int64_t ne = int64_t(conn.size()) / 8;
  std::vector<std::array<int32_t, 6>> gn(static_cast<size_t>(nn));
  std::vector<uint8_t> gdeg(static_cast<size_t>(nn), 0);
  for (int64_t e = 0; e < ne; ++e)
    for (const auto& ed : kEdge) {
      const int32_t a = conn[8 * e + ed[0]], b = conn[8 * e + ed[1]];
      for (int t = 0; t < 2; ++t) {
        const int32_t x = t ? b : a, y = t ? a : b;
        auto& L = gn[size_t(x)];
        const int d = gdeg[size_t(x)];
        if (std::find(L.begin(), L.begin() + d, y) != L.begin() + d) continue;
        if (d == 6) return {};  // more than 6 edge neighbours: no lattice
        L[size_t(d)] = y;
        gdeg[size_t(x)] = uint8_t(d + 1);
      }
    }
  auto elems_of = [&](int32_t n) { return node_elem.data() + 8 * size_t(n); };
  auto share_elem = [&](int32_t a, int32_t b) {
    const int32_t* E = elems_of(a);
    for (int s = 0; s < 8 && E[s] >= 0; ++s)
      for (int q = 0; q < 8; ++q)
        if (conn[8 * size_t(E[s]) + q] == b) return true;
    return false;
  };
  auto n_elems_of = [&](int32_t n) {
    int k = 0;
    while (k < 8 && elems_of(n)[k] >= 0) ++k;
    return k;
  };
  int32_t corner = -1;
  for (int32_t n = 0; n < nn && corner < 0; ++n)
    if (n_elems_of(n) == 1 && gdeg[size_t(n)] == 3) corner = n;
  if (corner < 0) return {};
  // the three axis lines from the corner
  std::vector<int32_t> axis[3];
  for (int a = 0; a < 3; ++a) {
    std::vector<int32_t>& line = axis[a];
    line.push_back(corner);
    line.push_back(gn[size_t(corner)][size_t(a)]);
    while (true) {
      const int32_t prev = line[line.size() - 2], cur = line.back();
      int32_t next = -1, cnt = 0;
      for (int q = 0; q < gdeg[size_t(cur)]; ++q) {
        const int32_t cand = gn[size_t(cur)][size_t(q)];
        if (cand != prev && !share_elem(prev, cand)) {
          next = cand;
          ++cnt;
        }
      }
      if (cnt == 0) break;
      if (cnt > 1 || int64_t(line.size()) > nn) return {};
      line.push_back(next);
    }
  }
  const int64_t NX = int64_t(axis[0].size()), NY = int64_t(axis[1].size()),
                NZ = int64_t(axis[2].size());
  if (NX * NY * NZ != nn || (NX - 1) * (NY - 1) * (NZ - 1) != ne) return {};
  std::vector<int32_t> lab(static_cast<size_t>(nn), -1);
  auto at = [&](int64_t i, int64_t j, int64_t k) -> int32_t& {
    return lab[size_t(i + NX * (j + NY * k))];
  };
  for (int64_t i = 0; i < NX; ++i) at(i, 0, 0) = axis[0][size_t(i)];
  for (int64_t j = 0; j < NY; ++j) at(0, j, 0) = axis[1][size_t(j)];
  for (int64_t k = 0; k < NZ; ++k) at(0, 0, k) = axis[2][size_t(k)];
  // the one edge neighbour common to a and b other than x
  auto common = [&](int32_t a, int32_t b, int32_t x) -> int32_t {
    int32_t r = -1;
    for (int p = 0; p < gdeg[size_t(a)]; ++p) {
      const int32_t cand = gn[size_t(a)][size_t(p)];
      if (cand == x) continue;
      for (int q = 0; q < gdeg[size_t(b)]; ++q)
        if (gn[size_t(b)][size_t(q)] == cand) {
          if (r >= 0 && r != cand) return -2;
          r = cand;
        }
    }
    return r;
  };
  for (int64_t k = 0; k < NZ; ++k)
    for (int64_t j = 0; j < NY; ++j)
      for (int64_t i = 0; i < NX; ++i) {
        if (at(i, j, k) >= 0) continue;
        int32_t v;
        if (k == 0)
          v = common(at(i - 1, j, 0), at(i, j - 1, 0), at(i - 1, j - 1, 0));
        else if (i > 0)
          v = common(at(i - 1, j, k), at(i, j, k - 1), at(i - 1, j, k - 1));
        else
          v = common(at(0, j - 1, k), at(0, j, k - 1), at(0, j - 1, k - 1));
        if (v < 0) return {};
        at(i, j, k) = v;
      }
  std::vector<int32_t> perm(static_cast<size_t>(nn), -1);
  for (int64_t p = 0; p < nn; ++p) {
    const int32_t n = lab[size_t(p)];
    if (perm[size_t(n)] >= 0) return {};  // a node labelled twice
    perm[size_t(n)] = int32_t(p);
  }
  // every element a unit cell of the lattice
  for (int64_t e = 0; e < ne; ++e) {
    int64_t lo[3] = {NX, NY, NZ};
    int64_t pos[8][3];
    for (int a = 0; a < 8; ++a) {
      const int64_t p = perm[size_t(conn[8 * e + a])];
      pos[a][0] = p % NX;
      pos[a][1] = (p / NX) % NY;
      pos[a][2] = p / (NX * NY);
      for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], pos[a][d]);
    }
    int seen = 0;
    for (int a = 0; a < 8; ++a) {
      int bit = 0;
      for (int d = 0; d < 3; ++d) {
        const int64_t r = pos[a][d] - lo[d];
        if (r > 1) return {};
        bit |= int(r) << d;
      }
      seen |= 1 << bit;
    }
    if (seen != 0xff) return {};
  }
  return perm;
}

// Offsets of the symmetric block-diagonal view: the distinct m - n > 0 over
// all neighbour pairs (n, m), in the node numbering or, when that has too
// many and the mesh is a structured box, in the box's lexicographic
// numbering (c->sym_permuted).  More than kMaxDiags of them: no such view.
// Chunks of nodes collect their offsets in parallel and stop early once a
// chunk alone has too many.
bool collect_sym_offsets(ufc_ctx* c, const std::vector<int32_t>& nb_idx,
                         const std::vector<int32_t>* perm, std::vector<int32_t>& all);

// The LDS x window of the symmetric SpMV (XWindow): the offsets {0, +-o_s}
// sorted and merged into runs wherever consecutive offsets are at most a tile
// apart; none when more than kMaxXClusters runs or kMaxXWin nodes are needed
// (the SpMV then gathers x from global memory).
void build_x_window(ufc_ctx* c) {
  c->have_xwin = false;
  const int32_t tile = c->diag_tile;
  std::vector<int32_t> offs{0};
  for (int s = 0; s < c->n_diag; ++s) {
    offs.push_back(c->h_diag_off[s]);
    offs.push_back(-c->h_diag_off[s]);
  }
  std::sort(offs.begin(), offs.end());
  ufc::XWindow w{};
  int32_t total = 0;
  size_t i = 0;
  while (i < offs.size()) {
    size_t j = i;
    while (j + 1 < offs.size() && offs[j + 1] - offs[j] <= tile) ++j;
    if (w.nc == ufc::kMaxXClusters) return;
    w.lo[w.nc] = offs[i];
    w.base[w.nc] = total;
    w.len[w.nc] = offs[j] - offs[i] + tile;
    total += w.len[w.nc];
    ++w.nc;
    i = j + 1;
  }
  if (total > ufc::max_xwin(tile)) return;
  auto slot = [&](int32_t o) {
    for (int k = 0; k < w.nc; ++k)
      if (o >= w.lo[k] && o < w.lo[k] + w.len[k] - tile + 1) return w.base[k] + o - w.lo[k];
    return -1;
  };
  w.self = slot(0);
  for (int s = 0; s < ufc::kMaxDiags; ++s) {
    w.up[s] = s < c->n_diag ? slot(c->h_diag_off[s]) : w.self;
    w.dn[s] = s < c->n_diag ? slot(-c->h_diag_off[s]) : w.self;
    if (w.up[s] < 0 || w.dn[s] < 0) return;
  }
  c->xwin = w;
  c->have_xwin = true;
}

void reset_sym_diagonals(ufc_ctx* c) {
  c->n_diag = 0;
  c->diag_tile = ufc::kDiagTile;
  c->sym_permuted = false;
  c->h_sym_perm.clear();
  c->sym_perm.release();
  c->sym_iperm.release();
  c->diag32p.release();
  c->diag_a.release();
  c->diag_b.release();
  c->diag_c.release();
  c->diag_from_assembly = false;
  c->have_xwin = false;
}

void build_sym_diagonals(ufc_ctx* c, const std::vector<int32_t>& nb_idx,
                         const std::vector<int32_t>& node_elem) {
  reset_sym_diagonals(c);
  const int64_t nn = int64_t(c->h_nb_ptr.size()) - 1;
  std::vector<int32_t> all;
  if (!collect_sym_offsets(c, nb_idx, nullptr, all)) {
    std::vector<int32_t> perm = box_numbering(int32_t(nn), c->h_conn, node_elem);
    if (perm.empty() || !collect_sym_offsets(c, nb_idx, &perm, all)) return;
    c->sym_permuted = true;
    c->h_sym_perm = std::move(perm);
  }
  if (all.empty()) return;
  std::sort(all.begin(), all.end());
  // tile-major entries with ceil(o_last / tile) zero tiles in front; the
  // SpMV indexes them with 32-bit integers
  const int64_t tile = nn >= ufc::kDiagWideMinNodes ? ufc::kDiagTileWide : ufc::kDiagTile;
  const int64_t pad_tiles = (int64_t(all.back()) + tile - 1) / tile;
  const int64_t tiles = (nn + tile - 1) / tile + pad_tiles;
  const int64_t entries = tiles * int64_t(all.size()) * tile;
  if (entries >= (int64_t(1) << 31)) {
    c->sym_permuted = false;
    c->h_sym_perm.clear();
    return;
  }
  c->n_diag = int32_t(all.size());
  for (int s = 0; s < ufc::kMaxDiags; ++s) c->h_diag_off[s] = s < c->n_diag ? all[size_t(s)] : 0;
  c->diag_pad_tiles = int32_t(pad_tiles);
  c->diag_tile = int32_t(tile);
  build_x_window(c);
  c->diag_entries = entries;
  if (c->sym_permuted) {
    std::vector<int32_t> iperm(static_cast<size_t>(nn));
    for (int64_t n = 0; n < nn; ++n) iperm[size_t(c->h_sym_perm[size_t(n)])] = int32_t(n);
    c->sym_perm.resize(size_t(nn));
    c->sym_iperm.resize(size_t(nn));
    UFC_HIP_CHECK(hipMemcpyAsync(c->sym_perm.get(), c->h_sym_perm.data(), c->sym_perm.bytes(),
                                 hipMemcpyHostToDevice, c->stream));
    UFC_HIP_CHECK(hipMemcpyAsync(c->sym_iperm.get(), iperm.data(), c->sym_iperm.bytes(),
                                 hipMemcpyHostToDevice, c->stream));
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
  }
}

bool collect_sym_offsets(ufc_ctx* c, const std::vector<int32_t>& nb_idx,
                         const std::vector<int32_t>* perm, std::vector<int32_t>& all) {
  all.clear();
  const int64_t nn = int64_t(c->h_nb_ptr.size()) - 1;
  constexpr int64_t kChunk = 32;  // enough chunks for parallel_for's threads on large meshes
  const int64_t nchunk = (nn + kChunk - 1) / kChunk;
  constexpr int kCap = ufc::kMaxDiags + 1;
  std::vector<std::array<int32_t, kCap>> found(static_cast<size_t>(nchunk));
  std::vector<int> count(static_cast<size_t>(nchunk), 0);
  parallel_for(nchunk, [&](int64_t ch) {
    auto& f = found[size_t(ch)];
    int& k = count[size_t(ch)];
    for (int64_t n = ch * kChunk; n < std::min(nn, (ch + 1) * kChunk) && k < kCap; ++n)
      for (int32_t q = c->h_nb_ptr[size_t(n)]; q < c->h_nb_ptr[size_t(n) + 1] && k < kCap; ++q) {
        const int64_t m = nb_idx[size_t(q)];
        const int64_t o = perm ? int64_t((*perm)[size_t(m)]) - (*perm)[size_t(n)] : m - n;
        if (o <= 0) continue;
        if (std::find(f.begin(), f.begin() + k, int32_t(o)) == f.begin() + k) f[size_t(k++)] = int32_t(o);
      }
  });
  for (int64_t ch = 0; ch < nchunk; ++ch) {
    for (int j = 0; j < count[size_t(ch)]; ++j)
      if (std::find(all.begin(), all.end(), found[size_t(ch)][size_t(j)]) == all.end())
        all.push_back(found[size_t(ch)][size_t(j)]);
    if (int(all.size()) > ufc::kMaxDiags) return false;
  }
  return true;
}

// Stencil-pattern dictionary of the 16-bit view (off16, CSR pair order):
// every node's (degree, offsets) list gets an id; with at most kMaxPatterns
// distinct lists the ids and the list table go to the device, otherwise
// n_pat = 0 and the SpMV keeps the per-neighbour offsets.  Lists are matched
// by a 64-bit hash and then compared element by element, so a hash collision
// only disables the view.
void build_stencil_patterns(ufc_ctx* c, const std::vector<int16_t>& off16) {
  c->n_pat = 0;
  c->nb_pat.release();
  c->pat_off.release();
  if (!c->have_off16) return;
  const int64_t nn = int64_t(c->h_nb_ptr.size()) - 1;
  auto list = [&](int64_t n, int* deg) {
    const int32_t q0 = c->h_nb_ptr[size_t(n)];
    *deg = c->h_nb_ptr[size_t(n) + 1] - q0;
    return off16.data() + q0;
  };
  std::vector<uint64_t> hash(static_cast<size_t>(nn));
  parallel_for(nn, [&](int64_t n) {
    int deg;
    const int16_t* o = list(n, &deg);
    uint64_t h = 1469598103934665603ull ^ uint64_t(deg);  // FNV-1a
    for (int k = 0; k < deg; ++k) h = (h ^ uint64_t(uint16_t(o[k]))) * 1099511628211ull;
    hash[size_t(n)] = h;
  });
  std::map<uint64_t, int> ids;
  std::vector<int64_t> rep;  // first node of each pattern
  const size_t padded = (size_t(nn) + 3) / 4 * 4 + 4;
  std::vector<uint8_t> pat(padded, 0);
  for (int64_t n = 0; n < nn; ++n) {
    auto it = ids.find(hash[size_t(n)]);
    if (it == ids.end()) {
      if (int(rep.size()) == ufc::kMaxPatterns) return;
      it = ids.emplace(hash[size_t(n)], int(rep.size())).first;
      rep.push_back(n);
    }
    pat[size_t(n)] = uint8_t(it->second);
  }
  std::atomic<int> mismatch{0};
  parallel_for(nn, [&](int64_t n) {
    int d0, d1;
    const int16_t* a = list(n, &d0);
    const int16_t* b = list(rep[pat[size_t(n)]], &d1);
    if (d0 != d1 || !std::equal(a, a + d0, b)) mismatch.store(1, std::memory_order_relaxed);
  });
  if (mismatch.load()) return;
  const int np = int(rep.size());
  // table row: the offsets of the neighbours other than the node itself in
  // ascending order, then 0 (the node itself) at position deg - 1 -- the lane
  // order of the off-diagonal values (val32m)
  std::vector<int16_t> table(size_t(np) * ufc::kMaxNodeNbrs, 0);
  for (int p = 0; p < np; ++p) {
    int deg;
    const int16_t* o = list(rep[size_t(p)], &deg);
    int16_t* row = table.data() + size_t(p) * ufc::kMaxNodeNbrs;
    int j = 0;
    for (int k = 0; k < deg; ++k)
      if (o[k] != 0) row[j++] = o[k];
    if (j != deg - 1) return;  // every list holds the node itself once
    row[deg - 1] = 0;
  }
  c->nb_pat.resize(padded);
  c->pat_off.resize(table.size());
  UFC_HIP_CHECK(hipMemcpyAsync(c->nb_pat.get(), pat.data(), c->nb_pat.bytes(),
                               hipMemcpyHostToDevice, c->stream));
  UFC_HIP_CHECK(hipMemcpyAsync(c->pat_off.get(), table.data(), c->pat_off.bytes(),
                               hipMemcpyHostToDevice, c->stream));
  UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
  c->n_pat = np;
}

}  // namespace

namespace ufc {

// The fastest view first: symmetric block diagonals (each off-diagonal block
// read once from HBM), then stencil patterns, 16-bit offsets, 4-byte ids.
// The off-diagonal copy of the pattern view is dropped when another view is
// chosen, and the block diagonals likewise.
void select_fp32_view(ufc_ctx* c) {
  int v = c->fp32_view_request;
  if (v < 0 || !fp32_view_available(c, v)) {
    v = kFp32ViewId32;
    for (int w : {kFp32ViewSymDiag, kFp32ViewPattern, kFp32ViewOff16})
      if (fp32_view_available(c, w)) {
        v = w;
        break;
      }
  }
  c->fp32_view = v;
  if (v != kFp32ViewPattern) c->val32m.release();
  if (v != kFp32ViewSymDiag) {
    c->diag_a.release();
    c->diag_b.release();
    c->diag_c.release();
    c->diag_from_assembly = false;
  }
}

}  // namespace ufc

extern "C" {

// ======================================================================= ctx
int ufc_ctx_create(int32_t device, ufc_ctx** ctx, ufc_status* st) {
  return guarded(st, "ufc_ctx_create", [&](std::string& msg) {
    UFC_REQUIRE(ctx != nullptr, UFC_STATUS_INVALID, "ctx out-pointer is null");
    int ndev = 0;
    UFC_HIP_CHECK(hipGetDeviceCount(&ndev));
    UFC_REQUIRE(device >= 0 && device < ndev, UFC_STATUS_INVALID, "no such HIP device");
    UFC_HIP_CHECK(hipSetDevice(device));
    ufc::noalloc_setup(device);
    auto* c = new ufc_ctx();
    c->device = device;
    try {
      UFC_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      UFC_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_scalars), sizeof(ufc::PcgScalars),
                                  hipHostMallocDefault));
      UFC_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_progress), 2 * sizeof(int32_t),
                                  hipHostMallocMapped | hipHostMallocCoherent));
      c->h_progress[0] = c->h_progress[1] = 0;
      UFC_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_progress),
                                            c->h_progress, 0));
      UFC_HIP_CHECK(hipEventCreate(&c->ev0));
      UFC_HIP_CHECK(hipEventCreate(&c->ev1));
      UFC_HIP_CHECK(hipEventCreate(&c->ev2));
      UFC_HIP_CHECK(hipEventCreate(&c->ev3));
    } catch (...) {
      delete c;
      throw;
    }
    *ctx = c;
    msg = "context created";
    return UFC_STATUS_OK;
  });
}

int ufc_ctx_destroy(ufc_ctx* c, ufc_status* st) {
  return guarded(st, "ufc_ctx_destroy", [&](std::string& msg) {
    if (!c) return UFC_STATUS_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->spmv_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : {c->ev0, c->ev1, c->ev2, c->ev3})
      if (e) (void)hipEventDestroy(e);
    if (c->h_scalars) (void)hipHostFree(c->h_scalars);
    if (c->h_progress) (void)hipHostFree(c->h_progress);
    hipStream_t s = c->stream;
    delete c;  // frees device buffers
    if (s) (void)hipStreamDestroy(s);
    msg = "context destroyed";
    return UFC_STATUS_OK;
  });
}

int ufc_ctx_info(ufc_ctx* c, ufc_rt_asm_info* info, ufc_status* st) {
  return guarded(st, "ufc_ctx_info", [&](std::string&) {
    check_ctx(c);
    UFC_REQUIRE(info != nullptr, UFC_STATUS_INVALID, "info is null");
    info->n_dof = c->n_dof;
    info->nnz = c->nnz;
    info->n_nodes = c->n_nodes;
    info->n_elems = c->n_elems;
    info->max_row_nnz = c->max_row_nnz;
    info->nb_index_bytes = nb_index_bytes(c);
    info->fp32_view = c->fp32_view;
    info->sym_tile = c->n_diag > 0 ? c->diag_tile : 0;
    return UFC_STATUS_OK;
  });
}

int ufc_rt_asm_set_fp32_view(ufc_ctx* c, int32_t view, ufc_status* st) {
  return guarded(st, "ufc_rt_asm_set_fp32_view", [&](std::string& msg) {
    check_mesh(c);
    UFC_REQUIRE(view >= -1 && view <= ufc::kFp32ViewSymDiag, UFC_STATUS_INVALID,
                "view must be -1 (fastest) or 0-3");
    UFC_REQUIRE(view < 0 || fp32_view_available(c, view), UFC_STATUS_UNSUPPORTED,
                "the mesh does not allow FP32 view " + std::to_string(view));
    UFC_HIP_CHECK(hipSetDevice(c->device));
    c->fp32_view_request = view;
    const int before = c->fp32_view;
    ufc::select_fp32_view(c);
    // FP32 values already assembled: build the new view's layout from them
    if (c->have32 && c->fp32_view != before) {
      ufc::launch_fp32_offdiag(c);
      UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    msg = "FP32 view " + std::to_string(c->fp32_view);
    return UFC_STATUS_OK;
  });
}

int ufc_rt_asm_set_fp64_view(ufc_ctx* c, int32_t view, ufc_status* st) {
  return guarded(st, "ufc_rt_asm_set_fp64_view", [&](std::string& msg) {
    check_mesh(c);
    UFC_REQUIRE(view == ufc::kFp64ViewCsr || view == ufc::kFp64ViewBlock, UFC_STATUS_INVALID,
                "view must be 0 (CSR) or 1 (node blocks)");
    c->fp64_view = view;
    msg = "FP64 view " + std::to_string(view);
    return UFC_STATUS_OK;
  });
}

int ufc_ctx_enable_timing(ufc_ctx* c, int32_t per_kernel, ufc_status* st) {
  return guarded(st, "ufc_ctx_enable_timing", [&](std::string&) {
    check_ctx(c);
    UFC_REQUIRE(per_kernel >= 0, UFC_STATUS_INVALID, "per_kernel < 0");
    c->per_kernel = per_kernel;
    return UFC_STATUS_OK;
  });
}

int ufc_ctx_timing(ufc_ctx* c, ufc_timing* t, ufc_status* st) {
  return guarded(st, "ufc_ctx_timing", [&](std::string&) {
    check_ctx(c);
    UFC_REQUIRE(t != nullptr, UFC_STATUS_INVALID, "timing out-pointer is null");
    t->assemble_ms = c->assemble_ms;
    t->solve_ms = c->solve_ms;
    t->spmv_ms_avg = c->spmv_ms_avg;
    t->spmv_calls = c->spmv_calls;
    t->n_replace = c->n_replace;
    t->j2_update_ms = c->j2_update_ms;
    return UFC_STATUS_OK;
  });
}

// ======================================================================= L4_PH
int ufc_ph_c3d8_eval_ke(ufc_ctx* c, const ufc_ph_ke_algo* algo, ufc_ph_ke_args* a,
                        ufc_status* st) {
  return guarded(st, "ufc_ph_c3d8_eval_ke", [&](std::string&) {
    check_ctx(c);
    UFC_REQUIRE(algo && a && a->coords && a->young && a->ke, UFC_STATUS_INVALID,
                "null argument");
    UFC_REQUIRE(a->n_elems >= 0, UFC_STATUS_INVALID, "n_elems < 0");
    UFC_REQUIRE(algo->precision == UFC_PREC_FP64 || algo->precision == UFC_PREC_FP32,
                UFC_STATUS_INVALID, "precision must be FP64 or FP32");
    UFC_REQUIRE(a->nu > -1.0 && a->nu < 0.5, UFC_STATUS_INVALID, "nu out of (-1, 0.5)");
    if (a->n_elems == 0) return UFC_STATUS_OK;
    UFC_HIP_CHECK(hipSetDevice(c->device));
    const size_t ne = size_t(a->n_elems);
    ufc::DBuf<double> tc, ty;
    const double* dc = dev_in(c, a->coords, ne * 24, tc);
    const double* dy = dev_in(c, a->young, ne, ty);
    const int variant = ufc::canonical_variant(algo->variant);
    UFC_REQUIRE(variant > 0, UFC_STATUS_INVALID, "unknown element variant");
    const bool fp32 = algo->precision == UFC_PREC_FP32;
    const size_t esz = fp32 ? sizeof(float) : sizeof(double);
    void* out = a->ke;
    ufc::DBuf<char> tk;
    if (!is_device_ptr(a->ke)) {
      tk.resize(ne * 576 * esz);
      out = tk.get();
    }
    ufc::DBuf<double> tcf;
    if (variant == ufc::kVarIncompat) tcf.resize(ne * ufc::kIncompatC);
    ufc::launch_eval_ke(a->n_elems, dc, dy, a->nu, out, fp32, variant, tcf.get(), c->stream);
    if (out != a->ke)
      UFC_HIP_CHECK(hipMemcpyAsync(a->ke, out, ne * 576 * esz, hipMemcpyDeviceToHost, c->stream));
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    return UFC_STATUS_OK;
  });
}

int ufc_ph_c3d8_sensitivity(ufc_ctx* c, const ufc_ph_sens_args* a, ufc_status* st) {
  return guarded(st, "ufc_ph_c3d8_sensitivity", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(a && a->u && a->lambda && a->dedrho && a->grad, UFC_STATUS_INVALID,
                "null argument");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    ufc::DBuf<double> tu, tl, td, tg;
    const double* du = dev_in(c, a->u, size_t(c->n_dof), tu);
    const double* dl = dev_in(c, a->lambda, size_t(c->n_dof), tl);
    const double* dd = dev_in(c, a->dedrho, size_t(c->n_elems), td);
    double* g = a->grad;
    if (!is_device_ptr(g)) {
      tg.resize(size_t(c->n_elems));
      g = tg.get();
    }
    ufc::launch_sensitivity(c, du, dl, dd, a->nu, g);
    if (g != a->grad) copy_out(c, a->grad, g, size_t(c->n_elems));
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    return UFC_STATUS_OK;
  });
}

int ufc_ph_j2_init(ufc_ctx* c, const ufc_ph_j2_props* p, ufc_status* st) {
  return guarded(st, "ufc_ph_j2_init", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(p != nullptr, UFC_STATUS_INVALID, "props is null");
    // PH_J2_Init validation, PH_Mat_Plast_J2_Iso_Core.f90:220-245
    UFC_REQUIRE(p->young > 0.0 && p->nu >= 0.0 && p->nu < 0.5, UFC_STATUS_ERROR,
                "Invalid elastic parameters (E<=0 or nu out of range)");
    UFC_REQUIRE(p->sigma_y0 > 0.0, UFC_STATUS_ERROR, "sigma_y0 must be positive");
    ufc_ph_j2_props law = *p;
    if (law.hardening_type == 0) law.hardening_type = UFC_J2_HARD_LINEAR;
    UFC_REQUIRE(law.hardening_type >= UFC_J2_HARD_LINEAR && law.hardening_type <= UFC_J2_HARD_AF,
                UFC_STATUS_INVALID, "unknown hardening_type");
    // the reference updates the Armstrong-Frederick backstress
    // (PH_Mat_Plast_J2_Iso_Core.f90:449-453) but never subtracts it from the
    // trial deviator (:254-317): it cannot change a result, so it is refused
    UFC_REQUIRE(law.hardening_type != UFC_J2_HARD_AF && law.af_c == 0.0 && law.af_gamma == 0.0,
                UFC_STATUS_UNSUPPORTED,
                "Armstrong-Frederick kinematic hardening is not supported (the reference's "
                "backstress never enters its yield function)");
    if (law.hardening_type == UFC_J2_HARD_SWIFT)
      UFC_REQUIRE(law.swift_k > 0.0 && law.swift_n >= 0.0 && law.swift_eps0 > 0.0,
                  UFC_STATUS_ERROR, "Swift hardening needs K > 0, n >= 0 and eps0 > 0");
    if (law.hardening_type == UFC_J2_HARD_VOCE)
      UFC_REQUIRE(law.voce_delta >= 0.0 && std::isfinite(law.voce_q), UFC_STATUS_ERROR,
                  "Voce hardening needs delta >= 0 and a finite sigma_inf");
    UFC_REQUIRE(law.max_local_iter <= 1000, UFC_STATUS_INVALID, "max_local_iter > 1000");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    const size_t ngp = size_t(c->n_elems) * 8;
    c->j2 = law;
    c->j2_commit.resize(ngp * 13);
    c->j2_trial.resize(ngp * 13);
    c->j2_tangent.resize(ngp * ufc::kJ2Tangent);
    c->j2_flags.resize(ngp);
    UFC_HIP_CHECK(hipMemsetAsync(c->j2_commit.get(), 0, c->j2_commit.bytes(), c->stream));
    UFC_HIP_CHECK(hipMemsetAsync(c->j2_trial.get(), 0, c->j2_trial.bytes(), c->stream));
    UFC_HIP_CHECK(hipMemsetAsync(c->j2_tangent.get(), 0, c->j2_tangent.bytes(), c->stream));
    UFC_HIP_CHECK(hipMemsetAsync(c->j2_flags.get(), 0, c->j2_flags.bytes(), c->stream));
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->have_j2 = true;
    return UFC_STATUS_OK;
  });
}

int ufc_ph_j2_update(ufc_ctx* c, const double* u, const double* u_n, int64_t* n_yield,
                     ufc_status* st) {
  return guarded(st, "ufc_ph_j2_update", [&](std::string& msg) {
    check_mesh(c);
    UFC_REQUIRE(c->have_j2, UFC_STATUS_INVALID, "call ufc_ph_j2_init first");
    UFC_REQUIRE(u && u_n, UFC_STATUS_INVALID, "null argument");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    ufc::DBuf<double> tu, tn;
    const double* du = dev_in(c, u, size_t(c->n_dof), tu);
    const double* dn = dev_in(c, u_n, size_t(c->n_dof), tn);
    UFC_HIP_CHECK(hipEventRecord(c->ev0, c->stream));
    ufc::launch_j2_update(c, du, dn);
    UFC_HIP_CHECK(hipEventRecord(c->ev1, c->stream));
    std::vector<int32_t> fl(static_cast<size_t>(c->n_elems) * 8);
    UFC_HIP_CHECK(hipMemcpyAsync(fl.data(), c->j2_flags.get(), fl.size() * sizeof(int32_t),
                                 hipMemcpyDeviceToHost, c->stream));
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    UFC_HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->j2_update_ms = ms;
    int64_t ny = 0, nfail = 0;
    for (int32_t f : fl) {
      ny += (f == 1);
      nfail += (f == 2);
    }
    if (n_yield) *n_yield = ny;
    if (nfail) {
      // local Newton failure -> the caller should cut the increment back
      // (pnewdt = 0.25, PH_Mat_Plast_J2_Iso_Core.f90:419-427)
      msg = "J2 local Newton failed at " + std::to_string(nfail) + " Gauss points";
      return UFC_STATUS_ERROR;
    }
    return UFC_STATUS_OK;
  });
}

int ufc_ph_j2_commit(ufc_ctx* c, ufc_status* st) {
  return guarded(st, "ufc_ph_j2_commit", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(c->have_j2, UFC_STATUS_INVALID, "call ufc_ph_j2_init first");
    UFC_HIP_CHECK(hipMemcpyAsync(c->j2_commit.get(), c->j2_trial.get(), c->j2_trial.bytes(),
                                 hipMemcpyDeviceToDevice, c->stream));
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    return UFC_STATUS_OK;
  });
}

int ufc_ph_j2_get_state(ufc_ctx* c, double* state, ufc_status* st) {
  return guarded(st, "ufc_ph_j2_get_state", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(c->have_j2 && state, UFC_STATUS_INVALID, "no J2 state or null argument");
    copy_out(c, state, c->j2_commit.get(), c->j2_commit.n);
    return UFC_STATUS_OK;
  });
}

int ufc_ph_j2_set_state(ufc_ctx* c, const double* state, ufc_status* st) {
  return guarded(st, "ufc_ph_j2_set_state", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(c->have_j2 && state, UFC_STATUS_INVALID, "no J2 state or null argument");
    UFC_HIP_CHECK(hipMemcpyAsync(c->j2_commit.get(), state, c->j2_commit.bytes(),
                                 hipMemcpyDefault, c->stream));
    UFC_HIP_CHECK(hipMemcpyAsync(c->j2_trial.get(), c->j2_commit.get(), c->j2_commit.bytes(),
                                 hipMemcpyDeviceToDevice, c->stream));
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    return UFC_STATUS_OK;
  });
}

// ======================================================================= L5_RT
int ufc_rt_asm_build_pattern(ufc_ctx* c, const ufc_mesh_desc* m, ufc_rt_asm_info* info,
                             ufc_status* st) {
  return guarded(st, "ufc_rt_asm_build_pattern", [&](std::string& msg) {
    check_ctx(c);
    UFC_REQUIRE(m && m->coords && m->conn, UFC_STATUS_INVALID, "null mesh argument");
    UFC_REQUIRE(m->n_nodes > 0 && m->n_elems > 0, UFC_STATUS_INVALID, "empty mesh");
    UFC_REQUIRE(int64_t(m->n_nodes) * 3 < INT32_MAX, UFC_STATUS_UNSUPPORTED,
                "3*n_nodes exceeds int32 equation ids");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    const int32_t nn = m->n_nodes, ne = m->n_elems;
    const int base = m->index_base;
    // mesh copies (0-based connectivity), validated
    c->h_coords.assign(m->coords, m->coords + size_t(nn) * 3);
    c->h_conn.resize(size_t(ne) * 8);
    for (size_t i = 0; i < c->h_conn.size(); ++i) {
      const int32_t v = m->conn[i] - base;
      UFC_REQUIRE(v >= 0 && v < nn, UFC_STATUS_INVALID,
                  "connectivity entry out of range at element " + std::to_string(i / 8 + base));
      c->h_conn[i] = v;
    }
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < 8; ++a)
        for (int b = a + 1; b < 8; ++b)
          UFC_REQUIRE(c->h_conn[8 * e + a] != c->h_conn[8 * e + b], UFC_STATUS_INVALID,
                      "element " + std::to_string(e + base) + " repeats a node");
    // node -> (element, local index), elements ascending
    std::vector<int32_t> cnt(static_cast<size_t>(nn), 0);
    for (int32_t v : c->h_conn) cnt[size_t(v)]++;
    int32_t maxcnt = 0;
    for (int32_t n = 0; n < nn; ++n) {
      // a node in no element has an empty CSR row and no diagonal: its
      // equations could be neither solved nor constrained
      UFC_REQUIRE(cnt[size_t(n)] > 0, UFC_STATUS_INVALID,
                  "node " + std::to_string(n + base) + " belongs to no element");
      maxcnt = std::max(maxcnt, cnt[size_t(n)]);
    }
    std::vector<int64_t> ne_ptr(size_t(nn) + 1, 0);
    for (int32_t n = 0; n < nn; ++n) ne_ptr[size_t(n) + 1] = ne_ptr[size_t(n)] + cnt[size_t(n)];
    std::vector<int32_t> ne_elem(size_t(ne) * 8);
    std::vector<uint8_t> ne_loc(size_t(ne) * 8);
    {
      std::vector<int64_t> at(ne_ptr.begin(), ne_ptr.end() - 1);
      for (int32_t e = 0; e < ne; ++e)
        for (int a = 0; a < 8; ++a) {
          const int32_t n = c->h_conn[8 * size_t(e) + a];
          const int64_t k = at[size_t(n)]++;
          ne_elem[size_t(k)] = e;
          ne_loc[size_t(k)] = uint8_t(a);
        }
    }
    // the slot map: the first kMaxNodeElems elements of every node
    std::vector<int32_t> s_elem(static_cast<size_t>(nn) * 8, -1);
    std::vector<uint8_t> s_loc(static_cast<size_t>(nn) * 8, 0);
    parallel_for(nn, [&](int64_t n) {
      const int64_t q0 = ne_ptr[size_t(n)];
      const int k = int(std::min<int64_t>(ufc::kMaxNodeElems, ne_ptr[size_t(n) + 1] - q0));
      for (int s = 0; s < k; ++s) {
        s_elem[8 * size_t(n) + s] = ne_elem[size_t(q0 + s)];
        s_loc[8 * size_t(n) + s] = ne_loc[size_t(q0 + s)];
      }
    });
    // neighbour lists (sorted, unique): the union of the node's elements
    auto node_list = [&](int64_t n, std::vector<int32_t>& buf) {
      buf.clear();
      for (int64_t q = ne_ptr[size_t(n)]; q < ne_ptr[size_t(n) + 1]; ++q) {
        const int32_t* en = c->h_conn.data() + 8 * size_t(ne_elem[size_t(q)]);
        buf.insert(buf.end(), en, en + 8);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    };
    std::vector<int32_t> deg(static_cast<size_t>(nn));
    std::vector<int32_t> bad(1, -1);
    parallel_for(nn, [&](int64_t n) {
      thread_local std::vector<int32_t> buf;
      node_list(n, buf);
      deg[size_t(n)] = int32_t(buf.size());
      if (deg[size_t(n)] > ufc::kMaxGenNbrs) bad[0] = int32_t(n);
    });
    UFC_REQUIRE(bad[0] < 0, UFC_STATUS_UNSUPPORTED,
                "node " + std::to_string(bad[0] + base) + " has more than " +
                    std::to_string(ufc::kMaxGenNbrs) + " neighbours");
    // node-pair pointers are int32 (nnz / 9 < 2^31); value positions are
    // int64, so nnz itself may pass 2^31 (the CSR export, whose row_ptr is
    // the reference's int32 rowPtr, then refuses)
    int64_t npairs = 0;
    int32_t maxdeg = 0;
    for (int32_t n = 0; n < nn; ++n) {
      npairs += deg[size_t(n)];
      maxdeg = std::max(maxdeg, deg[size_t(n)]);
    }
    UFC_REQUIRE(npairs < INT32_MAX, UFC_STATUS_UNSUPPORTED,
                "more than 2^31 - 1 node pairs (nnz / 9)");
    c->h_nb_ptr.assign(size_t(nn) + 1, 0);
    for (int32_t n = 0; n < nn; ++n)
      c->h_nb_ptr[size_t(n) + 1] = c->h_nb_ptr[size_t(n)] + deg[size_t(n)];
    // general nodes: blocks of 8 nodes with one outside the slot map
    const int64_t nblk = (int64_t(nn) + 7) / 8;
    std::vector<uint8_t> gflag(static_cast<size_t>(nblk), 0);
    for (int32_t n = 0; n < nn; ++n)
      if (cnt[size_t(n)] > ufc::kMaxNodeElems || deg[size_t(n)] > ufc::kMaxNodeNbrs)
        gflag[size_t(n / 8)] = 1;
    const int64_t nnz = 9 * npairs;
    const bool csr32 = nnz < INT32_MAX;
    std::vector<int32_t> nb_idx(static_cast<size_t>(npairs));
    std::vector<int32_t> row_ptr(csr32 ? static_cast<size_t>(nn) * 3 + 1 : 0);
    std::vector<int64_t> diag_pos(static_cast<size_t>(nn) * 3);
    std::vector<uint64_t> s_code(static_cast<size_t>(nn) * 8, 0);
    parallel_for(nn, [&](int64_t n) {
      thread_local std::vector<int32_t> buf;
      node_list(n, buf);
      const int32_t q0 = c->h_nb_ptr[size_t(n)];
      const int dg = deg[size_t(n)];
      const int32_t* L = buf.data();
      std::copy(L, L + dg, nb_idx.begin() + q0);
      const int kself = int(std::lower_bound(L, L + dg, int32_t(n)) - L);
      for (int d = 0; d < 3; ++d) {
        if (csr32) row_ptr[3 * size_t(n) + d] = int32_t(9 * int64_t(q0) + d * 3 * dg);
        diag_pos[3 * size_t(n) + d] = 9 * int64_t(q0) + d * 3 * dg + 3 * kself + d;
      }
      if (gflag[size_t(n / 8)]) return;  // the general kernels find positions themselves
      for (int s = 0; s < 8; ++s) {
        const int32_t e = s_elem[8 * size_t(n) + s];
        if (e < 0) break;
        uint64_t code = uint64_t(s_loc[8 * size_t(n) + s]) << 40;
        for (int b = 0; b < 8; ++b) {
          const int32_t m2 = c->h_conn[8 * size_t(e) + b];
          const uint64_t pos = uint64_t(std::lower_bound(L, L + dg, m2) - L);
          code |= pos << (5 * b);
        }
        s_code[8 * size_t(n) + s] = code;
      }
    });
    if (csr32) row_ptr[size_t(nn) * 3] = int32_t(nnz);
    // 16-bit neighbour offsets when the numbering's bandwidth allows (and
    // every node fits a 32-lane group)
    int64_t maxoff = 0;
    for (int64_t n = 0; n < nn; ++n)
      for (int32_t q = c->h_nb_ptr[size_t(n)]; q < c->h_nb_ptr[size_t(n) + 1]; ++q)
        maxoff = std::max<int64_t>(maxoff, std::abs(int64_t(nb_idx[size_t(q)]) - n));
    c->wide = maxdeg > ufc::kMaxNodeNbrs;
    c->have_off16 = maxoff <= 32767 && !c->wide;

    c->n_nodes = nn;
    c->n_elems = ne;
    c->n_dof = int64_t(nn) * 3;
    c->nnz = nnz;
    c->max_row_nnz = 3 * maxdeg;
    c->have64 = c->have32 = false;
    c->h_f.assign(size_t(c->n_dof), 0.0);
    c->h_bc_eq.clear();
    c->h_bc_pen.clear();
    c->h_bc_rhs.clear();
    c->coords.resize(size_t(nn) * 3);
    c->conn.resize(size_t(ne) * 8);
    c->nb_ptr.resize(size_t(nn) + 1 + ufc::kNbPtrPad);
    c->nb_idx.resize(size_t(npairs));
    if (csr32)
      c->row_ptr.resize(size_t(nn) * 3 + 1);
    else
      c->row_ptr.release();
    c->diag_pos.resize(size_t(nn) * 3);
    c->slot_elem.resize(size_t(nn) * 8);
    c->slot_code.resize(size_t(nn) * 8);
    c->slot_loc.resize(size_t(nn));
    auto up = [&](void* d, const void* h, size_t bytes) {
      UFC_HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
    };
    up(c->coords.get(), c->h_coords.data(), c->coords.bytes());
    up(c->conn.get(), c->h_conn.data(), c->conn.bytes());
    {
      // padded with copies of the last entry: the FP32 SpMV reads the row
      // pointers of a node pair as one scalar load past the last node
      std::vector<int32_t> padded(c->h_nb_ptr);
      padded.resize(size_t(nn) + 1 + ufc::kNbPtrPad, c->h_nb_ptr[size_t(nn)]);
      up(c->nb_ptr.get(), padded.data(), c->nb_ptr.bytes());
      UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    up(c->nb_idx.get(), nb_idx.data(), c->nb_idx.bytes());
    std::vector<int16_t> off16;
    if (c->have_off16) {
      off16.resize(size_t(npairs));
      parallel_for(nn, [&](int64_t n) {
        for (int32_t q = c->h_nb_ptr[size_t(n)]; q < c->h_nb_ptr[size_t(n) + 1]; ++q)
          off16[size_t(q)] = int16_t(int64_t(nb_idx[size_t(q)]) - n);
      });
      c->nb_off16.resize(size_t(npairs));
      up(c->nb_off16.get(), off16.data(), c->nb_off16.bytes());
    } else {
      c->nb_off16.release();
    }
    build_stencil_patterns(c, off16);
    // general nodes: their lists of 8 e + a, and the block flags
    std::vector<int32_t> g_node, g_ptr(1, 0), g_ne;
    for (int64_t b = 0; b < nblk; ++b) {
      if (!gflag[size_t(b)]) continue;
      for (int64_t n = 8 * b; n < std::min<int64_t>(nn, 8 * b + 8); ++n) {
        g_node.push_back(int32_t(n));
        for (int64_t q = ne_ptr[size_t(n)]; q < ne_ptr[size_t(n) + 1]; ++q)
          g_ne.push_back(8 * ne_elem[size_t(q)] + ne_loc[size_t(q)]);
        g_ptr.push_back(int32_t(g_ne.size()));
      }
    }
    c->n_gen = int32_t(g_node.size());
    if (c->n_gen) {
      UFC_REQUIRE(int64_t(ne) * 8 < INT32_MAX, UFC_STATUS_UNSUPPORTED,
                  "general nodes need 8 n_elems < 2^31");
      c->gen_flag.resize(size_t(nblk));
      c->gen_node.resize(g_node.size());
      c->gen_ne_ptr.resize(g_ptr.size());
      c->gen_ne.resize(g_ne.size());
      up(c->gen_flag.get(), gflag.data(), c->gen_flag.bytes());
      up(c->gen_node.get(), g_node.data(), c->gen_node.bytes());
      up(c->gen_ne_ptr.get(), g_ptr.data(), c->gen_ne_ptr.bytes());
      up(c->gen_ne.get(), g_ne.data(), c->gen_ne.bytes());
      UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    } else {
      c->gen_flag.release();
      c->gen_node.release();
      c->gen_ne_ptr.release();
      c->gen_ne.release();
    }
    // the symmetric view needs a structured box, whose nodes all fit the slot
    // map (the FP32 assembly writes its diagonals)
    if (c->n_gen == 0)
      build_sym_diagonals(c, nb_idx, s_elem);
    else
      reset_sym_diagonals(c);
    c->fp32_view_request = -1;
    ufc::select_fp32_view(c);
    if (csr32) up(c->row_ptr.get(), row_ptr.data(), c->row_ptr.bytes());
    up(c->diag_pos.get(), diag_pos.data(), c->diag_pos.bytes());
    up(c->slot_elem.get(), s_elem.data(), c->slot_elem.bytes());
    up(c->slot_code.get(), s_code.data(), c->slot_code.bytes());
    std::vector<uint32_t> s_word(static_cast<size_t>(nn), 0);
    parallel_for(nn, [&](int64_t n) {
      uint32_t w = 0;
      for (int s = 0; s < 8; ++s) w |= uint32_t(s_loc[8 * size_t(n) + s] & 7) << (4 * s);
      s_word[size_t(n)] = w;
    });
    up(c->slot_loc.get(), s_word.data(), c->slot_loc.bytes());
    c->col_idx.release();
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    upload_rhs(c);
    if (info) {
      info->n_dof = c->n_dof;
      info->nnz = c->nnz;
      info->n_nodes = nn;
      info->n_elems = ne;
      info->max_row_nnz = c->max_row_nnz;
      info->nb_index_bytes = nb_index_bytes(c);
      info->fp32_view = c->fp32_view;
      info->sym_tile = c->n_diag > 0 ? c->diag_tile : 0;
    }
    msg = "pattern built: nnz " + std::to_string(nnz);
    return UFC_STATUS_OK;
  });
}

int ufc_rt_asm_get_pattern(ufc_ctx* c, int32_t index_base, int32_t* row_ptr, int32_t* col_idx,
                           ufc_status* st) {
  return guarded(st, "ufc_rt_asm_get_pattern", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(row_ptr || col_idx, UFC_STATUS_INVALID, "null argument");
    UFC_REQUIRE(index_base == 0 || index_base == 1, UFC_STATUS_INVALID, "index_base must be 0 or 1");
    UFC_REQUIRE(c->nnz + index_base <= INT32_MAX, UFC_STATUS_UNSUPPORTED,
                "nnz = " + std::to_string(c->nnz) + " does not fit the int32 rowPtr of "
                "RT_CSRMatrix (NM_Assem_Sparse.f90:45-56)");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    if (!c->col_idx.get()) ufc::build_col_idx(c);
    if (row_ptr) {
      std::vector<int32_t> h(static_cast<size_t>(c->n_dof) + 1);
      UFC_HIP_CHECK(hipMemcpyAsync(h.data(), c->row_ptr.get(), h.size() * 4,
                                   hipMemcpyDeviceToHost, c->stream));
      UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
      for (auto& v : h) v += index_base;
      UFC_HIP_CHECK(hipMemcpy(row_ptr, h.data(), h.size() * 4, hipMemcpyDefault));
    }
    if (col_idx) {
      std::vector<int32_t> h(static_cast<size_t>(c->nnz));
      UFC_HIP_CHECK(hipMemcpyAsync(h.data(), c->col_idx.get(), h.size() * 4,
                                   hipMemcpyDeviceToHost, c->stream));
      UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
      if (index_base)
        parallel_for(int64_t(h.size()), [&](int64_t i) { h[size_t(i)] += 1; });
      UFC_HIP_CHECK(hipMemcpy(col_idx, h.data(), h.size() * 4, hipMemcpyDefault));
    }
    return UFC_STATUS_OK;
  });
}

int ufc_rt_asm_assemble_k(ufc_ctx* c, const ufc_rt_asm_algo* algo, const ufc_rt_asm_args* a,
                          ufc_status* st) {
  return guarded(st, "ufc_rt_asm_assemble_k", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(algo != nullptr, UFC_STATUS_INVALID, "algo is null");
    const int prec = algo->precision;
    UFC_REQUIRE(prec == UFC_PREC_FP64 || prec == UFC_PREC_FP32 || prec == UFC_PREC_MIXED,
                UFC_STATUS_INVALID, "unknown precision");
    UFC_REQUIRE(algo->material == UFC_MAT_ELASTIC || algo->material == UFC_MAT_J2,
                UFC_STATUS_INVALID, "unknown material");
    const int variant = ufc::canonical_variant(algo->variant);
    UFC_REQUIRE(variant > 0, UFC_STATUS_INVALID, "unknown element variant");
    UFC_REQUIRE(algo->material == UFC_MAT_ELASTIC || variant == ufc::kVarFull,
                UFC_STATUS_UNSUPPORTED, "J2 assembly is available for the standard element only");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    if (algo->material == UFC_MAT_ELASTIC) {
      c->variant = variant;
      if (variant == ufc::kVarIncompat) c->inc_c.resize(size_t(c->n_elems) * ufc::kIncompatC);
      UFC_REQUIRE(a && a->young, UFC_STATUS_INVALID, "elastic assembly needs young[]");
      UFC_REQUIRE(a->nu > -1.0 && a->nu < 0.5, UFC_STATUS_INVALID, "nu out of (-1, 0.5)");
      c->young.resize(size_t(c->n_elems));
      UFC_HIP_CHECK(hipMemcpyAsync(c->young.get(), a->young, c->young.bytes(), hipMemcpyDefault,
                                   c->stream));
      c->nu = a->nu;
      c->have_young = true;
    } else {
      UFC_REQUIRE(c->have_j2, UFC_STATUS_INVALID, "J2 assembly needs ufc_ph_j2_init/update");
    }
    const bool fp32 = prec != UFC_PREC_FP64;
    UFC_HIP_CHECK(hipEventRecord(c->ev0, c->stream));
    if (fp32) {
      c->val32.resize(size_t(c->nnz));
      c->diag32.resize(size_t(c->n_dof));
      c->kdiag32.resize(size_t(c->n_dof));
    } else {
      c->val64.resize(size_t(c->nnz));
      c->diag64.resize(size_t(c->n_dof));
    }
    ufc::launch_assemble(c, fp32 ? UFC_PREC_FP32 : UFC_PREC_FP64, algo->material);
    if (fp32) {
      ufc::launch_fp32_difference_form(c);
    } else {
      ufc::launch_penalty(c, false, +1.0);
      ufc::launch_diag(c, false);
    }
    UFC_HIP_CHECK(hipEventRecord(c->ev1, c->stream));
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    UFC_HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->assemble_ms = ms;
    if (fp32) {
      c->have32 = true;
      c->have64 = false;  // a later FP64 solve must not use stale values
    } else {
      c->have64 = true;
      c->have32 = false;
    }
    c->last_material = algo->material;
    c->user_values = false;
    return UFC_STATUS_OK;
  });
}

int ufc_rt_asm_get_values(ufc_ctx* c, int32_t precision, void* values, ufc_status* st) {
  return guarded(st, "ufc_rt_asm_get_values", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(values != nullptr, UFC_STATUS_INVALID, "values is null");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    if (precision == UFC_PREC_FP64) {
      UFC_REQUIRE(c->have64, UFC_STATUS_INVALID, "no FP64 values assembled");
      copy_out(c, static_cast<double*>(values), c->val64.get(), size_t(c->nnz));
    } else {
      UFC_REQUIRE(c->have32, UFC_STATUS_INVALID, "no FP32 values assembled");
      copy_out(c, static_cast<float*>(values), c->val32.get(), size_t(c->nnz));
    }
    return UFC_STATUS_OK;
  });
}

int ufc_rt_asm_set_values(ufc_ctx* c, int32_t precision, const double* values,
                          ufc_status* st) {
  return guarded(st, "ufc_rt_asm_set_values", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(values != nullptr, UFC_STATUS_INVALID, "values is null");
    UFC_REQUIRE(precision == UFC_PREC_FP64, UFC_STATUS_UNSUPPORTED, "FP64 values only");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    c->val64.resize(size_t(c->nnz));
    c->diag64.resize(size_t(c->n_dof));
    UFC_HIP_CHECK(hipMemcpyAsync(c->val64.get(), values, c->val64.bytes(), hipMemcpyDefault,
                                 c->stream));
    ufc::launch_diag(c, false);
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->have64 = true;
    c->have32 = false;
    c->user_values = true;
    return UFC_STATUS_OK;
  });
}

int ufc_rt_asm_apply_bc(ufc_ctx* c, const ufc_rt_bc_args* a, ufc_status* st) {
  return guarded(st, "ufc_rt_asm_apply_bc", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(a != nullptr, UFC_STATUS_INVALID, "args is null");
    UFC_REQUIRE(a->n_bc >= 0 && (a->n_bc == 0 || a->eq), UFC_STATUS_INVALID, "bad BC list");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    // remove the previous set from assembled values
    if (c->have64) ufc::launch_penalty(c, false, -1.0);
    if (c->have32) ufc::launch_penalty(c, true, -1.0);
    std::map<int32_t, std::pair<double, double>> acc;  // eq -> (penalty sum, rhs sum)
    std::vector<int32_t> heq(static_cast<size_t>(a->n_bc));
    std::vector<double> hval(static_cast<size_t>(a->n_bc), 0.0);
    if (a->n_bc) {
      UFC_HIP_CHECK(hipMemcpy(heq.data(), a->eq, heq.size() * 4, hipMemcpyDefault));
      if (a->value)
        UFC_HIP_CHECK(hipMemcpy(hval.data(), a->value, hval.size() * 8, hipMemcpyDefault));
    }
    for (int32_t i = 0; i < a->n_bc; ++i) {
      const int32_t q = heq[size_t(i)] - a->index_base;
      UFC_REQUIRE(q >= 0 && q < c->n_dof, UFC_STATUS_INVALID,
                  "BC equation out of range: " + std::to_string(heq[size_t(i)]));
      auto& pr = acc[q];
      pr.first += a->penalty;
      pr.second += a->penalty * hval[size_t(i)];
    }
    c->h_bc_eq.clear();
    c->h_bc_pen.clear();
    c->h_bc_rhs.clear();
    for (auto& kv : acc) {
      c->h_bc_eq.push_back(kv.first);
      c->h_bc_pen.push_back(kv.second.first);
      c->h_bc_rhs.push_back(kv.second.second);
    }
    c->bc_eq.resize(std::max<size_t>(1, c->h_bc_eq.size()));
    c->bc_pen.resize(std::max<size_t>(1, c->h_bc_eq.size()));
    if (!c->h_bc_eq.empty()) {
      UFC_HIP_CHECK(hipMemcpyAsync(c->bc_eq.get(), c->h_bc_eq.data(), c->h_bc_eq.size() * 4,
                                   hipMemcpyHostToDevice, c->stream));
      UFC_HIP_CHECK(hipMemcpyAsync(c->bc_pen.get(), c->h_bc_pen.data(), c->h_bc_pen.size() * 8,
                                   hipMemcpyHostToDevice, c->stream));
    }
    if (c->have64) {
      ufc::launch_penalty(c, false, +1.0);
      ufc::launch_diag(c, false);
    }
    if (c->have32) {
      ufc::launch_penalty(c, true, +1.0);
      ufc::launch_diag32_penalty(c);
      ufc::launch_fp32_offdiag(c);
    }
    upload_rhs(c);
    return UFC_STATUS_OK;
  });
}

int ufc_rt_asm_load(ufc_ctx* c, const ufc_rt_load_args* a, ufc_status* st) {
  return guarded(st, "ufc_rt_asm_load", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(a != nullptr, UFC_STATUS_INVALID, "args is null");
    UFC_REQUIRE(a->kind == UFC_LOAD_TRACTION || a->kind == UFC_LOAD_PRESSURE, UFC_STATUS_INVALID,
                "unknown load kind");
    UFC_REQUIRE(a->n_faces >= 0 && (a->n_faces == 0 || (a->elem && a->face)), UFC_STATUS_INVALID,
                "bad face list");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    if (!a->accumulate) std::fill(c->h_f.begin(), c->h_f.end(), 0.0);
    std::vector<int32_t> he(static_cast<size_t>(a->n_faces)), hf(size_t(a->n_faces));
    if (a->n_faces) {
      UFC_HIP_CHECK(hipMemcpy(he.data(), a->elem, he.size() * 4, hipMemcpyDefault));
      UFC_HIP_CHECK(hipMemcpy(hf.data(), a->face, hf.size() * 4, hipMemcpyDefault));
    }
    for (int32_t i = 0; i < a->n_faces; ++i) {
      const int32_t e = he[size_t(i)] - a->index_base;
      const int f = hf[size_t(i)] - a->index_base;
      UFC_REQUIRE(e >= 0 && e < c->n_elems, UFC_STATUS_INVALID, "load element out of range");
      UFC_REQUIRE(f >= 0 && f < 6, UFC_STATUS_INVALID, "face id out of range");
      double x[8][3];
      for (int q = 0; q < 8; ++q)
        for (int d = 0; d < 3; ++d)
          x[q][d] = c->h_coords[3 * size_t(c->h_conn[8 * size_t(e) + q]) + d];
      if (a->kind == UFC_LOAD_TRACTION) {
        // LoadBC_DistributeLoad_ToSurface: traction * area / 4 per face node
        double area = 0.0;
        face_quadrature(x, f, [&](const double*, double dA, const double*) { area += dA; });
        for (int q = 0; q < 4; ++q) {
          const size_t n = size_t(c->h_conn[8 * size_t(e) + kFaceNodes[f][q]]);
          for (int d = 0; d < 3; ++d) c->h_f[3 * n + d] += a->traction[d] * (area / 4.0);
        }
      } else {
        // PH_Elem_C3D8_FormFacePressure: sum over points of N p n dA
        double fe[4][3] = {};
        face_quadrature(x, f, [&](const double* Nf, double dA, const double* n) {
          for (int q = 0; q < 4; ++q)
            for (int d = 0; d < 3; ++d) fe[q][d] += Nf[q] * a->magnitude * n[d] * dA;
        });
        for (int q = 0; q < 4; ++q) {
          const size_t n = size_t(c->h_conn[8 * size_t(e) + kFaceNodes[f][q]]);
          for (int d = 0; d < 3; ++d) c->h_f[3 * n + d] += fe[q][d];
        }
      }
    }
    upload_rhs(c);
    return UFC_STATUS_OK;
  });
}

int ufc_rt_set_rhs(ufc_ctx* c, const double* f, ufc_status* st) {
  return guarded(st, "ufc_rt_set_rhs", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(f != nullptr, UFC_STATUS_INVALID, "f is null");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    UFC_HIP_CHECK(hipMemcpy(c->h_f.data(), f, c->h_f.size() * 8, hipMemcpyDefault));
    upload_rhs(c);
    return UFC_STATUS_OK;
  });
}

int ufc_rt_get_rhs(ufc_ctx* c, double* f, ufc_status* st) {
  return guarded(st, "ufc_rt_get_rhs", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(f != nullptr, UFC_STATUS_INVALID, "f is null");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    copy_out(c, f, c->f.get(), size_t(c->n_dof));
    return UFC_STATUS_OK;
  });
}

int ufc_rt_residual(ufc_ctx* c, const double* u, double* r, double* norms, ufc_status* st) {
  return guarded(st, "ufc_rt_residual", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(u != nullptr, UFC_STATUS_INVALID, "u is null");
    UFC_REQUIRE(c->have_j2 || c->have_young, UFC_STATUS_INVALID,
                "no material: assemble or update J2 first");
    UFC_REQUIRE(!c->user_values, UFC_STATUS_INVALID,
                "values set by ufc_rt_asm_set_values: the element operator is not known");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    ufc::ensure_solver_buffers(c, true, false);
    ufc::DBuf<double> tu;
    const double* du = dev_in(c, u, size_t(c->n_dof), tu);
    if (c->have_j2)
      ufc::launch_apply_elem64(c, 2, nullptr, c->w_t.get());
    else
      ufc::launch_apply_elem64(c, 0, du, c->w_t.get());
    ufc::residual64(c, c->f.get(), c->w_t.get(), du, c->w_r.get());
    if (norms) {
      norms[0] = ufc::norm2_64(c, c->w_r.get());
      norms[1] = ufc::norm2_64(c, c->w_t.get());
    }
    if (r) copy_out(c, r, c->w_r.get(), size_t(c->n_dof));
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    return UFC_STATUS_OK;
  });
}

int ufc_mesh_box_numbering(int32_t n_nodes, int32_t n_elems, int32_t index_base,
                           const int32_t* conn, int32_t* perm, ufc_status* st) {
  return guarded(st, "ufc_mesh_box_numbering", [&](std::string& msg) {
    UFC_REQUIRE(conn && perm, UFC_STATUS_INVALID, "null argument");
    UFC_REQUIRE(n_nodes > 0 && n_elems > 0, UFC_STATUS_INVALID, "empty mesh");
    std::vector<int32_t> cv(static_cast<size_t>(n_elems) * 8);
    std::vector<int32_t> cnt(static_cast<size_t>(n_nodes), 0);
    for (size_t i = 0; i < cv.size(); ++i) {
      const int32_t v = conn[i] - index_base;
      UFC_REQUIRE(v >= 0 && v < n_nodes, UFC_STATUS_INVALID, "connectivity entry out of range");
      cv[i] = v;
      UFC_REQUIRE(++cnt[size_t(v)] <= ufc::kMaxNodeElems, UFC_STATUS_UNSUPPORTED,
                  "not a structured box (a node in more than 8 elements)");
    }
    std::vector<int32_t> node_elem(static_cast<size_t>(n_nodes) * 8, -1);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int32_t e = 0; e < n_elems; ++e)
      for (int a = 0; a < 8; ++a) {
        const int32_t n = cv[8 * size_t(e) + a];
        node_elem[8 * size_t(n) + size_t(cnt[size_t(n)]++)] = e;
      }
    const std::vector<int32_t> p = box_numbering(n_nodes, cv, node_elem);
    if (p.empty()) {
      msg = "not a structured box of hexes";
      return UFC_STATUS_UNSUPPORTED;
    }
    std::copy(p.begin(), p.end(), perm);
    msg = "box numbering found";
    return UFC_STATUS_OK;
  });
}

// ======================================================================= L2_NM
int ufc_nm_spmv(ufc_ctx* c, int32_t precision, const double* x, double* y, ufc_status* st) {
  return guarded(st, "ufc_nm_spmv", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(x && y, UFC_STATUS_INVALID, "null argument");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    ufc::ensure_solver_buffers(c, true, precision != UFC_PREC_FP64);
    ufc::DBuf<double> tx;
    const double* dx = dev_in(c, x, size_t(c->n_dof), tx);
    if (precision == UFC_PREC_FP64) {
      UFC_REQUIRE(c->have64, UFC_STATUS_INVALID, "no FP64 values assembled");
      if (!c->col_idx.get() && c->fp64_view == ufc::kFp64ViewCsr) ufc::build_col_idx(c);
      ufc::spmv64(c, dx, c->w_t.get());
      copy_out(c, y, c->w_t.get(), size_t(c->n_dof));
    } else {
      UFC_REQUIRE(c->have32, UFC_STATUS_INVALID, "no FP32 values assembled");
      ufc::to_view32(c, dx, c->s_x.get());
      ufc::spmv32(c, c->s_x.get(), c->s_ap.get());
      ufc::from_view32(c, c->s_ap.get(), c->w_t.get());
      copy_out(c, y, c->w_t.get(), size_t(c->n_dof));
    }
    ufc::reset_spmv_events(c);
    return UFC_STATUS_OK;
  });
}

int ufc_nm_spmv_t(ufc_ctx* c, int32_t precision, const double* x, double* y, ufc_status* st) {
  return guarded(st, "ufc_nm_spmv_t", [&](std::string&) {
    check_mesh(c);
    UFC_REQUIRE(x && y, UFC_STATUS_INVALID, "null argument");
    UFC_REQUIRE(precision == UFC_PREC_FP64, UFC_STATUS_UNSUPPORTED,
                "K^T x is formed from the FP64 values");
    UFC_REQUIRE(c->have64, UFC_STATUS_INVALID, "no FP64 values assembled");
    UFC_HIP_CHECK(hipSetDevice(c->device));
    ufc::ensure_solver_buffers(c, true, false);
    ufc::DBuf<double> tx;
    const double* dx = dev_in(c, x, size_t(c->n_dof), tx);
    ufc::spmv_t64(c, dx, c->w_t.get());
    copy_out(c, y, c->w_t.get(), size_t(c->n_dof));
    return UFC_STATUS_OK;
  });
}

static int solve_impl(ufc_ctx* c, const ufc_nm_solver_algo* algo, const double* b_user,
                      ufc_nm_solver_state* s, std::string& msg) {
  UFC_REQUIRE(algo && s && s->x, UFC_STATUS_INVALID, "null argument");
  UFC_REQUIRE(algo->max_iter >= 0, UFC_STATUS_INVALID, "max_iter < 0");
  UFC_HIP_CHECK(hipSetDevice(c->device));
  const int prec = algo->precision;
  const size_t n = size_t(c->n_dof);
  ufc::ensure_solver_buffers(c, true, prec != UFC_PREC_FP64);
  ufc::reset_spmv_events(c);
  // right-hand side into w_b
  const double* b = c->f.get();
  if (b_user) {
    UFC_HIP_CHECK(hipMemcpyAsync(c->w_b.get(), b_user, n * 8, hipMemcpyDefault, c->stream));
    b = c->w_b.get();
  }
  UFC_HIP_CHECK(hipEventRecord(c->ev2, c->stream));
  int code = UFC_STATUS_OK;
  int niter = 0, nouter = 0;
  double rnorm = 0.0;
  if (prec == UFC_PREC_FP64) {
    UFC_REQUIRE(c->have64, UFC_STATUS_INVALID, "no FP64 values assembled");
    if (!c->col_idx.get() && c->fp64_view == ufc::kFp64ViewCsr) ufc::build_col_idx(c);
    if (!s->x0_zero)
      UFC_HIP_CHECK(hipMemcpyAsync(c->w_x.get(), s->x, n * 8, hipMemcpyDefault, c->stream));
    code = ufc::pcg64(c, b, c->w_x.get(), s->x0_zero != 0, algo->max_iter, algo->rtol,
                      algo->check_every, &niter, &rnorm);
  } else if (prec == UFC_PREC_FP32) {
    UFC_REQUIRE(c->have32, UFC_STATUS_INVALID, "no FP32 values assembled");
    UFC_REQUIRE(s->x0_zero, UFC_STATUS_UNSUPPORTED, "FP32 solve starts from x0 = 0");
    code = ufc::pcg32(c, b, c->w_x.get(), true, false, algo->max_iter, algo->rtol,
                      algo->check_every, &niter, &rnorm);
  } else {
    // FP32 inner PCG in FP64 iterative refinement:
    //   r = b - K u (FP64, element by element), d = PCG32(K32, r), u += d
    // The inner iteration accumulates d into u in FP64: an FP32 d would carry
    // rounding noise that K amplifies to ~eps32 * cond(K) in the next residual.
    UFC_REQUIRE(c->have32, UFC_STATUS_INVALID, "no FP32 values assembled");
    UFC_REQUIRE(c->last_material == UFC_MAT_J2 ? c->have_j2 : c->have_young, UFC_STATUS_INVALID,
                "no material for the FP64 residual");
    double* u = c->w_x.get();
    if (s->x0_zero)
      UFC_HIP_CHECK(hipMemsetAsync(u, 0, n * 8, c->stream));
    else
      UFC_HIP_CHECK(hipMemcpyAsync(u, s->x, n * 8, hipMemcpyDefault, c->stream));
    const double bnorm0 = ufc::norm2_64(c, b);
    const double bnorm = bnorm0 < 1.0e-30 ? 1.0 : bnorm0;
    const double inner = algo->inner_rtol > 0.0 ? algo->inner_rtol : 1.0e-5;
    // residual replacement interval inside the inner solve; >= 1 switches it
    // off (each refinement step then restarts the FP32 PCG from p = z)
    const double delta = algo->replace_rtol > 0.0 ? algo->replace_rtol : 0.1;
    const bool reliable = delta < 1.0;
    const int max_outer = algo->max_outer > 0 ? algo->max_outer : 20;
    // bound on FP64 residual evaluations: a few per decade of the target
    const int max_eval = reliable ? max_outer * (2 + int(std::ceil(std::log(inner) /
                                                                    std::log(delta))))
                                  : max_outer;
    double* r = c->w_r.get();
    bool have_dir = false;
    double step_start = 0.0;  // ||r||/||b|| where the current refinement step began
    int nrepl = 0;
    code = UFC_STATUS_INVALID;
    for (int k = 0; k <= max_eval; ++k) {
      // r = b - K u in FP64, element by element (the refinement residual);
      // from u = 0 the first one is b itself
      if (k == 0 && s->x0_zero) {
        UFC_HIP_CHECK(hipMemcpyAsync(r, b, n * 8, hipMemcpyDeviceToDevice, c->stream));
        rnorm = bnorm0 / bnorm;
      } else {
        apply_fp64_operator(c, u, c->w_t.get());
        ufc::residual64(c, b, c->w_t.get(), u, r);
        rnorm = ufc::norm2_64(c, r) / bnorm;
      }
      if (!std::isfinite(rnorm)) {
        code = UFC_STATUS_ERROR;
        break;
      }
      if (rnorm < algo->rtol) {
        code = UFC_STATUS_OK;
        break;
      }
      // a refinement step begins at the first residual and whenever the
      // current one has reached inner_rtol of the step's starting residual
      const bool new_step = k == 0 || !reliable || rnorm <= inner * step_start;
      if (new_step) {
        if (nouter == max_outer) break;
        ++nouter;
        step_start = rnorm;
      } else {
        ++nrepl;
      }
      if (k == max_eval) break;
      double tol_k;
      if (reliable) {
        // FP32 PCG with residual replacement (reliable updates): the inner
        // CG runs until its FP32 recurrence residual has dropped by delta,
        // its residual is then replaced by the FP64 one and it continues
        // with the same search direction, re-orthogonalised against the new
        // residual.  With delta ~0.1 the replaced residual differs from the
        // recurrence one by ~1e-4 of its size, small enough to keep CG's
        // conjugacy, so the whole solve costs ~FP64 PCG's iteration count.
        // A segment stops early where the refinement step (inner_rtol) or
        // the solve (rtol) ends.
        const double need = std::max(inner * step_start, 0.5 * algo->rtol) / rnorm;
        tol_k = need > delta ? std::min(need, 0.5) : delta;
      } else {
        // Restarted refinement: inner tolerance inner_rtol, loosened to
        // 0.1 rtol/||r|| for a step that can finish the solve.  Tightening
        // below inner_rtol does not pay: a restarted FP32 inner solve
        // contracts the FP64 residual by ~1e-5 at best on these problems.
        tol_k = inner;
        const double loose = 0.1 * algo->rtol / rnorm;
        if (loose > inner) tol_k = std::min(loose, 0.1);
      }
      int it = 0;
      double rin = 0.0;
      // Reliable form: the segment's correction d is accumulated in FP32 and
      // added to u in FP64 at its end (a 0.1 segment tolerates FP32 rounding
      // in d; it saves 8 bytes per DOF and iteration).  Restarted form: d is
      // accumulated straight into u in FP64, and each step restarts from
      // p = z: continuing the old direction after a replacement as large as
      // inner_rtol allows (the FP64 residual sits ~10x above the FP32
      // recurrence one at 1e-5) stalls CG.
      const bool resume = reliable && have_dir;
      int ic;
      if (reliable) {
        ic = ufc::pcg32_segment(c, r, c->s_x.get(), resume, algo->max_iter, tol_k,
                                algo->check_every, &it, &rin);
        ufc::axpy_f32_to_f64(c, c->s_x.get(), u);
      } else {
        ic = ufc::pcg32(c, r, u, false, false, algo->max_iter, tol_k, algo->check_every, &it,
                        &rin);
      }
      have_dir = have_dir || it > 0;
      if (verbose())
        std::fprintf(stderr,
                     "[ufc_hip] refinement %d%s: ||r||/||b|| %.3e, segment tol %.2e, %d it\n",
                     nouter, new_step ? "" : " (replacement)", rnorm, tol_k, it);
      niter += it;
      if (ic == UFC_STATUS_ERROR) {
        code = UFC_STATUS_ERROR;
        break;
      }
    }
    c->n_replace = nrepl;
  }
  UFC_HIP_CHECK(hipEventRecord(c->ev3, c->stream));
  UFC_HIP_CHECK(hipMemcpyAsync(s->x, c->w_x.get(), n * 8, hipMemcpyDefault, c->stream));
  UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  UFC_HIP_CHECK(hipEventElapsedTime(&ms, c->ev2, c->ev3));
  c->solve_ms = ms;
  c->spmv_ms_avg = ufc::spmv_event_avg_ms(c, &c->spmv_calls);
  s->niter = niter;
  s->n_outer = nouter;
  s->rnorm = rnorm;
  s->converged = code == UFC_STATUS_OK;
  if (code == UFC_STATUS_OK)
    msg = "converged";
  else if (code == UFC_STATUS_ERROR)
    msg = "breakdown (p.Ap <= 0 or non-finite)";
  else
    msg = "[NM_Solver_PCG]: did not converge";
  return code;
}

int ufc_nm_pcg(ufc_ctx* c, const ufc_nm_solver_algo* algo, ufc_nm_solver_state* s,
               ufc_status* st) {
  return guarded(st, "ufc_nm_pcg", [&](std::string& msg) {
    check_mesh(c);
    return solve_impl(c, algo, s ? s->b : nullptr, s, msg);
  });
}

static int gmres_adjoint(ufc_ctx* c, const ufc_nm_solver_algo* algo, const double* dJdu,
                         ufc_nm_solver_state* s, std::string& msg) {
  UFC_REQUIRE(s->x != nullptr, UFC_STATUS_INVALID, "null argument");
  UFC_REQUIRE(algo->precision == UFC_PREC_FP64, UFC_STATUS_UNSUPPORTED,
              "GMRES adjoint runs on the FP64 values");
  UFC_REQUIRE(c->have64, UFC_STATUS_INVALID, "no FP64 values assembled");
  UFC_REQUIRE(algo->max_iter >= 0, UFC_STATUS_INVALID, "max_iter < 0");
  UFC_HIP_CHECK(hipSetDevice(c->device));
  const size_t n = size_t(c->n_dof);
  ufc::ensure_solver_buffers(c, true, false);
  UFC_HIP_CHECK(hipMemcpyAsync(c->w_b.get(), dJdu, n * 8, hipMemcpyDefault, c->stream));
  UFC_HIP_CHECK(hipEventRecord(c->ev2, c->stream));
  int niter = 0, nrest = 0;
  double rnorm = 0.0;
  const int restart = algo->restart > 0 ? algo->restart : 30;
  const int code = ufc::gmres_t64(c, c->w_b.get(), c->w_x.get(), restart, algo->max_iter,
                                  algo->rtol, &niter, &nrest, &rnorm, msg);
  UFC_HIP_CHECK(hipEventRecord(c->ev3, c->stream));
  UFC_HIP_CHECK(hipMemcpyAsync(s->x, c->w_x.get(), n * 8, hipMemcpyDefault, c->stream));
  UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  UFC_HIP_CHECK(hipEventElapsedTime(&ms, c->ev2, c->ev3));
  c->solve_ms = ms;
  c->spmv_ms_avg = 0.0;
  c->spmv_calls = 0;
  s->niter = niter;
  s->n_outer = nrest;
  s->rnorm = rnorm;
  s->converged = code == UFC_STATUS_OK;
  if (code == UFC_STATUS_OK) msg = "converged";
  return code;
}

int ufc_nm_adjoint_solve(ufc_ctx* c, const ufc_nm_solver_algo* algo, const double* dJdu,
                         ufc_nm_solver_state* s, ufc_status* st) {
  return guarded(st, "ufc_nm_adjoint_solve", [&](std::string& msg) {
    check_mesh(c);
    UFC_REQUIRE(dJdu != nullptr, UFC_STATUS_INVALID, "dJdu is null");
    UFC_REQUIRE(s != nullptr && algo != nullptr, UFC_STATUS_INVALID, "null argument");
    UFC_REQUIRE(algo->krylov == UFC_KRYLOV_PCG || algo->krylov == UFC_KRYLOV_GMRES,
                UFC_STATUS_INVALID, "unknown krylov method");
    if (algo->krylov == UFC_KRYLOV_GMRES) return gmres_adjoint(c, algo, dJdu, s, msg);
    // lambda_0 = 0 as in NM_CG_Solve_Transpose (NM_Solv_GMRESTranspose.f90:163-166)
    const int32_t keep = s->x0_zero;
    s->x0_zero = 1;
    const int code = solve_impl(c, algo, dJdu, s, msg);
    s->x0_zero = keep;
    return code;
  });
}

}  // extern "C"
