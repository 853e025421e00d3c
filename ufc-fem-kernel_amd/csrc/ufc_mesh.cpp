// ufc_mesh.cpp -- structured C3D8 box meshes and seeded per-element fields
// for the benchmark and test inputs.  The reference reads meshes from input
// decks and has no generator; these follow the conventions of its element
// (Abaqus node order, PH_Elem_C3D8.f90:1930-1933).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>

#include "ufc_hip.h"

namespace {

// SplitMix64 finaliser (Steele, Lea, Flood 2014), keyed (seed, stream, counter)
uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t counter) {
  const uint64_t G = 0x9E3779B97F4A7C15ULL;
  return mix64(mix64(seed + G * (stream + 1)) + G * (counter + 1));
}
// uniform on (0, 1]
double rng_u01(uint64_t seed, uint64_t stream, uint64_t counter) {
  return double((rng_u64(seed, stream, counter) >> 11) + 1) * 0x1.0p-53;
}

void ok(ufc_status* st, const char* src) {
  if (!st) return;
  std::memset(st, 0, sizeof *st);
  std::snprintf(st->source, sizeof st->source, "%s", src);
}
int fail(ufc_status* st, const char* src, const char* msg) {
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->status_code = UFC_STATUS_INVALID;
    st->category = UFC_STATUS_INVALID;
    st->severity = 3;
    std::snprintf(st->message, sizeof st->message, "%s", msg);
    std::snprintf(st->source, sizeof st->source, "%s", src);
  }
  return UFC_STATUS_INVALID;
}

}  // namespace

extern "C" {

int ufc_mesh_box(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                 double jitter, uint64_t seed, int32_t index_base, double* coords,
                 int32_t* conn, ufc_status* st) {
  if (nx < 1 || ny < 1 || nz < 1 || !coords || !conn)
    return fail(st, "ufc_mesh_box", "bad mesh arguments");
  const int64_t sx = nx + 1, sy = ny + 1;
  for (int64_t k = 0; k <= nz; ++k)
    for (int64_t j = 0; j <= ny; ++j)
      for (int64_t i = 0; i <= nx; ++i) {
        const int64_t n = i + sx * (j + sy * k);
        coords[3 * n + 0] = lx * double(i) / double(nx);
        coords[3 * n + 1] = ly * double(j) / double(ny);
        coords[3 * n + 2] = lz * double(k) / double(nz);
      }
  if (jitter > 0.0) {
    const double h[3] = {lx / nx, ly / ny, lz / nz};
    for (int64_t k = 1; k < nz; ++k)
      for (int64_t j = 1; j < ny; ++j)
        for (int64_t i = 1; i < nx; ++i) {
          const int64_t n = i + sx * (j + sy * k);
          for (int d = 0; d < 3; ++d)
            coords[3 * n + d] += (2.0 * rng_u01(seed, 2, 3 * uint64_t(n) + d) - 1.0) * jitter * h[d];
        }
  }
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t e = i + int64_t(nx) * (j + int64_t(ny) * k);
        const int64_t n0 = i + sx * (j + sy * k);
        const int64_t v[8] = {n0,           n0 + 1,           n0 + 1 + sx,           n0 + sx,
                              n0 + sx * sy, n0 + 1 + sx * sy, n0 + 1 + sx + sx * sy, n0 + sx + sx * sy};
        for (int a = 0; a < 8; ++a) conn[8 * e + a] = int32_t(v[a] + index_base);
      }
  ok(st, "ufc_mesh_box");
  return UFC_STATUS_OK;
}

int ufc_field(int32_t kind, int64_t n, uint64_t seed, double a, double b, double* out,
              ufc_status* st) {
  if (n < 0 || (n > 0 && !out)) return fail(st, "ufc_field", "bad field arguments");
  if (kind == 0) {
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t e = 0; e < n; ++e) {
      const double u1 = rng_u01(seed, 0, 2 * uint64_t(e));
      const double u2 = rng_u01(seed, 0, 2 * uint64_t(e) + 1);
      out[e] = std::exp(a + b * std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2));
    }
  } else if (kind == 1) {
    for (int64_t e = 0; e < n; ++e) out[e] = a + (b - a) * rng_u01(seed, 1, uint64_t(e));
  } else {
    return fail(st, "ufc_field", "unknown field kind");
  }
  ok(st, "ufc_field");
  return UFC_STATUS_OK;
}

}  // extern "C"
