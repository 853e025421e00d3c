// ufc_restart.cpp -- restart files: the record sequence of RT_Out_RestartSave
// (L5_RT/Output/RT_Out_Restart.f90:189-306) as Fortran unformatted sequential
// records, followed by the committed J2 state of the context.  Host code.
//
// A Fortran unformatted sequential record is [int32 n][n bytes][int32 n];
// IF_FileHandle opens IF_IO_FORMAT_BINARY files that way
// (L1_IF/IO/IF_IO_File.f90:219-221) and every WriteBinary / ReadBinary call is
// one record.  The file is written to "<path>.tmp" and renamed over <path>
// when complete, so an interrupted save leaves the previous checkpoint.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ufc_internal.hpp"

using ufc::guarded;

namespace {

constexpr int32_t kVersion = 1;
constexpr int32_t kStamp = 20260204;       // RT_Out_Restart.f90:243
constexpr int64_t kChunkElems = int64_t(1) << 21;
constexpr int kPoints = 8, kNtens = 6, kNstatv = 7, kVals = kNtens + kNstatv;
constexpr size_t kStageDoubles = size_t(1) << 23;  // 64 MiB host staging

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

void check_mesh(ufc_ctx* c) {
  UFC_REQUIRE(c != nullptr, UFC_STATUS_INVALID, "null context");
  UFC_REQUIRE(c->n_nodes > 0 && c->nb_ptr.get(), UFC_STATUS_INVALID,
              "no mesh: call ufc_rt_asm_build_pattern first");
}

// ------------------------------------------------------------------ writing
class Writer {
 public:
  explicit Writer(FILE* f) : f_(f) {}
  void begin(int64_t bytes) {
    UFC_REQUIRE(bytes >= 0 && bytes <= INT32_MAX, UFC_STATUS_UNSUPPORTED,
                "restart record over 2^31 - 1 bytes");
    open_ = int32_t(bytes);
    put(&open_, 4);
  }
  void put(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f_) != n)
      throw ufc::Error{UFC_STATUS_ERROR,
                       std::string("Restart save: Write error - ") + std::strerror(errno)};
  }
  void end() { put(&open_, 4); }
  template <typename... T>
  void packed(const T&... v) {  // one record of scalars, no padding
    begin(int64_t((sizeof(T) + ...)));
    (put(&v, sizeof(T)), ...);
    end();
  }

 private:
  FILE* f_;
  int32_t open_ = 0;
};

// ------------------------------------------------------------------ reading
class Reader {
 public:
  explicit Reader(FILE* f) : f_(f) {}
  // opens the next record and checks its length
  void begin(int64_t want, const char* what) {
    int32_t n = 0;
    get(&n, 4, what);
    if (int64_t(n) != want)
      throw ufc::Error{UFC_STATUS_ERROR, std::string("Restart restore: Read error - ") + what +
                                            " record holds " + std::to_string(n) +
                                            " bytes, expected " + std::to_string(want)};
    open_ = n;
  }
  void get(void* p, size_t n, const char* what) {
    if (n && std::fread(p, 1, n, f_) != n)
      throw ufc::Error{UFC_STATUS_ERROR,
                       std::string("Restart restore: Read error - truncated ") + what};
  }
  void skip(int64_t n, const char* what) {
    if (std::fseek(f_, long(n), SEEK_CUR) != 0)
      throw ufc::Error{UFC_STATUS_ERROR,
                       std::string("Restart restore: Read error - truncated ") + what};
  }
  void end(const char* what) {
    int32_t n = 0;
    get(&n, 4, what);
    if (n != open_)
      throw ufc::Error{UFC_STATUS_ERROR,
                       std::string("Restart restore: Read error - bad record marker after ") + what};
  }
  bool at_eof() {
    const int ch = std::fgetc(f_);
    if (ch == EOF) return true;
    std::ungetc(ch, f_);
    return false;
  }

 private:
  FILE* f_;
  int32_t open_ = 0;
};

template <typename T>
T take(const unsigned char*& p) {
  T v;
  std::memcpy(&v, p, sizeof v);
  p += sizeof v;
  return v;
}

}  // namespace

extern "C" {

int ufc_rt_restart_save(ufc_ctx* c, const char* path, const ufc_rt_restart_data* d,
                        int32_t with_state, ufc_status* st) {
  return guarded(st, "ufc_rt_restart_save", [&](std::string& msg) {
    UFC_REQUIRE(path && path[0], UFC_STATUS_INVALID, "Restart save: Empty filename");
    UFC_REQUIRE(d && d->u, UFC_STATUS_INVALID, "Restart save: Invalid restart data");
    int64_t n = d->n_dof;
    if (c) {
      check_mesh(c);
      UFC_REQUIRE(!with_state || c->have_j2, UFC_STATUS_INVALID,
                  "Restart save: no J2 state (ufc_ph_j2_init)");
      n = c->n_dof;
    } else {
      UFC_REQUIRE(!with_state, UFC_STATUS_INVALID, "Restart save: the J2 state needs a context");
      UFC_REQUIRE(n > 0, UFC_STATUS_INVALID, "Restart save: Invalid restart data");
    }
    std::vector<double> u(size_t(n), 0.0);
    if (c) {
      UFC_HIP_CHECK(hipSetDevice(c->device));
      UFC_HIP_CHECK(hipMemcpy(u.data(), d->u, size_t(n) * sizeof(double), hipMemcpyDefault));
    } else {
      std::memcpy(u.data(), d->u, size_t(n) * sizeof(double));
    }

    const std::string tmp = std::string(path) + ".tmp";
    File f(std::fopen(tmp.c_str(), "wb"));
    UFC_REQUIRE(f, UFC_STATUS_ERROR,
                std::string("Restart save: Cannot open file - ") + path + " - " +
                    std::strerror(errno));
    Writer w(f.get());
    w.packed(kVersion, kStamp, int32_t(n));
    w.begin(8 * n);
    w.put(u.data(), size_t(n) * 8);
    w.end();
    std::fill(u.begin(), u.end(), 0.0);  // v and a of a static step
    for (int k = 0; k < 2; ++k) {
      w.begin(8 * n);
      w.put(u.data(), size_t(n) * 8);
      w.end();
    }
    w.packed(d->time, d->increment, d->lambda);
    w.packed(d->residual_norm, d->iterations, int32_t(d->converged ? 1 : 0));
    w.packed(int32_t(n));
    if (with_state) {
      const int64_t ne = c->n_elems;
      const int32_t n_chunks = int32_t((ne + kChunkElems - 1) / kChunkElems);
      w.packed(int32_t(ne), int32_t(kPoints), int32_t(kNtens), int32_t(kNstatv), n_chunks);
      std::vector<double> stage(kStageDoubles);
      for (int64_t e0 = 0; e0 < ne; e0 += kChunkElems) {
        const int64_t vals = std::min(kChunkElems, ne - e0) * kPoints * kVals;
        const double* src = c->j2_commit.get() + e0 * kPoints * kVals;
        w.begin(vals * 8);
        for (int64_t q = 0; q < vals; q += int64_t(kStageDoubles)) {
          const size_t m = size_t(std::min<int64_t>(int64_t(kStageDoubles), vals - q));
          UFC_HIP_CHECK(hipMemcpy(stage.data(), src + q, m * 8, hipMemcpyDeviceToHost));
          w.put(stage.data(), m * 8);
        }
        w.end();
      }
    }
    if (std::fflush(f.get()) != 0)
      throw ufc::Error{UFC_STATUS_ERROR,
                       std::string("Restart save: Write error - ") + std::strerror(errno)};
    f.reset();
    if (std::rename(tmp.c_str(), path) != 0)
      throw ufc::Error{UFC_STATUS_ERROR, std::string("Restart save: Cannot open file - ") +
                                             path + " - " + std::strerror(errno)};
    msg = "Restart data saved to " + std::string(path) + " (" + std::to_string(n) + " DOFs)";
    return UFC_STATUS_OK;
  });
}

int ufc_rt_restart_load(ufc_ctx* c, const char* path, ufc_rt_restart_data* d,
                        int32_t with_state, ufc_status* st) {
  return guarded(st, "ufc_rt_restart_load", [&](std::string& msg) {
    UFC_REQUIRE(path && path[0] && d && d->u, UFC_STATUS_INVALID, "null argument");
    int64_t expect = d->n_dof;
    if (c) {
      check_mesh(c);
      UFC_REQUIRE(!with_state || c->have_j2, UFC_STATUS_INVALID,
                  "Restart restore: no J2 state to load into (ufc_ph_j2_init)");
      UFC_HIP_CHECK(hipSetDevice(c->device));
      expect = c->n_dof;
    } else {
      UFC_REQUIRE(!with_state, UFC_STATUS_INVALID,
                  "Restart restore: the J2 state needs a context");
      UFC_REQUIRE(expect > 0, UFC_STATUS_INVALID, "Restart restore: n_dof gives no capacity");
    }
    File f(std::fopen(path, "rb"));
    UFC_REQUIRE(f, UFC_STATUS_INVALID, std::string("Restart restore: File not found: ") + path);
    Reader r(f.get());

    unsigned char hdr[20];
    r.begin(12, "header");
    r.get(hdr, 12, "header");
    r.end("header");
    const unsigned char* p = hdr;
    const int32_t version = take<int32_t>(p);
    (void)take<int32_t>(p);  // date stamp
    const int32_t n = take<int32_t>(p);
    UFC_REQUIRE(version == kVersion, UFC_STATUS_INVALID,
                "Restart restore: Incompatible file version: " + std::to_string(version));
    UFC_REQUIRE(n > 0, UFC_STATUS_INVALID, "Restart restore: Invalid data size");
    UFC_REQUIRE(int64_t(n) == expect, UFC_STATUS_INVALID,
                "Restart restore: file holds " + std::to_string(n) + " DOFs, expected " +
                    std::to_string(expect));

    std::vector<double> u(static_cast<size_t>(n));
    r.begin(8 * int64_t(n), "u");
    r.get(u.data(), size_t(n) * 8, "u");
    r.end("u");
    for (const char* what : {"v", "a"}) {
      r.begin(8 * int64_t(n), what);
      r.skip(8 * int64_t(n), what);
      r.end(what);
    }
    r.begin(20, "step state");
    r.get(hdr, 20, "step state");
    r.end("step state");
    p = hdr;
    const double time = take<double>(p);
    const int32_t increment = take<int32_t>(p);
    const double lambda = take<double>(p);
    r.begin(16, "convergence");
    r.get(hdr, 16, "convergence");
    r.end("convergence");
    p = hdr;
    const double rnorm = take<double>(p);
    const int32_t iterations = take<int32_t>(p);
    const int32_t converged = take<int32_t>(p);
    int32_t checksum = 0;
    r.begin(4, "checksum");
    r.get(&checksum, 4, "checksum");
    r.end("checksum");
    UFC_REQUIRE(checksum == n, UFC_STATUS_ERROR,
                "Restart restore: Checksum mismatch (file corrupted)");

    if (with_state) {
      UFC_REQUIRE(!r.at_eof(), UFC_STATUS_ERROR, "Restart restore: the file holds no J2 state");
      const int64_t ne = c->n_elems;
      const int32_t n_chunks = int32_t((ne + kChunkElems - 1) / kChunkElems);
      int32_t sh[5];
      r.begin(20, "state header");
      r.get(sh, 20, "state header");
      r.end("state header");
      UFC_REQUIRE(sh[0] == ne && sh[1] == kPoints && sh[2] == kNtens && sh[3] == kNstatv &&
                      sh[4] == n_chunks,
                  UFC_STATUS_INVALID,
                  "Restart restore: J2 state of " + std::to_string(sh[0]) +
                      " elements does not fit the mesh of " + std::to_string(ne));
      // The records go to the trial copy first and replace the committed
      // state only once every one of them has been read: a file that ends
      // (or is damaged) inside the state leaves the committed state as it
      // was, and the trial copy is reset to it.
      std::vector<double> stage(kStageDoubles);
      try {
        for (int64_t e0 = 0; e0 < ne; e0 += kChunkElems) {
          const int64_t vals = std::min(kChunkElems, ne - e0) * kPoints * kVals;
          double* dst = c->j2_trial.get() + e0 * kPoints * kVals;
          r.begin(vals * 8, "state");
          for (int64_t q = 0; q < vals; q += int64_t(kStageDoubles)) {
            const size_t m = size_t(std::min<int64_t>(int64_t(kStageDoubles), vals - q));
            r.get(stage.data(), m * 8, "state");
            UFC_HIP_CHECK(hipMemcpy(dst + q, stage.data(), m * 8, hipMemcpyHostToDevice));
          }
          r.end("state");
        }
      } catch (...) {
        (void)hipMemcpy(c->j2_trial.get(), c->j2_commit.get(), c->j2_commit.bytes(),
                        hipMemcpyDeviceToDevice);
        throw;
      }
      UFC_HIP_CHECK(hipMemcpy(c->j2_commit.get(), c->j2_trial.get(), c->j2_trial.bytes(),
                              hipMemcpyDeviceToDevice));
    }
    if (c)
      UFC_HIP_CHECK(hipMemcpy(d->u, u.data(), size_t(n) * 8, hipMemcpyDefault));
    else
      std::memcpy(d->u, u.data(), size_t(n) * 8);
    d->n_dof = n;
    d->time = time;
    d->increment = increment;
    d->lambda = lambda;
    d->residual_norm = rnorm;
    d->iterations = iterations;
    d->converged = converged != 0;
    msg = "Restart data restored from " + std::string(path) + " (" + std::to_string(n) + " DOFs)";
    return UFC_STATUS_OK;
  });
}

}  // extern "C"
