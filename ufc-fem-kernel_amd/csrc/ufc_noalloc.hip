// ufc_noalloc.hip -- the allocation-free variant of the batched element
// stiffness entry point (ufc_ph_c3d8_eval_ke_noalloc).
//
// ufc_ph_c3d8_eval_ke allocates device memory per call: staging copies of
// host arrays and, for the incompatible-mode element, the condensation
// factor C (24 x 9 doubles per element) that k_incompat_c writes and
// k_eval_ke<INCOMPAT> reads.  Callers that cannot allocate -- code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc -- call this
// variant instead: every array is the caller's device memory, and C lives in
// `scratch`, a file-scope device array of kNoallocScratchBytes that each GPU
// holds for the life of the process.
//
// `scratch` names a host-side stand-in in host code, not a device address:
// its device address on each GPU comes from hipGetSymbolAddress, taken once
// per device (with that device current) when the first context is created
// on it, and only that address is passed to kernels.  Every call on a GPU
// shares its array, so calls that overlap on one GPU (several contexts or
// host threads) race; k_incompat_c writes every element's factor before the
// element kernel reads it, so nothing in the array has to start at zero.
#include <mutex>

#include "ufc_internal.hpp"

namespace {

constexpr size_t kNoallocScratchBytes = size_t(16) << 20;  // 16 MiB per GPU
constexpr size_t kNoallocScratchDoubles = kNoallocScratchBytes / sizeof(double);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` per GPU

bool on_device(const void* p, int device) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice && a.device == device;
}

}  // namespace

namespace ufc {

// Resolves this GPU's address of `scratch` once; `device` must be current.
void noalloc_setup(int device) {
  UFC_REQUIRE(device >= 0 && device < kMaxDevices, UFC_STATUS_UNSUPPORTED,
              "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  UFC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  UFC_REQUIRE(p != nullptr, UFC_STATUS_ERROR, "no device address for the scratch array");
  g_scratch_dev[device] = static_cast<double*>(p);
}

}  // namespace ufc

extern "C" int ufc_ph_c3d8_eval_ke_noalloc(ufc_ctx* c, const ufc_ph_ke_algo* algo,
                                           ufc_ph_ke_args* a, ufc_status* st) {
  return ufc::guarded(st, "ufc_ph_c3d8_eval_ke_noalloc", [&](std::string&) {
    UFC_REQUIRE(c != nullptr, UFC_STATUS_INVALID, "null context");
    UFC_REQUIRE(algo && a && a->coords && a->young && a->ke, UFC_STATUS_INVALID,
                "null argument");
    UFC_REQUIRE(a->n_elems >= 0, UFC_STATUS_INVALID, "n_elems < 0");
    UFC_REQUIRE(algo->precision == UFC_PREC_FP64 || algo->precision == UFC_PREC_FP32,
                UFC_STATUS_INVALID, "precision must be FP64 or FP32");
    UFC_REQUIRE(a->nu > -1.0 && a->nu < 0.5, UFC_STATUS_INVALID, "nu out of (-1, 0.5)");
    const int variant = ufc::canonical_variant(algo->variant);
    UFC_REQUIRE(variant > 0, UFC_STATUS_INVALID, "unknown element variant");
    if (a->n_elems == 0) return UFC_STATUS_OK;
    UFC_HIP_CHECK(hipSetDevice(c->device));
    UFC_REQUIRE(on_device(a->coords, c->device) && on_device(a->young, c->device) &&
                    on_device(a->ke, c->device),
                UFC_STATUS_INVALID,
                "coords, young and ke must be device memory of the context's GPU");
    double* cf = nullptr;
    if (variant == ufc::kVarIncompat) {
      const int64_t need = int64_t(a->n_elems) * ufc::kIncompatC;
      UFC_REQUIRE(need <= int64_t(kNoallocScratchDoubles), UFC_STATUS_INVALID,
                  "incompatible-mode batch of " + std::to_string(a->n_elems) +
                      " elements exceeds the scratch (at most " +
                      std::to_string(kNoallocScratchDoubles / ufc::kIncompatC) +
                      " elements per call)");
      {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        cf = g_scratch_dev[c->device];
      }
      UFC_REQUIRE(cf != nullptr, UFC_STATUS_ERROR, "scratch not set up on this device");
    }
    ufc::launch_eval_ke(a->n_elems, a->coords, a->young, a->nu, a->ke,
                        algo->precision == UFC_PREC_FP32, variant, cf, c->stream);
    UFC_HIP_CHECK(hipStreamSynchronize(c->stream));
    return UFC_STATUS_OK;
  });
}
