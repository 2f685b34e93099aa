// The operator entry point that allocates nothing (globe_vx_dev_noalloc).
//
// Its workspaces -- segment sums, the genetic product's digit matrix, split partials and
// scale factors, GRM product partials -- are slices of one file-scope device array of
// kScratchBytes. The runtime places that array on every device the library runs on and
// keeps it until the process ends. Its device address comes from hipGetSymbolAddress,
// called once per device (with that device current) when the first context on it is
// created; host code passes that address, never the array's name, to kernels.
#include <mutex>
#include <vector>

#include "context.h"

namespace {

constexpr size_t kScratchBytes = size_t(128) << 20;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

}  // namespace

namespace globe {

void set_noalloc(bool on);

// Device address of this device's scratch (the current device), looked up once per device.
void* scratch_address(int device, size_t* bytes) {
    static std::mutex mu;
    static std::vector<void*> addr;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= device) addr.resize((size_t)device + 1, nullptr);
    if (!addr[(size_t)device]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, GLOBE_E_DEVICE, "no device address for the static scratch");
        addr[(size_t)device] = p;
    }
    if (bytes) *bytes = kScratchBytes;
    return addr[(size_t)device];
}

}  // namespace globe

using namespace globe;

extern "C" {

int globe_vx_dev_noalloc(globe_ctx* c, const double* X, int64_t ldx, int k, double* Y, int64_t ldy, int mode,
                         int precision) {
    API_BEGIN(c)
    require(c->scratch != nullptr, GLOBE_E_STATE, "context has no static scratch");
    require(X && Y && k >= 1 && k <= kMaxCols, GLOBE_E_INVALID, "globe_vx_dev_noalloc: bad arguments");
    require(mode == GLOBE_VX_FULL || mode == GLOBE_VX_GRM_ONLY, GLOBE_E_INVALID, "unknown mode");
    require(precision >= GLOBE_PREC_FP64 && precision <= GLOBE_PREC_FP64_VALU, GLOBE_E_INVALID,
            "unknown precision");
    require(c->has_cohort && c->gen != 0, GLOBE_E_STATE, "cohort and genetics required");
    require(c->has_theta && c->theta_k >= k, GLOBE_E_STATE, "variance components not set (globe_set_varcomp)");
    require(c->serve == nullptr, GLOBE_E_STATE, "serving loop open");
    const int64_t rows_out = mode == GLOBE_VX_FULL ? c->nrec : c->n;
    require(ldx >= c->nrec && ldy >= rows_out, GLOBE_E_INVALID, "leading dimension too small");
    // every workspace of the call is a slice of the scratch; an allocation would throw
    struct Scope {
        globe_ctx* c;
        Scope(globe_ctx* c_) : c(c_) {
            c->arena = globe_ctx::Arena{};
            c->arena.base = static_cast<uint8_t*>(c->scratch);
            c->arena.cap = c->scratch_bytes;
            set_noalloc(true);
        }
        ~Scope() {
            set_noalloc(false);
            c->arena = globe_ctx::Arena{};
        }
    } scope(c);
    op_vx(c, X, ldx, k, Y, ldy, mode, precision, nullptr, nullptr);
    HIP_CHECK(hipGetLastError());
    API_END(c)
}

}  // extern "C"
