// gnt_mul_noalloc: the product of gnt_mul_buf (device operands, device
// result) for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): every scratch buffer the product's kernels
// use lives in one file-scope device array, `scratch`, sized for transforms of
// up to 2^GNT_NOALLOC_MAX_LOG_N elements.  The caller passes the operands and
// the output (and the word that receives the trimmed length) in device memory.
//
// Each GPU the library runs on holds its own copy of `scratch` until the
// process ends.  Its device address is looked up once per device at context
// set-up (gnt_init, with that device current) and kept in the context; host
// code never passes the array's own name to a kernel (in host code that name
// is a host-side stand-in, not a device address).  Every call on one GPU uses
// that GPU's array, so calls whose work overlaps on the device race; the
// kernels write every scratch element before they read it, so nothing needs
// clearing between calls.
#include <hip/hip_runtime.h>

#include "../../include/gnt.h"
#include "gnt_carry.hpp"
#include "gnt_kernels.hpp"
#include "gnt_runtime.hpp"

namespace gnt {
namespace {

constexpr size_t NA_N = (size_t)1 << GNT_NOALLOC_MAX_LOG_N;           // coefficient array length
constexpr size_t NA_LIMBS = (size_t)1 << (GNT_NOALLOC_MAX_LOG_N - 2);  // na + nb at most
constexpr size_t NA_GP = 2 * ((NA_LIMBS + CARRY_TILE - 1) / CARRY_TILE);
// xa | xb | limb sums (2 * NA_LIMBS) | carry flags
constexpr size_t NA_WORDS = 2 * NA_N + 2 * NA_LIMBS + (NA_GP + 7) / 8;
static_assert(NA_WORDS * 8 == GNT_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

static __device__ __align__(256) uint64_t scratch[NA_WORDS];

__global__ void k_set_len(uint32_t *len) { *len = 0; }

}  // namespace

// The device address of this device's `scratch` (current device = c->device).
int noalloc_setup(Ctx *c) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        c->noalloc_scratch = nullptr;
        return set_error(GNT_E_DEVICE, "hipGetSymbolAddress(scratch) failed");
    }
    c->noalloc_scratch = static_cast<uint64_t *>(p);
    return GNT_OK;
}

}  // namespace gnt

using gnt::Ctx;

extern "C" int gnt_mul_noalloc(gnt_ctx *ctx, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb,
                               uint64_t *d_out, size_t out_cap, uint32_t *d_len, void *stream) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return gnt::set_error(GNT_E_ARG, "null context");
    if (!d_len) return gnt::set_error(GNT_E_ARG, "null length word");
    if (hipSetDevice(c->device) != hipSuccess) return gnt::set_error(GNT_E_DEVICE, "hipSetDevice failed");
    if (!c->noalloc_scratch) return gnt::set_error(GNT_E_DEVICE, "no scratch address for this device");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const bool sqr = d_b == nullptr;
    if (sqr) nb = na;
    if (na == 0 || nb == 0) {
        hipLaunchKernelGGL(gnt::k_set_len, dim3(1), dim3(1), 0, st, d_len);
        return hipGetLastError() == hipSuccess ? GNT_OK : gnt::set_error(GNT_E_DEVICE, "launch failed");
    }
    if (!d_a || !d_out) return gnt::set_error(GNT_E_ARG, "null operand or output");
    if (out_cap < na + nb) return gnt::set_error(GNT_E_ARG, "out_cap < na + nb");
    gnt::Plan pl;
    if (!gnt::plan_for(na, nb, pl) || pl.log_n > GNT_NOALLOC_MAX_LOG_N)
        return gnt::set_error(GNT_E_TOO_LARGE, "gnt_mul_noalloc: 4(na+nb) exceeds 2^22 (its scratch size)");
    const gnt::ProductScratch s = gnt::product_scratch_layout(na, nb, sqr, pl.log_n);
    uint64_t *xa = c->noalloc_scratch;
    uint64_t *xb = xa + gnt::NA_N;
    uint64_t *lh = xb + gnt::NA_N;
    unsigned char *gp = reinterpret_cast<unsigned char *>(lh + 2 * gnt::NA_LIMBS);
    // no pass twiddle tables (they are allocated on first use): every strided
    // pass steps its twiddles; the product is the same
    const gnt::PassTables none;
    const int rc = gnt::launch_product(pl, none, c->tw(), d_a, na, sqr ? nullptr : d_b, nb, d_out, d_len, s, xa,
                                       sqr ? nullptr : xb, lh, gp, st, nullptr);
    if (rc) return gnt::set_error(GNT_E_DEVICE, "kernel launch failed");
    return GNT_OK;
}
