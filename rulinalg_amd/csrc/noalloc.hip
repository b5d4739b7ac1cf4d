// An LU factorization entry point that allocates no device memory: the same
// kernels as rlg_lu_decompose_f64_dev, with their scratch (the panel
// exchange slots, the inverses of the diagonal blocks and the exchange
// timeout flag) in a static device array instead of pooled memory.
//
// Each GPU holds its own copy of `scratch` for the life of the process.  Its
// device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and kept; host code never passes the array's name
// itself to a kernel.  Every call on one GPU shares that GPU's copy: calls
// that overlap on one GPU race, so a caller runs them on one stream or
// orders them.  The kernels clear what must start at zero (exchange tags,
// timeout flag) at the start of every factorization, with fill kernels.
#include "kernels.hpp"

namespace rlg {

static __device__ __align__(256) unsigned char scratch[RLG_LU_NOALLOC_SCRATCH];

namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {nullptr};

// The device address of this GPU's scratch array (looked up on first use
// with the device current).
int scratch_of(Device *d, void **out) {
    if (d->id < 0 || d->id >= 64) { set_error("device index %d out of range", d->id); return RLG_E_DEVICE; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[d->id]) {
        void *p = nullptr;
        RLG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) { set_error("scratch array has no device address"); return RLG_E_DEVICE; }
        g_scratch_ptr[d->id] = p;
    }
    *out = g_scratch_ptr[d->id];
    return RLG_OK;
}

// info = -1 when the panel exchange gave up waiting.
__global__ void k_fold_timeout(int64_t *info, const int *timeout) {
    if (threadIdx.x == 0 && *timeout) *info = -1;
}

}  // namespace
}  // namespace rlg

using namespace rlg;

extern "C" {

int rlg_lu_factor_f64_noalloc(size_t n, double *a, size_t lda, int64_t *ipiv, int64_t *info,
                              void *stream) {
    if (lda < n || (n && (!a || !ipiv)) || !info) { set_error("bad operand"); return RLG_E_INVALID_ARG; }
    const size_t need = lu_work_bytes(n, 8);
    if (need > (size_t)RLG_LU_NOALLOC_SCRATCH) {
        set_error("rlg_lu_factor_f64_noalloc needs %zu bytes of scratch; the static array holds %zu",
                  need, (size_t)RLG_LU_NOALLOC_SCRATCH);
        return RLG_E_INVALID_ARG;
    }
    Device *d;
    RLG_TRY(current_device(&d));
    void *work = nullptr;
    RLG_TRY(scratch_of(d, &work));
    hipStream_t s = pick_stream(d, stream);
    RLG_TRY(lu_factor<double>(s, n, a, lda, ipiv, info, work));
    hipLaunchKernelGGL(k_fold_timeout, dim3(1), dim3(64), 0, s, info, lu_timeout_flag(work, n, 8));
    RLG_LAUNCHED();
    return RLG_OK;
}

}  // extern "C"
