// noalloc.hip -- the allocation-free matvec (hodlr_matvec_device_noalloc).
//
// hodlr_matvec_device keeps its intermediates (the projection's V^T x values
// and partial-sum slots, the skeleton phase's slot values) in buffers the
// operator allocates on the first product of each (precision, RHS count).
// Callers that cannot allocate -- code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc -- use this entry point
// instead: the same kernels, with the intermediates carved out of one
// fixed device array, `scratch`, of kScratchBytes per GPU.  Nothing here
// calls hipMalloc, hipFree or any allocating library routine.
//
// The array is a file-scope __device__ object: each GPU the library runs on
// holds its own copy (at its own address) until the process ends.  In host
// code its name is only a host-side stand-in, so its device address is
// looked up once per device with hipGetSymbolAddress (at context creation,
// with that device current) and only that address is ever handed to
// kernels.  Every call on one GPU shares that GPU's array: calls that
// overlap on one GPU (different streams or threads) race and must be
// serialised by the caller.  No kernel reads an intermediate it has not
// written in the same product (the skeleton phases' FP64 result vector is
// zeroed by a kernel of the product itself), so nothing needs clearing
// between calls.
#include <mutex>
#include <vector>

#include "op.h"

namespace hodlr {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
constexpr size_t kScratchAlign = 256;

}  // namespace hodlr

static __device__ __align__(256) unsigned char scratch[hodlr::kScratchBytes];

namespace hodlr {

namespace {
std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // per device; nullptr until looked up
}  // namespace

// Looks up this device's copy of the array once (the device must be
// current).  Called from hodlr_ctx_create and, for a device no context was
// created on in this process, from the entry point's first call there.
void *noalloc_scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0) throw ArgError{HODLR_E_STATE, "host-only context has no device scratch"};
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) throw HipError{e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                    "hipGetSymbolAddress(scratch)"};
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

}  // namespace hodlr

using namespace hodlr;

extern "C" int hodlr_matvec_device_noalloc(hodlr_op *op, int prec, int nrhs, const void *x, void *y,
                                           void *stream) {
    if (!op || !x || !y) return HODLR_E_ARG;
    hodlr_ctx *ctx = op->ctx;
    try {
        if (nrhs != 1 && nrhs != 2 && nrhs != 4 && nrhs != 8 && nrhs != 16 && nrhs != 32)
            throw ArgError{HODLR_E_ARG, "nrhs must be 1, 2, 4, 8, 16 or 32"};
        if (prec != HODLR_PREC_F64 && prec != HODLR_PREC_F32) throw ArgError{HODLR_E_ARG, "bad precision"};
        if (!(op->prec_flags & prec)) throw ArgError{HODLR_E_STATE, "operator has no copy in that precision"};
        if (prec == HODLR_PREC_F64 && nrhs > 16 && op->n_skel > 0)
            throw ArgError{HODLR_E_ARG, "FP64 skeleton operators support at most 16 RHS"};
        if (ctx->device < 0) throw ArgError{HODLR_E_STATE, "host-only context (device -1) cannot run device work"};
        HIP_CHECK(hipSetDevice(ctx->device));
        size_t sz[6];
        op->matvec_work_sizes(prec, nrhs, sz);
        size_t off[6], total = 0;
        for (int i = 0; i < 6; i++) {
            off[i] = total;
            total += (sz[i] + kScratchAlign - 1) / kScratchAlign * kScratchAlign;
        }
        if (total > kScratchBytes)
            throw ArgError{HODLR_E_NOMEM, "allocation-free matvec needs " + std::to_string(total) +
                                              " bytes of scratch for this operator and RHS count, more than the " +
                                              std::to_string(kScratchBytes) + " bytes it holds per GPU"};
        unsigned char *base = static_cast<unsigned char *>(noalloc_scratch_setup(ctx->device));
        MatvecWork w;
        w.t = base + off[0];
        w.part = base + off[1];
        if (sz[2]) {
            w.skt1 = reinterpret_cast<double *>(base + off[2]);
            w.skt2 = reinterpret_cast<double *>(base + off[3]);
            w.skpart = reinterpret_cast<double *>(base + off[4]);
            w.skys = reinterpret_cast<double *>(base + off[5]);
        }
        // the skeleton phases stay on the caller's stream (w.overlap false):
        // the operator's second stream is created on first use, which this
        // entry point must not do
        op->matvec_dev_on(prec, nrhs, x, y, stream ? static_cast<hipStream_t>(stream) : ctx->stream, w);
        return HODLR_OK;
    } catch (const HipError &e) {
        ctx->last_error = std::string(hipGetErrorString(e.code)) + " in " + e.where;
        return HODLR_E_DEVICE;
    } catch (const ArgError &e) {
        ctx->last_error = e.msg;
        return e.code;
    }
}
