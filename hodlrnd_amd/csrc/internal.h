// internal.h -- shared declarations of libhodlrnd (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hodlrnd.h"

namespace hodlr {

constexpr int kMaxDim = 8;

struct HipError {
    hipError_t code;
    std::string where;
};
struct ArgError {
    int code;
    std::string msg;
};

#define HIP_CHECK(expr)                                                             \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            throw ::hodlr::HipError{e_, std::string(#expr) + " @" + __FILE__ + ":" + \
                                            std::to_string(__LINE__)};              \
    } while (0)

// One device allocation carved up by a first-fit sub-allocator (256-byte
// aligned, coalescing, thread-safe).  A solve server takes its device
// memory from one slab allocated at start-up, so a stream of builds,
// engines and evictions never hands memory back to the driver: the driver
// clears freed device memory before it can be allocated again (~60 GB/s
// measured), and allocations that waited for that clearing slowed a
// serving stream ~3x.
struct DevSlab {
    char *base = nullptr;
    size_t size = 0;
    DevSlab(const DevSlab &) = delete;
    DevSlab &operator=(const DevSlab &) = delete;
    explicit DevSlab(size_t bytes);
    ~DevSlab();
    void *alloc(size_t nb);   // nullptr when no free run fits
    void free(void *p);
    size_t largest_free();
    size_t in_use();

  private:
    struct Impl;
    Impl *impl;
};

// the slab that DevBuf allocations made on this thread come from (nullptr:
// the driver); a buffer that does not fit in it falls back to the driver
extern thread_local DevSlab *tl_slab;

// set while a thread releases buffers after it has synchronised the device
// itself (one wait for a whole operator or engine instead of one per buffer)
extern thread_local bool tl_slab_synced;

// free device memory for sizing decisions: the slab's largest free run on a
// slab thread, else the driver's free memory
size_t device_free_bytes();

// Device buffer owned by an object; freed in the destructor.  A buffer
// allocated with a stream is stream-ordered (hipMallocAsync / hipFreeAsync
// on that stream): freeing it does not synchronise the device, so
// temporaries of one assembly never stall work on other streams.  On a
// slab thread the buffer comes from the slab; its release waits for its
// stream (or, without one, the device, as hipFree does) before the range
// is reused.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t fs = nullptr;  // stream-ordered when set
    DevSlab *slab = nullptr;   // the slab p came from
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), fs(o.fs), slab(o.slab) {
        o.p = nullptr;
        o.bytes = 0;
        o.fs = nullptr;
        o.slab = nullptr;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            fs = o.fs;
            slab = o.slab;
            o.p = nullptr;
            o.bytes = 0;
            o.fs = nullptr;
            o.slab = nullptr;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t nb, hipStream_t stream = nullptr) {
        release();
        if (nb == 0) nb = 8;
        if (tl_slab && (p = tl_slab->alloc(nb)) != nullptr) {
            bytes = nb;
            fs = stream;
            slab = tl_slab;
            return;
        }
        hipError_t e = stream ? hipMallocAsync(&p, nb, stream) : hipMalloc(&p, nb);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw ArgError{HODLR_E_NOMEM, "device allocation of " + std::to_string(nb) + " bytes failed"};
        }
        bytes = nb;
        fs = stream;
    }
    void release() {
        if (p && slab) {
            if (!tl_slab_synced) (void)(fs ? hipStreamSynchronize(fs) : hipDeviceSynchronize());
            slab->free(p);
        } else if (p) {
            (void)(fs ? hipFreeAsync(p, fs) : hipFree(p));
        }
        p = nullptr;
        bytes = 0;
        fs = nullptr;
        slab = nullptr;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// v to a new device buffer.  With a stream the copy is ordered on it (and
// waited for, so v may go away); temp = true makes the buffer itself
// stream-ordered.
template <class T>
void upload(DevBuf &b, const T *v, size_t n, hipStream_t s = nullptr, bool temp = false) {
    b.alloc(n * sizeof(T), temp ? s : nullptr);
    if (n == 0) return;
    if (s) {
        HIP_CHECK(hipMemcpyAsync(b.p, v, n * sizeof(T), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
    } else {
        HIP_CHECK(hipMemcpy(b.p, v, n * sizeof(T), hipMemcpyHostToDevice));
    }
}
template <class T>
void upload(DevBuf &b, const std::vector<T> &v, hipStream_t s = nullptr, bool temp = false) {
    upload(b, v.data(), v.size(), s, temp);
}

}  // namespace hodlr

struct hodlr_ctx {
    int device = 0;
    std::string last_error;
    hipStream_t stream = nullptr;
    // assembly staging of the rank pass's factors: -1 default room, 0 off,
    // > 0 at most that many bytes
    int64_t stage_bytes = -1;
};

// Uniform-depth 2^d tree (HDD_tree.hpp:23-74).  Nodes are numbered level by
// level; node g = level_off[l] + j is node j of level l, whose children are
// nodes 2^d j .. 2^d j + 2^d - 1 of level l + 1.  Every node owns the range
// [start, start + count) of the leaf-order permutation.
struct hodlr_tree {
    hodlr_ctx *ctx = nullptr;
    int ndim = 0, nlevels = 0, nmax = 0, dprime = 0;
    int64_t n = 0;
    std::vector<double> pts;          // point order, n x ndim
    std::vector<int64_t> perm;        // tree position -> point id
    std::vector<int64_t> level_off;   // nlevels + 1
    std::vector<int64_t> start, count;
    std::vector<double> lo, hi, len, centre;  // per node: box, max side, centre
    std::vector<int64_t> nbr_ptr, nbr, il_ptr, il;
    int64_t node(int l, int64_t j) const { return level_off[l] + j; }
    int64_t level_size(int l) const { return level_off[l + 1] - level_off[l]; }
};

namespace hodlr {
void build_tree(hodlr_tree &t);
}
