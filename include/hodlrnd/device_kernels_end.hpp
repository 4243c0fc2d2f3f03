// device_kernels_end.hpp -- closes the region opened by
// device_kernels_begin.hpp and gives every kernel class with a Kernel_Fun
// its entry kernels: f(r) = Kernel_Fun(r) called on the device copy of the
// caller's kernel object (HDD_matrix.hpp:39-41 -> LowRank_matrix.hpp:64-377
// with kernel_function.hpp:20-53 evaluating user_kernel.hpp's entries).
#pragma clang force_cuda_host_device end

#ifndef HODLRND_DEVICE_KERNELS_END_ONCE
#define HODLRND_DEVICE_KERNELS_END_ONCE

#include <type_traits>

#include "device/entry_kernels.hpp"

namespace hodlr_detail {

// f(r) of kernel class K: its Kernel_Fun, called without virtual dispatch
template <class K>
struct kernel_fun_entry {
    __device__ __forceinline__ static double f(const hodlr::KernelDev &k, double r) {
        return (double)static_cast<K *>(const_cast<void *>(k.obj))->K::Kernel_Fun(r);
    }
};

template <class K>
struct device_kernel_fun<K, std::void_t<decltype(&K::Kernel_Fun)>> {
    static constexpr bool value = true;
    static const void *launchers() { return hodlr::entry_launchers<kernel_fun_entry<K>>(); }
};

}  // namespace hodlr_detail

#endif
