// bkx_sumfact.hpp -- C++ drop-in for the reference's SumFactorization<T, nq>.
//
// The reference's kernels are templates with the order fixed at compile time
// and run synchronously (the OpenMP target region completes before return):
//   BK1 bk::SumFactorization<T, nq>(nelmt, basis, JxW, in, out)        openmp/bk1_sumfact.h:16-22
//   BK3 bk::SumFactorization<T, nq>(nelmt, basis, dbasis, G, in, out)  openmp/bk3_sumfact.h:17-24
//   BK5 bk::SumFactorization<T, nq>(nelmt, dbasis, G, in, out)         openmp/BK5.cpp:18-24
// The wrappers below keep those argument lists and that synchronous
// behaviour for T = double, with device pointers (hipMalloc'd), calling the
// C ABI on the null stream and waiting for it.  BK1 and BK5 share one
// signature, so each operator has its own namespace.  A failure throws
// bkx::error (the reference asserts or exits instead).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "bkx.h"

namespace bkx {

struct error : std::runtime_error {
    int status;
    error(const char* what, int s) : std::runtime_error(std::string(what) + ": " + bkx_status_string(s)), status(s) {}
};

inline void check(const char* what, int status)
{
    if (status != BKX_SUCCESS) {
        throw error(what, status);
    }
    if (hipStreamSynchronize(nullptr) != hipSuccess) {
        throw error(what, BKX_ERROR_DEVICE);
    }
}

namespace bk1 {
// nq = p + 2 Gauss points per direction
template <typename T, int nq>
void SumFactorization(std::size_t nelmt, const T* basis, const T* JxW, const T* in, T* out)
{
    static_assert(std::is_same_v<T, double>, "the drop-in is the FP64 apply");
    check("bkx_bk1_apply_f64", bkx_bk1_apply_f64(nq - 2, nelmt, basis, JxW, in, out, nullptr));
}
}  // namespace bk1

namespace bk3 {
// nq = p + 2 Gauss points per direction
template <typename T, int nq>
void SumFactorization(std::size_t nelmt, const T* basis, const T* dbasis, const T* G, const T* in, T* out)
{
    static_assert(std::is_same_v<T, double>, "the drop-in is the FP64 apply");
    check("bkx_bk3_apply_f64", bkx_bk3_apply_f64(nq - 2, nelmt, basis, dbasis, G, in, out, nullptr));
}
}  // namespace bk3

namespace bk5 {
// nq = p + 1 GLL points per direction
template <typename T, int nq>
void SumFactorization(std::size_t nelmt, const T* dbasis, const T* G, const T* in, T* out)
{
    static_assert(std::is_same_v<T, double>, "the drop-in is the FP64 apply");
    check("bkx_bk5_apply_f64", bkx_bk5_apply_f64(nq - 1, nelmt, dbasis, G, in, out, nullptr));
}
}  // namespace bk5

}  // namespace bkx
