// bkx_api.hip -- the C ABI of include/bkx.h: argument checks, then the
// launch of the operator's kernel on the caller's stream.
#include <hip/hip_runtime.h>

#include "bkx.h"
#include "bkx_launch.hpp"

namespace bkx {
int bk1_apply_f64(int p, size_t nelmt, const double* B, const double* JxW, const double* u, double* v,
                  hipStream_t s);
int bk3_apply_f64(int p, size_t nelmt, const double* B, const double* D, const double* G, const double* u,
                  double* v, hipStream_t s);
int bk5_apply_f64(int p, size_t nelmt, const double* D, const double* G, const double* u, double* v,
                  hipStream_t s);
}  // namespace bkx

namespace {

bool order_ok(int op, int p) { return p >= bkx_min_order(op) && p <= bkx_max_order(op); }

hipStream_t as_stream(bkx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

const char* bkx_status_string(int status)
{
    switch (status) {
    case BKX_SUCCESS:
        return "success";
    case BKX_ERROR_ORDER:
        return "unsupported polynomial order p";
    case BKX_ERROR_NULL:
        return "null pointer argument";
    case BKX_ERROR_LAUNCH:
        return "kernel launch failed";
    case BKX_ERROR_SIZE:
        return "too many elements for one launch";
    case BKX_ERROR_ARG:
        return "invalid argument";
    case BKX_ERROR_DEVICE:
        return "HIP device error";
    default:
        return "unknown status";
    }
}

int bkx_version(void) { return 100; }

int bkx_min_order(int op) { return (op == 1 || op == 3 || op == 5) ? 1 : -1; }

int bkx_max_order(int op)
{
    if (op == 1 || op == 3) {
        return 8;
    }
    return op == 5 ? 7 : -1;
}

int bkx_bk1_apply_f64(int p, size_t nelmt, const double* basis, const double* JxW, const double* in, double* out,
                      bkx_stream_t stream)
{
    if (!order_ok(1, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !JxW || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk1_apply_f64(p, nelmt, basis, JxW, in, out, as_stream(stream));
}

int bkx_bk3_apply_f64(int p, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                      const double* in, double* out, bkx_stream_t stream)
{
    if (!order_ok(3, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!basis || !dbasis || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk3_apply_f64(p, nelmt, basis, dbasis, G, in, out, as_stream(stream));
}

int bkx_bk5_apply_f64(int p, size_t nelmt, const double* dbasis, const double* G, const double* in, double* out,
                      bkx_stream_t stream)
{
    if (!order_ok(5, p)) {
        return BKX_ERROR_ORDER;
    }
    if (nelmt == 0) {
        return BKX_SUCCESS;
    }
    if (!dbasis || !G || !in || !out) {
        return BKX_ERROR_NULL;
    }
    return bkx::bk5_apply_f64(p, nelmt, dbasis, G, in, out, as_stream(stream));
}

}  // extern "C"
