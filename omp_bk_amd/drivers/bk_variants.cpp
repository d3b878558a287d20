// bk_variants.cpp -- the variant-table harnesses BK1_variants, BK3_variants
// and BK5_variants (-DBK_OP=1/3/5), the counterparts of the reference's
// openmp/offload/bk{1,3,5}_offload.cpp: every launch variant of the FP64
// operator at one order, on the reference harness's data (cos test basis,
// LCG data: in in [-1, 1), JxW / G = 1.5 + [-1, 1)), one line each in the
// reference's format
//   <name> GDoF/s = x  GB/s = y  time = t ms  relerr = r[  FAIL]
// so that its pick scripts (scripts/study/*_pick.sh) parse them unchanged.
//   BKn_variants <p> [nelmt] [ntests]
// Variant names are v0, v1, ... (bkx_tune_count); BK_KERNEL="a b -c" runs
// the variants named, all if none is named, never those named with a
// leading "-" (include/bk_harness.h:20-40).  The time is the minimum over
// ntests applies after a warm-up, host wall clock around a synchronised
// call (include/bk_common.h:181-192).  relerr is max|out - ref| / max|ref|
// against variant 0 -- a different mapping from the tuned defaults -- where
// the reference compares with its host FP64 reference; FAIL marks >= 1e-5
// (include/bk_harness.h:63-85).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "bkx.h"

#ifndef BK_OP
#error "compile with -DBK_OP=1, 3 or 5"
#endif

namespace {

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::cerr << what << ": " << hipGetErrorString(e) << "\n";
        std::exit(1);
    }
}

bool selected(const std::string& name)
{
    const char* env = std::getenv("BK_KERNEL");
    std::istringstream ss(env ? env : "");
    std::string w;
    bool any = false;
    bool named = false;
    while (ss >> w) {
        if (w[0] == '-') {
            if (w.compare(1, std::string::npos, name) == 0) {
                return false;
            }
        } else {
            any = true;
            named = named || w == name;
        }
    }
    return !any || named;
}

std::vector<double> test_basis(int rows, int cols)
{
    std::vector<double> m(std::size_t(rows) * cols);
    for (int a = 0; a < rows; ++a) {
        for (int b = 0; b < cols; ++b) {
            m[std::size_t(a) * cols + b] = std::cos(double(a * cols + b));
        }
    }
    return m;
}

void fill_random(std::vector<double>& in, std::vector<double>& q)
{
    uint32_t s = 12345u;
    auto next = [&] {
        s = 1664525u * s + 1013904223u;
        return double(s >> 8) / double(1u << 23) - 1.0;
    };
    for (auto& v : in) {
        v = next();
    }
    for (auto& v : q) {
        v = 1.5 + next();
    }
}

double* to_device(const std::vector<double>& h)
{
    double* d = nullptr;
    if (hipMalloc(&d, std::max<std::size_t>(1, h.size()) * sizeof(double)) != hipSuccess ||
        hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        std::cerr << "device allocation failed\n";
        std::exit(1);
    }
    return d;
}

}  // namespace

int main(int argc, char** argv)
{
    const int op = BK_OP;
    const int p = argc > 1 ? std::atoi(argv[1]) : (op == 5 ? 3 : 2);
    const std::size_t nelmt = argc > 2 ? std::size_t(std::atoll(argv[2])) : (std::size_t(1) << 17);
    const int ntests = argc > 3 ? std::atoi(argv[3]) : 10;
    if (p < bkx_min_order(op) || p > bkx_max_order(op)) {
        std::cerr << "unsupported polynomial order p = " << p << " (supported: " << bkx_min_order(op) << ".."
                  << bkx_max_order(op) << ")\n";
        return 1;
    }
    const int nq = op == 5 ? p + 1 : p + 2;
    const int nm = op == 5 ? nq : p + 1;
    const std::size_t nin = nelmt * nm * nm * nm;
    const std::size_t nqq = nelmt * (op == 1 ? 1 : 6) * nq * nq * nq;
    std::vector<double> in(nin);
    std::vector<double> q(nqq);
    fill_random(in, q);
    double* d_in = to_device(in);
    double* d_q = to_device(q);
    double* d_b = to_device(test_basis(nm, nq));
    double* d_d = to_device(test_basis(nq, nq));
    double* d_out = to_device(std::vector<double>(nin, 0.0));
    std::printf("BK%d hip: p = %d, nq = %d, nelmt = %zu (%zu DoFs), ntests = %d\n", op, p, nq, nelmt, nin, ntests);

    auto apply = [&] {
        if (op == 1) {
            return bkx_bk1_apply_f64(p, nelmt, d_b, d_q, d_in, d_out, nullptr);
        }
        if (op == 3) {
            return bkx_bk3_apply_f64(p, nelmt, d_b, d_d, d_q, d_in, d_out, nullptr);
        }
        return bkx_bk5_apply_f64(p, nelmt, d_d, d_q, d_in, d_out, nullptr);
    };
    const int count = bkx_tune_count(op, BKX_FORM_F64);
    const int prev = bkx_tune_get(op, BKX_FORM_F64, p);
    const double bytes = sizeof(double) * (2.0 * double(nin) + double(nqq));
    std::vector<double> ref;
    double big = 0.0;
    std::vector<double> out(nin);
    int status = BKX_SUCCESS;
    for (int v = 0; v < count && status == BKX_SUCCESS; ++v) {
        const std::string name = "v" + std::to_string(v);
        if (v > 0 && !selected(name)) {
            continue;   // variant 0 always runs: it is the comparison
        }
        bkx_tune_set(op, BKX_FORM_F64, p, v);
        hip_check(hipMemset(d_out, 0, nin * sizeof(double)), "hipMemset");
        status = apply();   // warm-up
        double best = std::numeric_limits<double>::max();
        for (int t = 0; t < ntests && status == BKX_SUCCESS; ++t) {
            hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            const auto t0 = std::chrono::steady_clock::now();
            status = apply();
            hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
        if (status != BKX_SUCCESS) {
            break;
        }
        hip_check(hipMemcpy(out.data(), d_out, nin * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        if (v == 0) {
            ref = out;
            for (const double x : ref) {
                big = std::max(big, std::fabs(x));
            }
        }
        double diff = 0.0;
        for (std::size_t i = 0; i < nin; ++i) {
            diff = std::max(diff, std::fabs(out[i] - ref[i]));
        }
        const double relerr = big > 0.0 ? diff / big : diff;
        if (v == 0 && !selected(name)) {
            continue;
        }
        std::printf("%-10s GDoF/s = %7.3f  GB/s = %6.1f  time = %8.3f ms  relerr = %.1e%s\n", name.c_str(),
                    1e-9 * double(nin) / best, 1e-9 * bytes / best, 1e3 * best, relerr,
                    relerr >= 1e-5 ? "  FAIL" : "");
    }
    bkx_tune_set(op, BKX_FORM_F64, p, prev);
    for (double* d : {d_in, d_q, d_b, d_d, d_out}) {
        (void)hipFree(d);
    }
    if (status != BKX_SUCCESS) {
        std::cerr << "bkx error: " << bkx_status_string(status) << "\n";
        return 1;
    }
    return 0;
}
