// bk_driver.cpp -- per-order drivers BK1, BK3 and BK5 (one binary each,
// -DBK_OP=1/3/5), the MI355X counterparts of openmp/BK1.cpp, openmp/BK3.cpp
// and openmp/BK5.cpp:
//   BKn <p> [nelmt] [ntests]      (defaults p = 2, 2, 3; nelmt = 524288; ntests = 5)
// Same data (cos test basis, JxW = 1 / G = 2 and in = 3, or the LCG data
// with BK_RANDOM=1), same minimum-wall-time loop and the same two stdout
// lines, so the reference's sweep and pick scripts parse them unchanged:
//   SumFactorization -> nelmt = N GDoF/s = x GB/s = y
//   norm = z
// The arithmetic is FP64 (the reference drivers run float).  The arrays are
// copied to the device once, as the reference's `omp target data` does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <limits>
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

// LCG of the reference's test data: in first, then 1.5 + draw
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
    const std::size_t nelmt = argc > 2 ? std::size_t(std::atoll(argv[2])) : (std::size_t(1) << 19);
    const int ntests = argc > 3 ? std::atoi(argv[3]) : 5;
    if (p < bkx_min_order(op) || p > bkx_max_order(op)) {
        std::cerr << "unsupported polynomial order p = " << p << " (supported: " << bkx_min_order(op) << ".."
                  << bkx_max_order(op) << ")\n";
        return 1;
    }
    const bool random = std::getenv("BK_RANDOM") != nullptr;
    const int nq = op == 5 ? p + 1 : p + 2;
    const int nm = op == 5 ? nq : p + 1;
    const std::size_t nin = nelmt * nm * nm * nm;
    const std::size_t nqq = nelmt * (op == 1 ? 1 : 6) * nq * nq * nq;

    std::vector<double> in(nin, 3.0);
    std::vector<double> q(nqq, op == 1 ? 1.0 : 2.0);
    if (random) {
        fill_random(in, q);
    }
    const std::vector<double> basis = test_basis(nm, nq);
    const std::vector<double> dbasis = test_basis(nq, nq);
    double* d_in = to_device(in);
    double* d_q = to_device(q);
    double* d_b = to_device(basis);
    double* d_d = to_device(dbasis);
    double* d_out = to_device(std::vector<double>(nin, 0.0));

    auto apply = [&] {
        if (op == 1) {
            return bkx_bk1_apply_f64(p, nelmt, d_b, d_q, d_in, d_out, nullptr);
        }
        if (op == 3) {
            return bkx_bk3_apply_f64(p, nelmt, d_b, d_d, d_q, d_in, d_out, nullptr);
        }
        return bkx_bk5_apply_f64(p, nelmt, d_d, d_q, d_in, d_out, nullptr);
    };
    int status = apply();   // warm-up
    double best = std::numeric_limits<double>::max();
    for (int t = 0; t < ntests && status == BKX_SUCCESS; ++t) {
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const auto t0 = std::chrono::steady_clock::now();
        status = apply();
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const auto t1 = std::chrono::steady_clock::now();
        best = std::min(best, std::chrono::duration<double>(t1 - t0).count());
    }
    if (status != BKX_SUCCESS) {
        std::cerr << "bkx error: " << bkx_status_string(status) << "\n";
        return 1;
    }
    std::vector<double> out(nin);
    hip_check(hipMemcpy(out.data(), d_out, nin * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    double ss = 0.0;
    for (const double v : out) {
        ss += v * v;
    }
    const std::size_t bytes = sizeof(double) * (2 * nin + nqq);
    std::cout << "SumFactorization -> nelmt = " << nelmt << " GDoF/s = " << 1.0e-9 * double(nin) / best
              << " GB/s = " << 1.0e-9 * double(bytes) / best << "\n";
    std::cout << "norm = " << std::sqrt(ss) << "\n";
    for (double* d : {d_in, d_q, d_b, d_d, d_out}) {
        (void)hipFree(d);
    }
    return 0;
}
