// bk_ff_driver.cpp -- self-checking precision drivers BK1_ff, BK3_ff and
// BK5_ff (one binary each, -DBK_OP=1/3/5), the MI355X counterparts of the
// reference's BK1_ff driver (openmp/BK1_ff.cpp:749-859):
//   BKn_ff <p> [nelmt] [ntests]    (defaults p = 2, nelmt = 524288, ntests = 5)
// Same data as the per-order drivers (cos test basis, JxW = 1 / G = 2 and
// in = 3, or the LCG data with BK_RANDOM=1).  Every arithmetic form of the
// operator runs on the same data, one row each in the reference's format
//   <label> GDoF/s = x GB/s = y norm = z
// and then the error lines against the FP64 apply:
//   <form> rel. error vs double = e          (informational rows)
//   ok|FAIL: <gated forms> ... (tol = t)
// The relative error is the L2 one, ||x - ref|| / ||ref||, accumulated in
// double (openmp/BK1_ff.cpp:749-763).  Gated against BK_TOL (default
// 1e-10, as the reference): the float-float forms on FP64 storage (ff,
// BK1/BK3) against the FP64 apply, and the FP32-storage forms (f32w, f32c)
// against the FP64 apply of the same FP32-rounded inputs.  Informational:
// the FP32-storage forms against the FP64 apply of the unrounded inputs
// (their storage rounding, ~1e-8) and split-BF16 (BK5 at p = 7, ~1e-5).
// Exit status 1 when a gated form fails or p is unsupported, as the
// reference driver's (openmp/BK1_ff.cpp:858).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <iomanip>
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

template <typename T>
T* to_device(const std::vector<T>& h)
{
    T* d = nullptr;
    hip_check(hipMalloc(&d, std::max<std::size_t>(1, h.size()) * sizeof(T)), "hipMalloc");
    hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}

template <typename T>
std::vector<T> to_host(const T* d, std::size_t n)
{
    std::vector<T> h(n);
    hip_check(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
}

template <typename T>
double norm2(const std::vector<T>& x)
{
    double s = 0.0;
    for (const T v : x) {
        s += double(v) * double(v);
    }
    return std::sqrt(s);
}

template <typename T>
double relative_error(const std::vector<T>& x, const std::vector<double>& ref)
{
    double num = 0.0, den = 0.0;
    for (std::size_t i = 0; i < ref.size(); ++i) {
        const double d = double(x[i]) - ref[i];
        num += d * d;
        den += ref[i] * ref[i];
    }
    return std::sqrt(num) / std::sqrt(den);
}

// minimum wall time of ntests applies after one warm-up, or -1 on an error
double time_min(const std::function<int()>& apply, int ntests)
{
    int status = apply();
    double best = std::numeric_limits<double>::max();
    for (int t = 0; t < ntests && status == BKX_SUCCESS; ++t) {
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const auto t0 = std::chrono::steady_clock::now();
        status = apply();
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    if (status != BKX_SUCCESS) {
        std::cerr << "bkx error: " << bkx_status_string(status) << "\n";
        return -1.0;
    }
    return best;
}

void row(const char* label, double t, std::size_t dofs, double bytes, double norm)
{
    std::cout << std::left << std::setw(12) << label << std::right << " GDoF/s = " << std::setw(9)
              << 1.0e-9 * double(dofs) / t << " GB/s = " << std::setw(9) << 1.0e-9 * bytes / t
              << " norm = " << std::setprecision(10) << norm << std::setprecision(6) << "\n";
}

}  // namespace

int main(int argc, char** argv)
{
    const int op = BK_OP;
    const int p = argc > 1 ? std::atoi(argv[1]) : 2;
    const std::size_t nelmt = argc > 2 ? std::size_t(std::atoll(argv[2])) : (std::size_t(1) << 19);
    const int ntests = argc > 3 ? std::atoi(argv[3]) : 5;
    if (p < bkx_min_order(op) || p > bkx_max_order(op)) {
        std::cerr << "unsupported polynomial order p = " << p << " (supported: " << bkx_min_order(op) << ".."
                  << bkx_max_order(op) << ")\n";
        return 1;
    }
    const bool random = std::getenv("BK_RANDOM") != nullptr;
    const char* tol_env = std::getenv("BK_TOL");
    const double tol = tol_env ? std::atof(tol_env) : 1e-10;
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
    const std::vector<float> in32(in.begin(), in.end()), q32(q.begin(), q.end());
    const std::vector<double> in_r(in32.begin(), in32.end()), q_r(q32.begin(), q32.end());

    std::cout << "SumFactorization -> nelmt = " << nelmt << " nq = " << nq
              << " data = " << (random ? "random" : "constant") << " op = BK" << op << "\n";

    double* d_b = to_device(basis);
    double* d_d = to_device(dbasis);
    double* d_in = to_device(in);
    double* d_q = to_device(q);
    double* d_in_r = to_device(in_r);
    double* d_q_r = to_device(q_r);
    float* d_in32 = to_device(in32);
    float* d_q32 = to_device(q32);
    double* d_out = to_device(std::vector<double>(nin, 0.0));
    float* d_out32 = to_device(std::vector<float>(nin, 0.0f));
    const double b64 = 8.0 * (2 * nin + nqq), b32 = 4.0 * (nin + nqq) + 8.0 * nin;

    // the FP64 operator on FP64 arrays (x, Q) into out
    auto f64 = [&](const double* x, const double* Q) {
        if (op == 1) {
            return bkx_bk1_apply_f64(p, nelmt, d_b, Q, x, d_out, nullptr);
        }
        if (op == 3) {
            return bkx_bk3_apply_f64(p, nelmt, d_b, d_d, Q, x, d_out, nullptr);
        }
        return bkx_bk5_apply_f64(p, nelmt, d_d, Q, x, d_out, nullptr);
    };
    // an FP32-storage form (BKX_FORM_F32W or BKX_FORM_F32C)
    auto f32 = [&](int form) {
        const bool c = form == BKX_FORM_F32C;
        if (op == 1) {
            return c ? bkx_bk1_apply_f32c(p, nelmt, d_b, d_q32, d_in32, d_out, nullptr)
                     : bkx_bk1_apply_f32w(p, nelmt, d_b, d_q32, d_in32, d_out, nullptr);
        }
        if (op == 3) {
            return c ? bkx_bk3_apply_f32c(p, nelmt, d_b, d_d, d_q32, d_in32, d_out, nullptr)
                     : bkx_bk3_apply_f32w(p, nelmt, d_b, d_d, d_q32, d_in32, d_out, nullptr);
        }
        return c ? bkx_bk5_apply_f32c(p, nelmt, d_d, d_q32, d_in32, d_out, nullptr)
                 : bkx_bk5_apply_f32w(p, nelmt, d_d, d_q32, d_in32, d_out, nullptr);
    };

    double t = time_min([&] { return f64(d_in, d_q); }, ntests);
    if (t < 0) {
        return 1;
    }
    const std::vector<double> ref = to_host(d_out, nin);
    row("double", t, nin, b64, norm2(ref));
    if (time_min([&] { return f64(d_in_r, d_q_r); }, 0) < 0) {
        return 1;
    }
    const std::vector<double> ref_r = to_host(d_out, nin);

    struct result {
        std::string name;
        double err;
        bool gated;
    };
    std::vector<result> res;
    auto form_row = [&](const char* label, const std::function<int()>& apply, double bytes, bool rounded,
                        bool gated) {
        const double tt = time_min(apply, ntests);
        if (tt < 0) {
            return false;
        }
        const std::vector<double> out = to_host(d_out, nin);
        row(label, tt, nin, bytes, norm2(out));
        res.push_back({label, relative_error(out, rounded ? ref_r : ref), gated});
        if (rounded) {
            res.push_back({std::string(label) + " (unrounded inputs)", relative_error(out, ref), false});
        }
        return true;
    };
    if (op != 5 && !form_row("float-float", [&] {
            return op == 1 ? bkx_bk1_apply_ff(p, nelmt, d_b, d_q, d_in, d_out, nullptr)
                           : bkx_bk3_apply_ff(p, nelmt, d_b, d_d, d_q, d_in, d_out, nullptr);
        }, b64, false, true)) {
        return 1;
    }
    if (!form_row("f32c", [&] { return f32(BKX_FORM_F32C); }, b32, true, true) ||
        !form_row("f32w", [&] { return f32(BKX_FORM_F32W); }, b32, true, true)) {
        return 1;
    }
    if (op == 5 && p == 7) {
        const double tt = time_min([&] { return bkx_bk5_apply_bf16s(p, nelmt, d_d, d_q32, d_in32, d_out32, nullptr); },
                                   ntests);
        if (tt < 0) {
            return 1;
        }
        const std::vector<float> out = to_host(d_out32, nin);
        row("bf16s", tt, nin, 4.0 * (2 * nin + nqq), norm2(out));
        res.push_back({"bf16s (rounded inputs)", relative_error(out, ref_r), false});
    }

    bool ok = true;
    for (const result& r : res) {
        ok = ok && (!r.gated || r.err <= tol);
        if (!r.gated) {
            std::cout << std::left << std::setw(12) << r.name << " rel. error vs double = " << r.err << "\n";
        }
    }
    std::cout << (ok ? "ok" : "FAIL") << ":";
    const char* sep = " ";
    for (const result& r : res) {
        if (r.gated) {
            std::cout << sep << r.name << " rel. error vs double = " << r.err;
            sep = ", ";
        }
    }
    std::cout << " (tol = " << tol << ")\n";
    for (void* d : {(void*)d_b, (void*)d_d, (void*)d_in, (void*)d_q, (void*)d_in_r, (void*)d_q_r, (void*)d_in32,
                    (void*)d_q32, (void*)d_out, (void*)d_out32}) {
        (void)hipFree(d);
    }
    return ok ? 0 : 1;
}
