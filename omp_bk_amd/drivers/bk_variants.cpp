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
// against the harness's host FP64 reference of the operator, as the
// reference's harnesses compare with theirs (reference_bk1/bk3/bk5,
// include/bk_harness.h:104-354); FAIL marks >= 1e-5 (include/bk_harness.h:
// 63-85).  The host reference covers every element of batches up to 8192
// and every ceil(nelmt / 8192)-th element of larger ones (the header line
// says how many).
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

// y = x (extents n[0] x n[1] x n[2], row-major) contracted along axis ax
// with M: y(.., c, ..) = sum_m M[m * ld + c] x(.., m, ..), or M[c * ld + m]
// when tr is set; that extent becomes nout.
std::vector<double> contract(const std::vector<double>& x, const int (&n)[3], int ax, const double* M, int ld,
                             int nout, bool tr)
{
    int m[3] = {n[0], n[1], n[2]};
    m[ax] = nout;
    std::vector<double> y(std::size_t(m[0]) * m[1] * m[2], 0.0);
    for (int a = 0; a < m[0]; ++a) {
        for (int b = 0; b < m[1]; ++b) {
            for (int c = 0; c < m[2]; ++c) {
                const int o[3] = {a, b, c};
                double s = 0.0;
                for (int k = 0; k < n[ax]; ++k) {
                    int i[3] = {a, b, c};
                    i[ax] = k;
                    const double w = tr ? M[o[ax] * ld + k] : M[k * ld + o[ax]];
                    s += w * x[(std::size_t(i[0]) * n[1] + i[1]) * n[2] + i[2]];
                }
                y[(std::size_t(a) * m[1] + b) * m[2] + c] = s;
            }
        }
    }
    return y;
}

// The operator on element e in plain FP64 loops, in the C ABI's layouts:
// BK1 B^T (JxW * B u); BK3 the interpolation by B, the collocated gradient
// by D(n, p) at the Gauss points, the six factors G(e, f, p, q, r), the
// transposed gradient and interpolation; BK5 the gradient by D(i, n), the
// factors interleaved G(e, i, j, f, k) and its transpose.
std::vector<double> host_element(int op, int p, std::size_t e, const std::vector<double>& B,
                                 const std::vector<double>& D, const std::vector<double>& Q,
                                 const std::vector<double>& in)
{
    const int nq = op == 5 ? p + 1 : p + 2;
    const int nm = op == 5 ? nq : p + 1;
    const std::size_t nm3 = std::size_t(nm) * nm * nm, nq3 = std::size_t(nq) * nq * nq;
    std::vector<double> u(in.begin() + e * nm3, in.begin() + (e + 1) * nm3);
    const int Q3[3] = {nq, nq, nq};
    auto interp = [&](std::vector<double> x) {   // modes -> points
        const int n0[3] = {nm, nm, nm}, n1[3] = {nq, nm, nm}, n2[3] = {nq, nq, nm};
        x = contract(x, n0, 0, B.data(), nq, nq, false);
        x = contract(x, n1, 1, B.data(), nq, nq, false);
        return contract(x, n2, 2, B.data(), nq, nq, false);
    };
    auto interp_back = [&](std::vector<double> x) {   // points -> modes
        const int n0[3] = {nq, nq, nq}, n1[3] = {nq, nq, nm}, n2[3] = {nq, nm, nm};
        x = contract(x, n0, 2, B.data(), nq, nm, true);
        x = contract(x, n1, 1, B.data(), nq, nm, true);
        return contract(x, n2, 0, B.data(), nq, nm, true);
    };
    if (op == 1) {
        std::vector<double> w = interp(u);
        for (std::size_t x = 0; x < nq3; ++x) {
            w[x] *= Q[e * nq3 + x];
        }
        return interp_back(w);
    }
    const std::vector<double> w = op == 3 ? interp(u) : u;
    const bool tr = op == 5;   // BK5 differentiates with D(i, n), BK3 with D(n, p)
    std::vector<double> g[3];
    for (int a = 0; a < 3; ++a) {
        g[a] = contract(w, Q3, a, D.data(), nq, nq, tr);
    }
    std::vector<double> f[3] = {std::vector<double>(nq3), std::vector<double>(nq3), std::vector<double>(nq3)};
    const int sym[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int i = 0; i < nq; ++i) {
        for (int j = 0; j < nq; ++j) {
            for (int k = 0; k < nq; ++k) {
                const std::size_t x = (std::size_t(i) * nq + j) * nq + k;
                for (int a = 0; a < 3; ++a) {
                    double s = 0.0;
                    for (int b = 0; b < 3; ++b) {
                        const int fa = sym[a][b];
                        const double gf = op == 3 ? Q[e * 6 * nq3 + fa * nq3 + x]
                                                  : Q[e * 6 * nq3 + ((std::size_t(i) * nq + j) * 6 + fa) * nq + k];
                        s += gf * g[b][x];
                    }
                    f[a][x] = s;
                }
            }
        }
    }
    std::vector<double> t(nq3, 0.0);
    for (int a = 0; a < 3; ++a) {
        const std::vector<double> d = contract(f[a], Q3, a, D.data(), nq, nq, !tr);
        for (std::size_t x = 0; x < nq3; ++x) {
            t[x] += d[x];
        }
    }
    return op == 3 ? interp_back(t) : t;
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
    const std::vector<double> basis = test_basis(nm, nq), dbasis = test_basis(nq, nq);
    double* d_in = to_device(in);
    double* d_q = to_device(q);
    double* d_b = to_device(basis);
    double* d_d = to_device(dbasis);
    double* d_out = to_device(std::vector<double>(nin, 0.0));

    // the host reference on the checked elements
    const std::size_t nm3 = std::size_t(nm) * nm * nm;
    const std::size_t stride = nelmt > 8192 ? (nelmt + 8191) / 8192 : 1;
    std::vector<std::size_t> checked;
    for (std::size_t e = 0; e < nelmt; e += stride) {
        checked.push_back(e);
    }
    std::vector<double> ref(checked.size() * nm3);
    double big = 0.0;
    for (std::size_t c = 0; c < checked.size(); ++c) {
        const std::vector<double> v = host_element(op, p, checked[c], basis, dbasis, q, in);
        for (std::size_t x = 0; x < nm3; ++x) {
            ref[c * nm3 + x] = v[x];
            big = std::max(big, std::fabs(v[x]));
        }
    }
    std::printf("BK%d hip: p = %d, nq = %d, nelmt = %zu (%zu DoFs), ntests = %d, host reference on %zu elements\n",
                op, p, nq, nelmt, nin, ntests, checked.size());

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
    std::vector<double> out(nin);
    int status = BKX_SUCCESS;
    for (int v = 0; v < count && status == BKX_SUCCESS; ++v) {
        const std::string name = "v" + std::to_string(v);
        if (!selected(name)) {
            continue;
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
        double diff = 0.0;
        for (std::size_t c = 0; c < checked.size(); ++c) {
            for (std::size_t x = 0; x < nm3; ++x) {
                diff = std::max(diff, std::fabs(out[checked[c] * nm3 + x] - ref[c * nm3 + x]));
            }
        }
        const double relerr = big > 0.0 ? diff / big : diff;
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
