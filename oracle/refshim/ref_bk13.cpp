// ref_bk13.cpp -- C entry points into the reference's own BK1 and BK3
// sum-factorised kernels and its FP64 harness references, compiled from the
// reference sources where they lie (include/ and openmp/ on the include
// path).  Test infrastructure only: built into oracle/_ref/ by
// oracle/Makefile, loaded by tests/ and tests/golden/ as the checker.
#include <cstddef>
#include <cstdint>
#include <vector>

#include "bk_common.h"
#include "bk_harness.h"
#include "bk1_sumfact.h"   // bk::SumFactorization<T, nq>(nelmt, basis, JxW, in, out)
#include "bk3_sumfact.h"   // bk::SumFactorization<T, nq>(nelmt, basis, dbasis, G, in, out)

namespace {
std::vector<float> to_float(const double* x, std::size_t n) { return std::vector<float>(x, x + n); }
}

extern "C" {

// p = 1..8, nq = p + 2 (openmp/BK1.cpp:99-106)
int bkref_bk1_sumfact(int p, std::size_t nelmt, const double* basis, const double* JxW, const double* in,
                      double* out)
{
    return bk::with_order<1, 8>("p", p, [&](auto P) {
        bk::SumFactorization<double, decltype(P)::value + 2>(nelmt, basis, JxW, in, out);
    }) ? 0 : 1;
}

// p = 1..8, nq = p + 2 (openmp/BK3.cpp:55-63)
int bkref_bk3_sumfact(int p, std::size_t nelmt, const double* basis, const double* dbasis, const double* G,
                      const double* in, double* out)
{
    return bk::with_order<1, 8>("p", p, [&](auto P) {
        bk::SumFactorization<double, decltype(P)::value + 2>(nelmt, basis, dbasis, G, in, out);
    }) ? 0 : 1;
}

// The harness references take float inputs: the double arrays passed here
// must hold float-representable values.
void bkref_reference_bk1(int nq, std::size_t nelmt, const double* basis, const double* JxW, const double* in,
                         double* out)
{
    const std::size_t nm3 = std::size_t(nq - 1) * (nq - 1) * (nq - 1);
    const std::size_t nq3 = std::size_t(nq) * nq * nq;
    std::vector<double> o(nelmt * nm3);
    bk::reference_bk1(nq, nelmt, to_float(basis, std::size_t(nq - 1) * nq), to_float(JxW, nelmt * nq3),
                      to_float(in, nelmt * nm3), o);
    for (std::size_t x = 0; x < o.size(); ++x) {
        out[x] = o[x];
    }
}

void bkref_reference_bk3(int nq, std::size_t nelmt, const double* basis, const double* D, const double* G,
                         const double* in, double* out)
{
    const std::size_t nm3 = std::size_t(nq - 1) * (nq - 1) * (nq - 1);
    const std::size_t nq3 = std::size_t(nq) * nq * nq;
    std::vector<double> o(nelmt * nm3);
    bk::reference_bk3(nq, nelmt, to_float(basis, std::size_t(nq - 1) * nq), to_float(D, std::size_t(nq) * nq),
                      to_float(G, nelmt * 6 * nq3), to_float(in, nelmt * nm3), o);
    for (std::size_t x = 0; x < o.size(); ++x) {
        out[x] = o[x];
    }
}

void bkref_reference_bk5(int nq, std::size_t nelmt, const double* D, const double* G, const double* in,
                         double* out)
{
    const std::size_t nq3 = std::size_t(nq) * nq * nq;
    std::vector<double> o(nelmt * nq3);
    bk::reference_bk5(nq, nelmt, to_float(D, std::size_t(nq) * nq), to_float(G, nelmt * 6 * nq3),
                      to_float(in, nelmt * nq3), o);
    for (std::size_t x = 0; x < o.size(); ++x) {
        out[x] = o[x];
    }
}

// G6i -> layout 0/1/2 with F = 6 or 9 factors (bk_harness.h:220-244)
void bkref_make_bk5_G(const double* G6i, std::size_t nelmt, int nq, int F, int layout, double* G)
{
    const std::size_t nq3 = std::size_t(nq) * nq * nq;
    const std::vector<float> g = bk::make_bk5_G(to_float(G6i, nelmt * 6 * nq3), nelmt, nq, F, layout);
    for (std::size_t x = 0; x < g.size(); ++x) {
        G[x] = g[x];
    }
}

// BK1's transposed JxW (bk_harness.h:87-102); float-representable values
void bkref_transpose_jxw(const double* JxW, std::size_t nelmt, int nq, double* JxW_t)
{
    const std::size_t nq3 = std::size_t(nq) * nq * nq;
    const std::vector<float> t = bk::transpose_jxw(to_float(JxW, nelmt * nq3), nelmt, nq);
    for (std::size_t x = 0; x < t.size(); ++x) {
        JxW_t[x] = t[x];
    }
}

// Test data (bk_common.h:63-107)
void bkref_make_test_basis_f64(int rows, int cols, double* m)
{
    const std::vector<double> b = bk::make_test_basis<double>(rows, cols);
    for (std::size_t x = 0; x < b.size(); ++x) {
        m[x] = b[x];
    }
}

void bkref_make_test_basis_f32(int rows, int cols, double* m)
{
    const std::vector<float> b = bk::make_test_basis<float>(rows, cols);
    for (std::size_t x = 0; x < b.size(); ++x) {
        m[x] = b[x];
    }
}

void bkref_fill_random(double* in, std::size_t n_in, double* jxw, std::size_t n_jxw, std::uint32_t seed)
{
    std::vector<double> a(n_in);
    std::vector<double> b(n_jxw);
    bk::fill_random_test_data(a, b, seed);
    for (std::size_t x = 0; x < n_in; ++x) {
        in[x] = a[x];
    }
    for (std::size_t x = 0; x < n_jxw; ++x) {
        jxw[x] = b[x];
    }
}

double bkref_rel_maxdiff(const float* out, const double* ref, std::size_t n)
{
    const std::vector<float> o(out, out + n);
    const std::vector<double> r(ref, ref + n);
    return bk::rel_maxdiff(o, r, bk::max_abs(r));
}

}  // extern "C"
