// bk_shim_check.cpp -- a reference-style caller of the C++ drop-in
// (include/bkx_sumfact.hpp): the shim's SumFactorization<double, nq>
// templates are stored in the reference's own kernel pointer types, in a
// {name, fn} table like the reference's variant tables, and called on
// hipMalloc'd arrays.
//
//   BK1, BK5: sumfact_kernel<T> = void (*)(std::size_t, const T*, const T*, const T*, T*)
//             (openmp/bk1_sumfact.h:162-163; BK5's kernel_fn, openmp/offload/bk5_offload.cpp:716)
//   BK3:      kernel_fn = void (*)(std::size_t, const T*, const T*, const T*, const T*, T*)
//             (openmp/offload/bk3_offload.cpp:226)
//
//   bk_shim_check <op> <p> <nelmt> <in.bin> <out.bin>
//
// in.bin holds the FP64 arrays of the call in the reference's argument order
// (BK1: basis, JxW, in; BK3: basis, dbasis, G, in; BK5: dbasis, G, in), raw
// and back to back; out.bin receives out.  Exit status 0 on success, 1 on a
// bad argument or a failed call (the message on stderr).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <utility>
#include <vector>

#include "bkx_sumfact.hpp"

namespace {

template <typename T>
using sumfact_kernel = void (*)(std::size_t, const T*, const T*, const T*, T*);
template <typename T>
using bk3_kernel_fn = void (*)(std::size_t, const T*, const T*, const T*, const T*, T*);

struct entry1 {
    const char* name;
    int p;
    sumfact_kernel<double> fn;
};
struct entry3 {
    const char* name;
    int p;
    bk3_kernel_fn<double> fn;
};

template <int... P>
std::vector<entry1> table1(std::integer_sequence<int, P...>)
{
    return {{"bkx::bk1", P + 1, &bkx::bk1::SumFactorization<double, P + 3>}...};
}
template <int... P>
std::vector<entry3> table3(std::integer_sequence<int, P...>)
{
    return {{"bkx::bk3", P + 1, &bkx::bk3::SumFactorization<double, P + 3>}...};
}
template <int... P>
std::vector<entry1> table5(std::integer_sequence<int, P...>)
{
    return {{"bkx::bk5", P + 1, &bkx::bk5::SumFactorization<double, P + 2>}...};
}

double* to_device(const double* h, std::size_t n)
{
    double* d = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&d), (n ? n : 1) * sizeof(double)) != hipSuccess) {
        throw bkx::error("hipMalloc", BKX_ERROR_DEVICE);
    }
    if (h && n && hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        throw bkx::error("hipMemcpy", BKX_ERROR_DEVICE);
    }
    return d;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s <op 1|3|5> <p> <nelmt> <in.bin> <out.bin>\n", argv[0]);
        return 1;
    }
    const int op = std::atoi(argv[1]);
    const int p = std::atoi(argv[2]);
    const std::size_t nelmt = std::strtoull(argv[3], nullptr, 10);
    if ((op != 1 && op != 3 && op != 5) || p < bkx_min_order(op) || p > bkx_max_order(op)) {
        std::fprintf(stderr, "unsupported operator %d or polynomial order %d\n", op, p);
        return 1;
    }
    const std::size_t nm = p + 1, nq = op == 5 ? p + 1 : p + 2;
    const std::size_t nu = nelmt * nm * nm * nm, nq3 = nq * nq * nq;
    std::vector<std::size_t> sizes;   // the arrays in the reference's argument order
    if (op == 1) {
        sizes = {nm * nq, nelmt * nq3, nu};
    } else if (op == 3) {
        sizes = {nm * nq, nq * nq, nelmt * 6 * nq3, nu};
    } else {
        sizes = {nq * nq, nelmt * 6 * nq3, nu};
    }
    std::size_t total = 0;
    for (std::size_t s : sizes) {
        total += s;
    }
    std::vector<double> host(total), out(nu);
    std::FILE* f = std::fopen(argv[4], "rb");
    if (!f || std::fread(host.data(), sizeof(double), total, f) != total) {
        std::fprintf(stderr, "cannot read %zu doubles from %s\n", total, argv[4]);
        return 1;
    }
    std::fclose(f);

    try {
        std::vector<double*> d;
        std::size_t off = 0;
        for (std::size_t s : sizes) {
            d.push_back(to_device(host.data() + off, s));
            off += s;
        }
        double* dout = to_device(nullptr, nu);   // allocated only
        const char* name = nullptr;
        if (op == 3) {
            for (const entry3& e : table3(std::make_integer_sequence<int, 8>{})) {
                if (e.p == p) {
                    e.fn(nelmt, d[0], d[1], d[2], d[3], dout);
                    name = e.name;
                }
            }
        } else {
            const auto table = op == 1 ? table1(std::make_integer_sequence<int, 8>{})
                                       : table5(std::make_integer_sequence<int, 7>{});
            for (const entry1& e : table) {
                if (e.p == p) {
                    e.fn(nelmt, d[0], d[1], d[2], dout);
                    name = e.name;
                }
            }
        }
        if (hipMemcpy(out.data(), dout, nu * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
            throw bkx::error("hipMemcpy", BKX_ERROR_DEVICE);
        }
        for (double* x : d) {
            (void)hipFree(x);
        }
        (void)hipFree(dout);
        std::printf("%s::SumFactorization<double, %zu> p = %d nelmt = %zu\n", name, nq, p, nelmt);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "%s\n", ex.what());
        return 1;
    }
    f = std::fopen(argv[5], "wb");
    if (!f || std::fwrite(out.data(), sizeof(double), nu, f) != nu) {
        std::fprintf(stderr, "cannot write %s\n", argv[5]);
        return 1;
    }
    std::fclose(f);
    return 0;
}
