// bk_stream.cpp -- bin/stream, the counterpart of the reference's
// roofline/stream_omp.cpp for the pick scripts' roofs: the HBM bandwidth of
// a copy and of the read-heavy mixes the operators stream (read9: nine bytes
// read per byte written, the reference's BK3 roof; read7: BK5 in FP64, 56 B
// read and 8 B written per DoF), in FP64 on the GPU through the library's
// probes (bkx_probe_copy, bkx_probe_read_mix).  One line each in the
// reference's format (roofline/stream_omp.cpp:135-136)
//   <name> bytes = b  time = t ms  GB/s = x
// so that `awk '$1 == "copy" {print $NF}'` reads the roof.
//   stream [n_doubles] [ntests]
// bytes counts what is read plus what is written; the time is the minimum
// over ntests launches after a warm-up, timed by HIP events.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <limits>
#include <vector>

#include "bkx.h"

namespace {

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::cerr << what << ": " << hipGetErrorString(e) << "\n";
        std::exit(1);
    }
}

double time_min(int ntests, const std::function<int()>& f)
{
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
    double best = std::numeric_limits<double>::max();
    int st = f();   // warm-up
    for (int t = 0; t < ntests && st == BKX_SUCCESS; ++t) {
        hip_check(hipEventRecord(a, nullptr), "hipEventRecord");
        st = f();
        hip_check(hipEventRecord(b, nullptr), "hipEventRecord");
        hip_check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        best = std::min(best, double(ms));
    }
    hip_check(hipEventDestroy(a), "hipEventDestroy");
    hip_check(hipEventDestroy(b), "hipEventDestroy");
    if (st != BKX_SUCCESS) {
        std::cerr << "bkx error: " << bkx_status_string(st) << "\n";
        std::exit(1);
    }
    return best;
}

void report(const char* name, double bytes, double ms)
{
    std::printf("%-6s bytes = %10.3e  time = %9.3f ms  GB/s = %8.2f\n", name, bytes, ms, 1e-6 * bytes / ms);
}

}  // namespace

int main(int argc, char** argv)
{
    std::size_t n = argc > 1 ? std::size_t(std::atoll(argv[1])) : (std::size_t(1) << 28);   // 2 GiB per array
    const int ntests = argc > 2 ? std::atoi(argv[2]) : 20;
    n = std::max<std::size_t>(n / 9 / 256, 1) * 9 * 256;   // whole read9 blocks, and even for the copy
    double* a = nullptr;
    double* c = nullptr;
    if (hipMalloc(&a, n * sizeof(double)) != hipSuccess || hipMalloc(&c, n * sizeof(double)) != hipSuccess) {
        std::cerr << "hipMalloc failed for n = " << n << "\n";
        return 1;
    }
    const std::vector<double> ones(n, 1.0);
    hip_check(hipMemcpy(a, ones.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    int dev = 0;
    hipDeviceProp_t prop;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::printf("device %d (%s), n = %zu doubles (%.2f GB per array)\n", dev, prop.gcnArchName, n,
                1e-9 * double(n) * sizeof(double));
    const double w = sizeof(double);
    report("copy", 2.0 * w * double(n), time_min(ntests, [&] { return bkx_probe_copy(a, c, n, nullptr); }));
    for (const int r : {7, 9}) {
        const std::size_t nout = n / std::size_t(r) / 256 * 256;
        const char* name = r == 7 ? "read7" : "read9";
        report(name, w * double(nout) * (r + 1),
               time_min(ntests, [&] { return bkx_probe_read_mix(r, a, c, nout, nullptr); }));
    }
    std::vector<double> back(16);
    hip_check(hipMemcpy(back.data(), c, back.size() * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    std::printf("check: c[0] = %.1f (expect 9.0 after read9)\n", back[0]);
    (void)hipFree(a);
    (void)hipFree(c);
    return back[0] == 9.0 ? 0 : 1;
}
