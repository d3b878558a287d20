// bkx_ff.hpp -- float-float ("double-float") arithmetic for the compensated
// FP32 kernels: a value is the unevaluated sum h + l of two binary32 numbers.
//
// The error-free transformations are the published ones: TwoSum (Knuth),
// FastTwoSum (Dekker, valid for |a| >= |b|) and TwoProd through a fused
// multiply-add; they need IEEE round-to-nearest FP32 with no reassociation
// and no contraction.  HIP compiles with -ffp-contract=fast, and LLVM fuses
// across statements: the product p = a * x of TwoProd would be folded into
// the add s + p of the following TwoSum (fma(a, x, s)), which is no longer
// the sum of the rounded product and breaks the identity s + e = a + b
// (the result degrades to plain-FP32 accuracy, ~1e-8).  Every helper below
// therefore turns contraction off for its own operations; the explicit
// __builtin_fmaf calls are the only fused operations.  The same
// transformations are restated on the CPU by the reference's
// openmp/float_float.h:85-105.
//
// Dot products use the Ogita-Rump-Oishi "Dot2" scheme: the running sum s of
// the products' high parts is carried with TwoSum, and the TwoSum errors, the
// TwoProd errors and the products of the low parts are summed in plain FP32
// into c; (s, c) is renormalised once, when read.  The result is as accurate
// as if the dot product were computed in about twice the working precision
// and then rounded.
#pragma once

#include <hip/hip_runtime.h>

namespace bkx {

struct alignas(8) ff {
    float h;
    float l;
};

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
#pragma clang fp contract(off)
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void fast_two_sum(float a, float b, float& s, float& e)
{
#pragma clang fp contract(off)
    s = a + b;
    e = b - (s - a);
}

// split a double into its leading float and the float of the remainder
__device__ __forceinline__ ff ff_split(double x)
{
#pragma clang fp contract(off)
    const float h = float(x);
    return ff{h, float(x - double(h))};
}

__device__ __forceinline__ double ff_join(ff a)
{
#pragma clang fp contract(off)
    return double(a.h) + double(a.l);
}

// Dot2 accumulator: value s + c (unnormalised)
struct dot2 {
    float s = 0.0f;
    float c = 0.0f;

    // += a * x, a float-float, x an exact float
    __device__ __forceinline__ void add(ff a, float x)
    {
#pragma clang fp contract(off)
        const float p = a.h * x;
        float pe = __builtin_fmaf(a.h, x, -p);
        pe = __builtin_fmaf(a.l, x, pe);
        float t, e;
        two_sum(s, p, t, e);
        s = t;
        c += e + pe;
    }

    // += a * b, both float-float (the l * l term is below the precision)
    __device__ __forceinline__ void add(ff a, ff b)
    {
#pragma clang fp contract(off)
        const float p = a.h * b.h;
        float pe = __builtin_fmaf(a.h, b.h, -p);
        pe = __builtin_fmaf(a.h, b.l, pe);
        pe = __builtin_fmaf(a.l, b.h, pe);
        float t, e;
        two_sum(s, p, t, e);
        s = t;
        c += e + pe;
    }

    // += another accumulator's normalised value
    __device__ __forceinline__ void add(ff a)
    {
#pragma clang fp contract(off)
        float t, e;
        two_sum(s, a.h, t, e);
        s = t;
        c += e + a.l;
    }

    __device__ __forceinline__ ff value() const
    {
        ff r;
        fast_two_sum(s, c, r.h, r.l);
        return r;
    }
};

// a * x for a float-float a and an exact float x, normalised
__device__ __forceinline__ ff ff_mul(ff a, float x)
{
    dot2 d;
    d.add(a, x);
    return d.value();
}

}  // namespace bkx
