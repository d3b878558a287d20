// ref_ff.cpp -- C entry points into the reference's float-float BK1 kernels
// and error-free transformations (openmp/BK1_ff.cpp, openmp/float_float.h),
// compiled from those sources as part of this translation unit with the
// driver's main() renamed.  Test infrastructure only.
#define main bk1ff_driver_main
#include "BK1_ff.cpp"
#undef main

extern "C" {

// variant 0: SumFactorizationFF, 1: ...FFTeams, 2: ...FFPlane<nq, 1>,
// 3: ...FFPlane<nq, ff_plane_batch(nq)>; p = 1..8, nq = p + 2
int bkref_bk1_ff(int variant, int p, std::size_t nelmt, const double* basis, const double* JxW,
                 const double* in, double* out)
{
    return bk::with_order<1, 8>("p", p, [&](auto P) {
        constexpr int nq = decltype(P)::value + 2;
        if (variant == 0) {
            SumFactorizationFF<nq>(nelmt, basis, JxW, in, out);
        } else if (variant == 1) {
            SumFactorizationFFTeams<nq>(nelmt, basis, JxW, in, out);
        } else if (variant == 2) {
            SumFactorizationFFPlane<nq, 1>(nelmt, basis, JxW, in, out);
        } else {
            SumFactorizationFFPlane<nq, ff_plane_batch(nq)>(nelmt, basis, JxW, in, out);
        }
    }) ? 0 : 1;
}

void bkref_ff_split(double x, float* hi, float* lo) { bk::ff::split(x, *hi, *lo); }
double bkref_ff_join(float hi, float lo) { return bk::ff::join(hi, lo); }
void bkref_ff_two_sum(float a, float b, float* s, float* e) { bk::ff::two_sum(a, b, *s, *e); }
void bkref_ff_quick_two_sum(float a, float b, float* s, float* e) { bk::ff::quick_two_sum(a, b, *s, *e); }
void bkref_ff_two_prod(float a, float b, float* p, float* e) { bk::ff::two_prod(a, b, *p, *e); }
void bkref_ff_mul(float* ah, float* al, float bh, float bl) { bk::ff::mul(*ah, *al, bh, bl); }
void bkref_ff_mul_add(float* th, float* tl, float ah, float al, float bh, float bl)
{
    bk::ff::mul_add(*th, *tl, ah, al, bh, bl);
}

}  // extern "C"
