// ref_bk5.cpp -- C entry point into the reference's BK5 kernel, which lives in
// the driver source openmp/BK5.cpp together with its main().  That file is
// compiled here as part of this translation unit with its kernel template
// renamed (it has the same name and signature as BK1's, so the two would
// collide at link time) and its main() renamed.  Test infrastructure only.
#define SumFactorization SumFactorizationBK5
#define main bk5_driver_main
#include "BK5.cpp"
#undef main
#undef SumFactorization

extern "C" {

// p = 1..7, nq = p + 1 (openmp/BK5.cpp:148-156)
int bkref_bk5_sumfact(int p, std::size_t nelmt, const double* dbasis, const double* G, const double* in,
                      double* out)
{
    return bk::with_order<1, 7>("p", p, [&](auto P) {
        bk::SumFactorizationBK5<double, decltype(P)::value + 1>(nelmt, dbasis, G, in, out);
    }) ? 0 : 1;
}

}  // extern "C"
