/*
 * bk_oracle.c -- CPU oracle of the sum-factorised bake-off operators.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it; the product
 * (omp_bk_amd/, include/) never does.
 *
 * It is a plain-C restatement, in double, of the reference's CPU path:
 *   - the test-data generators of include/bk_common.h (cos basis, LCG);
 *   - the element loops of SumFactorization<T, nq> for BK1
 *     (openmp/bk1_sumfact.h:16-159), BK3 (openmp/bk3_sumfact.h:17-187) and
 *     BK5 (openmp/BK5.cpp:18-103), with the same loop nests, the same
 *     contraction order (index ascending from 0) and the same work-array
 *     layouts, so that with -ffp-contract=off the results are bit-identical
 *     to SumFactorization<double, nq> built with the same flag;
 *   - the harness references reference_bk1/3/5 (include/bk_harness.h:104-354),
 *     whose divergence sums interleave the three directions in one loop;
 *   - the float-float error-free transformations (openmp/float_float.h:73-134)
 *     and the float-float BK1 kernel (openmp/BK1_ff.cpp:73-241);
 *   - the two error metrics (bk_harness.h:63-83, BK1_ff.cpp:751-763);
 *   - the layout helpers make_bk5_G and transpose_jxw (bk_harness.h:87-102,
 *     207-244).
 * It adds what the reference does not have: the BK3 operator with a
 * per-quadrature-point coefficient kappa scaling the six factors and the
 * adjoint of L = 1/2 |A(kappa) u - b|^2 (gradients in u and kappa).
 *
 * Conventions (all arrays row-major, last index fastest):
 *   BK1/BK3: nm = p + 1 modes, nq = p + 2 points per direction;
 *            in/out(e, i, j, k) nm^3, B(i, p) = basis[i * nq + p];
 *            BK1 JxW(e, p, q, r) nq^3;
 *            BK3 D(n, p) = dbasis[n * nq + p], G(e, f, p, q, r) factor-major.
 *   BK5:     nq = p + 1 points; in/out(e, i, j, k) nq^3;
 *            D(i, n) = dbasis[i * nq + n]; G(e, i, j, f, k) interleaved.
 *   Factor order f: rr rs rt ss st tt.
 *
 * Parity status: every function below is pinned by a test in tests/
 * (golden vectors from the compiled reference, dense Kronecker assembly,
 * invariants, the EFT identities, finite differences).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define W3(n, a, b, c) ((((a) * (n)) + (b)) * (n) + (c))

/* ------------------------------------------------------------------------ */
/* Test data (include/bk_common.h)                                          */
/* ------------------------------------------------------------------------ */

/* M(a, b) = cos(a * cols + b) in double (bk_common.h:63-86, T = double). */
void bko_make_test_basis(double* m, int rows, int cols)
{
    for (int a = 0; a < rows; ++a) {
        for (int b = 0; b < cols; ++b) {
            m[a * cols + b] = cos((double)(a * cols + b));
        }
    }
}

/* The same in float (T = float: std::cos(float) == cosf), widened. */
void bko_make_test_basis_f32(double* m, int rows, int cols)
{
    for (int a = 0; a < rows; ++a) {
        for (int b = 0; b < cols; ++b) {
            m[a * cols + b] = (double)cosf((float)(a * cols + b));
        }
    }
}

/* LCG of bk_common.h:92-107: s <- 1664525 s + 1013904223 (mod 2^32),
 * value (s >> 8) / 2^23 - 1; `in` first, then JxW/G = 1.5 + value. */
void bko_fill_random(double* in, size_t n_in, double* jxw, size_t n_jxw, uint32_t seed)
{
    uint32_t s = seed;
    for (size_t x = 0; x < n_in; ++x) {
        s = 1664525u * s + 1013904223u;
        in[x] = (double)(s >> 8) / (double)(1u << 23) - 1.0;
    }
    for (size_t x = 0; x < n_jxw; ++x) {
        s = 1664525u * s + 1013904223u;
        jxw[x] = 1.5 + ((double)(s >> 8) / (double)(1u << 23) - 1.0);
    }
}

/* ------------------------------------------------------------------------ */
/* BK1: openmp/bk1_sumfact.h:16-159                                          */
/* ------------------------------------------------------------------------ */

static void bk1_element(int nq, const double* B, const double* eJ, const double* ein, double* eout,
                        double* wsp0, double* wsp1)
{
    const int nm = nq - 1;
    /* step-1 (bk1_sumfact.h:52-59) */
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j)
            for (int k = 0; k < nm; ++k)
                wsp0[W3(nq, i, j, k)] = ein[W3(nm, i, j, k)];
    /* step-2 (:61-72) */
    for (int p = 0; p < nq; ++p)
        for (int k = 0; k < nm; ++k)
            for (int j = 0; j < nm; ++j) {
                double t = 0;
                for (int i = 0; i < nm; ++i) t += wsp0[W3(nq, i, j, k)] * B[i * nq + p];
                wsp1[W3(nq, p, j, k)] = t;
            }
    /* step-3 (:74-85) */
    for (int q = 0; q < nq; ++q)
        for (int p = 0; p < nq; ++p)
            for (int k = 0; k < nm; ++k) {
                double t = 0;
                for (int j = 0; j < nm; ++j) t += wsp1[W3(nq, p, j, k)] * B[j * nq + q];
                wsp0[W3(nq, q, p, k)] = t;
            }
    /* step-4 (:87-98) */
    for (int r = 0; r < nq; ++r)
        for (int q = 0; q < nq; ++q)
            for (int p = 0; p < nq; ++p) {
                double t = 0;
                for (int k = 0; k < nm; ++k) t += wsp0[W3(nq, q, p, k)] * B[k * nq + r];
                wsp1[W3(nq, p, q, r)] = t;
            }
    /* step-5 (:102-109) */
    for (int r = 0; r < nq; ++r)
        for (int q = 0; q < nq; ++q)
            for (int p = 0; p < nq; ++p) wsp1[W3(nq, p, q, r)] *= eJ[W3(nq, p, q, r)];
    /* step-6 (:111-122) */
    for (int k = 0; k < nm; ++k)
        for (int q = 0; q < nq; ++q)
            for (int p = 0; p < nq; ++p) {
                double t = 0;
                for (int r = 0; r < nq; ++r) t += wsp1[W3(nq, p, q, r)] * B[k * nq + r];
                wsp0[W3(nq, q, p, k)] = t;
            }
    /* step-7 (:124-135) */
    for (int j = 0; j < nm; ++j)
        for (int k = 0; k < nm; ++k)
            for (int p = 0; p < nq; ++p) {
                double t = 0;
                for (int q = 0; q < nq; ++q) t += wsp0[W3(nq, q, p, k)] * B[j * nq + q];
                wsp1[W3(nq, p, j, k)] = t;
            }
    /* step-8 (:137-148) */
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j)
            for (int k = 0; k < nm; ++k) {
                double t = 0;
                for (int p = 0; p < nq; ++p) t += wsp1[W3(nq, p, j, k)] * B[i * nq + p];
                wsp0[W3(nq, i, j, k)] = t;
            }
    /* step-9 (:150-157) */
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j)
            for (int k = 0; k < nm; ++k) eout[W3(nm, i, j, k)] = wsp0[W3(nq, i, j, k)];
}

void bko_bk1_sumfact(int nq, size_t nelmt, const double* basis, const double* JxW, const double* in,
                     double* out)
{
    const int nm = nq - 1;
    const size_t nm3 = (size_t)nm * nm * nm;
    const size_t nq3 = (size_t)nq * nq * nq;
#pragma omp parallel
    {
        double* scratch = (double*)malloc(2 * nq3 * sizeof(double));
#pragma omp for schedule(static)
        for (size_t e = 0; e < nelmt; ++e) {
            bk1_element(nq, basis, JxW + e * nq3, in + e * nm3, out + e * nm3, scratch, scratch + nq3);
        }
        free(scratch);
    }
}

/* ------------------------------------------------------------------------ */
/* BK3: openmp/bk3_sumfact.h:17-187 (optionally with kappa)                 */
/* ------------------------------------------------------------------------ */

/* Steps 1-4: in(i,j,k) -> u = wsp1(p,q,r) (bk3_sumfact.h:55-96). */
static void bk3_interp(int nq, const double* B, const double* ein, double* wsp0, double* wsp1)
{
    const int nm = nq - 1;
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j)
            for (int k = 0; k < nm; ++k) wsp0[W3(nq, i, j, k)] = ein[W3(nm, i, j, k)];
    for (int p = 0; p < nq; ++p)
        for (int k = 0; k < nm; ++k)
            for (int j = 0; j < nm; ++j) {
                double t = 0;
                for (int i = 0; i < nm; ++i) t += wsp0[W3(nq, i, j, k)] * B[i * nq + p];
                wsp1[W3(nq, p, j, k)] = t;
            }
    for (int q = 0; q < nq; ++q)
        for (int p = 0; p < nq; ++p)
            for (int k = 0; k < nm; ++k) {
                double t = 0;
                for (int j = 0; j < nm; ++j) t += wsp1[W3(nq, p, j, k)] * B[j * nq + q];
                wsp0[W3(nq, q, p, k)] = t;
            }
    for (int r = 0; r < nq; ++r)
        for (int q = 0; q < nq; ++q)
            for (int p = 0; p < nq; ++p) {
                double t = 0;
                for (int k = 0; k < nm; ++k) t += wsp0[W3(nq, q, p, k)] * B[k * nq + r];
                wsp1[W3(nq, p, q, r)] = t;
            }
}

/* Steps 9-12: wsp1(p,q,r) -> out(i,j,k) (bk3_sumfact.h:144-185). */
static void bk3_interp_back(int nq, const double* B, double* wsp0, double* wsp1, double* eout)
{
    const int nm = nq - 1;
    for (int k = 0; k < nm; ++k)
        for (int q = 0; q < nq; ++q)
            for (int p = 0; p < nq; ++p) {
                double t = 0;
                for (int r = 0; r < nq; ++r) t += wsp1[W3(nq, p, q, r)] * B[k * nq + r];
                wsp0[W3(nq, q, p, k)] = t;
            }
    for (int j = 0; j < nm; ++j)
        for (int k = 0; k < nm; ++k)
            for (int p = 0; p < nq; ++p) {
                double t = 0;
                for (int q = 0; q < nq; ++q) t += wsp0[W3(nq, q, p, k)] * B[j * nq + q];
                wsp1[W3(nq, p, j, k)] = t;
            }
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j)
            for (int k = 0; k < nm; ++k) {
                double t = 0;
                for (int p = 0; p < nq; ++p) t += wsp1[W3(nq, p, j, k)] * B[i * nq + p];
                wsp0[W3(nq, i, j, k)] = t;
            }
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j)
            for (int k = 0; k < nm; ++k) eout[W3(nm, i, j, k)] = wsp0[W3(nq, i, j, k)];
}

/* Steps 5-8 (bk3_sumfact.h:97-143).  With ekap != NULL each flux is scaled by
 * kappa(p,q,r) after the G product: r = kappa * (G q). */
static void bk3_flux(int nq, const double* D, const double* eG, const double* ekap, double* wsp1,
                     double* rqr, double* rqs, double* rqt)
{
    const size_t nq3 = (size_t)nq * nq * nq;
    for (int p = 0; p < nq; ++p)
        for (int q = 0; q < nq; ++q)
            for (int r = 0; r < nq; ++r) {
                const int x = W3(nq, p, q, r);
                const double Grr = eG[0 * nq3 + x];
                const double Grs = eG[1 * nq3 + x];
                const double Grt = eG[2 * nq3 + x];
                const double Gss = eG[3 * nq3 + x];
                const double Gst = eG[4 * nq3 + x];
                const double Gtt = eG[5 * nq3 + x];
                double qr = 0, qs = 0, qt = 0;
                for (int n = 0; n < nq; ++n) qr += wsp1[W3(nq, n, q, r)] * D[n * nq + p];
                for (int n = 0; n < nq; ++n) qs += wsp1[W3(nq, p, n, r)] * D[n * nq + q];
                for (int n = 0; n < nq; ++n) qt += wsp1[W3(nq, p, q, n)] * D[n * nq + r];
                double a = Grr * qr + Grs * qs + Grt * qt;
                double b = Grs * qr + Gss * qs + Gst * qt;
                double c = Grt * qr + Gst * qs + Gtt * qt;
                if (ekap) {
                    a *= ekap[x];
                    b *= ekap[x];
                    c *= ekap[x];
                }
                rqr[x] = a;
                rqs[x] = b;
                rqt[x] = c;
            }
    for (int p = 0; p < nq; ++p)
        for (int q = 0; q < nq; ++q)
            for (int r = 0; r < nq; ++r) {
                double t = 0;
                for (int n = 0; n < nq; ++n) t += rqr[W3(nq, n, q, r)] * D[p * nq + n];
                for (int n = 0; n < nq; ++n) t += rqs[W3(nq, p, n, r)] * D[q * nq + n];
                for (int n = 0; n < nq; ++n) t += rqt[W3(nq, p, q, n)] * D[r * nq + n];
                wsp1[W3(nq, p, q, r)] = t;
            }
}

void bko_bk3_sumfact_kappa(int nq, size_t nelmt, const double* basis, const double* dbasis,
                           const double* G, const double* kappa, const double* in, double* out)
{
    const int nm = nq - 1;
    const size_t nm3 = (size_t)nm * nm * nm;
    const size_t nq3 = (size_t)nq * nq * nq;
#pragma omp parallel
    {
        double* s = (double*)malloc(5 * nq3 * sizeof(double));
#pragma omp for schedule(static)
        for (size_t e = 0; e < nelmt; ++e) {
            bk3_interp(nq, basis, in + e * nm3, s, s + nq3);
            bk3_flux(nq, dbasis, G + e * 6 * nq3, kappa ? kappa + e * nq3 : NULL, s + nq3, s + 2 * nq3,
                     s + 3 * nq3, s + 4 * nq3);
            bk3_interp_back(nq, basis, s, s + nq3, out + e * nm3);
        }
        free(s);
    }
}

void bko_bk3_sumfact(int nq, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                     const double* in, double* out)
{
    bko_bk3_sumfact_kappa(nq, nelmt, basis, dbasis, G, NULL, in, out);
}

/* ------------------------------------------------------------------------ */
/* BK5: openmp/BK5.cpp:18-103                                                */
/* ------------------------------------------------------------------------ */

void bko_bk5_sumfact(int nq, size_t nelmt, const double* D, const double* G, const double* in, double* out)
{
    const size_t nq3 = (size_t)nq * nq * nq;
#pragma omp parallel
    {
        double* s = (double*)malloc(3 * nq3 * sizeof(double));
        double* rqr = s;
        double* rqs = s + nq3;
        double* rqt = s + 2 * nq3;
#pragma omp for schedule(static)
        for (size_t e = 0; e < nelmt; ++e) {
            const double* ei = in + e * nq3;
            const double* eG = G + e * 6 * nq3;
            double* eo = out + e * nq3;
            for (int i = 0; i < nq; ++i)
                for (int j = 0; j < nq; ++j)
                    for (int k = 0; k < nq; ++k) {
                        const double* g = eG + ((size_t)(i * nq + j) * 6) * nq + k; /* G(i,j,f,k) */
                        const double Grr = g[0 * nq], Grs = g[1 * nq], Grt = g[2 * nq];
                        const double Gss = g[3 * nq], Gst = g[4 * nq], Gtt = g[5 * nq];
                        double qr = 0, qs = 0, qt = 0;
                        for (int n = 0; n < nq; ++n) qr += ei[W3(nq, n, j, k)] * D[i * nq + n];
                        for (int n = 0; n < nq; ++n) qs += ei[W3(nq, i, n, k)] * D[j * nq + n];
                        for (int n = 0; n < nq; ++n) qt += ei[W3(nq, i, j, n)] * D[k * nq + n];
                        const int x = W3(nq, i, j, k);
                        rqr[x] = Grr * qr + Grs * qs + Grt * qt;
                        rqs[x] = Grs * qr + Gss * qs + Gst * qt;
                        rqt[x] = Grt * qr + Gst * qs + Gtt * qt;
                    }
            for (int i = 0; i < nq; ++i)
                for (int j = 0; j < nq; ++j)
                    for (int k = 0; k < nq; ++k) {
                        double t = 0;
                        for (int n = 0; n < nq; ++n) t += rqr[W3(nq, n, j, k)] * D[n * nq + i];
                        for (int n = 0; n < nq; ++n) t += rqs[W3(nq, i, n, k)] * D[n * nq + j];
                        for (int n = 0; n < nq; ++n) t += rqt[W3(nq, i, j, n)] * D[n * nq + k];
                        eo[W3(nq, i, j, k)] = t;
                    }
        }
        free(s);
    }
}

/* ------------------------------------------------------------------------ */
/* Harness references (include/bk_harness.h:104-354): one n-loop for the    */
/* three directions of each derivative and of the divergence.              */
/* ------------------------------------------------------------------------ */

/* ref_interp (bk_harness.h:106-143) */
static void ref_interp(int nq, const double* basis, const double* in, double* a, double* b, double* u)
{
    const int nm = nq - 1;
    for (int p = 0; p < nq; ++p)
        for (int j = 0; j < nm; ++j)
            for (int k = 0; k < nm; ++k) {
                double s = 0;
                for (int i = 0; i < nm; ++i) s += in[(i * nm + j) * nm + k] * basis[i * nq + p];
                a[(p * nm + j) * nm + k] = s;
            }
    for (int p = 0; p < nq; ++p)
        for (int q = 0; q < nq; ++q)
            for (int k = 0; k < nm; ++k) {
                double s = 0;
                for (int j = 0; j < nm; ++j) s += a[(p * nm + j) * nm + k] * basis[j * nq + q];
                b[(p * nq + q) * nm + k] = s;
            }
    for (int p = 0; p < nq; ++p)
        for (int q = 0; q < nq; ++q)
            for (int r = 0; r < nq; ++r) {
                double s = 0;
                for (int k = 0; k < nm; ++k) s += b[(p * nq + q) * nm + k] * basis[k * nq + r];
                u[(p * nq + q) * nq + r] = s;
            }
}

/* ref_interp_back (bk_harness.h:146-183) */
static void ref_interp_back(int nq, const double* basis, const double* u, double* a, double* b, double* out)
{
    const int nm = nq - 1;
    for (int p = 0; p < nq; ++p)
        for (int q = 0; q < nq; ++q)
            for (int k = 0; k < nm; ++k) {
                double s = 0;
                for (int r = 0; r < nq; ++r) s += u[(p * nq + q) * nq + r] * basis[k * nq + r];
                b[(p * nq + q) * nm + k] = s;
            }
    for (int p = 0; p < nq; ++p)
        for (int j = 0; j < nm; ++j)
            for (int k = 0; k < nm; ++k) {
                double s = 0;
                for (int q = 0; q < nq; ++q) s += b[(p * nq + q) * nm + k] * basis[j * nq + q];
                a[(p * nm + j) * nm + k] = s;
            }
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j)
            for (int k = 0; k < nm; ++k) {
                double s = 0;
                for (int p = 0; p < nq; ++p) s += a[(p * nm + j) * nm + k] * basis[i * nq + p];
                out[(i * nm + j) * nm + k] = s;
            }
}

/* reference_bk1 (bk_harness.h:186-205) */
void bko_reference_bk1(int nq, size_t nelmt, const double* basis, const double* JxW, const double* in,
                       double* out)
{
    const int nm = nq - 1;
    const size_t nm3 = (size_t)nm * nm * nm;
    const size_t nq3 = (size_t)nq * nq * nq;
#pragma omp parallel
    {
        double* s = (double*)malloc(3 * nq3 * sizeof(double));
#pragma omp for schedule(static)
        for (size_t e = 0; e < nelmt; ++e) {
            ref_interp(nq, basis, in + e * nm3, s, s + nq3, s + 2 * nq3);
            for (size_t x = 0; x < nq3; ++x) s[2 * nq3 + x] *= JxW[e * nq3 + x];
            ref_interp_back(nq, basis, s + 2 * nq3, s, s + nq3, out + e * nm3);
        }
        free(s);
    }
}

/* reference_bk5 (bk_harness.h:247-293) */
void bko_reference_bk5(int nq, size_t nelmt, const double* D, const double* G, const double* in, double* out)
{
    const size_t nq3 = (size_t)nq * nq * nq;
#pragma omp parallel
    {
        double* s = (double*)malloc(3 * nq3 * sizeof(double));
        double* rr = s;
        double* rs = s + nq3;
        double* rt = s + 2 * nq3;
#pragma omp for schedule(static)
        for (size_t e = 0; e < nelmt; ++e) {
            const double* ei = in + e * nq3;
            const double* eG = G + e * 6 * nq3;
            for (int i = 0; i < nq; ++i)
                for (int j = 0; j < nq; ++j)
                    for (int k = 0; k < nq; ++k) {
                        double qr = 0, qs = 0, qt = 0;
                        for (int n = 0; n < nq; ++n) {
                            qr += ei[(n * nq + j) * nq + k] * D[i * nq + n];
                            qs += ei[(i * nq + n) * nq + k] * D[j * nq + n];
                            qt += ei[(i * nq + j) * nq + n] * D[k * nq + n];
                        }
#define GF(f) eG[((i * nq + j) * 6 + (f)) * nq + k]
                        const int x = (i * nq + j) * nq + k;
                        rr[x] = GF(0) * qr + GF(1) * qs + GF(2) * qt;
                        rs[x] = GF(1) * qr + GF(3) * qs + GF(4) * qt;
                        rt[x] = GF(2) * qr + GF(4) * qs + GF(5) * qt;
#undef GF
                    }
            for (int i = 0; i < nq; ++i)
                for (int j = 0; j < nq; ++j)
                    for (int k = 0; k < nq; ++k) {
                        double t = 0;
                        for (int n = 0; n < nq; ++n) {
                            t += rr[(n * nq + j) * nq + k] * D[n * nq + i];
                            t += rs[(i * nq + n) * nq + k] * D[n * nq + j];
                            t += rt[(i * nq + j) * nq + n] * D[n * nq + k];
                        }
                        out[e * nq3 + (i * nq + j) * nq + k] = t;
                    }
        }
        free(s);
    }
}

/* reference_bk3 (bk_harness.h:299-354) */
void bko_reference_bk3(int nq, size_t nelmt, const double* basis, const double* D, const double* G,
                       const double* in, double* out)
{
    const int nm = nq - 1;
    const size_t nm3 = (size_t)nm * nm * nm;
    const size_t nq3 = (size_t)nq * nq * nq;
#pragma omp parallel
    {
        double* s = (double*)malloc(6 * nq3 * sizeof(double));
        double *a = s, *b = s + nq3, *u = s + 2 * nq3, *fr = s + 3 * nq3, *fs = s + 4 * nq3, *ft = s + 5 * nq3;
#pragma omp for schedule(static)
        for (size_t e = 0; e < nelmt; ++e) {
            const double* eG = G + e * 6 * nq3;
            ref_interp(nq, basis, in + e * nm3, a, b, u);
            for (int p = 0; p < nq; ++p)
                for (int q = 0; q < nq; ++q)
                    for (int r = 0; r < nq; ++r) {
                        double qr = 0, qs = 0, qt = 0;
                        for (int n = 0; n < nq; ++n) {
                            qr += u[(n * nq + q) * nq + r] * D[n * nq + p];
                            qs += u[(p * nq + n) * nq + r] * D[n * nq + q];
                            qt += u[(p * nq + q) * nq + n] * D[n * nq + r];
                        }
                        const int x = (p * nq + q) * nq + r;
                        fr[x] = eG[0 * nq3 + x] * qr + eG[1 * nq3 + x] * qs + eG[2 * nq3 + x] * qt;
                        fs[x] = eG[1 * nq3 + x] * qr + eG[3 * nq3 + x] * qs + eG[4 * nq3 + x] * qt;
                        ft[x] = eG[2 * nq3 + x] * qr + eG[4 * nq3 + x] * qs + eG[5 * nq3 + x] * qt;
                    }
            for (int p = 0; p < nq; ++p)
                for (int q = 0; q < nq; ++q)
                    for (int r = 0; r < nq; ++r) {
                        double t = 0;
                        for (int n = 0; n < nq; ++n) {
                            t += fr[(n * nq + q) * nq + r] * D[p * nq + n];
                            t += fs[(p * nq + n) * nq + r] * D[q * nq + n];
                            t += ft[(p * nq + q) * nq + n] * D[r * nq + n];
                        }
                        u[(p * nq + q) * nq + r] = t;
                    }
            ref_interp_back(nq, basis, u, a, b, out + e * nm3);
        }
        free(s);
    }
}

/* ------------------------------------------------------------------------ */
/* BK3 adjoint (no reference counterpart).                                   */
/* A(kappa) u = sum_x (grad B u)_x^T kappa_x G_x (grad B u)_x, symmetric, so */
/*   r = A(kappa) u - b,  L = 1/2 r.r,  dL/du = A(kappa) r,                   */
/*   dL/dkappa_x = (grad B r)_x^T G_x (grad B u)_x                           */
/* with grad = (qr, qs, qt) of bk3_sumfact.h:111-119 at quadrature point x. */
/* ------------------------------------------------------------------------ */

/* Reference gradient (qr, qs, qt) at every point of one element. */
static void bk3_grad(int nq, const double* B, const double* D, const double* ein, double* s, double* g)
{
    const size_t nq3 = (size_t)nq * nq * nq;
    bk3_interp(nq, B, ein, s, s + nq3);
    const double* w = s + nq3;
    for (int p = 0; p < nq; ++p)
        for (int q = 0; q < nq; ++q)
            for (int r = 0; r < nq; ++r) {
                double qr = 0, qs = 0, qt = 0;
                for (int n = 0; n < nq; ++n) qr += w[W3(nq, n, q, r)] * D[n * nq + p];
                for (int n = 0; n < nq; ++n) qs += w[W3(nq, p, n, r)] * D[n * nq + q];
                for (int n = 0; n < nq; ++n) qt += w[W3(nq, p, q, n)] * D[n * nq + r];
                const int x = W3(nq, p, q, r);
                g[3 * x + 0] = qr;
                g[3 * x + 1] = qs;
                g[3 * x + 2] = qt;
            }
}

/* Returns L; writes r (optional), dL/du and dL/dkappa (either may be NULL). */
double bko_bk3_adjoint(int nq, size_t nelmt, const double* basis, const double* dbasis, const double* G,
                       const double* kappa, const double* u, const double* b, double* resid, double* dldu,
                       double* dldk)
{
    const int nm = nq - 1;
    const size_t nm3 = (size_t)nm * nm * nm;
    const size_t nq3 = (size_t)nq * nq * nq;
    const size_t ndof = nelmt * nm3;
    double* r = (double*)malloc(ndof * sizeof(double));
    bko_bk3_sumfact_kappa(nq, nelmt, basis, dbasis, G, kappa, u, r);
    double loss = 0;
    for (size_t x = 0; x < ndof; ++x) {
        r[x] -= b[x];
        loss += r[x] * r[x];
    }
    loss *= 0.5;
    if (resid) memcpy(resid, r, ndof * sizeof(double));
    if (dldu) bko_bk3_sumfact_kappa(nq, nelmt, basis, dbasis, G, kappa, r, dldu);
    if (dldk) {
#pragma omp parallel
        {
            double* s = (double*)malloc(8 * nq3 * sizeof(double));
            double* gu = s + 2 * nq3;
            double* gr = s + 5 * nq3;
#pragma omp for schedule(static)
            for (size_t e = 0; e < nelmt; ++e) {
                bk3_grad(nq, basis, dbasis, u + e * nm3, s, gu);
                bk3_grad(nq, basis, dbasis, r + e * nm3, s, gr);
                const double* eG = G + e * 6 * nq3;
                for (size_t x = 0; x < nq3; ++x) {
                    const double g0 = eG[0 * nq3 + x], g1 = eG[1 * nq3 + x], g2 = eG[2 * nq3 + x];
                    const double g3 = eG[3 * nq3 + x], g4 = eG[4 * nq3 + x], g5 = eG[5 * nq3 + x];
                    const double a = g0 * gu[3 * x] + g1 * gu[3 * x + 1] + g2 * gu[3 * x + 2];
                    const double c = g1 * gu[3 * x] + g3 * gu[3 * x + 1] + g4 * gu[3 * x + 2];
                    const double d = g2 * gu[3 * x] + g4 * gu[3 * x + 1] + g5 * gu[3 * x + 2];
                    dldk[e * nq3 + x] = gr[3 * x] * a + gr[3 * x + 1] * c + gr[3 * x + 2] * d;
                }
            }
            free(s);
        }
    }
    free(r);
    return loss;
}

/* ------------------------------------------------------------------------ */
/* Float-float (openmp/float_float.h:73-134)                                 */
/* ------------------------------------------------------------------------ */

void bko_ff_split(double x, float* hi, float* lo)
{
    *hi = (float)x;
    *lo = (float)(x - (double)*hi);
}

double bko_ff_join(float hi, float lo) { return (double)hi + (double)lo; }

void bko_ff_two_sum(float a, float b, float* s, float* e)
{
    const float ss = a + b;
    const float bb = ss - a;
    *s = ss;
    *e = (a - (ss - bb)) + (b - bb);
}

void bko_ff_quick_two_sum(float a, float b, float* s, float* e)
{
    const float ss = a + b;
    *s = ss;
    *e = b - (ss - a);
}

void bko_ff_two_prod(float a, float b, float* p, float* e)
{
    const float pp = a * b;
    *p = pp;
    *e = fmaf(a, b, -pp);
}

void bko_ff_mul(float* ah, float* al, float bh, float bl)
{
    float p, e;
    bko_ff_two_prod(*ah, bh, &p, &e);
    e = fmaf(*ah, bl, fmaf(*al, bh, e));
    bko_ff_quick_two_sum(p, e, ah, al);
}

void bko_ff_mul_add(float* th, float* tl, float ah, float al, float bh, float bl)
{
    float ph, pl, sh, sl;
    bko_ff_two_prod(ah, bh, &ph, &pl);
    pl = fmaf(ah, bl, fmaf(al, bh, pl));
    bko_ff_two_sum(*th, ph, &sh, &sl);
    sl += *tl + pl;
    bko_ff_quick_two_sum(sh, sl, th, tl);
}

/* SumFactorizationFF (openmp/BK1_ff.cpp:73-241): BK1 with every value split
 * into a float pair at load, the nine steps in float-float, joined at store. */
void bko_bk1_ff_sumfact(int nq, size_t nelmt, const double* basis, const double* JxW, const double* in,
                        double* out)
{
    const int nm = nq - 1;
    const size_t nm3 = (size_t)nm * nm * nm;
    const size_t nq3 = (size_t)nq * nq * nq;
#pragma omp parallel
    {
        float* Bh = (float*)malloc((size_t)nm * nq * sizeof(float));
        float* Bl = (float*)malloc((size_t)nm * nq * sizeof(float));
        float* w = (float*)malloc(4 * nq3 * sizeof(float));
        float *w0h = w, *w0l = w + nq3, *w1h = w + 2 * nq3, *w1l = w + 3 * nq3;
        for (int x = 0; x < nm * nq; ++x) bko_ff_split(basis[x], &Bh[x], &Bl[x]);
#pragma omp for schedule(static)
        for (size_t e = 0; e < nelmt; ++e) {
            const double* ei = in + e * nm3;
            const double* eJ = JxW + e * nq3;
            double* eo = out + e * nm3;
            for (int i = 0; i < nm; ++i)
                for (int j = 0; j < nm; ++j)
                    for (int k = 0; k < nm; ++k)
                        bko_ff_split(ei[W3(nm, i, j, k)], &w0h[W3(nq, i, j, k)], &w0l[W3(nq, i, j, k)]);
            for (int p = 0; p < nq; ++p)
                for (int k = 0; k < nm; ++k)
                    for (int j = 0; j < nm; ++j) {
                        float th = 0.0f, tl = 0.0f;
                        for (int i = 0; i < nm; ++i)
                            bko_ff_mul_add(&th, &tl, w0h[W3(nq, i, j, k)], w0l[W3(nq, i, j, k)], Bh[i * nq + p],
                                           Bl[i * nq + p]);
                        w1h[W3(nq, p, j, k)] = th;
                        w1l[W3(nq, p, j, k)] = tl;
                    }
            for (int q = 0; q < nq; ++q)
                for (int p = 0; p < nq; ++p)
                    for (int k = 0; k < nm; ++k) {
                        float th = 0.0f, tl = 0.0f;
                        for (int j = 0; j < nm; ++j)
                            bko_ff_mul_add(&th, &tl, w1h[W3(nq, p, j, k)], w1l[W3(nq, p, j, k)], Bh[j * nq + q],
                                           Bl[j * nq + q]);
                        w0h[W3(nq, q, p, k)] = th;
                        w0l[W3(nq, q, p, k)] = tl;
                    }
            for (int r = 0; r < nq; ++r)
                for (int q = 0; q < nq; ++q)
                    for (int p = 0; p < nq; ++p) {
                        float th = 0.0f, tl = 0.0f;
                        for (int k = 0; k < nm; ++k)
                            bko_ff_mul_add(&th, &tl, w0h[W3(nq, q, p, k)], w0l[W3(nq, q, p, k)], Bh[k * nq + r],
                                           Bl[k * nq + r]);
                        w1h[W3(nq, p, q, r)] = th;
                        w1l[W3(nq, p, q, r)] = tl;
                    }
            for (int r = 0; r < nq; ++r)
                for (int q = 0; q < nq; ++q)
                    for (int p = 0; p < nq; ++p) {
                        float jh, jl;
                        bko_ff_split(eJ[W3(nq, p, q, r)], &jh, &jl);
                        bko_ff_mul(&w1h[W3(nq, p, q, r)], &w1l[W3(nq, p, q, r)], jh, jl);
                    }
            for (int k = 0; k < nm; ++k)
                for (int q = 0; q < nq; ++q)
                    for (int p = 0; p < nq; ++p) {
                        float th = 0.0f, tl = 0.0f;
                        for (int r = 0; r < nq; ++r)
                            bko_ff_mul_add(&th, &tl, w1h[W3(nq, p, q, r)], w1l[W3(nq, p, q, r)], Bh[k * nq + r],
                                           Bl[k * nq + r]);
                        w0h[W3(nq, q, p, k)] = th;
                        w0l[W3(nq, q, p, k)] = tl;
                    }
            for (int j = 0; j < nm; ++j)
                for (int k = 0; k < nm; ++k)
                    for (int p = 0; p < nq; ++p) {
                        float th = 0.0f, tl = 0.0f;
                        for (int q = 0; q < nq; ++q)
                            bko_ff_mul_add(&th, &tl, w0h[W3(nq, q, p, k)], w0l[W3(nq, q, p, k)], Bh[j * nq + q],
                                           Bl[j * nq + q]);
                        w1h[W3(nq, p, j, k)] = th;
                        w1l[W3(nq, p, j, k)] = tl;
                    }
            for (int i = 0; i < nm; ++i)
                for (int j = 0; j < nm; ++j)
                    for (int k = 0; k < nm; ++k) {
                        float th = 0.0f, tl = 0.0f;
                        for (int p = 0; p < nq; ++p)
                            bko_ff_mul_add(&th, &tl, w1h[W3(nq, p, j, k)], w1l[W3(nq, p, j, k)], Bh[i * nq + p],
                                           Bl[i * nq + p]);
                        w0h[W3(nq, i, j, k)] = th;
                        w0l[W3(nq, i, j, k)] = tl;
                    }
            for (int i = 0; i < nm; ++i)
                for (int j = 0; j < nm; ++j)
                    for (int k = 0; k < nm; ++k)
                        eo[W3(nm, i, j, k)] = bko_ff_join(w0h[W3(nq, i, j, k)], w0l[W3(nq, i, j, k)]);
        }
        free(w);
        free(Bl);
        free(Bh);
    }
}

/* ------------------------------------------------------------------------ */
/* Layouts of the geometric factors (include/bk_harness.h)                   */
/* ------------------------------------------------------------------------ */

/* The full-tensor entry (a, b), row-major index 3a + b, of the six symmetric
 * factors rr rs rt ss st tt (bk_harness.h:215). */
static const int bko_sym9[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};

/* Index of point (i, j, k), factor f of F, within one element's factors in
 * layout 0 (factor-major G(f, i, j, k)), 1 (interleaved G(i, j, f, k)) or 2
 * (column G(f, k, i, j)). */
static size_t bko_g_index(int nq, int F, int layout, int i, int j, int k, int f)
{
    if (layout == 1) return ((size_t)(i * nq + j) * F + f) * nq + k;
    if (layout == 2) return ((size_t)(f * nq + k) * nq + i) * nq + j;
    return ((size_t)(f * nq + i) * nq + j) * nq + k;
}

/* make_bk5_G (bk_harness.h:217-244): from the interleaved six factors
 * G6i(e, i, j, s, k), the factors with F = 6 (symmetric) or 9 (the full
 * tensor, row-major) per point in the given layout.  Pure data movement. */
void bko_make_bk5_G(int nq, size_t nelmt, const double* G6i, int F, int layout, double* G)
{
    const size_t nq3 = (size_t)nq * nq * nq;
    for (size_t e = 0; e < nelmt; ++e) {
        const double* src = G6i + e * 6 * nq3;
        double* dst = G + e * (size_t)F * nq3;
        for (int f = 0; f < F; ++f) {
            const int s = F == 6 ? f : bko_sym9[f];
            for (int i = 0; i < nq; ++i)
                for (int j = 0; j < nq; ++j)
                    for (int k = 0; k < nq; ++k)
                        dst[bko_g_index(nq, F, layout, i, j, k, f)] = src[bko_g_index(nq, 6, 1, i, j, k, s)];
        }
    }
}

/* transpose_jxw (bk_harness.h:87-102): JxW(e, p, q, r) -> JxW_t(e, r, q, p). */
void bko_transpose_jxw(int nq, size_t nelmt, const double* JxW, double* JxW_t)
{
    const size_t nq3 = (size_t)nq * nq * nq;
    for (size_t e = 0; e < nelmt; ++e)
        for (int r = 0; r < nq; ++r)
            for (int q = 0; q < nq; ++q)
                for (int p = 0; p < nq; ++p)
                    JxW_t[e * nq3 + W3(nq, r, q, p)] = JxW[e * nq3 + W3(nq, p, q, r)];
}

/* ------------------------------------------------------------------------ */
/* Error metrics                                                             */
/* ------------------------------------------------------------------------ */

/* max|x - r| / max|r| (bk_harness.h:63-83) */
double bko_rel_maxdiff(const double* x, const double* r, size_t n)
{
    double big = 0, diff = 0;
    for (size_t i = 0; i < n; ++i) {
        if (fabs(r[i]) > big) big = fabs(r[i]);
        if (fabs(x[i] - r[i]) > diff) diff = fabs(x[i] - r[i]);
    }
    return diff / big;
}

/* |x - r|_2 / |r|_2 (BK1_ff.cpp:751-763) */
double bko_rel_l2(const double* x, const double* r, size_t n)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < n; ++i) {
        const double d = x[i] - r[i];
        num += d * d;
        den += r[i] * r[i];
    }
    return sqrt(num) / sqrt(den);
}

/* Euclidean norm accumulated in double (bk_common.h:44-58) */
double bko_norm2(const double* v, size_t n)
{
    double s = 0;
    for (size_t i = 0; i < n; ++i) s += v[i] * v[i];
    return sqrt(s);
}
