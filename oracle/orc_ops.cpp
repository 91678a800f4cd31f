// CPU oracle: banded solvers and single-line derivative / filter operators.
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the
// reference lines whose arithmetic it restates; evaluation order of sums is
// kept so that, built without FP contraction, results match the reference.
#include "oracle.h"

#include <omp.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Banded solvers
// ---------------------------------------------------------------------------

// Thomas algorithm without pivoting; a[i] is row i's sub-diagonal (a[0]
// unused), c[i] row i's super-diagonal.  src/Utils.cpp:4-20
void tri(const double *a, const double *b, const double *c, const double *d, double *x,
         double *w, int n) {
    for (int i = 0; i < n; ++i) { w[i] = b[i]; x[i] = d[i]; }
    for (int i = 1; i < n; ++i) {
        double m = a[i] / w[i - 1];
        w[i] = w[i] - m * c[i - 1];
        x[i] = x[i] - m * x[i - 1];
    }
    x[n - 1] /= w[n - 1];
    for (int i = n - 2; i >= 0; --i) x[i] = (x[i] - c[i] * x[i + 1]) / w[i];
}

// Sherman-Morrison periodic tridiagonal (two Thomas solves plus a rank-one
// correction).  src/Utils.cpp:22-46
void tri_cyclic(const double *a, const double *b, const double *c, double alpha, double beta,
                const double *r, int n, double *x) {
    std::vector<double> bb(n), u(n, 0.0), z(n), w(n);
    const double gamma = -b[0];
    for (int i = 0; i < n; ++i) bb[i] = b[i];
    bb[0] = b[0] - gamma;
    bb[n - 1] = b[n - 1] - alpha * beta / gamma;
    tri(a, bb.data(), c, r, x, w.data(), n);
    u[0] = gamma;
    u[n - 1] = alpha;
    tri(a, bb.data(), c, u.data(), z.data(), w.data(), n);
    const double fact = (x[0] + beta * x[n - 1] / gamma) / (1.0 + z[0] + beta * z[n - 1] / gamma);
    for (int i = 0; i < n; ++i) x[i] -= fact * z[i];
}

// Pentadiagonal elimination.  Row i of the matrix reads
//   e[i-2] x[i-2] + a[i-1] x[i-1] + d[i] x[i] + c[i] x[i+1] + f[i] x[i+2]
// i.e. the two lower bands are indexed by column.  src/Utils.cpp:48-83
void penta(const double *e, const double *a, const double *d, const double *c, const double *f,
           const double *b, double *x, int n) {
    std::vector<double> B(b, b + n), C(c, c + n), D(d, d + n);
    for (int i = 0; i < n; ++i) x[i] = a[i];
    for (int i = 1; i < n - 1; ++i) {
        double m = x[i - 1] / D[i - 1];
        D[i] = D[i] - m * C[i - 1];
        C[i] = C[i] - m * f[i - 1];
        B[i] = B[i] - m * B[i - 1];
        m = e[i - 1] / D[i - 1];
        x[i] = x[i] - m * C[i - 1];
        D[i + 1] = D[i + 1] - m * f[i - 1];
        B[i + 1] = B[i + 1] - m * B[i - 1];
    }
    double m = x[n - 2] / D[n - 2];
    D[n - 1] = D[n - 1] - m * C[n - 2];
    x[n - 1] = (B[n - 1] - m * B[n - 2]) / D[n - 1];
    x[n - 2] = (B[n - 2] - C[n - 2] * x[n - 1]) / D[n - 2];
    for (int i = n - 3; i >= 0; --i) x[i] = (B[i] - f[i] * x[i + 2] - C[i] * x[i + 1]) / D[i];
}

// 3x3 minor of a row-major 4x4 matrix.  src/Utils.cpp:193-198
double minor3(const double m[16], int r0, int r1, int r2, int c0, int c1, int c2) {
    return m[4 * r0 + c0] * (m[4 * r1 + c1] * m[4 * r2 + c2] - m[4 * r2 + c1] * m[4 * r1 + c2]) -
           m[4 * r0 + c1] * (m[4 * r1 + c0] * m[4 * r2 + c2] - m[4 * r2 + c0] * m[4 * r1 + c2]) +
           m[4 * r0 + c2] * (m[4 * r1 + c0] * m[4 * r2 + c1] - m[4 * r2 + c0] * m[4 * r1 + c1]);
}

}  // namespace

extern "C" {

int orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"

extern "C" {

void orc_solve_tri(const double *a, const double *b, const double *c, const double *d, double *x,
                   double *work, int n) {
    tri(a, b, c, d, x, work, n);
}

void orc_cyclic_tri(const double *a, const double *b, const double *c, double alpha, double beta,
                    const double *r, int n, double *x) {
    tri_cyclic(a, b, c, alpha, beta, r, n, x);
}

void orc_solve_penta(const double *e, const double *a, const double *d, const double *c,
                     const double *f, const double *b, double *x, int n) {
    penta(e, a, d, c, f, b, x, n);
}

// Adjugate / determinant inverse of a row-major 4x4.  src/Utils.cpp:201-224
void orc_invert4(const double m[16], double out[16]) {
    static const int R[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            // adj[i][j] = (-1)^(i+j) * minor(rows != j, cols != i)
            const int *rr = R[j], *cc = R[i];
            double v = minor3(m, rr[0], rr[1], rr[2], cc[0], cc[1], cc[2]);
            out[4 * i + j] = ((i + j) & 1) ? -v : v;
        }
    double det = m[0] * minor3(m, 1, 2, 3, 1, 2, 3) - m[1] * minor3(m, 1, 2, 3, 0, 2, 3) +
                 m[2] * minor3(m, 1, 2, 3, 0, 1, 3) - m[3] * minor3(m, 1, 2, 3, 0, 1, 2);
    double inv_det = 1.0 / det;  // the reference writes 1.0f, which is exactly 1
    for (int i = 0; i < 16; ++i) out[i] = out[i] * inv_det;
}

// Woodbury rank-4 corner correction of a pentadiagonal solve: five penta
// solves, p = I + V^T Z, h = p^{-1}, x = y - Z h V^T y.  The corner vector
// cp[6] fills (0,N-2),(0,N-1),(1,N-1),(N-2,0),(N-1,0),(N-1,1).
// src/Utils.cpp:85-191
void orc_cyclic_penta(const double *a, const double *b, const double *c, const double *d,
                      const double *e, const double *f, const double cp[6], double *x, int n) {
    std::vector<double> u(4 * n, 0.0), v(4 * n, 0.0), z(4 * n, 0.0), y(n);
    u[0 * n + 0] = 1.0;
    u[1 * n + 1] = 1.0;
    u[2 * n + n - 2] = 1.0;
    u[3 * n + n - 1] = 1.0;
    v[0 * n + n - 2] = cp[0];
    v[0 * n + n - 1] = cp[1];
    v[1 * n + n - 1] = cp[2];
    v[2 * n + 0] = cp[3];
    v[3 * n + 0] = cp[4];
    v[3 * n + 1] = cp[5];
    for (int k = 0; k < 4; ++k) penta(a, b, c, d, e, &u[k * n], &z[k * n], n);
    penta(a, b, c, d, e, f, y.data(), n);
    double p[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += v[i * n + k] * z[j * n + k];
            p[4 * i + j] = s;
        }
    for (int i = 0; i < 4; ++i) p[5 * i] += 1.0;
    double h[16];
    orc_invert4(p, h);
    double r[4], s[4];
    for (int i = 0; i < 4; ++i) {
        r[i] = 0.0;
        for (int k = 0; k < n; ++k) r[i] += v[i * n + k] * y[k];
    }
    for (int i = 0; i < 4; ++i) {
        s[i] = 0.0;
        for (int j = 0; j < 4; ++j) s[i] += h[4 * i + j] * r[j];
    }
    for (int k = 0; k < n; ++k) {
        double sum = 0.0;
        for (int j = 0; j < 4; ++j) sum += z[j * n + k] * s[j];
        x[k] = y[k] - sum;
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Derivative schemes
// ---------------------------------------------------------------------------
namespace {

// Line accessor: periodic wrap, or transformed-periodic wrap values.
struct Line {
    const double *p;
    int n;
    const double *tpb;  // {Nm3,Nm2,Nm1,Np1,Np2,Np3} or null
    double operator()(int i) const {
        if (i >= 0 && i < n) return p[i];
        if (tpb) {
            if (i < 0) return tpb[3 + i];       // -1 -> Nm1 (idx 2), -3 -> Nm3 (idx 0)
            return tpb[3 + (i - n)];            //  n -> Np1 (idx 3)
        }
        return p[(i % n + n) % n];
    }
};

// Pade6: alpha=1/3, a=14/9, b=1/9 (src/Pade6.hpp:73-75); Dirichlet closures
// row 0 third order (:85-91), row 1 Visbal BC5 (:107-112).  e2_1D is never
// assigned under boundaryScheme 1 (only in commented-out alternatives), so the
// BC5 stencil is used with that coefficient equal to 0.
void pade6_line(int bc, int n, double dd, const Line &f, double *out) {
    const double alpha = 1.0 / 3.0, a = 14.0 / 9.0, b = 1.0 / 9.0;
    const double cc2 = -b / 4.0, cc3 = -a / 2.0, cc4 = a / 2.0, cc5 = b / 4.0;
    std::vector<double> rhs(n), lo(n, alpha), di(n, 1.0), up(n, alpha);
    if (bc == ORC_PERIODIC) {
        // src/Pade6.cpp:187-218 (and the TPB form :220-251)
        for (int i = 0; i < n; ++i)
            rhs[i] = cc2 * f(i - 2) + cc3 * f(i - 1) + cc4 * f(i + 1) + cc5 * f(i + 2);
        for (int i = 0; i < n; ++i) rhs[i] /= dd;
        // src/Pade6.cpp:351-365
        tri_cyclic(lo.data(), di.data(), up.data(), alpha, alpha, rhs.data(), n, out);
        return;
    }
    const double a1 = -5.0 / 2.0, b1 = 2.0, c1 = 1.0 / 2.0, d1 = 0.0, e1 = 0.0, f1 = 0.0;
    const double al21 = 1.0 / 6.0, al22 = 1.0 / 2.0, al11 = 2.0;
    const double a2 = -5.0 / 9.0, b2 = -1.0 / 2.0, c2 = 1.0, d2 = 1.0 / 18.0, e2 = 0.0;
    const double *p = f.p;
    // src/Pade6.cpp:284-313
    rhs[0] = a1 * p[0] + b1 * p[1] + c1 * p[2] + d1 * p[3] + e1 * p[4] + f1 * p[5];
    rhs[1] = a2 * p[0] + b2 * p[1] + c2 * p[2] + d2 * p[3] + e2 * p[4];
    for (int i = 2; i < n - 2; ++i)
        rhs[i] = cc2 * p[i - 2] + cc3 * p[i - 1] + cc4 * p[i + 1] + cc5 * p[i + 2];
    rhs[n - 2] = -e2 * p[n - 5] - d2 * p[n - 4] - c2 * p[n - 3] - b2 * p[n - 2] - a2 * p[n - 1];
    rhs[n - 1] = -f1 * p[n - 6] - e1 * p[n - 5] - d1 * p[n - 4] - c1 * p[n - 3] - b1 * p[n - 2] -
                 a1 * p[n - 1];
    for (int i = 0; i < n; ++i) rhs[i] /= dd;
    // band edits src/Pade6.hpp:213-221 (lower band indexed by row)
    up[0] = al11; up[1] = al22; lo[1] = al21;
    lo[n - 1] = al11; lo[n - 2] = al22; up[n - 2] = al21;
    std::vector<double> w(n);
    tri(lo.data(), di.data(), up.data(), rhs.data(), out, w.data(), n);  // :376-383
}

// Penta10 coefficient set, src/Penta10.hpp:70-175
struct P10 {
    double beta = 1.0 / 20.0, alpha = 1.0 / 2.0, a = 17.0 / 12.0, b = 101.0 / 150.0, c = 1.0 / 100.0;
    double beta1 = 0.0, alpha1 = 5.0, a1 = -197.0 / 60.0, b1 = -5.0 / 12.0, c1 = 5.0 / 1.0,
           d1 = -5.0 / 3.0, e1 = 5.0 / 12.0, f1 = -1.0 / 20.0, g1 = 0.0, h1 = 0.0, i1 = 0.0;
    double beta2 = 0.0, alpha2_1 = 1.0 / 12.0, alpha2_2 = 5.0 / 4.0, a2 = -79.0 / 240.0,
           b2 = -77.0 / 60.0, c2 = 55.0 / 48.0, d2 = 5.0 / 9.0, e2 = -5.0 / 48.0, f2 = 1.0 / 60.0,
           g2 = -1.0 / 720.0, h2 = 0.0;
    double beta3_1 = 1.0 / 36.0, beta3_2 = 1.0 / 36.0, alpha3_1 = 4.0 / 9.0, alpha3_2 = 4.0 / 9.0,
           a3 = -25.0 / 54.0 / 4.0, b3 = -40.0 / 27.0 / 2.0, c3 = 0.0, d3 = 40.0 / 27.0 / 2.0,
           e3 = 25.0 / 54.0 / 4.0, f3 = 0.0, g3 = 0.0;
};

void penta10_line(int bc, int n, double dd, const Line &f, double *out) {
    const P10 k;
    const double cc1 = -k.c / 6.0, cc2 = -k.b / 4.0, cc3 = -k.a / 2.0;
    const double cc4 = k.a / 2.0, cc5 = k.b / 4.0, cc6 = k.c / 6.0;
    std::vector<double> rhs(n), di(n, 1.0), lo(n, k.alpha), up(n, k.alpha), lo2(n, k.beta),
        up2(n, k.beta);
    if (bc == ORC_PERIODIC) {
        // src/Penta10.cpp:141-173 / TPB :175-208
        for (int i = 0; i < n; ++i)
            rhs[i] = cc1 * f(i - 3) + cc2 * f(i - 2) + cc3 * f(i - 1) + cc4 * f(i + 1) +
                     cc5 * f(i + 2) + cc6 * f(i + 3);
        for (int i = 0; i < n; ++i) rhs[i] /= dd;
        const double cp[6] = {k.beta, k.alpha, k.beta, k.beta, k.alpha, k.beta};  // :170-175
        // :254-269
        orc_cyclic_penta(lo2.data(), lo.data(), di.data(), up.data(), up2.data(), rhs.data(), cp,
                         out, n);
        return;
    }
    const double *p = f.p;
    // src/Penta10.cpp:211-252
    rhs[0] = k.a1 * p[0] + k.b1 * p[1] + k.c1 * p[2] + k.d1 * p[3] + k.e1 * p[4] + k.f1 * p[5] +
             k.g1 * p[6] + k.h1 * p[7] + k.i1 * p[8];
    rhs[1] = k.a2 * p[0] + k.b2 * p[1] + k.c2 * p[2] + k.d2 * p[3] + k.e2 * p[4] + k.f2 * p[5] +
             k.g2 * p[6] + k.h2 * p[7];
    rhs[2] = k.a3 * p[0] + k.b3 * p[1] + k.c3 * p[2] + k.d3 * p[3] + k.e3 * p[4] + k.f3 * p[5] +
             k.g3 * p[6];
    for (int i = 3; i < n - 3; ++i)
        rhs[i] = cc1 * p[i - 3] + cc2 * p[i - 2] + cc3 * p[i - 1] + cc4 * p[i + 1] +
                 cc5 * p[i + 2] + cc6 * p[i + 3];
    rhs[n - 3] = -k.g3 * p[n - 7] + -k.f3 * p[n - 6] - k.e3 * p[n - 5] - k.d3 * p[n - 4] +
                 -k.c3 * p[n - 3] - k.b3 * p[n - 2] - k.a3 * p[n - 1];
    rhs[n - 2] = -k.h2 * p[n - 8] - k.g2 * p[n - 7] + -k.f2 * p[n - 6] - k.e2 * p[n - 5] -
                 k.d2 * p[n - 4] + -k.c2 * p[n - 3] - k.b2 * p[n - 2] - k.a2 * p[n - 1];
    rhs[n - 1] = -k.i1 * p[n - 9] - k.h1 * p[n - 8] - k.g1 * p[n - 7] + -k.f1 * p[n - 6] -
                 k.e1 * p[n - 5] - k.d1 * p[n - 4] + -k.c1 * p[n - 3] - k.b1 * p[n - 2] -
                 k.a1 * p[n - 1];
    for (int i = 0; i < n; ++i) rhs[i] /= dd;
    // band edits, lower bands indexed by column: src/Penta10.hpp:135-167
    up[0] = k.alpha1; up[1] = k.alpha2_2; up[2] = k.alpha3_2;
    lo[0] = k.alpha2_1; lo[1] = k.alpha3_1;
    up2[0] = k.beta1; up2[1] = k.beta2; up2[2] = k.beta3_2;
    lo2[0] = k.beta3_1;
    up[n - 2] = k.alpha2_1; up[n - 3] = k.alpha3_1;
    lo[n - 2] = k.alpha1; lo[n - 3] = k.alpha2_2; lo[n - 4] = k.alpha3_2;
    up2[n - 3] = k.beta3_1;
    lo2[n - 3] = k.beta1; lo2[n - 4] = k.beta2; lo2[n - 5] = k.beta3_2;
    penta(lo2.data(), lo.data(), di.data(), up.data(), up2.data(), rhs.data(), out, n);  // :271-282
}

// CD2: a=-1/2, c=+1/2 applied as (a*phi[i+1] + c*phi[i-1])/dd in the interior,
// which is the negated central difference; the wrap rows 0 and N-1 of the
// periodic form carry the usual sign.  src/CD2.hpp:54-61, src/CD2.cpp:129-181
void cd2_line(int bc, int n, double dd, const Line &f, double *out) {
    const double a = -1.0 / 2.0, c = 1.0 / 2.0;
    const double aa = -3.0 / 2.0, bb = 4.0 / 2.0, cc = -1.0 / 2.0;
    const double *p = f.p;
    if (bc == ORC_PERIODIC) {
        out[0] = (a * f(-1) + c * p[1]) / dd;
        out[n - 1] = (a * p[n - 2] + c * f(n)) / dd;
    } else {
        out[0] = (aa * p[0] + bb * p[1] + cc * p[2]) / dd;
        out[n - 1] = (aa * p[n - 1] + bb * p[n - 2] + cc * p[n - 3]) / dd;
    }
    for (int i = 1; i < n - 1; ++i) out[i] = (a * p[i + 1] + c * p[i - 1]) / dd;
}

// ---------------------------------------------------------------------------
// Filters
// ---------------------------------------------------------------------------

// Compact8: src/Compact8Filter.hpp:55-126, src/Compact8Filter.cpp:5-130
void compact8_line(int bc, int n, double af, const double *p, double *out) {
    const double a0 = (93.0 + 70.0 * af) / 128.0, a1 = (7.0 + 18.0 * af) / 16.0,
                 a2 = (-7.0 + 14.0 * af) / 32.0, a3 = (1.0 - 2.0 * af) / 16.0,
                 a4 = (-1.0 + 2.0 * af) / 128.0;
    const double c0 = a0, c1 = a1 / 2.0, c2 = a2 / 2.0, c3 = a3 / 2.0, c4 = a4 / 2.0;
    std::vector<double> rhs(n), lo(n, af), di(n, 1.0), up(n, af);
    auto P = [&](int i) { return p[(i % n + n) % n]; };
    if (bc == ORC_PERIODIC) {
        for (int i = 0; i < n; ++i)
            rhs[i] = c4 * P(i - 4) + c3 * P(i - 3) + c2 * P(i - 2) + c1 * P(i - 1) + c0 * p[i] +
                     c1 * P(i + 1) + c2 * P(i + 2) + c3 * P(i + 3) + c4 * P(i + 4);
        tri_cyclic(lo.data(), di.data(), up.data(), af, af, rhs.data(), n, out);  // :111-117
        return;
    }
    const double B2[9] = {1.0 / 256.0 + 127.0 * af / 128.0, 31.0 / 32.0 + 1.0 * af / 16.0,
                          7.0 / 64.0 + 25.0 * af / 32.0,    -7.0 / 32.0 + 7.0 * af / 16.0,
                          7.0 * (5.0 - 10.0 * af) / 128.0,  -7.0 / 32.0 + 7.0 * af / 16.0,
                          7.0 / 64.0 - 7.0 * af / 32.0,     -1.0 / 32.0 + 1.0 * af / 16.0,
                          1.0 / 256.0 - 1.0 * af / 128.0};
    const double B3[9] = {-1.0 / 256.0 + 1.0 * af / 128.0, 1.0 / 32.0 + 15.0 * af / 16.0,
                          57.0 / 64.0 + 7.0 * af / 32.0,    7.0 / 32.0 + 9.0 * af / 16.0,
                          7.0 * (-5.0 + 10.0 * af) / 128.0, 7.0 / 32.0 - 7.0 * af / 16.0,
                          -7.0 / 64.0 + 7.0 * af / 32.0,    1.0 / 32.0 - 1.0 * af / 16.0,
                          -1.0 / 256.0 + 1.0 * af / 128.0};
    const double B4[9] = {1.0 / 256.0 - 1.0 * af / 128.0, -1.0 / 32.0 + 1.0 * af / 16.0,
                          7.0 / 64.0 + 25.0 * af / 32.0,  25.0 / 32.0 + 7.0 * af / 16.0,
                          35.0 / 128.0 + 29.0 * af / 64.0, -7.0 / 32.0 + 7.0 * af / 16.0,
                          7.0 / 64.0 - 7.0 * af / 32.0,   -1.0 / 32.0 + 1.0 * af / 16.0,
                          1.0 / 256.0 - 1.0 * af / 128.0};
    auto dotL = [&](const double *w, int row) {
        double s = w[0] * p[0];
        for (int k = 1; k < 9; ++k) s = s + w[k] * p[k];
        rhs[row] = s;
    };
    auto dotR = [&](const double *w, int row) {
        double s = w[0] * p[n - 1];
        for (int k = 1; k < 9; ++k) s = s + w[k] * p[n - 1 - k];
        rhs[row] = s;
    };
    // src/Compact8Filter.cpp:64-109
    rhs[0] = p[0];
    dotL(B2, 1); dotL(B3, 2); dotL(B4, 3);
    rhs[n - 1] = p[n - 1];
    dotR(B2, n - 2); dotR(B3, n - 3); dotR(B4, n - 4);
    for (int i = 4; i < n - 4; ++i)
        rhs[i] = c4 * p[i - 4] + c3 * p[i - 3] + c2 * p[i - 2] + c1 * p[i - 1] + c0 * p[i] +
                 c1 * p[i + 1] + c2 * p[i + 2] + c3 * p[i + 3] + c4 * p[i + 4];
    // identity end rows: src/Compact8Filter.hpp:120-126
    lo[0] = 0.0; up[0] = 0.0; lo[n - 1] = 0.0; up[n - 1] = 0.0;
    std::vector<double> w(n);
    tri(lo.data(), di.data(), up.data(), rhs.data(), out, w.data(), n);  // :119-130
}

// Compact10 with boundaryScheme 1: src/Compact10Filter.hpp:58-134,283-295,
// src/Compact10Filter.cpp:5-154
void compact10_line(int bc, int n, double af, const double *p, double *out) {
    const double a0 = (193.0 + 126.0 * af) / 256.0, a1 = (105.0 + 302.0 * af) / 256.0,
                 a2 = (-15.0 + 30.0 * af) / 64.0, a3 = (45.0 - 90.0 * af) / 512.0,
                 a4 = (-5.0 + 10.0 * af) / 256.0, a5 = (1.0 - 2.0 * af) / 512.0;
    const double c0 = a0, c1 = a1 / 2.0, c2 = a2 / 2.0, c3 = a3 / 2.0, c4 = a4 / 2.0,
                 c5 = a5 / 2.0;
    std::vector<double> rhs(n), lo(n, af), di(n, 1.0), up(n, af);
    auto P = [&](int i) { return p[(i % n + n) % n]; };
    if (bc == ORC_PERIODIC) {
        for (int i = 0; i < n; ++i)
            rhs[i] = c5 * P(i - 5) + c4 * P(i - 4) + c3 * P(i - 3) + c2 * P(i - 2) +
                     c1 * P(i - 1) + c0 * p[i] + c1 * P(i + 1) + c2 * P(i + 2) + c3 * P(i + 3) +
                     c4 * P(i + 4) + c5 * P(i + 5);
        tri_cyclic(lo.data(), di.data(), up.data(), af, af, rhs.data(), n, out);
        return;
    }
    const double B[4][11] = {
        {(1.0 + 1022.0 * af) / 1024.0, (507.0 + 10.0 * af) / 512.0, (45.0 + 934.0 * af) / 1024.0,
         (-15.0 + 30.0 * af) / 128.0, (105.0 - 210.0 * af) / 512.0, (-63.0 + 126.0 * af) / 256.0,
         (105.0 - 210.0 * af) / 512.0, (-15.0 + 30.0 * af) / 128.0, (45.0 - 90.0 * af) / 1024.0,
         (-5.0 + 10.0 * af) / 512.0, (1.0 - 2.0 * af) / 1024.0},
        {(-1.0 + 2.0 * af) / 1024.0, (5.0 + 502.0 * af) / 512.0, (979.0 + 90.0 * af) / 1024.0,
         (15.0 + 98.0 * af) / 128.0, (-105.0 + 210.0 * af) / 512.0, (63.0 - 126.0 * af) / 256.0,
         (-105.0 + 210.0 * af) / 512.0, (15.0 - 30.0 * af) / 128.0, (-45.0 + 90.0 * af) / 1024.0,
         (5.0 - 10.0 * af) / 512.0, (-1.0 + 2.0 * af) / 1024.0},
        {(1.0 - 2.0 * af) / 1024.0, (-5.0 + 10.0 * af) / 512.0, (45.0 + 934.0 * af) / 1024.0,
         (113.0 + 30.0 * af) / 128.0, (105.0 + 302.0 * af) / 512.0, (-63.0 + 126.0 * af) / 256.0,
         (105.0 - 210.0 * af) / 512.0, (-15.0 + 30.0 * af) / 128.0, (45.0 - 90.0 * af) / 1024.0,
         (-5.0 + 10.0 * af) / 512.0, (1.0 - 2.0 * af) / 1024.0},
        {(-1.0 + 2.0 * af) / 1024.0, (5.0 - 10.0 * af) / 512.0, (-45.0 + 90.0 * af) / 1024.0,
         (15.0 + 98.0 * af) / 128.0, (407.0 + 210.0 * af) / 512.0, (63.0 + 130.0 * af) / 256.0,
         (-105.0 + 210.0 * af) / 512.0, (15.0 - 30.0 * af) / 128.0, (-45.0 + 90.0 * af) / 1024.0,
         (5.0 - 10.0 * af) / 512.0, (-1.0 + 2.0 * af) / 1024.0}};
    rhs[0] = p[0];
    rhs[n - 1] = p[n - 1];
    for (int r = 0; r < 4; ++r) {
        double s = B[r][0] * p[0];
        for (int k = 1; k < 11; ++k) s = s + B[r][k] * p[k];
        rhs[1 + r] = s;
        s = B[r][0] * p[n - 1];
        for (int k = 1; k < 11; ++k) s = s + B[r][k] * p[n - 1 - k];
        rhs[n - 2 - r] = s;
    }
    for (int i = 5; i < n - 5; ++i)
        rhs[i] = c5 * p[i - 5] + c4 * p[i - 4] + c3 * p[i - 3] + c2 * p[i - 2] + c1 * p[i - 1] +
                 c0 * p[i] + c1 * p[i + 1] + c2 * p[i + 2] + c3 * p[i + 3] + c4 * p[i + 4] +
                 c5 * p[i + 5];
    lo[0] = 0.0; up[0] = 0.0; lo[n - 1] = 0.0; up[n - 1] = 0.0;
    std::vector<double> w(n);
    tri(lo.data(), di.data(), up.data(), rhs.data(), out, w.data(), n);
}

// Vasilyev commutative Pade test filter: src/CommN5PadeTestFilter.hpp:55-186,
// src/CommN5PadeTestFilter.cpp:5-143
void commn5pade_line(int bc, int n, const double *p, double *out) {
    const double w0_n1 = 31.0 / 32.0, w1_n1 = 5.0 / 32.0, w2_n1 = -5.0 / 16.0, w3_n1 = 5.0 / 16.0,
                 w4_n1 = -5.0 / 32.0, w5_n1 = 1.0 / 32.0;
    const double wn1_n2 = 1.0 / 32.0, w0_n2 = 27.0 / 32.0, w1_n2 = 5.0 / 16.0,
                 w2_n2 = -5.0 / 16.0, w3_n2 = 5.0 / 32.0, w4_n2 = -1.0 / 32.0;
    const double wn2_n3 = -1.0 / 32.0, wn1_n3 = 5.0 / 32.0, w0_n3 = 11.0 / 16.0,
                 w1_n3 = 5.0 / 16.0, w2_n3 = -5.0 / 32.0, w3_n3 = 1.0 / 32.0;
    const double w0 = 11.0 / 16.0, w1 = 15.0 / 64.0, w2 = -3.0 / 32.0, w3 = 1.0 / 64.0;
    const double v0 = 7.0 / 12.0, v1 = 0.0, v2 = 5.0 / 24.0;
    const double c0 = 7.0 / 24.0, c1 = 175.0 / 768.0, c2 = 5.0 / 48.0, c3 = 35.0 / 1536.0,
                 c4 = 0.0, c5 = -1.0 / 1536.0;
    std::vector<double> rhs(n), di(n, v0), lo(n, v1), lo2(n, v2), up(n, v1), up2(n, v2);
    auto P = [&](int i) { return p[(i % n + n) % n]; };
    if (bc == ORC_PERIODIC) {
        for (int i = 0; i < n; ++i)
            rhs[i] = c5 * P(i - 5) + c4 * P(i - 4) + c3 * P(i - 3) + c2 * P(i - 2) +
                     c1 * P(i - 1) + c0 * p[i] + c1 * P(i + 1) + c2 * P(i + 2) + c3 * P(i + 3) +
                     c4 * P(i + 4) + c5 * P(i + 5);
        const double cp[6] = {v2, v1, v2, v2, v1, v2};
        orc_cyclic_penta(lo2.data(), lo.data(), di.data(), up.data(), up2.data(), rhs.data(), cp,
                         out, n);
        return;
    }
    rhs[0] = w0_n1 * p[0] + w1_n1 * p[1] + w2_n1 * p[2] + w3_n1 * p[3] + w4_n1 * p[4] + w5_n1 * p[5];
    rhs[1] = wn1_n2 * p[0] + w0_n2 * p[1] + w1_n2 * p[2] + w2_n2 * p[3] + w3_n2 * p[4] + w4_n2 * p[5];
    rhs[2] = wn2_n3 * p[0] + wn1_n3 * p[1] + w0_n3 * p[2] + w1_n3 * p[3] + w2_n3 * p[4] + w3_n3 * p[5];
    rhs[3] = w3 * p[0] + w2 * p[1] + w1 * p[2] + w0 * p[3] + w1 * p[4] + w2 * p[5] + w3 * p[6];
    rhs[4] = w3 * p[1] + w2 * p[2] + w1 * p[3] + w0 * p[4] + w1 * p[5] + w2 * p[6] + w3 * p[7];
    for (int i = 5; i < n - 5; ++i)
        rhs[i] = c5 * p[i - 5] + c4 * p[i - 4] + c3 * p[i - 3] + c2 * p[i - 2] + c1 * p[i - 1] +
                 c0 * p[i] + c1 * p[i + 1] + c2 * p[i + 2] + c3 * p[i + 3] + c4 * p[i + 4] +
                 c5 * p[i + 5];
    rhs[n - 1] = w0_n1 * p[n - 1] + w1_n1 * p[n - 2] + w2_n1 * p[n - 3] + w3_n1 * p[n - 4] +
                 w4_n1 * p[n - 5] + w5_n1 * p[n - 6];
    rhs[n - 2] = wn1_n2 * p[n - 1] + w0_n2 * p[n - 2] + w1_n2 * p[n - 3] + w2_n2 * p[n - 4] +
                 w3_n2 * p[n - 5] + w4_n2 * p[n - 6];
    rhs[n - 3] = wn2_n3 * p[n - 1] + wn1_n3 * p[n - 2] + w0_n3 * p[n - 3] + w1_n3 * p[n - 4] +
                 w2_n3 * p[n - 5] + w3_n3 * p[n - 6];
    rhs[n - 4] = w3 * p[n - 1] + w2 * p[n - 2] + w1 * p[n - 3] + w0 * p[n - 4] + w1 * p[n - 5] +
                 w2 * p[n - 6] + w3 * p[n - 7];
    rhs[n - 5] = w3 * p[n - 2] + w2 * p[n - 3] + w1 * p[n - 4] + w0 * p[n - 5] + w1 * p[n - 6] +
                 w2 * p[n - 7] + w3 * p[n - 8];
    // Dirichlet band edits: identity rows 0..4 and N-5..N-1,
    // src/CommN5PadeTestFilter.hpp:100-179 (lower bands indexed by column)
    di[0] = 1.0; up[0] = 0.0; up2[0] = 0.0;
    lo[0] = 0.0; di[1] = 1.0; up[1] = 0.0; up2[1] = 0.0;
    lo2[0] = 0.0; lo[1] = 0.0; di[2] = 1.0; up[2] = 0.0; up2[2] = 0.0;
    lo2[1] = 0.0; lo[2] = 0.0; di[3] = 1.0; up[3] = 0.0; up2[3] = 0.0;
    lo2[2] = 0.0; lo[3] = 0.0; di[4] = 1.0; up[4] = 0.0; up2[4] = 0.0;
    lo2[n - 3] = 0.0; lo[n - 2] = 0.0; di[n - 1] = 1.0;
    lo2[n - 4] = 0.0; lo[n - 3] = 0.0; di[n - 2] = 1.0; up[n - 2] = 0.0;
    lo2[n - 5] = 0.0; lo[n - 4] = 0.0; di[n - 3] = 1.0; up[n - 3] = 0.0; up2[n - 3] = 0.0;
    lo2[n - 6] = 0.0; lo[n - 5] = 0.0; di[n - 4] = 1.0; up[n - 4] = 0.0; up2[n - 4] = 0.0;
    lo2[n - 7] = 0.0; lo[n - 6] = 0.0; di[n - 5] = 1.0; up[n - 5] = 0.0; up2[n - 5] = 0.0;
    penta(lo2.data(), lo.data(), di.data(), up.data(), up2.data(), rhs.data(), out, n);
}

// Explicit commutative test filter: src/CommN5ExpTestFilter.hpp:46-70,
// src/CommN5ExpTestFilter.cpp:5-70
void commn5exp_line(int bc, int n, const double *p, double *out) {
    const double w0_n1 = 31.0 / 32.0, w1_n1 = 5.0 / 32.0, w2_n1 = -5.0 / 16.0, w3_n1 = 5.0 / 16.0,
                 w4_n1 = -5.0 / 32.0, w5_n1 = 1.0 / 32.0;
    const double wn1_n2 = 1.0 / 32.0, w0_n2 = 27.0 / 32.0, w1_n2 = 5.0 / 16.0,
                 w2_n2 = -5.0 / 16.0, w3_n2 = 5.0 / 32.0, w4_n2 = -1.0 / 32.0;
    const double wn2_n3 = -1.0 / 32.0, wn1_n3 = 5.0 / 32.0, w0_n3 = 11.0 / 16.0,
                 w1_n3 = 5.0 / 16.0, w2_n3 = -5.0 / 32.0, w3_n3 = 1.0 / 32.0;
    const double w0 = 11.0 / 16.0, w1 = 15.0 / 64.0, w2 = -3.0 / 32.0, w3 = 1.0 / 64.0;
    auto P = [&](int i) { return p[(i % n + n) % n]; };
    if (bc == ORC_PERIODIC) {
        for (int i = 0; i < n; ++i)
            out[i] = w3 * P(i - 3) + w2 * P(i - 2) + w1 * P(i - 1) + w0 * p[i] + w1 * P(i + 1) +
                     w2 * P(i + 2) + w3 * P(i + 3);
        return;
    }
    out[0] = w0_n1 * p[0] + w1_n1 * p[1] + w2_n1 * p[2] + w3_n1 * p[3] + w4_n1 * p[4] + w5_n1 * p[5];
    out[1] = wn1_n2 * p[0] + w0_n2 * p[1] + w1_n2 * p[2] + w2_n2 * p[3] + w3_n2 * p[4] + w4_n2 * p[5];
    out[2] = wn2_n3 * p[0] + wn1_n3 * p[1] + w0_n3 * p[2] + w1_n3 * p[3] + w2_n3 * p[4] + w3_n3 * p[5];
    for (int i = 3; i < n - 3; ++i)
        out[i] = w3 * p[i - 3] + w2 * p[i - 2] + w1 * p[i - 1] + w0 * p[i] + w1 * p[i + 1] +
                 w2 * p[i + 2] + w3 * p[i + 3];
    out[n - 1] = w0_n1 * p[n - 1] + w1_n1 * p[n - 2] + w2_n1 * p[n - 3] + w3_n1 * p[n - 4] +
                 w4_n1 * p[n - 5] + w5_n1 * p[n - 6];
    out[n - 2] = wn1_n2 * p[n - 1] + w0_n2 * p[n - 2] + w1_n2 * p[n - 3] + w2_n2 * p[n - 4] +
                 w3_n2 * p[n - 5] + w4_n2 * p[n - 6];
    out[n - 3] = wn2_n3 * p[n - 1] + wn1_n3 * p[n - 2] + w0_n3 * p[n - 3] + w1_n3 * p[n - 4] +
                 w2_n3 * p[n - 5] + w3_n3 * p[n - 6];
}

// Line extraction for the x-fastest layout.
struct Geo {
    int Nx, Ny, Nz;
    int n_line(int dir) const { return dir == 0 ? Nx : dir == 1 ? Ny : Nz; }
    long stride(int dir) const { return dir == 0 ? 1L : dir == 1 ? (long)Nx : (long)Nx * Ny; }
    int n_lines(int dir) const { return Nx * Ny * Nz / n_line(dir); }
    // line q enumerates the two other indices with the lower one fastest
    long base(int dir, int q) const {
        if (dir == 0) return (long)q * Nx;
        if (dir == 1) return (long)(q / Nx) * Nx * Ny + q % Nx;
        return q;
    }
};

}  // namespace

extern "C" {

int orc_deriv_line(int scheme, int bc, int n, double dd, const double *phi, double *dphi,
                   const double *tpb) {
    if (n <= 10) return -1;
    Line f{phi, n, tpb};
    switch (scheme) {
        case ORC_PADE6: pade6_line(bc, n, dd, f, dphi); break;
        case ORC_PENTA10: penta10_line(bc, n, dd, f, dphi); break;
        case ORC_CD2: cd2_line(bc, n, dd, f, dphi); break;
        default: return -2;
    }
    return 0;
}

int orc_filter_line(int kind, int bc, int n, double alphaF, const double *phi, double *phiF) {
    if (n <= 10) return -1;
    switch (kind) {
        case ORC_COMPACT8: compact8_line(bc, n, alphaF, phi, phiF); break;
        case ORC_COMPACT10: compact10_line(bc, n, alphaF, phi, phiF); break;
        case ORC_COMMN5PADE: commn5pade_line(bc, n, phi, phiF); break;
        case ORC_COMMN5EXP: commn5exp_line(bc, n, phi, phiF); break;
        default: return -2;
    }
    return 0;
}

// One-sided wall value for the Neumann temperature condition.
// src/Pade6.cpp:395-397, src/Penta10.cpp:483-485, src/CD2.cpp:183-185
double orc_neumann(int scheme, const double *f) {
    if (scheme == ORC_CD2) return (f[0] * 4.0 - f[1] * 1.0) / 3.0;
    return (f[0] * 360.0 - f[1] * 450.0 + f[2] * 400.0 - f[3] * 225.0 + f[4] * 72.0 -
            f[5] * 10.0) / 147.0;
}

// Field loops: src/Pade6.cpp:3-45 (and the identical loops of Penta10/CD2);
// here every direction walks its lines in the single x-fastest layout.
int orc_deriv_field(int scheme, int dir, int bc, int Nx, int Ny, int Nz, double dd,
                    const double *in, double *out) {
    Geo g{Nx, Ny, Nz};
    const int n = g.n_line(dir);
    const long s = g.stride(dir);
    int err = 0;
#pragma omp parallel
    {
        std::vector<double> li(n), lo(n);
#pragma omp for schedule(static)
        for (int q = 0; q < g.n_lines(dir); ++q) {
            const long b = g.base(dir, q);
            for (int i = 0; i < n; ++i) li[i] = in[b + i * s];
            int rc = orc_deriv_line(scheme, bc, n, dd, li.data(), lo.data(), nullptr);
            if (rc) {
#pragma omp atomic write
                err = rc;
            }
            for (int i = 0; i < n; ++i) out[b + i * s] = lo[i];
        }
    }
    return err;
}

int orc_deriv_field_tpb(int scheme, int dir, int Nx, int Ny, int Nz, double dd, const double *in,
                        double *out, const double *const wrap[6]) {
    Geo g{Nx, Ny, Nz};
    const int n = g.n_line(dir);
    const long s = g.stride(dir);
    std::vector<double> li(n), lo(n);
    double t[6];
    for (int q = 0; q < g.n_lines(dir); ++q) {
        const long b = g.base(dir, q);
        for (int i = 0; i < n; ++i) li[i] = in[b + i * s];
        for (int k = 0; k < 6; ++k) t[k] = wrap[k] ? wrap[k][q] : 0.0;
        int rc = orc_deriv_line(scheme, ORC_PERIODIC, n, dd, li.data(), lo.data(), t);
        if (rc) return rc;
        for (int i = 0; i < n; ++i) out[b + i * s] = lo[i];
    }
    return 0;
}

// Filter field loop with the pass-through of lines lying on a Dirichlet face
// of another direction.  src/Compact8Filter.cpp:142-230 (the same loop in
// Compact10Filter.cpp:166-252, CommN5PadeTestFilter.cpp:155-241,
// CommN5ExpTestFilter.cpp:72-158)
int orc_filter_field(int kind, double alphaF, int dir, const int bc3[3], int Nx, int Ny, int Nz,
                     const double *in, double *out) {
    Geo g{Nx, Ny, Nz};
    const int n = g.n_line(dir);
    const long s = g.stride(dir);
    int err = 0;
#pragma omp parallel
    {
        std::vector<double> li(n), lo(n);
#pragma omp for schedule(static)
        for (int q = 0; q < g.n_lines(dir); ++q) {
            const long b = g.base(dir, q);
            // coordinates of the line in the two other directions
            int ii = 0, jj = 0, kk = 0;
            if (dir == 0) { jj = q % Ny; kk = q / Ny; }
            if (dir == 1) { ii = q % Nx; kk = q / Nx; }
            if (dir == 2) { ii = q % Nx; jj = q / Nx; }
            bool pass = false;
            if (dir != 0 && bc3[0] == ORC_DIRICHLET && (ii == 0 || ii == Nx - 1)) pass = true;
            if (dir != 1 && bc3[1] == ORC_DIRICHLET && (jj == 0 || jj == Ny - 1)) pass = true;
            if (dir != 2 && bc3[2] == ORC_DIRICHLET && (kk == 0 || kk == Nz - 1)) pass = true;
            for (int i = 0; i < n; ++i) li[i] = in[b + i * s];
            if (pass) {
                lo = li;
            } else {
                int rc = orc_filter_line(kind, bc3[dir], n, alphaF, li.data(), lo.data());
                if (rc) {
#pragma omp atomic write
                    err = rc;
                }
            }
            for (int i = 0; i < n; ++i) out[b + i * s] = lo[i];
        }
    }
    return err;
}

}  // extern "C"
