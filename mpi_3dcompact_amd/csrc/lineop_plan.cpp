// Host-side construction of partitioned line operators (see lineop.h).
//
// Coefficient tables follow the reference schemes; each block cites the lines
// it takes its numbers from.  Matrices are assembled densely in long double from
// the reference's band layouts (tridiagonal lower band indexed by row,
// pentadiagonal lower bands indexed by column, src/Utils.cpp:4-83) and the
// partition factors are derived from them.
#include "lineop_plan.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace mc3d {
namespace {

using ld = long double;

struct Mat {
    int n = 0;
    std::vector<ld> a;
    explicit Mat(int n_ = 0) : n(n_), a((size_t)n_ * n_, 0.0L) {}
    ld &operator()(int i, int j) { return a[(size_t)i * n + j]; }
    ld operator()(int i, int j) const { return a[(size_t)i * n + j]; }
};

// tridiagonal with lower band indexed by row (a[i] = A[i][i-1])
void put_tri(Mat &A, const std::vector<ld> &lo, const std::vector<ld> &di,
             const std::vector<ld> &up) {
    int n = A.n;
    for (int i = 0; i < n; ++i) {
        A(i, i) = di[i];
        if (i > 0) A(i, i - 1) = lo[i];
        if (i < n - 1) A(i, i + 1) = up[i];
    }
}

// pentadiagonal with lower bands indexed by column (e[i] = A[i+2][i], a[i] = A[i+1][i])
void put_penta(Mat &A, const std::vector<ld> &e, const std::vector<ld> &a,
               const std::vector<ld> &d, const std::vector<ld> &c, const std::vector<ld> &f) {
    int n = A.n;
    for (int i = 0; i < n; ++i) {
        A(i, i) = d[i];
        if (i + 1 < n) { A(i, i + 1) = c[i]; A(i + 1, i) = a[i]; }
        if (i + 2 < n) { A(i, i + 2) = f[i]; A(i + 2, i) = e[i]; }
    }
}

// Special row: taps in the reference's own summation order; tap t reads
// phi[p + off + t*step] for the row at point p.
struct Row {
    int off;
    int step;
    std::vector<double> c;
};

// RHS coefficients are kept in double and formed with the reference's own
// expressions so that the explicit stage is bitwise reproducible; the matrix
// is assembled in long double from the same double band values.
struct Scheme {
    int W = 0, H = 0;
    int int_off = 0;
    std::vector<double> interior;  // taps p+int_off .. summed left to right
    std::vector<Row> left, right;  // left[r] is row r; right[r] is row N-nbr+r
    Mat A;
    bool divide = false;
};

std::vector<ld> L(const std::vector<double> &v) { return std::vector<ld>(v.begin(), v.end()); }

// rows summed in increasing index from phi[0]
Row inc_row(int r, std::vector<double> c) { return Row{-r, 1, std::move(c)}; }
// right-end row N-1-r summed downward from phi[N-1]
Row dec_row(int r, std::vector<double> c) { return Row{r, -1, std::move(c)}; }
// right-end row N-1-r summed upward, ending at phi[N-1]
Row incR_row(int r, std::vector<double> c) {
    int k = (int)c.size();
    return Row{r - (k - 1), 1, std::move(c)};
}

void finish_right(Scheme &s) {
    // right rows are generated for r = 0 (row N-1), 1 (row N-2), ...; store in row order
    std::reverse(s.right.begin(), s.right.end());
}

// Pade6: src/Pade6.hpp:73-112,201-229; RHS src/Pade6.cpp:187-218,284-313.
// e2_1D is not assigned under boundaryScheme 1 and enters as 0.
Scheme pade6(int n, bool periodic) {
    Scheme s;
    s.W = 1; s.H = 2; s.divide = true;
    const double al = 1.0 / 3.0, a = 14.0 / 9.0, b = 1.0 / 9.0;
    const double cc2 = -b / 4.0, cc3 = -a / 2.0, cc4 = a / 2.0, cc5 = b / 4.0;
    s.interior = {cc2, cc3, 0.0, cc4, cc5};
    s.int_off = -2;
    std::vector<double> lo(n, al), di(n, 1.0), up(n, al);
    s.A = Mat(n);
    if (periodic) {
        put_tri(s.A, L(lo), L(di), L(up));
        s.A(0, n - 1) = al;
        s.A(n - 1, 0) = al;
        return s;
    }
    const double a1 = -5.0 / 2.0, b1 = 2.0, c1 = 1.0 / 2.0, d1 = 0.0, e1 = 0.0, f1 = 0.0;
    const double a2 = -5.0 / 9.0, b2 = -1.0 / 2.0, c2 = 1.0, d2 = 1.0 / 18.0, e2 = 0.0;
    s.left = {inc_row(0, {a1, b1, c1, d1, e1, f1}), inc_row(1, {a2, b2, c2, d2, e2})};
    s.right = {incR_row(0, {-f1, -e1, -d1, -c1, -b1, -a1}), incR_row(1, {-e2, -d2, -c2, -b2, -a2})};
    finish_right(s);
    up[0] = 2.0; up[1] = 1.0 / 2.0; lo[1] = 1.0 / 6.0;
    lo[n - 1] = 2.0; lo[n - 2] = 1.0 / 2.0; up[n - 2] = 1.0 / 6.0;
    put_tri(s.A, L(lo), L(di), L(up));
    return s;
}

// Penta10: src/Penta10.hpp:70-175; RHS src/Penta10.cpp:141-173,211-252.
Scheme penta10(int n, bool periodic) {
    Scheme s;
    s.W = 2; s.H = 3; s.divide = true;
    const double be = 1.0 / 20.0, al = 1.0 / 2.0, a = 17.0 / 12.0, b = 101.0 / 150.0,
                 c = 1.0 / 100.0;
    const double cc1 = -c / 6.0, cc2 = -b / 4.0, cc3 = -a / 2.0, cc4 = a / 2.0, cc5 = b / 4.0,
                 cc6 = c / 6.0;
    s.interior = {cc1, cc2, cc3, 0.0, cc4, cc5, cc6};
    s.int_off = -3;
    std::vector<double> di(n, 1.0), lo(n, al), up(n, al), lo2(n, be), up2(n, be);
    s.A = Mat(n);
    if (periodic) {
        put_penta(s.A, L(lo2), L(lo), L(di), L(up), L(up2));
        s.A(0, n - 2) = be; s.A(0, n - 1) = al; s.A(1, n - 1) = be;
        s.A(n - 2, 0) = be; s.A(n - 1, 0) = al; s.A(n - 1, 1) = be;
        return s;
    }
    const double a1 = -197.0 / 60.0, b1 = -5.0 / 12.0, c1 = 5.0 / 1.0, d1 = -5.0 / 3.0,
                 e1 = 5.0 / 12.0, f1 = -1.0 / 20.0, g1 = 0.0, h1 = 0.0, i1 = 0.0;
    const double a2 = -79.0 / 240.0, b2 = -77.0 / 60.0, c2 = 55.0 / 48.0, d2 = 5.0 / 9.0,
                 e2 = -5.0 / 48.0, f2 = 1.0 / 60.0, g2 = -1.0 / 720.0, h2 = 0.0;
    const double a3 = -25.0 / 54.0 / 4.0, b3 = -40.0 / 27.0 / 2.0, c3 = 0.0,
                 d3 = 40.0 / 27.0 / 2.0, e3 = 25.0 / 54.0 / 4.0, f3 = 0.0, g3 = 0.0;
    s.left = {inc_row(0, {a1, b1, c1, d1, e1, f1, g1, h1, i1}),
              inc_row(1, {a2, b2, c2, d2, e2, f2, g2, h2}),
              inc_row(2, {a3, b3, c3, d3, e3, f3, g3})};
    s.right = {incR_row(0, {-i1, -h1, -g1, -f1, -e1, -d1, -c1, -b1, -a1}),
               incR_row(1, {-h2, -g2, -f2, -e2, -d2, -c2, -b2, -a2}),
               incR_row(2, {-g3, -f3, -e3, -d3, -c3, -b3, -a3})};
    finish_right(s);
    const double beta1 = 0.0, alpha1 = 5.0, beta2 = 0.0, a21 = 1.0 / 12.0, a22 = 5.0 / 4.0;
    const double b31 = 1.0 / 36.0, b32 = 1.0 / 36.0, a31 = 4.0 / 9.0, a32 = 4.0 / 9.0;
    up[0] = alpha1; up[1] = a22; up[2] = a32;
    lo[0] = a21; lo[1] = a31;
    up2[0] = beta1; up2[1] = beta2; up2[2] = b32;
    lo2[0] = b31;
    up[n - 2] = a21; up[n - 3] = a31;
    lo[n - 2] = alpha1; lo[n - 3] = a22; lo[n - 4] = a32;
    up2[n - 3] = b31;
    lo2[n - 3] = beta1; lo2[n - 4] = beta2; lo2[n - 5] = b32;
    put_penta(s.A, L(lo2), L(lo), L(di), L(up), L(up2));
    return s;
}

// CD2: src/CD2.hpp:54-61, src/CD2.cpp:129-181.  The interior evaluates
// (a*phi[i+1] + c*phi[i-1]) with a=-1/2, c=+1/2, i.e. the negated central
// difference; the periodic wrap rows evaluate (a*phi[i-1] + c*phi[i+1]).
Scheme cd2(int n, bool periodic) {
    Scheme s;
    s.W = 0; s.H = 1; s.divide = true;
    const double a = -1.0 / 2.0, c = 1.0 / 2.0;
    const double aa = -3.0 / 2.0, bb = 4.0 / 2.0, cc = -1.0 / 2.0;
    s.interior = {c, 0.0, a};  // c*phi[i-1] + 0 + a*phi[i+1] == a*phi[i+1] + c*phi[i-1]
    s.int_off = -1;
    s.A = Mat(0);
    if (periodic) {
        s.left = {Row{-1, 2, {a, c}}};
        s.right = {Row{-1, 2, {a, c}}};
    } else {
        s.left = {Row{0, 1, {aa, bb, cc}}};
        s.right = {Row{0, -1, {aa, bb, cc}}};
    }
    return s;
}

// Compact8 filter: src/Compact8Filter.hpp:55-126, src/Compact8Filter.cpp:5-109.
Scheme compact8(int n, bool periodic, double af) {
    Scheme s;
    s.W = 1; s.H = 4;
    const double a0 = (93.0 + 70.0 * af) / 128.0, a1 = (7.0 + 18.0 * af) / 16.0,
                 a2 = (-7.0 + 14.0 * af) / 32.0, a3 = (1.0 - 2.0 * af) / 16.0,
                 a4 = (-1.0 + 2.0 * af) / 128.0;
    const double c0 = a0, c1 = a1 / 2.0, c2 = a2 / 2.0, c3 = a3 / 2.0, c4 = a4 / 2.0;
    s.interior = {c4, c3, c2, c1, c0, c1, c2, c3, c4};
    s.int_off = -4;
    std::vector<double> lo(n, af), di(n, 1.0), up(n, af);
    s.A = Mat(n);
    if (periodic) {
        put_tri(s.A, L(lo), L(di), L(up));
        s.A(0, n - 1) = af; s.A(n - 1, 0) = af;
        return s;
    }
    std::vector<double> B2 = {1.0 / 256.0 + 127.0 * af / 128.0, 31.0 / 32.0 + 1.0 * af / 16.0,
                              7.0 / 64.0 + 25.0 * af / 32.0, -7.0 / 32.0 + 7.0 * af / 16.0,
                              7.0 * (5.0 - 10.0 * af) / 128.0, -7.0 / 32.0 + 7.0 * af / 16.0,
                              7.0 / 64.0 - 7.0 * af / 32.0, -1.0 / 32.0 + 1.0 * af / 16.0,
                              1.0 / 256.0 - 1.0 * af / 128.0};
    std::vector<double> B3 = {-1.0 / 256.0 + 1.0 * af / 128.0, 1.0 / 32.0 + 15.0 * af / 16.0,
                              57.0 / 64.0 + 7.0 * af / 32.0, 7.0 / 32.0 + 9.0 * af / 16.0,
                              7.0 * (-5.0 + 10.0 * af) / 128.0, 7.0 / 32.0 - 7.0 * af / 16.0,
                              -7.0 / 64.0 + 7.0 * af / 32.0, 1.0 / 32.0 - 1.0 * af / 16.0,
                              -1.0 / 256.0 + 1.0 * af / 128.0};
    std::vector<double> B4 = {1.0 / 256.0 - 1.0 * af / 128.0, -1.0 / 32.0 + 1.0 * af / 16.0,
                              7.0 / 64.0 + 25.0 * af / 32.0, 25.0 / 32.0 + 7.0 * af / 16.0,
                              35.0 / 128.0 + 29.0 * af / 64.0, -7.0 / 32.0 + 7.0 * af / 16.0,
                              7.0 / 64.0 - 7.0 * af / 32.0, -1.0 / 32.0 + 1.0 * af / 16.0,
                              1.0 / 256.0 - 1.0 * af / 128.0};
    s.left = {inc_row(0, {1.0}), inc_row(1, B2), inc_row(2, B3), inc_row(3, B4)};
    s.right = {dec_row(0, {1.0}), dec_row(1, B2), dec_row(2, B3), dec_row(3, B4)};
    finish_right(s);
    lo[0] = 0; up[0] = 0; lo[n - 1] = 0; up[n - 1] = 0;
    put_tri(s.A, L(lo), L(di), L(up));
    return s;
}

// Compact10 filter (boundaryScheme 1): src/Compact10Filter.hpp:58-134,283-295,
// src/Compact10Filter.cpp:5-133.
Scheme compact10(int n, bool periodic, double af) {
    Scheme s;
    s.W = 1; s.H = 5;
    const double a0 = (193.0 + 126.0 * af) / 256.0, a1 = (105.0 + 302.0 * af) / 256.0,
                 a2 = (-15.0 + 30.0 * af) / 64.0, a3 = (45.0 - 90.0 * af) / 512.0,
                 a4 = (-5.0 + 10.0 * af) / 256.0, a5 = (1.0 - 2.0 * af) / 512.0;
    const double c0 = a0, c1 = a1 / 2.0, c2 = a2 / 2.0, c3 = a3 / 2.0, c4 = a4 / 2.0,
                 c5 = a5 / 2.0;
    s.interior = {c5, c4, c3, c2, c1, c0, c1, c2, c3, c4, c5};
    s.int_off = -5;
    std::vector<double> lo(n, af), di(n, 1.0), up(n, af);
    s.A = Mat(n);
    if (periodic) {
        put_tri(s.A, L(lo), L(di), L(up));
        s.A(0, n - 1) = af; s.A(n - 1, 0) = af;
        return s;
    }
    std::vector<std::vector<double>> B = {
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
    s.left = {inc_row(0, {1.0})};
    s.right = {dec_row(0, {1.0})};
    for (int r = 0; r < 4; ++r) {
        s.left.push_back(inc_row(r + 1, B[r]));
        s.right.push_back(dec_row(r + 1, B[r]));
    }
    finish_right(s);
    lo[0] = 0; up[0] = 0; lo[n - 1] = 0; up[n - 1] = 0;
    put_tri(s.A, L(lo), L(di), L(up));
    return s;
}

// Commutative Pade test filter: src/CommN5PadeTestFilter.hpp:55-186,
// src/CommN5PadeTestFilter.cpp:5-122.
Scheme commn5pade(int n, bool periodic) {
    Scheme s;
    s.W = 2; s.H = 5;
    const double v0 = 7.0 / 12.0, v1 = 0.0, v2 = 5.0 / 24.0;
    const double wi0 = 7.0 / 24.0, wi1 = 175.0 / 768.0, wi2 = 5.0 / 48.0, wi3 = 35.0 / 1536.0,
                 wi4 = 0.0, wi5 = -1.0 / 1536.0;
    s.interior = {wi5, wi4, wi3, wi2, wi1, wi0, wi1, wi2, wi3, wi4, wi5};
    s.int_off = -5;
    std::vector<double> di(n, v0), lo(n, v1), lo2(n, v2), up(n, v1), up2(n, v2);
    s.A = Mat(n);
    if (periodic) {
        put_penta(s.A, L(lo2), L(lo), L(di), L(up), L(up2));
        s.A(0, n - 2) = v2; s.A(0, n - 1) = v1; s.A(1, n - 1) = v2;
        s.A(n - 2, 0) = v2; s.A(n - 1, 0) = v1; s.A(n - 1, 1) = v2;
        return s;
    }
    std::vector<double> n1 = {31.0 / 32.0, 5.0 / 32.0, -5.0 / 16.0, 5.0 / 16.0, -5.0 / 32.0, 1.0 / 32.0};
    std::vector<double> n2 = {1.0 / 32.0, 27.0 / 32.0, 5.0 / 16.0, -5.0 / 16.0, 5.0 / 32.0, -1.0 / 32.0};
    std::vector<double> n3 = {-1.0 / 32.0, 5.0 / 32.0, 11.0 / 16.0, 5.0 / 16.0, -5.0 / 32.0, 1.0 / 32.0};
    const double w0 = 11.0 / 16.0, w1 = 15.0 / 64.0, w2 = -3.0 / 32.0, w3 = 1.0 / 64.0;
    std::vector<double> w7 = {w3, w2, w1, w0, w1, w2, w3};
    s.left = {inc_row(0, n1), inc_row(1, n2), inc_row(2, n3), inc_row(3, w7), Row{-3, 1, w7}};
    // rows N-1..N-3 and N-4, N-5 are summed downward from their last tap
    s.right = {dec_row(0, n1), dec_row(1, n2), dec_row(2, n3), dec_row(3, w7), Row{3, -1, w7}};
    finish_right(s);
    di[0] = 1; up[0] = 0; up2[0] = 0;
    lo[0] = 0; di[1] = 1; up[1] = 0; up2[1] = 0;
    lo2[0] = 0; lo[1] = 0; di[2] = 1; up[2] = 0; up2[2] = 0;
    lo2[1] = 0; lo[2] = 0; di[3] = 1; up[3] = 0; up2[3] = 0;
    lo2[2] = 0; lo[3] = 0; di[4] = 1; up[4] = 0; up2[4] = 0;
    lo2[n - 3] = 0; lo[n - 2] = 0; di[n - 1] = 1;
    lo2[n - 4] = 0; lo[n - 3] = 0; di[n - 2] = 1; up[n - 2] = 0;
    lo2[n - 5] = 0; lo[n - 4] = 0; di[n - 3] = 1; up[n - 3] = 0; up2[n - 3] = 0;
    lo2[n - 6] = 0; lo[n - 5] = 0; di[n - 4] = 1; up[n - 4] = 0; up2[n - 4] = 0;
    lo2[n - 7] = 0; lo[n - 6] = 0; di[n - 5] = 1; up[n - 5] = 0; up2[n - 5] = 0;
    put_penta(s.A, L(lo2), L(lo), L(di), L(up), L(up2));
    return s;
}

// Explicit commutative test filter: src/CommN5ExpTestFilter.hpp:46-70,
// src/CommN5ExpTestFilter.cpp:5-70.
Scheme commn5exp(int n, bool periodic) {
    Scheme s;
    s.W = 0; s.H = 3;
    const double w0 = 11.0 / 16.0, w1 = 15.0 / 64.0, w2 = -3.0 / 32.0, w3 = 1.0 / 64.0;
    s.interior = {w3, w2, w1, w0, w1, w2, w3};
    s.int_off = -3;
    s.A = Mat(0);
    if (periodic) return s;
    std::vector<double> n1 = {31.0 / 32.0, 5.0 / 32.0, -5.0 / 16.0, 5.0 / 16.0, -5.0 / 32.0, 1.0 / 32.0};
    std::vector<double> n2 = {1.0 / 32.0, 27.0 / 32.0, 5.0 / 16.0, -5.0 / 16.0, 5.0 / 32.0, -1.0 / 32.0};
    std::vector<double> n3 = {-1.0 / 32.0, 5.0 / 32.0, 11.0 / 16.0, 5.0 / 16.0, -5.0 / 32.0, 1.0 / 32.0};
    s.left = {inc_row(0, n1), inc_row(1, n2), inc_row(2, n3)};
    s.right = {dec_row(0, n1), dec_row(1, n2), dec_row(2, n3)};
    finish_right(s);
    return s;
}

// ---- dense helpers (long double) ----
bool lu_band(const Mat &D, int W, std::vector<ld> &l1, std::vector<ld> &l2,
             std::vector<ld> &u1, std::vector<ld> &u2, std::vector<ld> &dd, ld &minpiv) {
    // Doolittle without pivoting, restricted to the band.
    const int m = D.n;
    Mat U = D;
    l1.assign(m, 0); l2.assign(m, 0); u1.assign(m, 0); u2.assign(m, 0); dd.assign(m, 0);
    minpiv = 1e300L;
    for (int k = 0; k < m; ++k) {
        ld piv = U(k, k);
        minpiv = std::min(minpiv, std::fabs(piv));
        if (std::fabs(piv) < 1e-14L) return false;
        for (int i = k + 1; i <= std::min(m - 1, k + W); ++i) {
            ld f = U(i, k) / piv;
            if (i == k + 1) l1[i] = f; else l2[i] = f;
            for (int j = k; j <= std::min(m - 1, k + W); ++j) U(i, j) -= f * U(k, j);
        }
    }
    for (int k = 0; k < m; ++k) {
        dd[k] = U(k, k);
        if (k + 1 < m) u1[k] = U(k, k + 1);
        if (k + 2 < m && W == 2) u2[k] = U(k, k + 2);
    }
    return true;
}

void lu_solve(int W, const std::vector<ld> &l1, const std::vector<ld> &l2,
              const std::vector<ld> &u1, const std::vector<ld> &u2, const std::vector<ld> &dd,
              std::vector<ld> &x) {
    const int m = (int)x.size();
    for (int i = 0; i < m; ++i) {
        if (i >= 1) x[i] -= l1[i] * x[i - 1];
        if (W == 2 && i >= 2) x[i] -= l2[i] * x[i - 2];
    }
    for (int i = m - 1; i >= 0; --i) {
        if (i + 1 < m) x[i] -= u1[i] * x[i + 1];
        if (W == 2 && i + 2 < m) x[i] -= u2[i] * x[i + 2];
        x[i] /= dd[i];
    }
}

bool invert(Mat &A) {  // Gauss-Jordan with partial pivoting
    const int n = A.n;
    Mat I(n);
    for (int i = 0; i < n; ++i) I(i, i) = 1;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A(i, k)) > std::fabs(A(p, k))) p = i;
        if (std::fabs(A(p, k)) < 1e-30L) return false;
        if (p != k)
            for (int j = 0; j < n; ++j) { std::swap(A(k, j), A(p, j)); std::swap(I(k, j), I(p, j)); }
        ld inv = 1.0L / A(k, k);
        for (int j = 0; j < n; ++j) { A(k, j) *= inv; I(k, j) *= inv; }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            ld f = A(i, k);
            if (f == 0) continue;
            for (int j = 0; j < n; ++j) { A(i, j) -= f * A(k, j); I(i, j) -= f * I(k, j); }
        }
    }
    A = I;
    return true;
}

}  // namespace

int build_plan(const PlanRequest &rq, Plan &plan, std::string &err) {
    const int n = rq.N;
    if (n <= 10) { err = "line length must exceed 10 (src/Options.hpp:241)"; return -1; }
    if (n > 4096) { err = "line length above 4096 is not supported"; return -1; }
    const bool per = rq.bc == 0;
    Scheme s;
    if (rq.cls == kDeriv) {
        if (rq.kind == 0) s = cd2(n, per);
        else if (rq.kind == 1) s = pade6(n, per);
        else if (rq.kind == 2) s = penta10(n, per);
        else { err = "unknown derivative scheme"; return -1; }
    } else {
        if (rq.kind == 0) s = compact8(n, per, rq.alphaF);
        else if (rq.kind == 1) s = compact10(n, per, rq.alphaF);
        else if (rq.kind == 2) s = commn5pade(n, per);
        else if (rq.kind == 3) s = commn5exp(n, per);
        else { err = "unknown filter kind"; return -1; }
    }
    const int W = s.W;
    // chunking: 32-row chunks when the line splits evenly into at least two,
    // 16-row ones when it splits evenly into those, else the largest of 28, 24,
    // 20 (and, periodic lines only, 12 and 8) rows that splits it evenly into
    // 2..16 chunks, so that the fast kernels take it; else 16-row chunks (32
    // above 256 points, so that at most 16 chunks reach 512 points) and one
    // shorter chunk (below)
    int mm = rq.mm;
    if (mm == 0) {
        if (n % 32 == 0 && n / 32 >= 2) mm = 32;
        else if (n % 16 == 0 && n / 16 <= 16) mm = 16;
        else {
            for (int c : {28, 24, 20, 12, 8})
                if (n % c == 0 && n / c >= 2 && n / c <= 16 && (per || c >= 16)) { mm = c; break; }
            if (!mm) mm = n > 256 ? 32 : 16;
        }
    }
    if (!chunk_rows_ok(mm)) { err = "chunk rows must be 8, 12, 16, 20, 24, 28 or 32"; return -1; }
    int P = (n + mm - 1) / mm;
    if (per && W > 0 && P < 2) P = 2;
    int M = (n + P - 1) / P;
    int last = n - (P - 1) * M;
    if (last < std::max(2 * W, 1) || M > mm) { err = "cannot partition line"; return -1; }
    // a line that mm-row chunks do not split evenly: mm-row chunks and one
    // short chunk of the remaining rows, placed second (neither end chunk, which
    // hold the special rows and the ends' halo), when there are at least three
    // chunks and the short one holds both its interface blocks.  Else P chunks
    // of M rows and a shorter last one (the LDS-tiled kernel only).
    const int rem = n - (P - 1) * mm;
    plan.cshort = -1;
    plan.lshort = 0;
    if (n % mm != 0 && P >= 3 && rem >= std::max(2 * W, 1)) {
        plan.cshort = 1;
        plan.lshort = rem;
    }
    plan.N = n; plan.W = W; plan.H = s.H; plan.P = P; plan.periodic = per;
    plan.MM = mm;
    plan.uniform = (M == mm && last == mm);
    plan.NI = 2 * W * P;
    plan.ext = n + 2 * s.H;
    if ((plan.ext & 1) == 0) plan.ext += 1;
    // lines per workgroup of the LDS-tiled kernel: at most 256 lanes and 64 KiB
    auto tile_bytes = [&](int l) {
        return sizeof(double) * ((size_t)l * plan.ext + (size_t)l * std::max(plan.NI, 1));
    };
    int L = 1;
    while (L * 2 * P <= 256 && tile_bytes(2 * L) <= 65536) L *= 2;
    plan.L = L;
    plan.divide = s.divide;
    plan.dd = s.divide ? rq.dd : 1.0;
    plan.int_off = s.int_off;
    plan.int_c = s.interior;
    plan.nbl = (int)s.left.size();
    plan.nbr = (int)s.right.size();
    plan.rows.assign((size_t)(plan.nbl + plan.nbr) * kRowRec, 0.0);
    auto put_row = [&](int idx, const Row &r) {
        double *d = &plan.rows[(size_t)idx * kRowRec];
        d[0] = r.off;
        d[1] = r.step;
        d[2] = (double)r.c.size();
        for (size_t k = 0; k < r.c.size(); ++k) d[3 + k] = r.c[k];
    };
    for (int r = 0; r < plan.nbl; ++r) put_row(r, s.left[r]);
    for (int r = 0; r < plan.nbr; ++r) put_row(plan.nbl + r, s.right[r]);
    // dense form: every nonzero tap must fall inside the K = 2H + 1 points the
    // row's window gives it, and the taps must be summed in increasing or
    // decreasing position (zero taps, kept or dropped, add exact zeros)
    const int K = 2 * s.H + 1;
    plan.drows.assign((size_t)(plan.nbl + plan.nbr) * kDRowRec, 0.0);
    plan.dense_ok = n >= K && K < kDRowRec && plan.nbl <= kSpecRows && plan.nbr <= kSpecRows;
    auto put_dense = [&](int idx, const Row &r, int row, int origin) {
        double *d = &plan.drows[(size_t)idx * kDRowRec];
        d[0] = r.step > 0 ? 1.0 : -1.0;
        for (size_t t = 0; t < r.c.size(); ++t) {
            const int pos = row + r.off + (int)t * r.step - origin;
            if (r.c[t] == 0.0) continue;
            if (pos < 0 || pos >= K || (r.c.size() > 1 && r.step == 0)) { plan.dense_ok = false; return; }
            d[1 + pos] = r.c[t];
        }
    };
    for (int r = 0; r < plan.nbl; ++r) put_dense(r, s.left[r], r, per ? r - s.H : 0);
    for (int r = 0; r < plan.nbr; ++r) {
        const int row = n - plan.nbr + r;
        put_dense(plan.nbl + r, s.right[r], row, per ? row - s.H : n - K);
    }
    plan.cstart.resize(P);
    plan.clen.resize(P);
    for (int c = 0, cs = 0; c < P; ++c) {
        plan.cstart[c] = plan.cshort >= 0 ? cs : c * M;
        plan.clen[c] = plan.cshort >= 0 ? (c == plan.cshort ? plan.lshort : mm) : (c == P - 1) ? last : M;
        cs += plan.clen[c];
    }
    plan.lu.assign((size_t)P * mm * kLuRec, 0.0);
    plan.sp.assign((size_t)P * mm * std::max(2 * W, 1), 0.0);
    plan.G.assign((size_t)P * std::max(2 * W, 1) * std::max(plan.NI, 1), 0.0);
    for (int c = 0; c < P; ++c)
        for (int i = 0; i < mm; ++i) plan.lu[((size_t)c * mm + i) * kLuRec + 4] = 1.0;
    if (W == 0) return 0;

    const int NI = plan.NI, W2 = 2 * W;
    // interface index of chunk c, slot k (k < W: row k; k >= W: row len-2W+k)
    auto vidx = [&](int c, int k) { return c * W2 + k; };
    auto vrow = [&](int c, int k) { return k < W ? k : plan.clen[c] - W2 + k; };
    std::vector<std::vector<ld>> S(P);  // spikes, clen x 2W
    for (int c = 0; c < P; ++c) {
        const int cs = plan.cstart[c], cm = plan.clen[c];
        Mat D(cm);
        for (int i = 0; i < cm; ++i)
            for (int j = 0; j < cm; ++j) D(i, j) = s.A(cs + i, cs + j);
        std::vector<ld> l1, l2, u1, u2, dd;
        ld minpiv;
        if (!lu_band(D, W, l1, l2, u1, u2, dd, minpiv)) {
            err = "zero pivot in chunk factorisation";
            return -4;
        }
        for (int i = 0; i < cm; ++i) {
            double *e = &plan.lu[((size_t)c * mm + i) * kLuRec];
            e[0] = (double)l1[i]; e[1] = (double)l2[i];
            e[2] = (double)u1[i]; e[3] = (double)u2[i];
            e[4] = (double)(1.0L / dd[i]);
        }
        // coupling columns
        const int cl = (c - 1 + P) % P, cr = (c + 1) % P;
        const bool hasL = per || c > 0, hasR = per || c < P - 1;
        std::vector<ld> Cm((size_t)cm * W2, 0.0L);
        for (int i = 0; i < cm; ++i) {
            const int g = cs + i;
            for (int j = 0; j < n; ++j) {
                if (j >= cs && j < cs + cm) continue;
                ld v = s.A(g, j);
                if (v == 0) continue;
                bool placed = false;
                if (hasL) {
                    int lstart = plan.cstart[cl] + plan.clen[cl] - W;
                    if (j >= lstart && j < lstart + W) { Cm[(size_t)i * W2 + (j - lstart)] += v; placed = true; }
                }
                if (!placed && hasR) {
                    int rstart = plan.cstart[cr];
                    if (j >= rstart && j < rstart + W) { Cm[(size_t)i * W2 + W + (j - rstart)] += v; placed = true; }
                }
                if (!placed) { err = "coupling outside interface (chunk too short)"; return -1; }
            }
        }
        S[c].assign((size_t)cm * W2, 0.0L);
        for (int k = 0; k < W2; ++k) {
            std::vector<ld> col(cm);
            for (int i = 0; i < cm; ++i) col[i] = Cm[(size_t)i * W2 + k];
            lu_solve(W, l1, l2, u1, u2, dd, col);
            for (int i = 0; i < cm; ++i) {
                S[c][(size_t)i * W2 + k] = col[i];
                plan.sp[((size_t)c * mm + i) * W2 + k] = (double)col[i];
            }
        }
    }
    // (ragged lines: the full chunks 2 .. P - 2, after the short chunk 1)
    const int c1 = plan.cshort >= 0 ? 2 : 1;
    if ((plan.uniform || plan.cshort == 1) && P >= c1 + 2) {
        const size_t nlu = (size_t)mm * kLuRec, nsp = (size_t)mm * W2;
        plan.same_interior = true;
        for (int c = c1 + 1; c < P - 1 && plan.same_interior; ++c)
            plan.same_interior = std::memcmp(&plan.lu[c * nlu], &plan.lu[c1 * nlu], nlu * sizeof(double)) == 0 &&
                                 std::memcmp(&plan.sp[c * nsp], &plan.sp[c1 * nsp], nsp * sizeof(double)) == 0;
    }
    // interface matrix R t = yhat
    Mat R(NI);
    auto nb = [&](int c, int k, int &cn, int &kn) -> bool {  // owner of s_c[k]
        if (k < W) { if (!per && c == 0) return false; cn = (c - 1 + P) % P; kn = W + k; }
        else { if (!per && c == P - 1) return false; cn = (c + 1) % P; kn = k - W; }
        return true;
    };
    for (int c = 0; c < P; ++c)
        for (int k = 0; k < W2; ++k) {
            const int v = vidx(c, k), row = vrow(c, k);
            R(v, v) += 1.0L;
            for (int j = 0; j < W2; ++j) {
                int cn, kn;
                if (!nb(c, j, cn, kn)) continue;
                R(v, vidx(cn, kn)) += S[c][(size_t)row * W2 + j];
            }
        }
    if (!invert(R)) { err = "singular interface system"; return -4; }
    for (int c = 0; c < P; ++c)
        for (int j = 0; j < W2; ++j) {
            int cn, kn;
            if (!nb(c, j, cn, kn)) continue;
            const int v = vidx(cn, kn);
            for (int u = 0; u < NI; ++u)
                plan.G[((size_t)c * W2 + j) * NI + u] = (double)R(v, u);
        }
    // evenly chunked periodic lines: the interface system is block circulant, so
    // chunk c's rows are chunk 0's rotated by c chunks; store them exactly so
    // (kernels may read chunk 0's rows with a rotated index)
    if (per && plan.uniform)
        for (int c = 1; c < P; ++c)
            for (int j = 0; j < W2; ++j)
                for (int u = 0; u < NI; ++u)
                    plan.G[((size_t)c * W2 + j) * NI + u] = plan.G[(size_t)j * NI + (u - c * W2 + NI) % NI];
    // interface band: the interface matrix is the inverse of a diagonally
    // dominant banded operator, so G_c decays geometrically away from chunk c
    // (Penta10 by about 1e-4 per 16 rows); blocks below 1e-17 of G_c's largest
    // entry contribute less than the rounding of the kept sum
    plan.gb = 0;
    for (int c = 0; c < P; ++c) {
        double gmax = 0.0;
        for (size_t e = (size_t)c * W2 * NI; e < (size_t)(c + 1) * W2 * NI; ++e)
            gmax = std::max(gmax, std::fabs(plan.G[e]));
        for (int b = 0; b < P; ++b) {
            double bmax = 0.0;
            for (int j = 0; j < W2; ++j)
                for (int k = 0; k < W2; ++k)
                    bmax = std::max(bmax, std::fabs(plan.G[((size_t)c * W2 + j) * NI + b * W2 + k]));
            if (bmax <= 1e-17 * gmax) continue;
            int dist = std::abs(b - c);
            if (per) dist = std::min(dist, P - dist);
            plan.gb = std::max(plan.gb, dist);
        }
    }
    return 0;
}

}  // namespace mc3d
