// Oracle (test infrastructure only): metric terms from mesh coordinates,
// restating AbstractSingleBlockMesh::solveForJacobians
// (src/AbstractSingleBlockMesh.cpp:23-1439) and the translation set-up of
// handlePeriodicStuff (:2019-2085) on the x-fastest layout.
//
// The reference forms the metrics in conservative form: derivatives of x and
// y (:40-281), products V = xE * {y, z} (:298-310), their cross derivatives
// (:348-1001), J_ij / J as differences of those (:1011-1023), the Jacobian
// from the divergence of x J_i (:1057-1298), and J_ij *= J.  Periodic
// directions use the transformed-periodic derivative with wrap values shifted
// by the direction's translation vector.  Kept as the reference has them
// (T8): with the 7-point stencil the zeta derivative of y passes Nm2 in the
// Nm3 slot (:273) and the xi derivative of Vc3 passes Nm2c3 in the Nm3 slot
// (:706).
#include <cstdio>
#include <vector>

#include "oracle.h"

namespace {

using Vec = std::vector<double>;

struct Geo {
    int N[3];
    long n() const { return (long)N[0] * N[1] * N[2]; }
    int nl(int d) const { return N[d]; }
    int nlines(int d) const { return (int)(n() / N[d]); }
    long stride(int d) const { return d == 0 ? 1 : d == 1 ? N[0] : (long)N[0] * N[1]; }
    // line q of direction d, lower remaining index fastest (x-lines q = k*Ny + j, ...)
    long base(int d, int q) const {
        if (d == 0) return (long)q * N[0];
        if (d == 1) return (long)(q / N[0]) * N[0] * N[1] + q % N[0];
        return q;
    }
};

struct Ctx {
    Geo g;
    int fd[3], bc[3];
    double dd[3];
    double T[3][3];  // translation of each periodic direction
    int err = 0;

    // d/dxi_d of f; periodic directions with the wrap arrays w (6 slots)
    void deriv(int d, const Vec &f, Vec &out, const Vec *w) {
        if (err) return;
        if (bc[d] == ORC_PERIODIC) {
            const double *wp[6];
            for (int k = 0; k < 6; ++k) wp[k] = w[k].data();
            err = orc_deriv_field_tpb(fd[d], d, g.N[0], g.N[1], g.N[2], dd[d], f.data(), out.data(), wp);
        } else {
            err = orc_deriv_field(fd[d], d, bc[d], g.N[0], g.N[1], g.N[2], dd[d], f.data(), out.data());
        }
    }

    // wrap slots {Nm3, Nm2, Nm1, Np1, Np2, Np3}: value(point, sign) at the
    // line's points N-3, N-2, N-1 (sign -1) and 0, 1, 2 (sign +1)
    template <class F>
    void wraps(int d, Vec *w, F value) {
        const int n = g.nl(d);
        const long s = g.stride(d);
        for (int k = 0; k < 6; ++k) w[k].assign(g.nlines(d), 0.0);
        for (int q = 0; q < g.nlines(d); ++q) {
            const long b = g.base(d, q);
            for (int m = 0; m < 3; ++m) {
                w[m][q] = value(b + (n - 3 + m) * s, -1.0);
                w[3 + m][q] = value(b + m * s, +1.0);
            }
        }
    }
};

}  // namespace

extern "C" int orc_metrics(const int fd[3], const int bc[3], const int N[3], const double trans[3][3],
                           const double *x, const double *y, const double *z, double *const out[10]) {
    Ctx c;
    for (int d = 0; d < 3; ++d) {
        c.g.N[d] = N[d];
        c.fd[d] = fd[d];
        c.bc[d] = bc[d];
        c.dd[d] = bc[d] == ORC_PERIODIC ? 1.0 / N[d] : 1.0 / (N[d] - 1);
        for (int e = 0; e < 3; ++e) c.T[d][e] = trans[d][e];
        // the metric path knows only the 5- and 7-point stencils (:94-97)
        if (bc[d] == ORC_PERIODIC && fd[d] != ORC_PADE6 && fd[d] != ORC_PENTA10) return -1;
    }
    const long n = c.g.n();
    const Vec X(x, x + n), Y(y, y + n), Z(z, z + n);
    Vec w[6];
    auto bw7 = [&](int d) { return c.fd[d] == ORC_PENTA10; };
    // the Nm3 slot takes the Nm2 values (the two 7-point call-site typos)
    auto nm3_from_nm2 = [&](int d) {
        if (bw7(d)) w[0] = w[1];
    };

    // ---- derivatives of x and y, :40-281 ----
    Vec xE[2][3];
    for (auto &r : xE)
        for (auto &v : r) v.assign(n, 0.0);
    const Vec *coord[2] = {&X, &Y};
    for (int d : {1, 0, 2}) {  // eta, xi, zeta as the reference orders them
        for (int i = 0; i < 2; ++i) {
            const Vec &f = *coord[i];
            if (c.bc[d] == ORC_PERIODIC) {
                const double t = c.T[d][i];
                c.wraps(d, w, [&](long p, double s) { return s < 0 ? f[p] - t : f[p] + t; });
                if (d == 2 && i == 1) nm3_from_nm2(d);  // :273
            }
            c.deriv(d, f, xE[i][d], w);
        }
    }
    // ---- V products, :298-310 ----
    Vec Va[3], Vb[3], Vc[3];
    for (int j = 0; j < 3; ++j) {
        Va[j].resize(n); Vb[j].resize(n); Vc[j].resize(n);
        for (long p = 0; p < n; ++p) {
            Va[j][p] = xE[0][j][p] * Y[p];
            Vb[j][p] = xE[0][j][p] * Z[p];
            Vc[j][p] = xE[1][j][p] * Z[p];
        }
    }
    // ---- cross derivatives dV[j][k] = d V_j / d xi_k, :348-1001 ----
    // family a: coordinate y with xE1j; b: z with xE1j; c: z with xE2j
    Vec dVa[3][3], dVb[3][3], dVc[3][3];
    struct Fam { Vec *V; Vec (*dV)[3]; const Vec *coord; int comp; int row; };
    Fam fams[3] = {{Va, dVa, &Y, 1, 0}, {Vb, dVb, &Z, 2, 0}, {Vc, dVc, &Z, 2, 1}};
    for (int k : {1, 0, 2}) {  // eta, xi, zeta as the reference orders them
        for (Fam &f : fams) {
            for (int j = 0; j < 3; ++j) {
                if (j == k) continue;
                Vec &o = f.dV[j][k];
                o.assign(n, 0.0);
                if (c.bc[k] == ORC_PERIODIC) {
                    const double t = c.T[k][f.comp];
                    const Vec &co = *f.coord, &e = xE[f.row][j];
                    c.wraps(k, w, [&](long p, double s) { return (s < 0 ? co[p] - t : co[p] + t) * e[p]; });
                    if (k == 0 && f.V == Vc && j == 2) nm3_from_nm2(k);  // :706
                }
                c.deriv(k, f.V[j], o, w);
            }
        }
    }
    // ---- J_ij / J, :1011-1023 ----
    Vec J(n), Jm[3][3];
    for (auto &r : Jm)
        for (auto &v : r) v.resize(n);
    for (long p = 0; p < n; ++p) {
        Jm[0][0][p] = dVc[1][2][p] - dVc[2][1][p];
        Jm[0][1][p] = dVb[2][1][p] - dVb[1][2][p];
        Jm[0][2][p] = dVa[1][2][p] - dVa[2][1][p];
        Jm[1][0][p] = dVc[2][0][p] - dVc[0][2][p];
        Jm[1][1][p] = dVb[0][2][p] - dVb[2][0][p];
        Jm[1][2][p] = dVa[2][0][p] - dVa[0][2][p];
        Jm[2][0][p] = dVc[0][1][p] - dVc[1][0][p];
        Jm[2][1][p] = dVb[1][0][p] - dVb[0][1][p];
        Jm[2][2][p] = dVa[0][1][p] - dVa[1][0][p];
    }
    // ---- Jacobian determinant, :1057-1298 ----
    Vec pre[3], Jdet[3];
    for (int k = 0; k < 3; ++k) {
        pre[k].resize(n);
        Jdet[k].assign(n, 0.0);
        for (long p = 0; p < n; ++p)
            pre[k][p] = X[p] * Jm[k][0][p] + Y[p] * Jm[k][1][p] + Z[p] * Jm[k][2][p];
    }
    for (int k : {1, 0, 2}) {
        if (c.bc[k] == ORC_PERIODIC) {
            const double *t = c.T[k];
            c.wraps(k, w, [&](long p, double s) {
                return s < 0 ? (X[p] - t[0]) * Jm[k][0][p] + (Y[p] - t[1]) * Jm[k][1][p] + (Z[p] - t[2]) * Jm[k][2][p]
                             : (X[p] + t[0]) * Jm[k][0][p] + (Y[p] + t[1]) * Jm[k][1][p] + (Z[p] + t[2]) * Jm[k][2][p];
            });
        }
        c.deriv(k, pre[k], Jdet[k], w);
    }
    if (c.err) return c.err;
    for (long p = 0; p < n; ++p) {
        J[p] = 1 / ((1.0 / 3.0) * (Jdet[0][p] + Jdet[1][p] + Jdet[2][p]));
        out[0][p] = J[p];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) out[1 + 3 * a + b][p] = J[p] * Jm[a][b][p];
    }
    return 0;
}

// handlePeriodicStuff (:2019-2085): the translation of a periodic direction is
// its PERIODICDISP row when both faces of the direction are PERIODIC, zero for
// INTERNALLY_PERIODIC faces.  The x check reads the y faces (T8, :2034);
// face kinds: 0 PERIODIC, 1 INTERNALLY_PERIODIC, other values anything else.
// Returns, per direction, 1 when a translation was set, 0 when the faces were
// incompatible (the reference leaves the translation unset then).
extern "C" void orc_periodic_translations(const int bc[3], const int face_lo[3], const int face_hi[3],
                                          const double disp[3][3], double trans[3][3], int set[3]) {
    for (int d = 0; d < 3; ++d) {
        set[d] = 1;
        for (int e = 0; e < 3; ++e) trans[d][e] = 0.0;
        if (bc[d] != ORC_PERIODIC) continue;
        const int checked = d == 0 ? 1 : d;  // x looks at the y faces
        if (face_lo[checked] == 0 && face_hi[checked] == 0) {
            for (int e = 0; e < 3; ++e) trans[d][e] = disp[d][e];
        } else if (face_lo[d] != 1 || face_hi[d] != 1) {
            set[d] = 0;
        }
    }
}
