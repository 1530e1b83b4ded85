// Symmetry classes of the Hex8 element stiffness of an isotropic cube (Fem/Hex8Element.cs:98-178).
//
// The 48 symmetries of the cube (3 axis permutations x reflections) map element DOF pairs
// ((a,i),(b,j)) to each other with a sign s_i s_j from the reflected displacement components, and an
// isotropic Ke is invariant under them. The 576 entries therefore fall into 14 orbits:
// Ke[r][c] = kKeSign[r][c] * v[kKeClass[r][c]], with v[k] = Ke[kKeRep[k][0]][kKeRep[k][1]] for any h and nu.
// The orbits were enumerated by applying the group to every pair; the host checks the identity on the
// Ke it computes before using it.
#pragma once

#include <cmath>

namespace mycel {

constexpr int kKeClasses = 14;

constexpr int kKeRep[kKeClasses][2] = {{0, 0}, {0, 1}, {0, 3}, {0, 4}, {0, 6}, {0, 7}, {0, 8}, {0, 9}, {0, 10}, {0, 11}, {0, 18}, {0, 19}, {0, 21}, {0, 22}};

constexpr signed char kKeClass[24][24] = {
    {0, 1, 1, 2, 3, 3, 4, 5, 6, 7, 8, 9, 7, 9, 8, 4, 6, 5, 10, 11, 11, 12, 13, 13},
    {1, 0, 1, 8, 7, 9, 5, 4, 6, 3, 2, 3, 9, 7, 8, 13, 12, 13, 11, 10, 11, 6, 4, 5},
    {1, 1, 0, 8, 9, 7, 13, 13, 12, 9, 8, 7, 3, 3, 2, 5, 6, 4, 11, 11, 10, 6, 5, 4},
    {2, 3, 3, 0, 1, 1, 7, 8, 9, 4, 5, 6, 4, 6, 5, 7, 9, 8, 12, 13, 13, 10, 11, 11},
    {8, 7, 9, 1, 0, 1, 3, 2, 3, 5, 4, 6, 13, 12, 13, 9, 7, 8, 6, 4, 5, 11, 10, 11},
    {8, 9, 7, 1, 1, 0, 9, 8, 7, 13, 13, 12, 5, 6, 4, 3, 3, 2, 6, 5, 4, 11, 11, 10},
    {4, 5, 6, 7, 8, 9, 0, 1, 1, 2, 3, 3, 10, 11, 11, 12, 13, 13, 7, 9, 8, 4, 6, 5},
    {5, 4, 6, 3, 2, 3, 1, 0, 1, 8, 7, 9, 11, 10, 11, 6, 4, 5, 9, 7, 8, 13, 12, 13},
    {13, 13, 12, 9, 8, 7, 1, 1, 0, 8, 9, 7, 11, 11, 10, 6, 5, 4, 3, 3, 2, 5, 6, 4},
    {7, 8, 9, 4, 5, 6, 2, 3, 3, 0, 1, 1, 12, 13, 13, 10, 11, 11, 4, 6, 5, 7, 9, 8},
    {3, 2, 3, 5, 4, 6, 8, 7, 9, 1, 0, 1, 6, 4, 5, 11, 10, 11, 13, 12, 13, 9, 7, 8},
    {9, 8, 7, 13, 13, 12, 8, 9, 7, 1, 1, 0, 6, 5, 4, 11, 11, 10, 5, 6, 4, 3, 3, 2},
    {7, 9, 8, 4, 6, 5, 10, 11, 11, 12, 13, 13, 0, 1, 1, 2, 3, 3, 4, 5, 6, 7, 8, 9},
    {9, 7, 8, 13, 12, 13, 11, 10, 11, 6, 4, 5, 1, 0, 1, 8, 7, 9, 5, 4, 6, 3, 2, 3},
    {3, 3, 2, 5, 6, 4, 11, 11, 10, 6, 5, 4, 1, 1, 0, 8, 9, 7, 13, 13, 12, 9, 8, 7},
    {4, 6, 5, 7, 9, 8, 12, 13, 13, 10, 11, 11, 2, 3, 3, 0, 1, 1, 7, 8, 9, 4, 5, 6},
    {13, 12, 13, 9, 7, 8, 6, 4, 5, 11, 10, 11, 8, 7, 9, 1, 0, 1, 3, 2, 3, 5, 4, 6},
    {5, 6, 4, 3, 3, 2, 6, 5, 4, 11, 11, 10, 8, 9, 7, 1, 1, 0, 9, 8, 7, 13, 13, 12},
    {10, 11, 11, 12, 13, 13, 7, 9, 8, 4, 6, 5, 4, 5, 6, 7, 8, 9, 0, 1, 1, 2, 3, 3},
    {11, 10, 11, 6, 4, 5, 9, 7, 8, 13, 12, 13, 5, 4, 6, 3, 2, 3, 1, 0, 1, 8, 7, 9},
    {11, 11, 10, 6, 5, 4, 3, 3, 2, 5, 6, 4, 13, 13, 12, 9, 8, 7, 1, 1, 0, 8, 9, 7},
    {12, 13, 13, 10, 11, 11, 4, 6, 5, 7, 9, 8, 7, 8, 9, 4, 5, 6, 2, 3, 3, 0, 1, 1},
    {6, 4, 5, 11, 10, 11, 13, 12, 13, 9, 7, 8, 3, 2, 3, 5, 4, 6, 8, 7, 9, 1, 0, 1},
    {6, 5, 4, 11, 11, 10, 5, 6, 4, 3, 3, 2, 9, 8, 7, 13, 13, 12, 8, 9, 7, 1, 1, 0},
};

constexpr signed char kKeSign[24][24] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1},
    {-1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1},
    {-1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1},
    {1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1},
    {1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1},
    {-1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1},
    {1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1},
    {-1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1},
    {1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1},
    {1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1},
    {1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1},
    {-1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1},
    {1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1},
    {-1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1},
    {1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1},
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1},
    {-1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1},
    {-1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1},
};

// The 14 class values of one element stiffness, passed to kernels by value (they land in SGPRs).
template <typename T>
struct KeSym {
    T v[kKeClasses];
};

// The same stiffness in the monomial basis of the trilinear element. With xi, eta, zeta = +-1 at the corners
// (Fem/Hex8Element.cs:52-54), P[a][m] is the monomial m of {1, xi, eta, zeta, xi eta, xi zeta, eta zeta,
// xi eta zeta} at corner a; P^T P = 8 I, so with H = P on each displacement component,
// Ke = H M H^T / 64 and M = H^T Ke H / 64 ("Mc" below). For the isotropic cube the 24x24 M has 45 non-zeros
// and 8 distinct values:
//   the constant-strain block on (u_xi, v_eta, w_zeta): a on the diagonal, b off it;
//   the shear pairs (u_eta, v_xi), (u_zeta, w_xi), (v_zeta, w_eta): every entry c;
//   the pairs (u_xi.eta, w_eta.zeta), (u_xi.zeta, v_eta.zeta), (v_xi.eta, w_xi.zeta): d diagonal, e off;
//   the block on (u_eta.zeta, v_xi.zeta, w_xi.eta): f diagonal, h off;
//   the trilinear modes u/v/w_xi.eta.zeta: q.
// The constant (translation) modes have no stiffness. Ke u_e is then a Walsh-Hadamard transform of each
// component's 8 corner values, this sparse product and the inverse transform: about 200 operations instead of
// 576 multiply-adds.
template <typename T>
struct KeMono {
    T a, b, c, d, e, f, h, q;
};

// Monomial index of component comp (0..2) and monomial m (0..7) in M.
constexpr int mono(int comp, int m) { return 8 * comp + m; }

// M = H^T Ke H / 64 (row-major 24x24) from a row-major Ke.
inline void ke_monomial(const double *ke, double *M) {
    const int xs[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, ys[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
              zs[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double P[8][8];
    for (int a = 0; a < 8; a++) {
        const int x = xs[a], y = ys[a], z = zs[a];
        const int v[8] = {1, x, y, z, x * y, x * z, y * z, x * y * z};
        for (int m = 0; m < 8; m++) P[a][m] = v[m];
    }
    for (int r = 0; r < 24; r++)
        for (int c = 0; c < 24; c++) {
            const int cr = r / 8, mr = r % 8, cc = c / 8, mc = c % 8;
            double s = 0.0;
            for (int a = 0; a < 8; a++)
                for (int b = 0; b < 8; b++) s += P[a][mr] * ke[(3 * a + cr) * 24 + 3 * b + cc] * P[b][mc];
            M[r * 24 + c] = s / 64.0;
        }
}

// The 8 values of M and the largest deviation of M from the pattern above, relative to max |M|.
inline double ke_mono_values(const double *ke, KeMono<double> &k) {
    double M[576];
    ke_monomial(ke, M);
    auto at = [&](int r, int c) { return M[r * 24 + c]; };
    k.a = at(mono(0, 1), mono(0, 1));
    k.b = at(mono(0, 1), mono(1, 2));
    k.c = at(mono(0, 2), mono(0, 2));
    k.d = at(mono(0, 4), mono(0, 4));
    k.e = at(mono(0, 4), mono(2, 6));
    k.f = at(mono(0, 6), mono(0, 6));
    k.h = at(mono(0, 6), mono(1, 5));
    k.q = at(mono(0, 7), mono(0, 7));
    double P[576] = {};
    auto set = [&](int r, int c, double v) { P[r * 24 + c] = v; };
    const int lin[3] = {mono(0, 1), mono(1, 2), mono(2, 3)};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) set(lin[i], lin[j], i == j ? k.a : k.b);
    const int shear[3][2] = {{mono(0, 2), mono(1, 1)}, {mono(0, 3), mono(2, 1)}, {mono(1, 3), mono(2, 2)}};
    for (auto &p : shear)
        for (int i : p)
            for (int j : p) set(i, j, k.c);
    const int bil[3][2] = {{mono(0, 4), mono(2, 6)}, {mono(0, 5), mono(1, 6)}, {mono(1, 4), mono(2, 5)}};
    for (auto &p : bil)
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) set(p[i], p[j], i == j ? k.d : k.e);
    const int tw[3] = {mono(0, 6), mono(1, 5), mono(2, 4)};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) set(tw[i], tw[j], i == j ? k.f : k.h);
    for (int c = 0; c < 3; c++) set(mono(c, 7), mono(c, 7), k.q);
    double mx = 0.0, dev = 0.0;
    for (int i = 0; i < 576; i++) {
        mx = std::fabs(M[i]) > mx ? std::fabs(M[i]) : mx;
        const double dv = std::fabs(M[i] - P[i]);
        dev = dv > dev ? dv : dev;
    }
    return mx > 0.0 ? dev / mx : 0.0;
}

}  // namespace mycel
