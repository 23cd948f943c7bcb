// Oracle: cloth membrane (Baraff-Witkin stretch + shear) and dihedral bending.
//
// TEST INFRASTRUCTURE ONLY (see orc_math.h). Restates on the CPU:
//   Solver/Energy/stretch_energy.h:240-414   stretch/shear gradient and PSD-projected 6x6 Hessians
//   Solver/Energy/stretch_energy.h:166-200   convert_force / convert_hessian (dF/dx mapping to 9x9)
//   Solver/Energy/stretch_energy.h:418-533   convert_prop, get_Dm_inv, compute_gradient_hessian, energy
//   Solver/Energy/bending_energy.h:167-274   dihedral angle, its gradient, compute_d_theta_d_x
//   Solver/SimulationSolver/newton_solver.cpp:1102-1172  bending gradient and Gauss-Newton Hessian
//
// Conventions: F = [x1 - x0, x2 - x0] * Dm^-1 with Dm^-1 stored column-major (d0, d1, d2, d3):
// Fu = e0*d0 + e1*d1, Fv = e0*d2 + e1*d3. The 6-vector f = (Fu; Fv). The 9x9 Hessian block
// (a, b) is the 3x3 matrix d^2E / dx_a dx_b (row-major).
#pragma once
#include "orc_math.h"

namespace orc
{

template <class T>
struct FaceGH
{
    V3<T> g[3];      // dE/dx_a
    T     H[9][9];   // d2E/dx2, rows/cols (vertex a, dim r) = 3a + r
    T     H6[6][6];  // projected Hessian in f-space (before mapping), for tests
};

// stretch_energy.h:264-289 (mu-scaled PSD stretch Hessian) + 290-314 (shear, projected).
template <class T>
inline void bw_hessian6(V3<T> Fu, V3<T> Fv, T mu, T lmd, T H6[6][6])
{
    T Hs[6][6] = {}, Hh[6][6] = {};
    {
        T I5u = dot(Fu, Fu), I5v = dot(Fv, Fv);
        T su = std::sqrt(I5u), sv = std::sqrt(I5v);
        T iu = T(1) / su, iv = T(1) / sv;
        for (int i = 0; i < 3; i++)
        {
            Hs[i][i]         = std::max(T(0), T(1) - iu);
            Hs[i + 3][i + 3] = std::max(T(0), T(1) - iv);
        }
        V3<T> fu = Fu * iu, fv = Fv * iv;
        T     uc = std::min(iu, T(1)), vc = std::min(iv, T(1));
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
            {
                Hs[r][c]         = Hs[r][c] + uc * (fu[r] * fu[c]);
                Hs[r + 3][c + 3] = Hs[r + 3][c + 3] + vc * (fv[r] * fv[c]);
            }
        for (int r = 0; r < 6; r++)
            for (int c = 0; c < 6; c++) Hs[r][c] = mu * Hs[r][c];
    }
    {
        T I6   = dot(Fu, Fv);
        T sgn  = I6 < T(0) ? T(-1) : T(1);
        // H0 swaps the two 3-blocks: (H0 g) = (g_v; g_u)
        T g[6] = {Fv.x, Fv.y, Fv.z, Fu.x, Fu.y, Fu.z};
        T I2   = dot(Fu, Fu) + dot(Fv, Fv);
        T lam0 = T(0.5f) * (I2 + std::sqrt(I2 * I2 + T(12) * I6 * I6));
        T q[6];
        for (int i = 0; i < 6; i++)
        {
            T h0g = I6 * g[(i + 3) % 6];  // (I6 H0) g
            q[i]  = h0g + lam0 * g[i];
        }
        T qn = 0;
        for (int b = 0; b < 2; b++) qn = qn + ((q[3 * b] * q[3 * b] + q[3 * b + 1] * q[3 * b + 1]) + q[3 * b + 2] * q[3 * b + 2]);
        T qlen = std::sqrt(qn);
        for (int i = 0; i < 6; i++) q[i] = q[i] / qlen;
        // T = 0.5 (I + sign * H0); Tq = T q
        T Tm[6][6] = {};
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++)
            {
                T id  = i == j ? T(1) : T(0);
                T h0  = ((i + 3) % 6 == j) ? T(1) : T(0);
                Tm[i][j] = T(0.5f) * (id + sgn * h0);
            }
        T Tq[6];
        for (int i = 0; i < 6; i++)
        {
            T s = 0;
            for (int j = 0; j < 6; j++) s = s + Tm[i][j] * q[j];
            Tq[i] = s;
        }
        T nTq = 0;
        for (int b = 0; b < 2; b++) nTq = nTq + ((Tq[3 * b] * Tq[3 * b] + Tq[3 * b + 1] * Tq[3 * b + 1]) + Tq[3 * b + 2] * Tq[3 * b + 2]);
        T aI6 = std::fabs(I6);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++)
                Hh[i][j] = lmd * (aI6 * (Tm[i][j] - (Tq[i] * Tq[j]) / nTq) + lam0 * (q[i] * q[j]));
    }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) H6[i][j] = Hs[i][j] + Hh[i][j];
}

// stretch_energy.h:440-471 (gradient and Hessian of area * (stretch + shear)).
template <class T>
inline void bw_face(V3<T> x0, V3<T> x1, V3<T> x2, const T Dm[4], T mu, T lmd, T area, FaceGH<T>& out)
{
    V3<T> e0 = x1 - x0, e1 = x2 - x0;
    V3<T> Fu = e0 * Dm[0] + e1 * Dm[1];
    V3<T> Fv = e0 * Dm[2] + e1 * Dm[3];
    // dE/dF
    V3<T> gu, gv;
    {
        T su = std::sqrt(dot(Fu, Fu)), sv = std::sqrt(dot(Fv, Fv));
        T iu = T(1) / su, iv = T(1) / sv;
        V3<T> su_g = mu * (((su - T(1)) * iu) * Fu);
        V3<T> sv_g = mu * (((sv - T(1)) * iv) * Fv);
        T     w    = dot(Fu, Fv);
        V3<T> hu   = lmd * (w * Fv);
        V3<T> hv   = lmd * (w * Fu);
        gu         = su_g + hu;
        gv         = sv_g + hv;
    }
    // dF/dx coefficients per vertex (stretch_energy.h:64-143)
    T cx[3] = {-(Dm[0] + Dm[1]), Dm[0], Dm[1]};
    T cy[3] = {-(Dm[2] + Dm[3]), Dm[2], Dm[3]};
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) out.g[a][d] = area * (cx[a] * gu[d] + cy[a] * gv[d]);
    bw_hessian6(Fu, Fv, mu, lmd, out.H6);
    // convert_hessian sums H6(i, j) D_i D_j^T in loop order; per entry that leaves four terms.
    // Blocks below the block diagonal are taken as transposes of the blocks above it (the
    // reference evaluates them with the two middle terms swapped, a difference of at most one
    // rounding in the last place; the GPU path stores only the upper blocks).
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++)
            for (int r = 0; r < 3; r++)
                for (int s = 0; s < 3; s++)
                {
                    T t1 = out.H6[r][s] * (cx[a] * cx[b]);
                    T t2 = out.H6[r][s + 3] * (cx[a] * cy[b]);
                    T t3 = out.H6[r + 3][s] * (cy[a] * cx[b]);
                    T t4 = out.H6[r + 3][s + 3] * (cy[a] * cy[b]);
                    out.H[3 * a + r][3 * b + s] = area * (((t1 + t2) + t3) + t4);
                }
    for (int a = 1; a < 3; a++)
        for (int b = 0; b < a; b++)
            for (int r = 0; r < 3; r++)
                for (int s = 0; s < 3; s++) out.H[3 * a + r][3 * b + s] = out.H[3 * b + s][3 * a + r];
}

template <class T>
inline T bw_energy(V3<T> x0, V3<T> x1, V3<T> x2, const T Dm[4], T mu, T lmd, T area)
{
    V3<T> e0 = x1 - x0, e1 = x2 - x0;
    V3<T> Fu = e0 * Dm[0] + e1 * Dm[1];
    V3<T> Fv = e0 * Dm[2] + e1 * Dm[3];
    T     a  = std::sqrt(dot(Fu, Fu)) - T(1), b = std::sqrt(dot(Fv, Fv)) - T(1);
    T     i6 = dot(Fu, Fv);
    return area * (T(0.5f) * mu * (a * a + b * b) + T(0.5f) * lmd * (i6 * i6));
}

// stretch_energy.h:424-439
template <class T>
inline void dm_inv(V3<T> x0, V3<T> x1, V3<T> x2, T out[4])
{
    V3<T> r1 = x1 - x0, r2 = x2 - x0;
    V3<T> c  = cross(r1, r2);
    V3<T> a1 = normalize(r1);
    V3<T> a2 = normalize(cross(c, a1));
    T     u0 = dot(a1, x0), v0 = dot(a2, x0), u1 = dot(a1, x1), v1 = dot(a2, x1), u2 = dot(a1, x2), v2 = dot(a2, x2);
    // duv = [duv0, duv1] columns; inverse by adjugate times 1/det
    T m00 = u1 - u0, m10 = v1 - v0, m01 = u2 - u0, m11 = v2 - v0;
    T inv = T(1) / (m00 * m11 - m01 * m10);
    out[0] = m11 * inv;   // column 0, row 0
    out[1] = -m10 * inv;  // column 0, row 1
    out[2] = -m01 * inv;  // column 1, row 0
    out[3] = m00 * inv;   // column 1, row 1
}

inline float sign_pos(float x) { return x < 0.0f ? -1.0f : 1.0f; }
inline double sign_pos(double x) { return x < 0.0 ? -1.0 : 1.0; }

// bending_energy.h:217-231 with the argument order used by compute_theta / compute_d_theta_d_x:
// the hinge (p0, p1 edge; p2, p3 opposite) enters as face_dihedral_angle(p2, p1, p0, p3).
template <class T>
inline T dihedral_angle(V3<T> p0, V3<T> p1, V3<T> p2, V3<T> p3)
{
    V3<T> v0 = p2, v1 = p1, v2 = p0, v3 = p3;
    V3<T> n1 = cross(v1 - v0, v2 - v0);
    V3<T> n2 = cross(v2 - v3, v1 - v3);
    T     d  = dot(n1, n2) / std::sqrt(dot(n1, n1) * dot(n2, n2));
    T     an = std::acos(std::max(T(-1), std::min(T(1), d)));
    T     sg = sign_pos(dot(cross(n2, n1), v1 - v2));
    return sg * an;
}

// bending_energy.h:167-189 + 255-265: gradient of the angle w.r.t. p0..p3.
template <class T>
inline void dihedral_grad(V3<T> p0, V3<T> p1, V3<T> p2, V3<T> p3, V3<T> g[4])
{
    // face_dihedral_angle_grad(v2 = p2, v0 = p1, v1 = p0, v3 = p3)
    V3<T> v0 = p1, v1 = p0, v2 = p2, v3 = p3;
    V3<T> e0 = v1 - v0, e1 = v2 - v0, e2 = v3 - v0, e3 = v2 - v1, e4 = v3 - v1;
    V3<T> n1 = cross(e0, e1), n2 = cross(e2, e0);
    T     n1s = dot(n1, n1), n2s = dot(n2, n2), e0n = length(e0);
    V3<T> g0 = ((-e0n) / n1s) * n1;
    V3<T> g1 = ((-dot(e0, e3)) / (e0n * n1s)) * n1 - (dot(e0, e4) / (e0n * n2s)) * n2;
    V3<T> g2 = (dot(e0, e1) / (e0n * n1s)) * n1 + (dot(e0, e2) / (e0n * n2s)) * n2;
    V3<T> g3 = ((-e0n) / n2s) * n2;
    g[0] = g2;
    g[1] = g1;
    g[2] = g0;
    g[3] = g3;
}

}  // namespace orc
