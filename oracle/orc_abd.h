// Oracle: affine-body (ABD) terms of the Newton step.
//
// TEST INFRASTRUCTURE ONLY. Nothing under oracle/ is part of the product; it is the checker
// that tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use.
//
// An affine body has 12 DOFs q = (p, a0, a1, a2): a vertex with body-frame position X sits at
// x = p + X.x a0 + X.y a1 + X.z a2 (reference Solver/SimulationSolver/newton_solver.cpp:505-526,
// fn_apply_dx_affine_bodies), i.e. its Jacobian weights over the four 3-DOF blocks are (1, X).
//   mass properties   Solver/Initializer/init_sim_data.cpp:51-188 (compute_trimesh_dyadic_mass),
//                     591-704 (shell: vertex masses; solid: surface integrals; 12x12 mass matrix)
//   inertia           newton_solver.cpp:1176-1245 (fn_evaluate_abd_inertia)
//   orthogonality     newton_solver.cpp:1246-1335 (fn_evaluate_abd_orthogonality)
//   ground contact    newton_solver.cpp:1336-1413 (fn_evaluate_abd_ground_collision)
// Hessian blocks of one body are kept as 16 3x3 blocks in the reference's order: the four
// diagonal blocks, then (ii, jj != ii) row-major (newton_solver.cpp:1220-1240).
#pragma once
#include <vector>
#include "orc_collision.h"
#include "orc_math.h"

namespace orc
{

constexpr float kAbdEpsilon = 1.19209290E-7f;  // Solver/Core/constant_value.h:12

// compute_trimesh_dyadic_mass: mass, first and second moments of a closed triangle surface by the
// divergence theorem, in float as the reference.
inline void dyadic_mass(const std::vector<V3<float>>& P, const uint32_t* tris, size_t nt, float rho, float& m,
                        float mx[3], float mxx[3][3])
{
    m = 0.0f;
    for (int a = 0; a < 3; a++)
    {
        mx[a] = 0.0f;
        for (int b = 0; b < 3; b++) mxx[a][b] = 0.0f;
    }
    for (size_t f = 0; f < nt; f++)
    {
        const V3<float> p0 = P[tris[3 * f]], p1 = P[tris[3 * f + 1]], p2 = P[tris[3 * f + 2]];
        const V3<float> N  = cross(p1 - p0, p2 - p0);
        m += rho * dot(p0, N) / 6.0;
        for (int a = 0; a < 3; a++)
        {
            float V = 0.0;
            V += p0[a] * p0[a] / 12;
            V += p0[a] * p1[a] / 12;
            V += p0[a] * p2[a] / 12;
            V += p1[a] * p1[a] / 12;
            V += p1[a] * p2[a] / 12;
            V += p2[a] * p2[a] / 12;
            mx[a] += rho / 2 * N[a] * V;
        }
        for (int a = 0; a < 3; a++)
        {
            float V = 0.0;
            float p0a_2 = p0[a] * p0[a], p1a_2 = p1[a] * p1[a], p2a_2 = p2[a] * p2[a];
            float p0a_3 = p0a_2 * p0[a], p1a_3 = p1a_2 * p1[a], p2a_3 = p2a_2 * p2[a];
            V += p0a_3 / 20;
            V += p0a_2 * p1[a] / 20;
            V += p0a_2 * p2[a] / 20;
            V += p0[a] * p1a_2 / 20;
            V += p0[a] * p1[a] * p2[a] / 20;
            V += p0[a] * p2a_2 / 20;
            V += p1a_3 / 20;
            V += p1a_2 * p2[a] / 20;
            V += p1[a] * p2a_2 / 20;
            V += p2a_3 / 20;
            mxx[a][a] += rho / 3 * N[a] * V;
        }
        auto Q = [&](int a, int b) {
            float V = 0.0;
            float p0a_2 = p0[a] * p0[a], p1a_2 = p1[a] * p1[a], p2a_2 = p2[a] * p2[a];
            V += p0a_2 * p0[b] / 20;
            V += p0a_2 * p1[b] / 60;
            V += p0a_2 * p2[b] / 60;
            V += p0[a] * p0[b] * p1[a] / 30;
            V += p0[a] * p0[b] * p2[a] / 30;
            V += p0[a] * p1[a] * p1[b] / 30;
            V += p0[a] * p1[a] * p2[b] / 60;
            V += p0[a] * p1[b] * p2[a] / 60;
            V += p0[a] * p2[a] * p2[b] / 30;
            V += p0[b] * p1a_2 / 60;
            V += p0[b] * p1[a] * p2[a] / 60;
            V += p0[b] * p2a_2 / 60;
            V += p1a_2 * p1[b] / 20;
            V += p1a_2 * p2[b] / 60;
            V += p1[a] * p1[b] * p2[a] / 30;
            V += p1[a] * p2[a] * p2[b] / 30;
            V += p1[b] * p2a_2 / 60;
            V += p2a_2 * p2[b] / 20;
            return rho / 2 * N[a] * V;
        };
        mxx[0][1] += Q(0, 1);
        mxx[0][2] += Q(0, 2);
        mxx[1][2] += Q(1, 2);
        mxx[1][0] = mxx[0][1];
        mxx[2][0] = mxx[0][2];
        mxx[2][1] = mxx[1][2];
    }
}

// The 4x4 scalar mass matrix of the 12x12 one (each entry times I3), diagonal clamped to
// Epsilon (init_sim_data.cpp:681-701).
inline void body_mass_matrix(float m, const float mx[3], const float mxx[3][3], float M[4][4])
{
    M[0][0] = m;
    for (int i = 0; i < 3; i++)
    {
        M[1 + i][0] = mx[i];
        M[0][1 + i] = mx[i];
        for (int j = 0; j < 3; j++) M[1 + i][1 + j] = mxx[i][j];
    }
    for (int i = 0; i < 4; i++) M[i][i] = std::max(M[i][i], kAbdEpsilon);
}

// Body terms accumulated as the reference does into gradients (grad of the energy, 4 blocks)
// and Hessian blocks (16, the order above).
template <class T>
struct BodyGH
{
    V3<T> g[4];
    M3<T> H[16];
};

// Index of block (ii, jj) in the 16-block order.
inline int abd_block(int ii, int jj)
{
    if (ii == jj) return ii;
    return 4 + 3 * ii + (jj < ii ? jj : jj - 1);
}

// fn_evaluate_abd_inertia: alpha h^-2 M (q - q_tilde), alpha = 1 + k_D for a fixed body.
template <class T>
inline void abd_inertia(const V3<T> q[4], const V3<T> qt[4], const T M[4][4], T h2inv, bool fixed, T kD, BodyGH<T>& o)
{
    V3<T> dq[4] = {q[0] - qt[0], q[1] - qt[1], q[2] - qt[2], q[3] - qt[3]};
    V3<T> g[4];
    for (int ii = 0; ii < 4; ii++)
        for (int jj = 0; jj < 4; jj++) g[ii] = g[ii] + M[ii][jj] * dq[jj];
    T alpha = fixed ? (T(1) + kD) : T(1);
    for (int k = 0; k < 4; k++) o.g[k] = (alpha * h2inv) * g[k];
    for (int k = 0; k < 4; k++) o.H[k] = M3<T>::diag((alpha * h2inv) * M[k][k]);
    int idx = 4;
    for (int ii = 0; ii < 4; ii++)
        for (int jj = 0; jj < 4; jj++)
            if (ii != jj) o.H[idx++] = M3<T>::diag((alpha * h2inv) * M[ii][jj]);
}

// fn_evaluate_abd_orthogonality: E = kappa V sum_ij (a_i . a_j - delta_ij)^2, its gradient
// 4 kappa V (sum_j (a_i . a_j) a_j - a_i) and exact Hessian blocks, added to o.
template <class T>
inline void abd_orthogonality(const V3<T> q[4], T stiff, BodyGH<T>& o)
{
    const V3<T> A[3] = {q[1], q[2], q[3]};
    for (int ii = 0; ii < 3; ii++)
    {
        V3<T> grad = T(-1) * A[ii];
        for (int jj = 0; jj < 3; jj++) grad = grad + dot(A[ii], A[jj]) * A[jj];
        o.g[1 + ii] = o.g[1 + ii] + (T(4) * stiff) * grad;
    }
    M3<T> oh[6];
    int   idx = 0;
    for (int ii = 0; ii < 3; ii++)
        for (int jj = ii; jj < 3; jj++)
        {
            M3<T> h;
            if (ii == jj)
            {
                M3<T> qiqiT = outer(A[ii], A[ii]);
                M3<T> qiTqi = M3<T>::diag(dot(A[ii], A[ii]) - T(1));
                h           = qiqiT + qiTqi;
                for (int kk = 0; kk < 3; kk++) h = h + outer(A[kk], A[kk]);
            }
            else
                h = outer(A[jj], A[ii]) + M3<T>::diag(dot(A[ii], A[jj]));
            oh[idx] = oh[idx] + (T(4) * stiff) * h;
            idx++;
        }
    auto add = [&](int blk, int k, bool tr) {
        M3<T> b = oh[k];
        if (tr)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) b.m[r][c] = oh[k].m[c][r];
        o.H[blk] = o.H[blk] + b;
    };
    add(1, 0, false);
    add(8, 1, false);
    add(9, 2, false);
    add(11, 1, true);
    add(2, 3, false);
    add(12, 4, false);
    add(14, 2, true);
    add(15, 4, true);
    add(3, 5, false);
}

// fn_evaluate_abd_ground_collision for one rigid vertex at x (barrier on y - floor), lifted to
// the body with weights (1, X): o.g[ii] += w_ii k1 n, o.H[ii, jj] += w_ii w_jj k2 n n^T.
template <class T>
inline void abd_ground_vertex(V3<T> x, V3<T> X, T floor_y, T dhat, T thickness, T area, T kappa, BodyGH<T>& o,
                              int quad = 0)
{
    T dist = x.y - floor_y;
    if (!(dist - thickness < dhat)) return;
    const V3<T> n(T(0), T(1), T(0));
    T           stiff = kappa * area;
    T           k1, k2;
    ground_k(quad, dist, thickness, dhat, stiff, k1, k2);
    V3<T>       grad  = k1 * n;
    M3<T>       hess  = k2 * outer(n, n);
    const T     w[4]  = {T(1), X.x, X.y, X.z};
    int         idx   = 4;
    for (int ii = 0; ii < 4; ii++)
    {
        o.g[ii] = o.g[ii] + w[ii] * grad;
        for (int jj = 0; jj < 4; jj++)
        {
            M3<T> hh = (w[ii] * w[jj]) * hess;
            if (ii == jj) o.H[ii] = o.H[ii] + hh;
            else
            {
                o.H[idx] = o.H[idx] + hh;
                idx++;
            }
        }
    }
}

}  // namespace orc
