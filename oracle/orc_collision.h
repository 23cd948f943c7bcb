// Oracle: closest-point barycentrics, ACCD time of impact, IPC log barrier.
//
// TEST INFRASTRUCTURE ONLY (see orc_math.h). Restates on the CPU:
//   Solver/CollisionDetector/distance.hpp:40-56    2x2 solve by adjugate/determinant
//   Solver/CollisionDetector/distance.hpp:58-182   point-edge / point-triangle coefficients (unclassified)
//   Solver/CollisionDetector/distance.hpp:255-319  edge-edge coefficients with parallel/clamped fallback
//   Solver/CollisionDetector/accd.hpp:31-205       additive CCD (ppf-contact-solver), line_search_max_t = 1.25
//   Solver/CollisionDetector/cipc_kernel.hpp:12-38 b(d) = -(d-dhat)^2 ln(d/dhat) and its two derivatives
#pragma once
#include <cfloat>
#include "orc_math.h"

namespace orc
{

constexpr float kEpsilon        = 1.19209290E-7f;  // Core/constant_value.h:12
constexpr float kLineSearchMaxT = 1.25f;           // accd.hpp:47
constexpr float kCcdReduction   = 0.01f;           // accd.hpp:46

// distance.hpp:40-56. a is symmetric 2x2 given as (a00, a01, a11); returns false if det == 0.
template <class T>
inline bool solve2(T a00, T a01, T a10, T a11, T b0, T b1, T& x0, T& x1)
{
    T det = a00 * a11 - a10 * a01;
    if (det != T(0))
    {
        x0 = (a11 / det) * b0 + (-a01 / det) * b1;
        x1 = (-a10 / det) * b0 + (a00 / det) * b1;
        return true;
    }
    return false;
}

template <class T>
inline void pe_coeff(V3<T> p, V3<T> e0, V3<T> e1, T& c0, T& c1)
{
    V3<T> r = e1 - e0;
    T     d = dot(r, r);
    c0 = T(0.5f);
    c1 = T(0.5f);
    if (d > T(kEpsilon))
    {
        T x = dot(r, p - e0) / d;
        c0  = T(1) - x;
        c1  = x;
    }
}

template <class T>
inline V3<T> pt_coeff(V3<T> p, V3<T> t0, V3<T> t1, V3<T> t2)
{
    V3<T> r0 = t1 - t0, r1 = t2 - t0, w = p - t0;
    T     a00 = dot(r0, r0), a01 = dot(r1, r0), a10 = dot(r0, r1), a11 = dot(r1, r1);
    T     c0, c1;
    if (!solve2(a00, a01, a10, a11, dot(r0, w), dot(r1, w), c0, c1))
    {
        c0 = T(1.0f / 3.0f);
        c1 = T(1.0f / 3.0f);
    }
    return V3<T>((T(1) - c0) - c1, c0, c1);
}

// distance.hpp:109-182
template <class T>
inline V3<T> pt_coeff_unclassified(V3<T> p, V3<T> t0, V3<T> t1, V3<T> t2)
{
    V3<T> c = pt_coeff(p, t0, t1, t2);
    if (c.x >= T(0) && c.y >= T(0) && c.z >= T(0) && c.x <= T(1) && c.y <= T(1) && c.z <= T(1)) return c;
    T e0, e1;
    if (c.x < T(0))
    {
        pe_coeff(p, t1, t2, e0, e1);
        if (e0 >= T(0) && e0 <= T(1)) return V3<T>(T(0), e0, e1);
        return e0 > T(1) ? V3<T>(T(0), T(1), T(0)) : V3<T>(T(0), T(0), T(1));
    }
    else if (c.y < T(0))
    {
        pe_coeff(p, t0, t2, e0, e1);
        if (e0 >= T(0) && e0 <= T(1)) return V3<T>(e0, T(0), e1);
        return e0 > T(1) ? V3<T>(T(1), T(0), T(0)) : V3<T>(T(0), T(0), T(1));
    }
    else
    {
        pe_coeff(p, t0, t1, e0, e1);
        if (e0 >= T(0) && e0 <= T(1)) return V3<T>(e0, e1, T(0));
        return e0 > T(1) ? V3<T>(T(1), T(0), T(0)) : V3<T>(T(0), T(1), T(0));
    }
}

template <class T>
inline T pt_dist2(V3<T> p, V3<T> t0, V3<T> t1, V3<T> t2)
{
    V3<T> c = pt_coeff_unclassified(p, t0, t1, t2);
    V3<T> x = (c.x * (t0 - p) + c.y * (t1 - p)) + c.z * (t2 - p);
    return dot(x, x);
}

template <class T>
struct C4
{
    T v[4];
};

template <class T>
inline C4<T> ee_coeff(V3<T> ea0, V3<T> ea1, V3<T> eb0, V3<T> eb1)
{
    V3<T> r0 = ea1 - ea0, r1 = eb1 - eb0, nr1 = -r1, w = eb0 - ea0;
    T     a00 = dot(r0, r0), a01 = dot(nr1, r0), a10 = dot(r0, nr1), a11 = dot(nr1, nr1);
    T     x0, x1;
    C4<T> b = {{T(0.5f), T(0.5f), T(0.5f), T(0.5f)}};
    if (solve2(a00, a01, a10, a11, dot(r0, w), dot(nr1, w), x0, x1)) b = {{T(1) - x0, x0, T(1) - x1, x1}};
    return b;
}

template <class T>
inline C4<T> ee_coeff_unclassified(V3<T> ea0, V3<T> ea1, V3<T> eb0, V3<T> eb1)
{
    V3<T> r0 = ea1 - ea0, r1 = eb1 - eb0;
    T     len0 = dot(r0, r0), len1 = dot(r1, r1);
    V3<T> cr   = cross(r0, r1);
    T     par  = dot(cr, cr) / (len0 * len1 + T(1e-8f));
    C4<T> c    = ee_coeff(ea0, ea1, eb0, eb1);
    bool  valid = true;
    for (int i = 0; i < 4; i++) valid = valid && c.v[i] >= T(0) && c.v[i] <= T(1);
    if (par < T(1e-4f) || !valid)
    {
        T cand[4][4];
        T a, b;
        pe_coeff(ea0, eb0, eb1, a, b);
        cand[0][0] = 1; cand[0][1] = 0; cand[0][2] = clampv<T>(a, 0, 1); cand[0][3] = clampv<T>(b, 0, 1);
        pe_coeff(ea1, eb0, eb1, a, b);
        cand[1][0] = 0; cand[1][1] = 1; cand[1][2] = clampv<T>(a, 0, 1); cand[1][3] = clampv<T>(b, 0, 1);
        pe_coeff(eb0, ea0, ea1, a, b);
        cand[2][0] = clampv<T>(a, 0, 1); cand[2][1] = clampv<T>(b, 0, 1); cand[2][2] = 1; cand[2][3] = 0;
        pe_coeff(eb1, ea0, ea1, a, b);
        cand[3][0] = clampv<T>(a, 0, 1); cand[3][1] = clampv<T>(b, 0, 1); cand[3][2] = 0; cand[3][3] = 1;
        int best  = 0;
        T   min_d = T(FLT_MAX);
        for (int i = 0; i < 4; i++)
        {
            V3<T> x0 = cand[i][0] * ea0 + cand[i][1] * ea1;
            V3<T> x1 = cand[i][2] * eb0 + cand[i][3] * eb1;
            V3<T> d  = x0 - x1;
            T     dd = dot(d, d);
            if (dd < min_d)
            {
                min_d = dd;
                best  = i;
            }
        }
        for (int k = 0; k < 4; k++) c.v[k] = cand[best][k];
    }
    return c;
}

template <class T>
inline T ee_dist2(V3<T> ea0, V3<T> ea1, V3<T> eb0, V3<T> eb1)
{
    C4<T> c  = ee_coeff_unclassified(ea0, ea1, eb0, eb1);
    V3<T> x0 = c.v[0] * ea0 + c.v[1] * ea1;
    V3<T> x1 = c.v[2] * eb0 + c.v[3] * eb1;
    V3<T> d  = x1 - x0;
    return dot(d, d);
}

// accd.hpp:31-43
template <class T>
inline void centerize(V3<T> x[4])
{
    V3<T> mov;
    for (int k = 0; k < 4; k++) mov = mov + T(0.25f) * x[k];
    for (int k = 0; k < 4; k++) x[k] = x[k] - mov;
}
template <class T>
inline T max_relative_u(const V3<T> u[4])
{
    T m = 0;
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++)
        {
            V3<T> du = u[i] - u[j];
            m        = std::max(m, dot(du, du));
        }
    return std::sqrt(m);
}

// accd.hpp:49-111. kind 0: point-triangle (x[0] point), kind 1: edge-edge.
template <class T>
inline T ccd_helper(const V3<T> x0[4], const V3<T> dx[4], T u_max, T offset, int kind, int* iters = nullptr)
{
    auto dist2 = [&](const V3<T> y[4]) { return kind == 0 ? pt_dist2(y[0], y[1], y[2], y[3]) : ee_dist2(y[0], y[1], y[2], y[3]); };
    T toi       = 0;
    T max_t     = T(kLineSearchMaxT);
    T eps       = T(kCcdReduction) * (std::sqrt(dist2(x0)) - offset);
    T target    = eps + offset;
    T eps_sqr   = eps * eps;
    T inv_u_max = T(1) / u_max;
    int it      = 0;
    while (true)
    {
        if (it > 10000)
        {
            toi = T(0.001f);
            break;
        }
        it += 1;
        V3<T> y[4];
        for (int k = 0; k < 4; k++) y[k] = x0[k] + toi * dx[k];
        T d2 = dist2(y);
        T dmt = (d2 - target * target) / (std::sqrt(d2) + target);
        if ((max_t - toi) * u_max < dmt - eps)
        {
            toi = max_t;
            break;
        }
        else if (toi > T(0) && dmt * dmt < eps_sqr)
            break;
        T toi_next = toi + dmt * inv_u_max;
        if (toi_next != toi)
            toi = toi_next;
        else
            break;
        if (toi > max_t)
        {
            toi = max_t;
            break;
        }
    }
    if (iters) *iters = it;
    return toi;
}

// accd.hpp:151-205: returns the raw ACCD toi in (0, 1.25]; 1.25 means no impact.
template <class T>
inline T accd(const V3<T> a0[4], const V3<T> a1[4], T offset, int kind, int* iters = nullptr)
{
    V3<T> x0[4], dx[4];
    for (int k = 0; k < 4; k++)
    {
        x0[k] = a0[k];
        dx[k] = a1[k] - a0[k];
    }
    centerize(x0);
    centerize(dx);
    T u_max = max_relative_u(dx);
    T toi   = T(kLineSearchMaxT);
    if (iters) *iters = 0;
    if (u_max != T(0)) toi = ccd_helper(x0, dx, u_max, offset, kind, iters);
    return toi;
}

template <class T> inline T barrier(T d, T dh)
{
    T dm = d - dh;
    return -dm * dm * std::log(d / dh);
}
template <class T> inline T barrier_d1(T d, T dh) { return (dh - d) * ((T(2) * std::log(d / dh) - dh / d) + T(1)); }
template <class T> inline T barrier_d2(T d, T dh)
{
    T r = dh / d;
    return ((r + T(2)) * r - T(2) * std::log(d / dh)) - T(3);
}

// Contact energy forms (SceneParams::contact_energy_type, scene_params.h:89): 1 the log barrier on
// d - thickness, 0 the quadratic penalty E = 1/2 k C^2 with C = thickness + d_hat - d.
// Ground, stiffness kappa * area: k1 = dE/d(dist), k2 its derivative (newton_solver.cpp:916-925),
// energy solver_interface.cpp:577-584.
template <class T> inline void ground_k(int quad, T dist, T th, T dh, T stiff, T& k1, T& k2)
{
    if (quad)
    {
        k1 = stiff * ((dist - th) - dh);
        k2 = stiff;
        return;
    }
    k1 = stiff * barrier_d1(dist - th, dh);
    k2 = stiff * barrier_d2(dist - th, dh);
}
template <class T> inline T ground_e(int quad, T dist, T th, T dh, T stiff)
{
    if (quad)
    {
        T q = (dist - th) - dh;
        return (T(0.5f) * stiff) * (q * q);
    }
    return stiff * barrier(dist - th, dh);
}
// Pairs, stiffness aa * kappa: k1 = -dE/dd, k2 = d2E/dd2 (narrow_phase.cpp:873-885, 1031-1043),
// energy narrow_phase.cpp:3022-3030.
template <class T> inline void pair_k(int quad, T d, T th, T dh, T aa, T kap, T& k1, T& k2)
{
    if (quad)
    {
        T C = (th + dh) - d, s = aa * kap;
        k1  = s * C;
        k2  = s;
        return;
    }
    k1 = ((-aa) * kap) * barrier_d1(d - th, dh);
    k2 = (aa * kap) * barrier_d2(d - th, dh);
}
template <class T> inline T pair_e(int quad, T d, T th, T dh, T stiff)
{
    if (quad)
    {
        T C = (th + dh) - d;
        return ((T(0.5f) * stiff) * C) * C;
    }
    return stiff * barrier(d - th, dh);
}

}  // namespace orc
