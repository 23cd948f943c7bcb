// FP64 collision arithmetic of the mixed-precision mode (host + device): distances, the additive
// CCD and the log-barrier derivatives evaluated in double from FP32 positions.
//
//   distance / barycentrics  reference Solver/CollisionDetector/distance.hpp:40-319
//   additive CCD             reference Solver/CollisionDetector/accd.hpp:31-205
//   IPC log barrier          reference Solver/CollisionDetector/cipc_kernel.hpp:12-38
//
// The algorithms and their constants are the reference's (its FP32 constants widened exactly to
// double: 1/3, FLT_EPSILON, 1e-8, 1e-4, 0.01, 0.001), so a pair evaluated here equals the FP32
// path's algorithm run in double. Evaluation order is fixed (no contraction, see lcsx_common.h).
#pragma once
#include "lcsx_common.h"

namespace lcsx
{
namespace fp64
{

struct d3
{
    double x, y, z;
};
LCSX_HD d3 mkd(double a, double b, double c) { return d3{a, b, c}; }
LCSX_HD d3 wide(f3 a) { return d3{(double)a.x, (double)a.y, (double)a.z}; }
LCSX_HD d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LCSX_HD d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LCSX_HD d3 operator-(d3 a) { return {-a.x, -a.y, -a.z}; }
LCSX_HD d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
LCSX_HD d3 operator/(d3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
LCSX_HD double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LCSX_HD d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
LCSX_HD double dmin(double a, double b) { return (b < a) ? b : a; }
LCSX_HD double dmax(double a, double b) { return (a < b) ? b : a; }
LCSX_HD double dclamp(double v, double lo, double hi) { return dmax(dmin(v, hi), lo); }

// the reference's FP32 constants, widened
constexpr double kEps      = (double)kEpsilon;
constexpr double kThird    = (double)(1.0f / 3.0f);
constexpr double kHalf     = 0.5;
constexpr double kParEps   = (double)1e-8f;
constexpr double kParTol   = (double)1e-4f;
constexpr double kMaxT     = (double)kLineSearchMaxT;
constexpr double kRed      = (double)kCcdReduction;
constexpr double kFallback = (double)0.001f;
constexpr double kFltMax   = (double)3.40282347e+38f;

LCSX_HD bool solve2(double a00, double a01, double a10, double a11, double b0, double b1, double& x0, double& x1)
{
    double det = a00 * a11 - a10 * a01;
    if (det != 0.0)
    {
        x0 = (a11 / det) * b0 + (-a01 / det) * b1;
        x1 = (-a10 / det) * b0 + (a00 / det) * b1;
        return true;
    }
    return false;
}

LCSX_HD void pe_coeff(d3 p, d3 e0, d3 e1, double& c0, double& c1)
{
    d3     r = e1 - e0;
    double d = dot(r, r);
    c0       = kHalf;
    c1       = kHalf;
    if (d > kEps)
    {
        double x = dot(r, p - e0) / d;
        c0       = 1.0 - x;
        c1       = x;
    }
}

LCSX_HD d3 pt_coeff(d3 p, d3 t0, d3 t1, d3 t2)
{
    d3     r0 = t1 - t0, r1 = t2 - t0, w = p - t0;
    double c0, c1;
    if (!solve2(dot(r0, r0), dot(r1, r0), dot(r0, r1), dot(r1, r1), dot(r0, w), dot(r1, w), c0, c1))
    {
        c0 = kThird;
        c1 = kThird;
    }
    return mkd((1.0 - c0) - c1, c0, c1);
}

LCSX_HD d3 pt_coeff_unclassified(d3 p, d3 t0, d3 t1, d3 t2)
{
    d3 c = pt_coeff(p, t0, t1, t2);
    if (c.x >= 0.0 && c.y >= 0.0 && c.z >= 0.0 && c.x <= 1.0 && c.y <= 1.0 && c.z <= 1.0) return c;
    double e0, e1;
    if (c.x < 0.0)
    {
        pe_coeff(p, t1, t2, e0, e1);
        if (e0 >= 0.0 && e0 <= 1.0) return mkd(0.0, e0, e1);
        return e0 > 1.0 ? mkd(0.0, 1.0, 0.0) : mkd(0.0, 0.0, 1.0);
    }
    else if (c.y < 0.0)
    {
        pe_coeff(p, t0, t2, e0, e1);
        if (e0 >= 0.0 && e0 <= 1.0) return mkd(e0, 0.0, e1);
        return e0 > 1.0 ? mkd(1.0, 0.0, 0.0) : mkd(0.0, 0.0, 1.0);
    }
    pe_coeff(p, t0, t1, e0, e1);
    if (e0 >= 0.0 && e0 <= 1.0) return mkd(e0, e1, 0.0);
    return e0 > 1.0 ? mkd(1.0, 0.0, 0.0) : mkd(0.0, 1.0, 0.0);
}

LCSX_HD double pt_dist2(d3 p, d3 t0, d3 t1, d3 t2)
{
    d3 c = pt_coeff_unclassified(p, t0, t1, t2);
    d3 x = (c.x * (t0 - p) + c.y * (t1 - p)) + c.z * (t2 - p);
    return dot(x, x);
}

struct dc4
{
    double v[4];
};

LCSX_HD dc4 ee_coeff(d3 ea0, d3 ea1, d3 eb0, d3 eb1)
{
    d3     r0 = ea1 - ea0, nr1 = -(eb1 - eb0), w = eb0 - ea0;
    double x0, x1;
    dc4    b = {{kHalf, kHalf, kHalf, kHalf}};
    if (solve2(dot(r0, r0), dot(nr1, r0), dot(r0, nr1), dot(nr1, nr1), dot(r0, w), dot(nr1, w), x0, x1))
        b = {{1.0 - x0, x0, 1.0 - x1, x1}};
    return b;
}

LCSX_HD dc4 ee_coeff_unclassified(d3 ea0, d3 ea1, d3 eb0, d3 eb1)
{
    d3     r0 = ea1 - ea0, r1 = eb1 - eb0;
    double len0 = dot(r0, r0), len1 = dot(r1, r1);
    d3     cr   = cross(r0, r1);
    double par  = dot(cr, cr) / (len0 * len1 + kParEps);
    dc4    c    = ee_coeff(ea0, ea1, eb0, eb1);
    bool   ok   = true;
    for (int i = 0; i < 4; i++) ok = ok && c.v[i] >= 0.0 && c.v[i] <= 1.0;
    if (par < kParTol || !ok)
    {
        double cand[4][4];
        double a, b;
        pe_coeff(ea0, eb0, eb1, a, b);
        cand[0][0] = 1.0; cand[0][1] = 0.0; cand[0][2] = dclamp(a, 0.0, 1.0); cand[0][3] = dclamp(b, 0.0, 1.0);
        pe_coeff(ea1, eb0, eb1, a, b);
        cand[1][0] = 0.0; cand[1][1] = 1.0; cand[1][2] = dclamp(a, 0.0, 1.0); cand[1][3] = dclamp(b, 0.0, 1.0);
        pe_coeff(eb0, ea0, ea1, a, b);
        cand[2][0] = dclamp(a, 0.0, 1.0); cand[2][1] = dclamp(b, 0.0, 1.0); cand[2][2] = 1.0; cand[2][3] = 0.0;
        pe_coeff(eb1, ea0, ea1, a, b);
        cand[3][0] = dclamp(a, 0.0, 1.0); cand[3][1] = dclamp(b, 0.0, 1.0); cand[3][2] = 0.0; cand[3][3] = 1.0;
        int    best  = 0;
        double min_d = kFltMax;
        for (int i = 0; i < 4; i++)
        {
            d3     x0 = cand[i][0] * ea0 + cand[i][1] * ea1;
            d3     x1 = cand[i][2] * eb0 + cand[i][3] * eb1;
            d3     d  = x0 - x1;
            double dd = dot(d, d);
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

LCSX_HD double ee_dist2(d3 ea0, d3 ea1, d3 eb0, d3 eb1)
{
    dc4 c  = ee_coeff_unclassified(ea0, ea1, eb0, eb1);
    d3  x0 = c.v[0] * ea0 + c.v[1] * ea1;
    d3  x1 = c.v[2] * eb0 + c.v[3] * eb1;
    d3  d  = x1 - x0;
    return dot(d, d);
}

LCSX_HD double ccd_dist2(const d3 y[4], int kind)
{
    return kind == 0 ? pt_dist2(y[0], y[1], y[2], y[3]) : ee_dist2(y[0], y[1], y[2], y[3]);
}

// The ACCD iteration of accd_budget (lcsx_physics.h) in double, with the same budget split and
// the same early exit on a repeated toi.
LCSX_HD double accd_budget(const d3 a0[4], const d3 a1[4], double offset, int kind, int budget, bool* done)
{
    d3 x0[4], dx[4];
    d3 m0 = mkd(0.0, 0.0, 0.0), m1 = mkd(0.0, 0.0, 0.0);
    for (int k = 0; k < 4; k++)
    {
        x0[k] = a0[k];
        dx[k] = a1[k] - a0[k];
    }
    for (int k = 0; k < 4; k++)
    {
        m0 = m0 + 0.25 * x0[k];
        m1 = m1 + 0.25 * dx[k];
    }
    for (int k = 0; k < 4; k++)
    {
        x0[k] = x0[k] - m0;
        dx[k] = dx[k] - m1;
    }
    double um = 0.0;
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++)
        {
            d3 du = dx[i] - dx[j];
            um    = dmax(um, dot(du, du));
        }
    double u_max = sqrt(um);
    *done        = true;
    if (u_max == 0.0) return kMaxT;
    double toi = 0.0, max_t = kMaxT;
    double eps       = kRed * (sqrt(ccd_dist2(x0, kind)) - offset);
    double target    = eps + offset;
    double eps_sqr   = eps * eps;
    double inv_u_max = 1.0 / u_max;
    int    it        = 0;
    double saved     = -1.0;
    int    lap = 0, power = 1;
    while (true)
    {
        if (it > 10000 || toi == saved)
        {
            toi = kFallback;
            break;
        }
        if (++lap == power)
        {
            saved = toi;
            power <<= 1;
            lap = 0;
        }
        if (it >= budget)
        {
            *done = false;
            break;
        }
        it += 1;
        d3 y[4];
        for (int k = 0; k < 4; k++) y[k] = x0[k] + toi * dx[k];
        double d2  = ccd_dist2(y, kind);
        double dmt = (d2 - target * target) / (sqrt(d2) + target);
        if ((max_t - toi) * u_max < dmt - eps)
        {
            toi = max_t;
            break;
        }
        else if (toi > 0.0 && dmt * dmt < eps_sqr)
            break;
        double toi_next = toi + dmt * inv_u_max;
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
    return toi;
}

LCSX_HD double accd(const d3 a0[4], const d3 a1[4], double offset, int kind)
{
    bool done;
    return accd_budget(a0, a1, offset, kind, 0x7fffffff, &done);
}

// cipc_kernel.hpp:12-38 on d - thickness, in double: the barrier, its first and second derivatives
LCSX_HD double barrier(double d, double dh)
{
    double dm = d - dh;
    return -dm * dm * log(d / dh);
}
LCSX_HD double barrier_d1(double d, double dh) { return (dh - d) * ((2.0 * log(d / dh) - dh / d) + 1.0); }
LCSX_HD double barrier_d2(double d, double dh)
{
    double r = dh / d;
    return ((r + 2.0) * r - 2.0 * log(d / dh)) - 3.0;
}

// the contact energy forms of lcsx_physics.h (ground_k, ground_e, pair_k, pair_e), in double
LCSX_HD double ground_e(int quad, double dist, double th, double dh, double stiff)
{
    if (quad)
    {
        double q = (dist - th) - dh;
        return (0.5 * stiff) * (q * q);
    }
    return stiff * barrier(dist - th, dh);
}
LCSX_HD void pair_k(int quad, double d, double th, double dh, double aa, double kappa, double& k1, double& k2)
{
    if (quad)
    {
        double C = (th + dh) - d, s = aa * kappa;
        k1       = s * C;
        k2       = s;
    }
    else
    {
        k1 = ((-aa) * kappa) * barrier_d1(d - th, dh);
        k2 = (aa * kappa) * barrier_d2(d - th, dh);
    }
}
LCSX_HD double pair_e(int quad, double d, double th, double dh, double stiff)
{
    if (quad)
    {
        double C = (th + dh) - d;
        return ((0.5 * stiff) * C) * C;
    }
    return stiff * barrier(d - th, dh);
}

}  // namespace fp64
}  // namespace lcsx
