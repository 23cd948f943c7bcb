// Per-element physics used by the Newton-step kernels (host + device, FP32).
//
//   distance / barycentrics  reference Solver/CollisionDetector/distance.hpp:40-319
//   additive CCD             reference Solver/CollisionDetector/accd.hpp:31-205
//   IPC log barrier          reference Solver/CollisionDetector/cipc_kernel.hpp:12-38
//   Baraff-Witkin membrane   reference Solver/Energy/stretch_energy.h:240-471
//   dihedral bending         reference Solver/Energy/bending_energy.h:167-274
//
// Evaluation order is fixed by the expressions (no contraction), see lcsx_common.h.
#pragma once
#include "lcsx_common.h"

namespace lcsx
{

// ---------------------------------------------------------------- distances
// 2x2 solve by adjugate / determinant (a symmetric, given column-wise).
LCSX_HD bool solve2(float a00, float a01, float a10, float a11, float b0, float b1, float& x0, float& x1)
{
    float det = a00 * a11 - a10 * a01;
    if (det != 0.0f)
    {
        x0 = (a11 / det) * b0 + (-a01 / det) * b1;
        x1 = (-a10 / det) * b0 + (a00 / det) * b1;
        return true;
    }
    return false;
}

LCSX_HD void pe_coeff(f3 p, f3 e0, f3 e1, float& c0, float& c1)
{
    f3    r = e1 - e0;
    float d = dot(r, r);
    c0      = 0.5f;
    c1      = 0.5f;
    if (d > kEpsilon)
    {
        float x = dot(r, p - e0) / d;
        c0      = 1.0f - x;
        c1      = x;
    }
}

LCSX_HD f3 pt_coeff(f3 p, f3 t0, f3 t1, f3 t2)
{
    f3    r0 = t1 - t0, r1 = t2 - t0, w = p - t0;
    float c0, c1;
    if (!solve2(dot(r0, r0), dot(r1, r0), dot(r0, r1), dot(r1, r1), dot(r0, w), dot(r1, w), c0, c1))
    {
        c0 = 1.0f / 3.0f;
        c1 = 1.0f / 3.0f;
    }
    return mk((1.0f - c0) - c1, c0, c1);
}

// Closest-point barycentrics on the triangle, falling back to edges / vertices.
LCSX_HD f3 pt_coeff_unclassified(f3 p, f3 t0, f3 t1, f3 t2)
{
    f3 c = pt_coeff(p, t0, t1, t2);
    if (c.x >= 0.0f && c.y >= 0.0f && c.z >= 0.0f && c.x <= 1.0f && c.y <= 1.0f && c.z <= 1.0f) return c;
    float e0, e1;
    if (c.x < 0.0f)
    {
        pe_coeff(p, t1, t2, e0, e1);
        if (e0 >= 0.0f && e0 <= 1.0f) return mk(0.0f, e0, e1);
        return e0 > 1.0f ? mk(0.0f, 1.0f, 0.0f) : mk(0.0f, 0.0f, 1.0f);
    }
    else if (c.y < 0.0f)
    {
        pe_coeff(p, t0, t2, e0, e1);
        if (e0 >= 0.0f && e0 <= 1.0f) return mk(e0, 0.0f, e1);
        return e0 > 1.0f ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 0.0f, 1.0f);
    }
    pe_coeff(p, t0, t1, e0, e1);
    if (e0 >= 0.0f && e0 <= 1.0f) return mk(e0, e1, 0.0f);
    return e0 > 1.0f ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f);
}

LCSX_HD float pt_dist2(f3 p, f3 t0, f3 t1, f3 t2)
{
    f3 c = pt_coeff_unclassified(p, t0, t1, t2);
    f3 x = (c.x * (t0 - p) + c.y * (t1 - p)) + c.z * (t2 - p);
    return dot(x, x);
}

struct c4
{
    float v[4];
};

LCSX_HD c4 ee_coeff(f3 ea0, f3 ea1, f3 eb0, f3 eb1)
{
    f3    r0 = ea1 - ea0, nr1 = -(eb1 - eb0), w = eb0 - ea0;
    float x0, x1;
    c4    b = {{0.5f, 0.5f, 0.5f, 0.5f}};
    if (solve2(dot(r0, r0), dot(nr1, r0), dot(r0, nr1), dot(nr1, nr1), dot(r0, w), dot(nr1, w), x0, x1))
        b = {{1.0f - x0, x0, 1.0f - x1, x1}};
    return b;
}

// Edge-edge barycentrics; near-parallel or out-of-range cases take the closest of the four
// clamped point-edge candidates.
LCSX_HD c4 ee_coeff_unclassified(f3 ea0, f3 ea1, f3 eb0, f3 eb1)
{
    f3    r0 = ea1 - ea0, r1 = eb1 - eb0;
    float len0 = dot(r0, r0), len1 = dot(r1, r1);
    f3    cr   = cross(r0, r1);
    float par  = dot(cr, cr) / (len0 * len1 + 1e-8f);
    c4    c    = ee_coeff(ea0, ea1, eb0, eb1);
    bool  ok   = true;
    for (int i = 0; i < 4; i++) ok = ok && c.v[i] >= 0.0f && c.v[i] <= 1.0f;
    if (par < 1e-4f || !ok)
    {
        float cand[4][4];
        float a, b;
        pe_coeff(ea0, eb0, eb1, a, b);
        cand[0][0] = 1.0f; cand[0][1] = 0.0f; cand[0][2] = clampf(a, 0.0f, 1.0f); cand[0][3] = clampf(b, 0.0f, 1.0f);
        pe_coeff(ea1, eb0, eb1, a, b);
        cand[1][0] = 0.0f; cand[1][1] = 1.0f; cand[1][2] = clampf(a, 0.0f, 1.0f); cand[1][3] = clampf(b, 0.0f, 1.0f);
        pe_coeff(eb0, ea0, ea1, a, b);
        cand[2][0] = clampf(a, 0.0f, 1.0f); cand[2][1] = clampf(b, 0.0f, 1.0f); cand[2][2] = 1.0f; cand[2][3] = 0.0f;
        pe_coeff(eb1, ea0, ea1, a, b);
        cand[3][0] = clampf(a, 0.0f, 1.0f); cand[3][1] = clampf(b, 0.0f, 1.0f); cand[3][2] = 0.0f; cand[3][3] = 1.0f;
        int   best  = 0;
        float min_d = 3.40282347e+38f;
        for (int i = 0; i < 4; i++)
        {
            f3    x0 = cand[i][0] * ea0 + cand[i][1] * ea1;
            f3    x1 = cand[i][2] * eb0 + cand[i][3] * eb1;
            f3    d  = x0 - x1;
            float dd = dot(d, d);
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

LCSX_HD float ee_dist2(f3 ea0, f3 ea1, f3 eb0, f3 eb1)
{
    c4 c  = ee_coeff_unclassified(ea0, ea1, eb0, eb1);
    f3 x0 = c.v[0] * ea0 + c.v[1] * ea1;
    f3 x1 = c.v[2] * eb0 + c.v[3] * eb1;
    f3 d  = x1 - x0;
    return dot(d, d);
}

// ---------------------------------------------------------------- additive CCD
// kind 0: point (x[0]) vs triangle (x[1..3]); kind 1: edge (x[0], x[1]) vs edge (x[2], x[3]).
LCSX_HD float ccd_dist2(const f3 y[4], int kind)
{
    return kind == 0 ? pt_dist2(y[0], y[1], y[2], y[3]) : ee_dist2(y[0], y[1], y[2], y[3]);
}

// Raw ACCD time of impact in (0, 1.25]; 1.25 means no impact before the cap. With a budget
// below the reference's 10001 iterations, a pair that needs more returns with *done = false
// (its result is then obtained by running it again with the full budget: the iteration is
// deterministic, so the split changes nothing).
//
// The iteration's only state is toi and its exits depend only on toi, so once toi repeats a
// value it is periodic and ends only at the iteration cap, with toi = 0.001. A toi saved at
// power-of-two iteration counts (Brent's cycle detection) finds such a repeat and returns that
// result early; the result is the same as running to the cap.
LCSX_HD float accd_budget(const f3 a0[4], const f3 a1[4], float offset, int kind, int budget, bool* done)

{
    f3 x0[4], dx[4];
    f3 m0 = mk(0.0f, 0.0f, 0.0f), m1 = mk(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < 4; k++)
    {
        x0[k] = a0[k];
        dx[k] = a1[k] - a0[k];
    }
    for (int k = 0; k < 4; k++)
    {
        m0 = m0 + 0.25f * x0[k];
        m1 = m1 + 0.25f * dx[k];
    }
    for (int k = 0; k < 4; k++)
    {
        x0[k] = x0[k] - m0;
        dx[k] = dx[k] - m1;
    }
    float um = 0.0f;
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++)
        {
            f3 du = dx[i] - dx[j];
            um    = smax(um, dot(du, du));
        }
    float u_max = sqrtf(um);
    *done       = true;
    if (u_max == 0.0f) return kLineSearchMaxT;
    float toi = 0.0f, max_t = kLineSearchMaxT;
    float eps       = kCcdReduction * (sqrtf(ccd_dist2(x0, kind)) - offset);
    float target    = eps + offset;
    float eps_sqr   = eps * eps;
    float inv_u_max = 1.0f / u_max;
    int   it        = 0;
    float saved     = -1.0f;  // toi at the last checkpoint (toi is never negative)
    int   lap = 0, power = 1;
    while (true)
    {
        if (it > 10000 || toi == saved)
        {
            toi = 0.001f;
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
        f3 y[4];
        for (int k = 0; k < 4; k++) y[k] = x0[k] + toi * dx[k];
        float d2  = ccd_dist2(y, kind);
        float dmt = (d2 - target * target) / (sqrtf(d2) + target);
        if ((max_t - toi) * u_max < dmt - eps)
        {
            toi = max_t;
            break;
        }
        else if (toi > 0.0f && dmt * dmt < eps_sqr)
            break;
        float toi_next = toi + dmt * inv_u_max;
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

LCSX_HD float accd(const f3 a0[4], const f3 a1[4], float offset, int kind)
{
    bool done;
    return accd_budget(a0, a1, offset, kind, 0x7fffffff, &done);
}

// ---------------------------------------------------------------- log barrier
LCSX_HD float barrier_d1(float d, float dh) { return (dh - d) * ((2.0f * logf(d / dh) - dh / d) + 1.0f); }
LCSX_HD float barrier_d2(float d, float dh)
{
    float r = dh / d;
    return ((r + 2.0f) * r - 2.0f * logf(d / dh)) - 3.0f;
}
LCSX_HD float barrier_e(float d, float dh)
{
    float dm = d - dh;
    return -dm * dm * logf(d / dh);
}

// ---------------------------------------------------------------- contact energy forms
// SceneParams::contact_energy_type (scene_params.h:89): 1 the log barrier above on d - thickness,
// 0 the quadratic penalty E = 1/2 k C^2, C = thickness + d_hat - d.
// Ground (newton_solver.cpp:916-925): k1 = dE/d(dist), k2 = its derivative; stiff = kappa * area.
LCSX_HD void ground_k(int quad, float dist, float th, float dh, float stiff, float& k1, float& k2)
{
    if (quad)
    {
        k1 = stiff * ((dist - th) - dh);
        k2 = stiff;
    }
    else
    {
        k1 = stiff * barrier_d1(dist - th, dh);
        k2 = stiff * barrier_d2(dist - th, dh);
    }
}
// solver_interface.cpp:577-584
LCSX_HD float ground_e(int quad, float dist, float th, float dh, float stiff)
{
    if (quad)
    {
        float q = (dist - th) - dh;
        return (0.5f * stiff) * (q * q);
    }
    return stiff * barrier_e(dist - th, dh);
}
// Pairs (narrow_phase.cpp:873-885, 1031-1043): k1 = -dE/dd, k2 = d2E/dd2, stiffness aa * kappa.
LCSX_HD void pair_k(int quad, float d, float th, float dh, float aa, float kappa, float& k1, float& k2)
{
    if (quad)
    {
        float C = (th + dh) - d, s = aa * kappa;
        k1      = s * C;
        k2      = s;
    }
    else
    {
        k1 = ((-aa) * kappa) * barrier_d1(d - th, dh);
        k2 = (aa * kappa) * barrier_d2(d - th, dh);
    }
}
// narrow_phase.cpp:3022-3030 (stiff: the pair's stored aa * kappa)
LCSX_HD float pair_e(int quad, float d, float th, float dh, float stiff)
{
    if (quad)
    {
        float C = (th + dh) - d;
        return ((0.5f * stiff) * C) * C;
    }
    return stiff * barrier_e(d - th, dh);
}

// ---------------------------------------------------------------- Baraff-Witkin membrane
// F = [x1 - x0, x2 - x0] Dm^-1, Dm^-1 column-major (d0, d1, d2, d3): Fu = e0 d0 + e1 d1,
// Fv = e0 d2 + e1 d3. Output: gradient per vertex and the 9x9 Hessian as 3x3 blocks
// H[a][b] (row-major); blocks with a > b are not produced (they are transposes of (b, a)).
struct FaceOut
{
    f3 g[3];
    m3 H[3][3];
};

LCSX_HD void bw_hessian6(f3 Fu, f3 Fv, float mu, float lmd, float H6[6][6])
{
    float fu[3], fv[3];
    float su = sqrtf(dot(Fu, Fu)), sv = sqrtf(dot(Fv, Fv));
    float iu = 1.0f / su, iv = 1.0f / sv;
    f3    nu = Fu * iu, nv = Fv * iv;
    fu[0] = nu.x; fu[1] = nu.y; fu[2] = nu.z;
    fv[0] = nv.x; fv[1] = nv.y; fv[2] = nv.z;
    float uc = smin(iu, 1.0f), vc = smin(iv, 1.0f);
    float du = smax(0.0f, 1.0f - iu), dv = smax(0.0f, 1.0f - iv);
    // shear projection
    float I6  = dot(Fu, Fv);
    float sgn = I6 < 0.0f ? -1.0f : 1.0f;
    float g[6] = {Fv.x, Fv.y, Fv.z, Fu.x, Fu.y, Fu.z};
    float I2   = dot(Fu, Fu) + dot(Fv, Fv);
    float lam0 = 0.5f * (I2 + sqrtf(I2 * I2 + 12.0f * I6 * I6));
    float q[6];
    for (int i = 0; i < 6; i++) q[i] = I6 * g[(i + 3) % 6] + lam0 * g[i];
    float qn = 0.0f;
    qn       = qn + ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]);
    qn       = qn + ((q[3] * q[3] + q[4] * q[4]) + q[5] * q[5]);
    float ql = sqrtf(qn);
    for (int i = 0; i < 6; i++) q[i] = q[i] / ql;
    // T = 0.5 (I + sgn H0): row i has 0.5 at i and 0.5 sgn at (i + 3) % 6
    float Tq[6];
    float half_s = 0.5f * (0.0f + sgn * 1.0f);
    float half_1 = 0.5f * (1.0f + sgn * 0.0f);
    for (int i = 0; i < 6; i++)
    {
        int   j0 = i < 3 ? i : i - 3, j1 = j0 + 3;  // the two nonzero columns in ascending order
        float t0 = (j0 == i) ? half_1 : half_s;
        float t1 = (j1 == i) ? half_1 : half_s;
        Tq[i]    = t0 * q[j0] + t1 * q[j1];
    }
    float nTq = 0.0f;
    nTq       = nTq + ((Tq[0] * Tq[0] + Tq[1] * Tq[1]) + Tq[2] * Tq[2]);
    nTq       = nTq + ((Tq[3] * Tq[3] + Tq[4] * Tq[4]) + Tq[5] * Tq[5]);
    float aI6 = fabsf(I6);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++)
        {
            float tm = (i == j) ? half_1 : (((i + 3) % 6 == j) ? half_s : 0.5f * (0.0f + sgn * 0.0f));
            float hh = lmd * (aI6 * (tm - (Tq[i] * Tq[j]) / nTq) + lam0 * (q[i] * q[j]));
            float hs = 0.0f;
            if (i < 3 && j < 3) hs = ((i == j) ? du : 0.0f) + uc * (fu[i] * fu[j]);
            if (i >= 3 && j >= 3) hs = ((i == j) ? dv : 0.0f) + vc * (fv[i - 3] * fv[j - 3]);
            H6[i][j] = mu * hs + hh;
        }
}

LCSX_HD void bw_face(f3 x0, f3 x1, f3 x2, const float Dm[4], float mu, float lmd, float area, FaceOut& out)
{
    f3 e0 = x1 - x0, e1 = x2 - x0;
    f3 Fu = e0 * Dm[0] + e1 * Dm[1];
    f3 Fv = e0 * Dm[2] + e1 * Dm[3];
    f3 gu, gv;
    {
        float su = sqrtf(dot(Fu, Fu)), sv = sqrtf(dot(Fv, Fv));
        float iu = 1.0f / su, iv = 1.0f / sv;
        f3    sug = mu * (((su - 1.0f) * iu) * Fu);
        f3    svg = mu * (((sv - 1.0f) * iv) * Fv);
        float w   = dot(Fu, Fv);
        gu        = sug + lmd * (w * Fv);
        gv        = svg + lmd * (w * Fu);
    }
    float cx[3] = {-(Dm[0] + Dm[1]), Dm[0], Dm[1]};
    float cy[3] = {-(Dm[2] + Dm[3]), Dm[2], Dm[3]};
    for (int a = 0; a < 3; a++)
        out.g[a] = mk(area * (cx[a] * gu.x + cy[a] * gv.x), area * (cx[a] * gu.y + cy[a] * gv.y),
                      area * (cx[a] * gu.z + cy[a] * gv.z));
    float H6[6][6];
    bw_hessian6(Fu, Fv, mu, lmd, H6);
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++)
            for (int r = 0; r < 3; r++)
                for (int s = 0; s < 3; s++)
                {
                    float t1 = H6[r][s] * (cx[a] * cx[b]);
                    float t2 = H6[r][s + 3] * (cx[a] * cy[b]);
                    float t3 = H6[r + 3][s] * (cy[a] * cx[b]);
                    float t4 = H6[r + 3][s + 3] * (cy[a] * cy[b]);
                    out.H[a][b].m[3 * r + s] = area * (((t1 + t2) + t3) + t4);
                }
}

// Dm^-1 of the rest triangle in its own plane (reference stretch_energy.h:424-439).
LCSX_HD void dm_inv(f3 x0, f3 x1, f3 x2, float out[4])
{
    f3    r1 = x1 - x0, r2 = x2 - x0;
    f3    c  = cross(r1, r2);
    f3    a1 = normalize(r1);
    f3    a2 = normalize(cross(c, a1));
    float u0 = dot(a1, x0), v0 = dot(a2, x0), u1 = dot(a1, x1), v1 = dot(a2, x1), u2 = dot(a1, x2), v2 = dot(a2, x2);
    float m00 = u1 - u0, m10 = v1 - v0, m01 = u2 - u0, m11 = v2 - v0;
    float inv = 1.0f / (m00 * m11 - m01 * m10);
    out[0]    = m11 * inv;
    out[1]    = -m10 * inv;
    out[2]    = -m01 * inv;
    out[3]    = m00 * inv;
}

// ---------------------------------------------------------------- dihedral bending
// Hinge (p0, p1) edge with opposite vertices p2, p3; sign(0) taken as +1.
LCSX_HD float dihedral_angle(f3 p0, f3 p1, f3 p2, f3 p3)
{
    f3    n1 = cross(p1 - p2, p0 - p2);
    f3    n2 = cross(p0 - p3, p1 - p3);
    float d  = dot(n1, n2) / sqrtf(dot(n1, n1) * dot(n2, n2));
    float an = acosf(smax(-1.0f, smin(1.0f, d)));
    float sg = dot(cross(n2, n1), p1 - p0) < 0.0f ? -1.0f : 1.0f;
    return sg * an;
}

LCSX_HD void dihedral_grad(f3 p0, f3 p1, f3 p2, f3 p3, f3 g[4])
{
    f3    e0 = p0 - p1, e1 = p2 - p1, e2 = p3 - p1, e3 = p2 - p0, e4 = p3 - p0;
    f3    n1 = cross(e0, e1), n2 = cross(e2, e0);
    float n1s = dot(n1, n1), n2s = dot(n2, n2), e0n = length(e0);
    g[2]      = ((-e0n) / n1s) * n1;
    g[1]      = ((-dot(e0, e3)) / (e0n * n1s)) * n1 - (dot(e0, e4) / (e0n * n2s)) * n2;
    g[0]      = (dot(e0, e1) / (e0n * n1s)) * n1 + (dot(e0, e2) / (e0n * n2s)) * n2;
    g[3]      = ((-e0n) / n2s) * n2;
}

}  // namespace lcsx
