// Host-compiled check: the product's per-element physics (luisacomputesimulator_amd/csrc/lcsx_physics.h,
// the same source the HIP kernels run) against the oracle restatement (oracle/orc_*.h), bit for bit
// in FP32 on random inputs. Prints "OK <cases>" or the first mismatch and exits non-zero.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include "../../luisacomputesimulator_amd/csrc/lcsx_fp64.h"
#include "../../luisacomputesimulator_amd/csrc/lcsx_physics.h"
#include "../../oracle/orc_collision.h"
#include "../../oracle/orc_energy.h"
#include "../../oracle/orc_lbvh.h"

static int g_fail = 0;
static void same64(double a, double b, const char* what, int k)
{
    uint64_t ua, ub;
    std::memcpy(&ua, &a, 8);
    std::memcpy(&ub, &b, 8);
    if (ua != ub && !(a != a && b != b))
    {
        if (g_fail < 10) std::printf("MISMATCH %s case %d: %.17g vs %.17g\n", what, k, a, b);
        g_fail++;
    }
}
static void same(float a, float b, const char* what, int k)
{
    uint32_t ua, ub;
    std::memcpy(&ua, &a, 4);
    std::memcpy(&ub, &b, 4);
    if (ua != ub && !(a != a && b != b))
    {
        if (g_fail < 10) std::printf("MISMATCH %s case %d: %.9g vs %.9g\n", what, k, a, b);
        g_fail++;
    }
}

// ACCD inputs that run into the 10001-iteration cap (tests/golden/accd_stall_cases.txt): the
// product's early exit on a repeated toi must give the oracle's full-length result.
static int stall_cases(const char* path)
{
    std::ifstream f(path);
    if (!f)
    {
        std::printf("MISSING %s\n", path);
        g_fail++;
        return 0;
    }
    std::string line;
    int         n = 0;
    while (std::getline(f, line))
    {
        std::istringstream ss(line);
        int                kind;
        std::string        tok;
        ss >> kind;
        float v[25];
        for (int i = 0; i < 25; i++)
        {
            ss >> tok;
            v[i] = std::strtof(tok.c_str(), nullptr);
        }
        lcsx::f3       a[4], b[4];
        orc::V3<float> oa[4], ob[4];
        for (int k = 0; k < 4; k++)
        {
            a[k]  = lcsx::mk(v[1 + 3 * k], v[2 + 3 * k], v[3 + 3 * k]);
            b[k]  = lcsx::mk(v[13 + 3 * k], v[14 + 3 * k], v[15 + 3 * k]);
            oa[k] = orc::V3<float>(a[k].x, a[k].y, a[k].z);
            ob[k] = orc::V3<float>(b[k].x, b[k].y, b[k].z);
        }
        int   iters = 0;
        float ref   = orc::accd(oa, ob, v[0], kind, &iters);
        if (iters <= 10000)
        {
            std::printf("NOT A STALL case %d (%d iterations)\n", n, iters);
            g_fail++;
        }
        same(lcsx::accd(a, b, v[0], kind), ref, "accd_stall", n);
        {
            lcsx::fp64::d3  da[4], db[4];
            orc::V3<double> qa[4], qb[4];
            for (int k = 0; k < 4; k++)
            {
                da[k] = lcsx::fp64::wide(a[k]);
                db[k] = lcsx::fp64::wide(b[k]);
                qa[k] = orc::V3<double>(da[k].x, da[k].y, da[k].z);
                qb[k] = orc::V3<double>(db[k].x, db[k].y, db[k].z);
            }
            same64(lcsx::fp64::accd(da, db, (double)v[0], kind), orc::accd(qa, qb, (double)v[0], kind), "accd64_stall", n);
        }
        bool  done = true;
        float part = lcsx::accd_budget(a, b, v[0], kind, 32, &done);
        if (done && part != ref)
        {
            std::printf("BUDGET case %d finished early with %.9g\n", n, part);
            g_fail++;
        }
        n++;
    }
    return n;
}

int main(int argc, char** argv)
{
    int stalls = stall_cases(argc > 1 ? argv[1] : "tests/golden/accd_stall_cases.txt");
    std::mt19937_64                       rng(1234);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    auto v  = [&]() { return lcsx::mk(U(rng), U(rng), U(rng)); };
    auto ov = [](lcsx::f3 a) { return orc::V3<float>(a.x, a.y, a.z); };
    const int N = 20000;
    for (int k = 0; k < N; k++)
    {
        lcsx::f3 p[4] = {v(), v(), v(), v()};
        // distances / barycentrics
        lcsx::f3 c = lcsx::pt_coeff_unclassified(p[0], p[1], p[2], p[3]);
        auto     oc = orc::pt_coeff_unclassified(ov(p[0]), ov(p[1]), ov(p[2]), ov(p[3]));
        same(c.x, oc.x, "pt_coeff.x", k);
        same(c.y, oc.y, "pt_coeff.y", k);
        same(c.z, oc.z, "pt_coeff.z", k);
        lcsx::c4 e  = lcsx::ee_coeff_unclassified(p[0], p[1], p[2], p[3]);
        auto     oe = orc::ee_coeff_unclassified(ov(p[0]), ov(p[1]), ov(p[2]), ov(p[3]));
        for (int i = 0; i < 4; i++) same(e.v[i], oe.v[i], "ee_coeff", k);
        same(lcsx::pt_dist2(p[0], p[1], p[2], p[3]), orc::pt_dist2(ov(p[0]), ov(p[1]), ov(p[2]), ov(p[3])), "pt_dist2", k);
        same(lcsx::ee_dist2(p[0], p[1], p[2], p[3]), orc::ee_dist2(ov(p[0]), ov(p[1]), ov(p[2]), ov(p[3])), "ee_dist2", k);
        // FP64 path (mixed-precision mode) against the oracle's double instantiation
        {
            lcsx::fp64::d3  d[4];
            orc::V3<double> q[4];
            for (int i = 0; i < 4; i++)
            {
                d[i] = lcsx::fp64::wide(p[i]);
                q[i] = orc::V3<double>(d[i].x, d[i].y, d[i].z);
            }
            same64(lcsx::fp64::pt_dist2(d[0], d[1], d[2], d[3]), orc::pt_dist2(q[0], q[1], q[2], q[3]), "pt_dist2_64", k);
            same64(lcsx::fp64::ee_dist2(d[0], d[1], d[2], d[3]), orc::ee_dist2(q[0], q[1], q[2], q[3]), "ee_dist2_64", k);
            double dd = 1e-4 + 1e-3 * (0.5 + 0.5 * (double)U(rng)), dh = 1e-3;
            same64(lcsx::fp64::barrier_d1(dd, dh), orc::barrier_d1(dd, dh), "barrier_d1_64", k);
            same64(lcsx::fp64::barrier_d2(dd, dh), orc::barrier_d2(dd, dh), "barrier_d2_64", k);
            if (k < 2000)
            {
                lcsx::fp64::d3  e[4];
                orc::V3<double> r[4];
                for (int i = 0; i < 4; i++)
                {
                    e[i] = lcsx::fp64::wide(p[i] + 0.5f * v());
                    r[i] = orc::V3<double>(e[i].x, e[i].y, e[i].z);
                }
                for (int kind = 0; kind < 2; kind++)
                {
                    double off = (double)(0.01f * (1.0f + U(rng)));
                    same64(lcsx::fp64::accd(d, e, off, kind), orc::accd(q, r, off, kind), kind ? "accd64_ee" : "accd64_vf", k);
                }
            }
        }
        // ACCD on a displaced copy
        if (k < 4000)
        {
            lcsx::f3       q[4];
            orc::V3<float> oa[4], ob[4];
            for (int i = 0; i < 4; i++)
            {
                q[i]  = p[i] + 0.5f * v();
                oa[i] = ov(p[i]);
                ob[i] = ov(q[i]);
            }
            for (int kind = 0; kind < 2; kind++)
            {
                float off = 0.01f * (1.0f + U(rng));
                same(lcsx::accd(p, q, off, kind), orc::accd(oa, ob, off, kind), kind ? "accd_ee" : "accd_vf", k);
            }
        }
        // barrier
        float d = 1e-3f * (0.5f + 0.5f * U(rng)) + 1e-6f, dh = 1e-3f;
        same(lcsx::barrier_d1(d, dh), orc::barrier_d1(d, dh), "barrier_d1", k);
        same(lcsx::barrier_d2(d, dh), orc::barrier_d2(d, dh), "barrier_d2", k);
        // Baraff-Witkin on a perturbed rest triangle
        lcsx::f3 X0 = lcsx::mk(0, 0, 0), X1 = lcsx::mk(1.0f + 0.3f * U(rng), 0.2f * U(rng), 0), X2 = lcsx::mk(0.2f * U(rng), 1.0f + 0.3f * U(rng), 0);
        float    Dm[4], oDm[4];
        lcsx::dm_inv(X0, X1, X2, Dm);
        orc::dm_inv(ov(X0), ov(X1), ov(X2), oDm);
        for (int i = 0; i < 4; i++) same(Dm[i], oDm[i], "dm_inv", k);
        lcsx::f3 x0 = X0 + 0.2f * v(), x1 = X1 + 0.2f * v(), x2 = X2 + 0.2f * v();
        float    mu = 1e3f * (1.5f + U(rng)), lmd = 1e3f * (1.5f + U(rng)), area = 0.5f;
        lcsx::FaceOut     fo;
        orc::FaceGH<float> og;
        lcsx::bw_face(x0, x1, x2, Dm, mu, lmd, area, fo);
        orc::bw_face(ov(x0), ov(x1), ov(x2), oDm, mu, lmd, area, og);
        for (int a = 0; a < 3; a++)
        {
            same(fo.g[a].x, og.g[a].x, "bw_grad", k);
            same(fo.g[a].y, og.g[a].y, "bw_grad", k);
            same(fo.g[a].z, og.g[a].z, "bw_grad", k);
            for (int b = a; b < 3; b++)
                for (int r = 0; r < 3; r++)
                    for (int s = 0; s < 3; s++) same(fo.H[a][b].m[3 * r + s], og.H[3 * a + r][3 * b + s], "bw_hess", k);
        }
        // dihedral bending
        lcsx::f3 h0 = lcsx::mk(0, 0, 0), h1 = lcsx::mk(1, 0, 0), h2 = lcsx::mk(0.3f, 1, 0) + 0.3f * v(),
                 h3 = lcsx::mk(0.6f, -1, 0) + 0.3f * v();
        same(lcsx::dihedral_angle(h0, h1, h2, h3), orc::dihedral_angle(ov(h0), ov(h1), ov(h2), ov(h3)), "dihedral", k);
        lcsx::f3       g[4];
        orc::V3<float> og4[4];
        lcsx::dihedral_grad(h0, h1, h2, h3, g);
        orc::dihedral_grad(ov(h0), ov(h1), ov(h2), ov(h3), og4);
        for (int i = 0; i < 4; i++)
        {
            same(g[i].x, og4[i].x, "dihedral_grad", k);
            same(g[i].y, og4[i].y, "dihedral_grad", k);
            same(g[i].z, og4[i].z, "dihedral_grad", k);
        }
    }
    if (g_fail)
    {
        std::printf("FAIL %d mismatches\n", g_fail);
        return 1;
    }
    std::printf("OK %d random cases, %d ACCD stall cases\n", N, stalls);
    return 0;
}
