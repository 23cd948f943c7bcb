// Host-compiled check: the product's per-element physics (luisacomputesimulator_amd/csrc/lcsx_physics.h,
// the same source the HIP kernels run) against the oracle restatement (oracle/orc_*.h), bit for bit
// in FP32 on random inputs. Prints "OK <cases>" or the first mismatch and exits non-zero.
#include <cstdio>
#include <cstring>
#include <random>
#include "../../luisacomputesimulator_amd/csrc/lcsx_physics.h"
#include "../../oracle/orc_collision.h"
#include "../../oracle/orc_energy.h"
#include "../../oracle/orc_lbvh.h"

static int g_fail = 0;
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

int main()
{
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
    std::printf("OK %d\n", N);
    return 0;
}
