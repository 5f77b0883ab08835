// Drives code_mpm_auxetics_amd/host/GraphicsEngine.h the way the reference's
// game loop calls drawGame (GraphicsEngine_gameLoop.cpp:26-31) on a small
// world with known stresses, plastic strains and flags, and prints every
// draw instance of the three canvases:
//   line 1: points, MAIN instances, J2_STRESS instances, J2_PLASTICSTRAIN instances
//   then per point: x y z size, then r g b a of MAIN, J2_STRESS, J2_PLASTICSTRAIN
//   then one line per fixed-node instance of MAIN: x y z size r g b a
// Exit code 3 and "no device" when no GPU is visible.
#include <cstdio>

#include "../code_mpm_auxetics_amd/host/GraphicsEngine.h"

int main()
{
    const double h = 0.001;
    const int nc = 8;
    mpm_config cfg{};
    for (int a = 0; a < 3; ++a) {
        cfg.length[a] = nc * h;
        cfg.cells[a] = nc;
    }
    cfg.dt_max = 1.0e-7;
    cfg.time_end = 1.0e-6;
    cfg.precision = MPM_FP64;
    // bottom nodes fixed in y, the corner column in all three axes
    std::vector<uint8_t> fixed((size_t)(nc + 1) * (nc + 1) * (nc + 1), 0);
    for (int k = 0; k <= nc; ++k)
        for (int i = 0; i <= nc; ++i) fixed[i + (nc + 1) * (0 + (nc + 1) * k)] = MPM_FIX_Y;
    fixed[0] = MPM_FIX_X | MPM_FIX_Y | MPM_FIX_Z;
    cfg.node_fixed = fixed.data();
    try {
        mpm_host::PhysicsEngine eng(cfg);
        std::vector<mpm_host::MaterialPoint> pts;
        int id = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 2; ++k, ++id) {
                    mpm_host::MaterialPoint p;
                    p.d3_Position = {(2.25 + 0.5 * i) * h, (2.25 + 0.5 * j) * h, (2.25 + 0.5 * k) * h};
                    p.d_Volume = p.d_Volume_Initial = (0.125 + 0.01 * id) * h * h * h;
                    p.d_Mass = 1000.0 * p.d_Volume;
                    p.d_ElasticModulus = 1.0e9 * (1.0 + 0.1 * (id % 3));
                    p.d_PoissonRatio = 0.3;
                    p.d_YieldStress = 2.0e6;
                    p.i_MaterialType = mpm_host::_ELASTIC;
                    const double s = 1.0e5 * (id + 1);
                    const double st[6] = {s, -0.5 * s, 0.25 * s, 0.1 * s, -0.2 * s, 0.05 * s * (id % 5)};
                    for (int q = 0; q < 6; ++q) {
                        p.d6_Stress[q] = st[q];
                        p.d6_Strain_Plastic[q] = 1.0e-5 * (id % 7) * (q + 1) * (q % 2 ? -1.0 : 1.0);
                    }
                    p.b_Mark_Stress = (id % 5) == 3;
                    p.b_Surface = (id % 4) == 1;
                    p.b_DisplacementControl = id == 7;
                    pts.push_back(p);
                }
        eng.setMaterialPoints(pts);
        mpm_host::GraphicsEngine ge;
        ge.setPhysicsEngineReference(&eng);
        ge.drawGame();
        const auto &m = ge.instances(mpm_host::enum_Canvas::MAIN);
        const auto &s = ge.instances(mpm_host::enum_Canvas::J2_STRESS);
        const auto &p = ge.instances(mpm_host::enum_Canvas::J2_PLASTICSTRAIN);
        std::printf("%d %d %d %d\n", (int)pts.size(), (int)m.size(), (int)s.size(), (int)p.size());
        for (size_t i = 0; i < pts.size(); ++i)
            std::printf("%.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", m[i].x,
                        m[i].y, m[i].z, m[i].size, m[i].r, m[i].g, m[i].b, m[i].a, s[i].r, s[i].g, s[i].b, s[i].a,
                        p[i].r, p[i].g, p[i].b, p[i].a);
        for (size_t i = pts.size(); i < m.size(); ++i)
            std::printf("%.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", m[i].x, m[i].y, m[i].z, m[i].size, m[i].r, m[i].g,
                        m[i].b, m[i].a);
        return 0;
    } catch (const mpm_host::Error &e) {
        std::printf("%s\n", e.what());
        return e.code == MPM_E_NO_DEVICE ? 3 : 1;
    }
}
