// Drives the C++ host mirror (code_mpm_auxetics_amd/host/PhysicsEngine.h) the
// way the reference's game loop drives PhysicsEngine
// (GraphicsEngine_gameLoop.cpp:26-31): a small elastic block in free flight,
// runSimulation_Classic_SinglePass_MP(10 * dt) until finished, then prints
// the mean position and velocity, then the getScript header and first row.
// Exit code 3 and "no device" when no GPU is visible (the constructor throws
// MPM_E_NO_DEVICE).
#include <cstdio>

#include "../code_mpm_auxetics_amd/host/PhysicsEngine.h"

int main()
{
    const double h = 0.001;
    mpm_config cfg{};
    for (int a = 0; a < 3; ++a) {
        cfg.length[a] = 16 * h;
        cfg.cells[a] = 16;
    }
    cfg.dt_max = 1.0e-6;
    cfg.time_end = 40.0e-6;
    cfg.precision = MPM_FP64;
    try {
        mpm_host::PhysicsEngine eng(cfg);
        std::vector<mpm_host::MaterialPoint> pts;
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j)
                for (int k = 0; k < 4; ++k) {
                    mpm_host::MaterialPoint p;
                    p.d3_Position = {(6.0 + 0.5 * i + 0.25) * h + 1.0e-6 * k, (6.0 + 0.5 * j + 0.25) * h, (6.0 + 0.5 * k + 0.25) * h};
                    p.d3_Velocity = {1.5, -0.5, 0.25};
                    p.d_Volume = p.d_Volume_Initial = 0.125 * h * h * h;
                    p.d_Mass = 1000.0 * p.d_Volume;
                    p.d_ElasticModulus = 1.0e7;
                    p.d_PoissonRatio = 0.3;
                    p.i_MaterialType = mpm_host::_ELASTIC;
                    pts.push_back(p);
                }
        eng.setMaterialPoints(pts);
        int steps = 0;
        while (!eng.runSimulation_Classic_SinglePass_MP(10 * eng.getTime_Increment())) ++steps;
        double mx[3] = {0, 0, 0}, mv[3] = {0, 0, 0};
        std::vector<mpm_host::MaterialPoint *> mp = eng.getMaterialPoints();
        for (auto *p : mp)
            for (int a = 0; a < 3; ++a) {
                mx[a] += p->d3_Position[a] / mp.size();
                mv[a] += p->d3_Velocity[a] / mp.size();
            }
        std::printf("%d %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", (int)mp.size(), eng.getTime_Current(), mx[0], mx[1],
                    mx[2], mv[0], mv[1], mv[2]);
        std::printf("%s%s", mpm_host::getScript(nullptr).c_str(), mpm_host::getScript(mp[0]).c_str());
        return 0;
    } catch (const mpm_host::Error &e) {
        std::printf("%s\n", e.what());
        return e.code == MPM_E_NO_DEVICE ? 3 : 1;
    }
}
