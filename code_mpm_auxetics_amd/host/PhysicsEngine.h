// PhysicsEngine.h -- C++ host mirror of the reference PhysicsEngine's step
// interface, on top of the C ABI in include/mpm_auxetics.h.
//
// Reference: PhysicsEngine.h:34-93 (class PhysicsEngine), MaterialPoint.h:18-60
// (class MaterialPoint).  Method names, argument meaning and return values are
// the reference's; the state lives on the GPU and the MaterialPoint objects
// returned by getMaterialPoints() are host mirrors refreshed by a readback.
// glm is not used: dvec3 is std::array<double, 3>, dmat3 is a column-major
// std::array<double, 9> (m[col][row] at [3*col+row], as glm stores it).
//
// Errors: every C ABI status other than MPM_OK throws mpm_host::Error with
// the library's message (the reference exits through fatalError instead).
#pragma once

#include <array>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpm_auxetics.h"

namespace mpm_host {

enum MaterialType { _ELASTIC, _NEOHOOKEAN, _VISCOELASTIC, _PLASTIC, _VONMISESHARDENING, _GASS };

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error("mpm error " + std::to_string(c) + ": " + m), code(c) {}
};

// MaterialPoint.h:18-60 fields (monitor flags and viscosity are carried but unused by the step)
class MaterialPoint {
  public:
    bool b_DisplacementControl = false;
    bool b_Mark_Force = false;
    bool b_Mark_Stress = false;
    unsigned int i_ID = 0;
    unsigned int i_MaterialType = _ELASTIC;
    double d_Mass = 0.0, d_Volume = 0.0, d_Volume_Initial = 0.0;
    double d_ElasticModulus = 0.0, d_PoissonRatio = 0.0, d_YieldStress = 0.0, d_Viscosity = 0.0;
    bool b_Surface = false;
    double d_Kernel = 0.0;
    std::array<double, 3> d3_Kernel_Gradient{};
    std::array<double, 3> d3_Position{}, d3_Velocity{}, d3_Force_External{};
    std::array<double, 9> d33_DeformationGradient{1, 0, 0, 0, 1, 0, 0, 0, 1};
    double d6_Strain[6] = {0, 0, 0, 0, 0, 0};
    double d6_Stress[6] = {0, 0, 0, 0, 0, 0};
    double d6_Strain_Plastic[6] = {0, 0, 0, 0, 0, 0};
    double d_BackStress_Isotropic = 0.0;
    double d_Hardening_Isotropic_C0 = 0.0;
    double d_Hardening_Isotropic_C1 = 0.0;
};

// GridPoint.h:6-33 fields, filled by getGridPoints() (mpm_grid_nodes); the
// kernel grid (d_Kernel, d3_Kernel_Gradient) of the contact step is carried
// too.  b3_Fixed holds the fixity per axis.
class GridPoint {
  public:
    bool b_Active = false;
    std::array<int, 3> i3_Index{};
    std::array<bool, 3> b3_Fixed{};
    std::array<double, 3> d3_Position{};
    double d_Mass = 0.0;
    std::array<double, 3> d3_Velocity{}, d3_Force{}, d3_Force_Temp{};
    double d_Kernel = 0.0;
    std::array<double, 3> d3_Kernel_Gradient{};
};

// MaterialPoint_Factory::getScript (MaterialPoint_Factory_getScript.cpp:4-28):
// the header line for nullptr, else "ID mass position_x position_y position_z",
// tab-separated with a trailing tab; doubles as Script(d, 8) writes them
// (Definitions.h:43-58: scientific, with a sign, 8 digits after the point).
inline std::string getScript(const MaterialPoint *p)
{
    if (p == nullptr) return "ID\tmass\t\t\tposition_x\t\tposition_y\t\tposition_z\t\t\n";
    char buf[160];
    std::snprintf(buf, sizeof buf, "%u\t%+.8e\t%+.8e\t%+.8e\t%+.8e\t\n", p->i_ID, p->d_Mass, p->d3_Position[0],
                  p->d3_Position[1], p->d3_Position[2]);
    return buf;
}

class PhysicsEngine {
  public:
    // The protected members initializeWorld_* sets (PhysicsEngine.h:68-93)
    // come in through mpm_config; node fixity is cfg.node_fixed.
    explicit PhysicsEngine(const mpm_config &cfg)
    {
        mpm_handle *h = nullptr;
        check(mpm_create(&cfg, &h), nullptr);
        h_.reset(h);
        for (int a = 0; a < 3; ++a) cells_[a] = cfg.cells[a];
        t_end_ = cfg.time_end;
        track_strain_ = cfg.track_strain != 0;
    }

    // The point part of initializeWorld_* (…PolygonCell.cpp:137-191): every
    // point, displacement-controlled ones flagged by b_DisplacementControl.
    void setMaterialPoints(const std::vector<MaterialPoint> &pts)
    {
        const size_t n = pts.size();
        std::vector<double> x(3 * n), v(3 * n), fe(3 * n), F(9 * n), eps(6 * n), sig(6 * n), ep(6 * n);
        std::vector<double> al(n), m(n), vol(n), vol0(n), E(n), nu(n), sy(n), c0(n), c1(n);
        std::vector<int32_t> ty(n);
        std::vector<uint8_t> dc(n), surf(n);
        for (size_t i = 0; i < n; ++i) {
            const MaterialPoint &p = pts[i];
            for (int a = 0; a < 3; ++a) {
                x[3 * i + a] = p.d3_Position[a];
                v[3 * i + a] = p.d3_Velocity[a];
                fe[3 * i + a] = p.d3_Force_External[a];
            }
            for (int a = 0; a < 9; ++a) F[9 * i + a] = p.d33_DeformationGradient[a];
            for (int a = 0; a < 6; ++a) {
                eps[6 * i + a] = p.d6_Strain[a];
                sig[6 * i + a] = p.d6_Stress[a];
                ep[6 * i + a] = p.d6_Strain_Plastic[a];
            }
            al[i] = p.d_BackStress_Isotropic;
            m[i] = p.d_Mass;
            vol[i] = p.d_Volume;
            vol0[i] = p.d_Volume_Initial;
            E[i] = p.d_ElasticModulus;
            nu[i] = p.d_PoissonRatio;
            sy[i] = p.d_YieldStress;
            c0[i] = p.d_Hardening_Isotropic_C0;
            c1[i] = p.d_Hardening_Isotropic_C1;
            ty[i] = (int32_t)p.i_MaterialType;
            dc[i] = p.b_DisplacementControl ? 1 : 0;
            surf[i] = p.b_Surface ? 1 : 0;
        }
        mpm_points P{};
        P.n = (int32_t)n;
        P.x = x.data(); P.v = v.data(); P.fext = fe.data(); P.F = F.data();
        P.strain = eps.data(); P.stress = sig.data(); P.eps_p = ep.data(); P.alpha = al.data();
        P.mass = m.data(); P.vol = vol.data(); P.vol0 = vol0.data();
        P.E = E.data(); P.nu = nu.data(); P.sy = sy.data(); P.C0 = c0.data(); P.C1 = c1.data();
        P.type = ty.data(); P.dc = dc.data(); P.surface = surf.data();
        check(mpm_upload_points(h_.get(), &P), h_.get());
        mirror_ = pts;
        stale_ = false;
    }

    // PhysicsEngine.h:45-46: advance dTimeIncrement_Total in steps of at most
    // the maximum increment; 1 once d_Time >= d_TimeEnd, else 0.
    int runSimulation_Classic_SinglePass_MP(double dTimeIncrement_Total) { return step(dTimeIncrement_Total, MPM_PATH_PLAIN); }
    int runSimulation_Classic_SinglePass_MP_Contact(double dTimeIncrement_Total) { return step(dTimeIncrement_Total, MPM_PATH_CONTACT); }

    // PhysicsEngine.h:57-60
    double getTime_Runtime() const { double ms = 0; mpm_last_step_ms(h_.get(), &ms); return ms * 1e-3; }
    double getTime_Current() const { double t, te, dt; mpm_get_time(h_.get(), &t, &te, &dt); return t; }
    double getTime_End() const { return t_end_; }
    double getTime_Increment() const { double t, te, dt; mpm_get_time(h_.get(), &t, &te, &dt); return dt; }

    // PhysicsEngine.h:62-64: the caller reads the points' current state
    unsigned int getCount_MaterialPoint() const { return (unsigned int)mirror_.size(); }
    std::vector<MaterialPoint *> getMaterialPoints()
    {
        refresh();
        std::vector<MaterialPoint *> out;
        out.reserve(mirror_.size());
        for (MaterialPoint &p : mirror_) out.push_back(&p);
        return out;
    }

    // PhysicsEngine.h:65-66: the grid nodes (x fastest, GridPoint_Factory.h:23-32).
    // On the plain path the step keeps no node grid, so these are the next
    // step's nodes (the P2G of the current points and the grid update); after
    // a contact step, that step's first field and its kernel grid.
    std::vector<GridPoint *> getGridPoints()
    {
        refreshNodes();
        std::vector<GridPoint *> out;
        out.reserve(nodes_.size());
        for (GridPoint &g : nodes_) out.push_back(&g);
        return out;
    }
    std::vector<GridPoint *> getGridPoints_Kernel() { return getGridPoints(); }

    // ConstitutiveRelation::calculateState_J2 (ConstitutiveRelation_calculateState_J2.cpp:4-26)
    // of every point's stress and plastic strain, evaluated on the device
    // (MPM_F_J2), in point order: what the drawing colours by
    // (GraphicsEngine_drawGame.cpp:164,178,252)
    void getState_J2(std::vector<double> &j2_stress, std::vector<double> &j2_plastic)
    {
        const size_t n = mirror_.size();
        std::vector<double> j2(2 * n);
        mpm_points P{};
        P.n = (int32_t)n;
        P.j2 = j2.data();
        check(mpm_readback(h_.get(), MPM_F_J2, &P), h_.get());
        j2_stress.resize(n);
        j2_plastic.resize(n);
        for (size_t i = 0; i < n; ++i) {
            j2_stress[i] = j2[2 * i];
            j2_plastic[i] = j2[2 * i + 1];
        }
    }

    // reportConsole sums (PhysicsEngine.cpp:33-97): grid mass, reaction
    // (Force_y), sum m|v|^2, momentum, non-finite count, time, cycle, ...
    std::array<double, 12> checks() const
    {
        std::array<double, 12> c{};
        check(mpm_checks(h_.get(), c.data()), h_.get());
        return c;
    }

    // PhysicsEngine::reportConsole(std::string) (PhysicsEngine.cpp:33-120):
    // prints (when verbose) and logs the line; returns it as well
    std::string reportConsole(const std::string &sDescription = "")
    {
        // the fields after the description take well under 512 bytes
        std::vector<char> buf(sDescription.size() + 512);
        int32_t len = 0;
        check(mpm_report(h_.get(), sDescription.c_str(), buf.data(), (int32_t)buf.size(), &len), h_.get());
        return std::string(buf.data());
    }

    // the log file of every report line (the reference's ./bmp/log.txt,
    // Definitions.h:22); empty: none
    void setLogFile(const std::string &path) { check(mpm_set_log(h_.get(), path.c_str()), h_.get()); }

  private:
    struct Del {
        void operator()(mpm_handle *h) const { mpm_destroy(h); }
    };
    std::unique_ptr<mpm_handle, Del> h_;
    std::vector<MaterialPoint> mirror_;
    std::vector<GridPoint> nodes_;
    bool stale_ = true;
    std::array<int, 3> cells_{};
    double t_end_ = 0.0;
    bool track_strain_ = false;

    static void check(int rc, const mpm_handle *h)
    {
        if (rc != MPM_OK) throw Error(rc, mpm_last_error(h));
    }
    int step(double dt_total, int path)
    {
        int32_t fin = 0;
        check(mpm_step(h_.get(), dt_total, path, &fin), h_.get());
        stale_ = true;
        return fin;
    }
    void refreshNodes()
    {
        const size_t nx = (size_t)cells_[0] + 1, ny = (size_t)cells_[1] + 1, nz = (size_t)cells_[2] + 1;
        const size_t nn = nx * ny * nz;
        std::vector<uint8_t> act(nn), fix(nn);
        std::vector<double> x(3 * nn), m(nn), v(3 * nn), f(3 * nn), ft(3 * nn), k(nn), kg(3 * nn);
        mpm_nodes N{};
        N.n = (int64_t)nn;
        N.active = act.data(); N.fixed = fix.data(); N.x = x.data(); N.mass = m.data(); N.v = v.data();
        N.f = f.data(); N.f_temp = ft.data(); N.kernel = k.data(); N.kernel_grad = kg.data();
        check(mpm_grid_nodes(h_.get(), &N), h_.get());
        nodes_.resize(nn);
        for (size_t I = 0; I < nn; ++I) {
            GridPoint &g = nodes_[I];
            g.b_Active = act[I] != 0;
            g.i3_Index = {(int)(I % nx), (int)((I / nx) % ny), (int)(I / (nx * ny))};
            for (int a = 0; a < 3; ++a) {
                g.b3_Fixed[a] = ((fix[I] >> a) & 1u) != 0;
                g.d3_Position[a] = x[3 * I + a];
                g.d3_Velocity[a] = v[3 * I + a];
                g.d3_Force[a] = f[3 * I + a];
                g.d3_Force_Temp[a] = ft[3 * I + a];
                g.d3_Kernel_Gradient[a] = kg[3 * I + a];
            }
            g.d_Mass = m[I];
            g.d_Kernel = k[I];
        }
    }
    void refresh()
    {
        if (!stale_) return;
        const size_t n = mirror_.size();
        std::vector<double> x(3 * n), v(3 * n), F(9 * n), sig(6 * n), eps(6 * n), ep(6 * n), al(n), vol(n);
        mpm_points P{};
        P.n = (int32_t)n;
        P.x = x.data(); P.v = v.data(); P.F = F.data(); P.stress = sig.data(); P.strain = eps.data();
        P.eps_p = ep.data(); P.alpha = al.data(); P.vol = vol.data();
        const uint32_t mask = MPM_F_X | MPM_F_V | MPM_F_F | MPM_F_STRESS | MPM_F_EPSP | MPM_F_ALPHA | MPM_F_VOL |
                              (track_strain_ ? MPM_F_STRAIN : 0u);
        check(mpm_readback(h_.get(), mask, &P), h_.get());
        for (size_t i = 0; i < n; ++i) {
            MaterialPoint &p = mirror_[i];
            for (int a = 0; a < 3; ++a) {
                p.d3_Position[a] = x[3 * i + a];
                p.d3_Velocity[a] = v[3 * i + a];
            }
            for (int a = 0; a < 9; ++a) p.d33_DeformationGradient[a] = F[9 * i + a];
            for (int a = 0; a < 6; ++a) {
                p.d6_Stress[a] = sig[6 * i + a];
                if (track_strain_) p.d6_Strain[a] = eps[6 * i + a];
                p.d6_Strain_Plastic[a] = ep[6 * i + a];
            }
            p.d_BackStress_Isotropic = al[i];
            p.d_Volume = vol[i];
        }
        stale_ = false;
    }
};

}  // namespace mpm_host
