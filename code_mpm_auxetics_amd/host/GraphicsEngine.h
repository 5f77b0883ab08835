// GraphicsEngine.h -- the reference GraphicsEngine's drawing over device
// readbacks, on top of the C++ PhysicsEngine mirror (PhysicsEngine.h).
//
// Reference: GraphicsEngine_drawGame.cpp:4-260 (GraphicsEngine::drawGame),
// GraphicsEngine.h:77-82 (enum_Canvas), Definitions.h:15-20 (colours).  The
// reference issues one draw call per material point and evaluates
// calculateState_J2 on the host for each point of the two J2 canvases.  Here a
// frame is one readback (positions and volumes; the J2 of stress and plastic
// strain are computed on the device, MPM_F_J2) turned into per-point draw
// instances -- position, size and colour, by the reference's rules -- for all
// three canvases.  With MPM_WITH_GL defined (OpenGL 3.3 core, in a context the
// caller owns, as the reference's SDL window) InstancedRenderer draws a
// canvas's instances with one instanced call; without it the header needs no
// GL at all and the instances feed any other renderer (the headless BMP
// snapshots of code_mpm_auxetics_amd/run.py use the same rules).
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "PhysicsEngine.h"

#ifdef MPM_WITH_GL
#include <GL/gl.h>
#include <GL/glext.h>
#include <stdexcept>
#include <string>
#endif

namespace mpm_host {

// GraphicsEngine.h:77-82
enum class enum_Canvas : int { MAIN = 0, J2_STRESS, J2_PLASTICSTRAIN, COUNT };

struct Color {
    float r, g, b, a;
};
// Definitions.h:15-20
constexpr Color kRed{1.0f, 0.0f, 0.0f, 1.0f}, kGreen{0.0f, 1.0f, 0.0f, 1.0f}, kBlue{0.0f, 0.0f, 1.0f, 1.0f},
    kWhite{1.0f, 1.0f, 1.0f, 1.0f}, kGray{0.5f, 0.5f, 0.5f, 1.0f};

// One drawn particle: the reference's Transformation (position, uniform
// scale) and objectColor uniform of a draw call.
struct DrawInstance {
    float x, y, z, size;
    float r, g, b, a;
};

// Material point colour on the MAIN canvas (GraphicsEngine_drawGame.cpp:73-83):
// red, green if the point is stress-marked, blue if it is a surface point,
// white if it is displacement-controlled (later rules win).
inline Color mainColor(bool mark_stress, bool surface, bool displacement_control)
{
    Color c = kRed;
    if (mark_stress) c = kGreen;
    if (surface) c = kBlue;
    if (displacement_control) c = kWhite;
    return c;
}

// Colour of a J2 canvas (GraphicsEngine_drawGame.cpp:172, 257):
// (1 - J2 / J2max) * blue + J2 / J2max * red, in float as the reference
// (glm::vec4 with float weights).
inline Color j2Color(float j2, float j2_max)
{
    const float w = j2 / j2_max, u = 1.0f - w;
    return Color{u * kBlue.r + w * kRed.r, u * kBlue.g + w * kRed.g, u * kBlue.b + w * kRed.b,
                 u * kBlue.a + w * kRed.a};
}

// Material point size (GraphicsEngine_drawGame.cpp:72): 2 * 0.4 * V^(1/3)
inline float pointSize(double volume) { return (float)(2.0 * 0.4 * std::pow(volume, 1.0 / 3.0)); }

// Fixed grid nodes on the MAIN canvas (GraphicsEngine_drawGame.cpp:98-121):
// gray, size 1e-5, nodes with no fixed axis skipped.
constexpr float kFixedNodeSize = 0.00001f;

class GraphicsEngine {
  public:
    // GraphicsEngine_setPhysicsEngineReference.cpp
    void setPhysicsEngineReference(PhysicsEngine *engine) { mpm_PhysicsEngine = engine; }

    // The three canvases of drawGame() for the engine's current state: one
    // readback of positions and volumes, one of the device J2 values, and the
    // grid's fixed nodes (their positions do not change; read once).
    void drawGame()
    {
        for (auto &v : canvas_) v.clear();
        if (!mpm_PhysicsEngine) return;
        std::vector<MaterialPoint *> mp = mpm_PhysicsEngine->getMaterialPoints();
        if (!fixed_ready_) {
            for (GridPoint *gp : mpm_PhysicsEngine->getGridPoints())
                if (gp->b3_Fixed[0] || gp->b3_Fixed[1] || gp->b3_Fixed[2])
                    fixed_nodes_.push_back({(float)gp->d3_Position[0], (float)gp->d3_Position[1],
                                            (float)gp->d3_Position[2], kFixedNodeSize, kGray.r, kGray.g, kGray.b,
                                            kGray.a});
            fixed_ready_ = true;
        }
        std::vector<double> j2s, j2p;
        mpm_PhysicsEngine->getState_J2(j2s, j2p);
        // J2_STRESS: the largest J2 of the points, at least 1e-12 (:161-170);
        // J2_PLASTICSTRAIN: yield stress over Young's modulus of point 0 (:245)
        float j2s_max = 1.0e-12f;
        for (double v : j2s) j2s_max = std::max(j2s_max, (float)v);
        const float j2p_max =
            mp.empty() ? 1.0f : (float)(mp[0]->d_YieldStress / mp[0]->d_ElasticModulus);
        auto &main = canvas_[(int)enum_Canvas::MAIN];
        auto &cs = canvas_[(int)enum_Canvas::J2_STRESS];
        auto &cp = canvas_[(int)enum_Canvas::J2_PLASTICSTRAIN];
        main.reserve(mp.size() + fixed_nodes_.size());
        cs.reserve(mp.size());
        cp.reserve(mp.size());
        for (size_t i = 0; i < mp.size(); ++i) {
            const MaterialPoint *p = mp[i];
            const float x = (float)p->d3_Position[0], y = (float)p->d3_Position[1], z = (float)p->d3_Position[2];
            const float s = pointSize(p->d_Volume);
            const Color c0 = mainColor(p->b_Mark_Stress, p->b_Surface, p->b_DisplacementControl);
            main.push_back({x, y, z, s, c0.r, c0.g, c0.b, c0.a});
            const Color c1 = j2Color((float)j2s[i], j2s_max);
            cs.push_back({x, y, z, s, c1.r, c1.g, c1.b, c1.a});
            // a zero scale (elastic point 0: yield stress 0) would make the
            // weight 0/0; nothing yields there, so the points are drawn blue
            const Color c2 = j2p_max > 0.0f ? j2Color((float)j2p[i], j2p_max) : kBlue;
            cp.push_back({x, y, z, s, c2.r, c2.g, c2.b, c2.a});
        }
        main.insert(main.end(), fixed_nodes_.begin(), fixed_nodes_.end());
    }

    const std::vector<DrawInstance> &instances(enum_Canvas c) const { return canvas_[(int)c]; }

  private:
    PhysicsEngine *mpm_PhysicsEngine = nullptr;
    std::vector<DrawInstance> canvas_[(int)enum_Canvas::COUNT];
    std::vector<DrawInstance> fixed_nodes_;
    bool fixed_ready_ = false;
};

#ifdef MPM_WITH_GL
// One instanced draw of a canvas: every instance is a screen-facing square of
// its size at its position, in its colour, under the caller's view-projection
// matrix (column-major, as the reference's Camera::getViewProjection).  Needs
// a current OpenGL 3.3 core context; GL entry points come from the system GL
// library (GL_GLEXT_PROTOTYPES) or from the caller's loader.
class InstancedRenderer {
  public:
    InstancedRenderer()
    {
        static const char *vs =
            "#version 330 core\n"
            "layout(location = 0) in vec2 corner;\n"
            "layout(location = 1) in vec4 posSize;\n"
            "layout(location = 2) in vec4 color;\n"
            "uniform mat4 viewProjection;\n"
            "out vec4 vColor;\n"
            "void main() {\n"
            "    vec4 c = viewProjection * vec4(posSize.xyz, 1.0);\n"
            "    vec4 e = viewProjection * vec4(posSize.xyz + vec3(posSize.w, 0.0, 0.0), 1.0);\n"
            "    float r = length(e.xy / e.w - c.xy / c.w);\n"
            "    gl_Position = vec4(c.xy / c.w + corner * r, c.z / c.w, 1.0);\n"
            "    vColor = color;\n"
            "}\n";
        static const char *fs =
            "#version 330 core\n"
            "in vec4 vColor;\n"
            "out vec4 fragColor;\n"
            "void main() { fragColor = vColor; }\n";
        prog_ = glCreateProgram();
        const GLuint v = compile(GL_VERTEX_SHADER, vs), f = compile(GL_FRAGMENT_SHADER, fs);
        glAttachShader(prog_, v);
        glAttachShader(prog_, f);
        glLinkProgram(prog_);
        GLint ok = 0;
        glGetProgramiv(prog_, GL_LINK_STATUS, &ok);
        glDeleteShader(v);
        glDeleteShader(f);
        if (!ok) throw std::runtime_error("instanced point program did not link");
        loc_vp_ = glGetUniformLocation(prog_, "viewProjection");
        static const float quad[8] = {-1.0f, -1.0f, 1.0f, -1.0f, -1.0f, 1.0f, 1.0f, 1.0f};
        glGenVertexArrays(1, &vao_);
        glBindVertexArray(vao_);
        glGenBuffers(1, &quad_);
        glBindBuffer(GL_ARRAY_BUFFER, quad_);
        glBufferData(GL_ARRAY_BUFFER, sizeof quad, quad, GL_STATIC_DRAW);
        glEnableVertexAttribArray(0);
        glVertexAttribPointer(0, 2, GL_FLOAT, GL_FALSE, 0, nullptr);
        glGenBuffers(1, &inst_);
        glBindBuffer(GL_ARRAY_BUFFER, inst_);
        glEnableVertexAttribArray(1);
        glVertexAttribPointer(1, 4, GL_FLOAT, GL_FALSE, sizeof(DrawInstance), nullptr);
        glVertexAttribDivisor(1, 1);
        glEnableVertexAttribArray(2);
        glVertexAttribPointer(2, 4, GL_FLOAT, GL_FALSE, sizeof(DrawInstance),
                              reinterpret_cast<const void *>(4 * sizeof(float)));
        glVertexAttribDivisor(2, 1);
        glBindVertexArray(0);
    }
    ~InstancedRenderer()
    {
        glDeleteBuffers(1, &inst_);
        glDeleteBuffers(1, &quad_);
        glDeleteVertexArrays(1, &vao_);
        glDeleteProgram(prog_);
    }
    InstancedRenderer(const InstancedRenderer &) = delete;
    InstancedRenderer &operator=(const InstancedRenderer &) = delete;

    // into the bound render target (the reference binds its canvas texture)
    void draw(const std::vector<DrawInstance> &inst, const float viewProjection[16])
    {
        if (inst.empty()) return;
        glUseProgram(prog_);
        glUniformMatrix4fv(loc_vp_, 1, GL_FALSE, viewProjection);
        glBindVertexArray(vao_);
        glBindBuffer(GL_ARRAY_BUFFER, inst_);
        glBufferData(GL_ARRAY_BUFFER, (GLsizeiptr)(inst.size() * sizeof(DrawInstance)), inst.data(), GL_STREAM_DRAW);
        glDrawArraysInstanced(GL_TRIANGLE_STRIP, 0, 4, (GLsizei)inst.size());
        glBindVertexArray(0);
        glUseProgram(0);
    }

  private:
    GLuint prog_ = 0, vao_ = 0, quad_ = 0, inst_ = 0;
    GLint loc_vp_ = -1;
    static GLuint compile(GLenum type, const char *src)
    {
        const GLuint s = glCreateShader(type);
        glShaderSource(s, 1, &src, nullptr);
        glCompileShader(s);
        GLint ok = 0;
        glGetShaderiv(s, GL_COMPILE_STATUS, &ok);
        if (!ok) {
            char log[512] = {0};
            glGetShaderInfoLog(s, sizeof log, nullptr, log);
            glDeleteShader(s);
            throw std::runtime_error(std::string("shader: ") + log);
        }
        return s;
    }
};
#endif

}  // namespace mpm_host
