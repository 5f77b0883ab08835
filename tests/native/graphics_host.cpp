// CPU checks of the drawing rules in code_mpm_auxetics_amd/host/GraphicsEngine.h
// against GraphicsEngine_drawGame.cpp: MAIN-canvas colour precedence
// (:73-83), the blue-to-red J2 blend (:172, :257) and the point size (:72).
// Exit code 0 when every check holds; the failing check is printed.
#include <cmath>
#include <cstdio>

#include "../../code_mpm_auxetics_amd/host/GraphicsEngine.h"

using namespace mpm_host;

static int fails = 0;
static void expect(bool ok, const char *what)
{
    if (!ok) {
        std::printf("FAIL %s\n", what);
        ++fails;
    }
}
static bool same(const Color &a, const Color &b) { return a.r == b.r && a.g == b.g && a.b == b.b && a.a == b.a; }

int main()
{
    // later rules win: red < stress mark < surface < displacement control
    expect(same(mainColor(false, false, false), kRed), "plain point is red");
    expect(same(mainColor(true, false, false), kGreen), "stress-marked point is green");
    expect(same(mainColor(true, true, false), kBlue), "surface wins over the mark");
    expect(same(mainColor(true, true, true), kWhite), "platen point is white");
    expect(same(mainColor(false, false, true), kWhite), "platen point alone is white");
    // blend endpoints and midpoint; weights in float
    expect(same(j2Color(0.0f, 2.0f), Color{0.0f, 0.0f, 1.0f, 1.0f}), "J2 = 0 is blue");
    expect(same(j2Color(2.0f, 2.0f), Color{1.0f, 0.0f, 0.0f, 1.0f}), "J2 = max is red");
    expect(same(j2Color(1.0f, 2.0f), Color{0.5f, 0.0f, 0.5f, 1.0f}), "J2 = max / 2 is purple");
    {
        const float j2 = 0.3f, mx = 0.7f, w = j2 / mx;
        const Color c = j2Color(j2, mx);
        expect(c.r == w && c.b == 1.0f - w && c.g == 0.0f && c.a == (1.0f - w) + w, "blend weights");
    }
    // plastic strain beyond the yield strain of point 0 is not clamped
    expect(j2Color(3.0f, 1.0f).r == 3.0f && j2Color(3.0f, 1.0f).b == -2.0f, "no clamp above the maximum");
    // size 2 * 0.4 * V^(1/3)
    expect(std::fabs(pointSize(1.0e-9) - 0.8e-3f) <= 1e-9f, "size of a 1 mm^3 point");
    expect(pointSize(8.0) == (float)(0.8 * std::pow(8.0, 1.0 / 3.0)), "size formula");
    // a drawGame without an engine draws nothing
    GraphicsEngine ge;
    ge.drawGame();
    expect(ge.instances(enum_Canvas::MAIN).empty(), "no engine, no instances");
    expect(sizeof(DrawInstance) == 8 * sizeof(float), "instance layout");
    if (fails == 0) std::printf("ok\n");
    return fails == 0 ? 0 : 1;
}
