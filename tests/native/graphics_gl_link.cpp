// Compiles the OpenGL path of code_mpm_auxetics_amd/host/GraphicsEngine.h
// (MPM_WITH_GL) and links it against the system GL library.  The program
// never creates a context: with an argument it would construct the renderer,
// which needs one, so it is only built and run without arguments.
#include "../../code_mpm_auxetics_amd/host/GraphicsEngine.h"

int main(int argc, char **)
{
    if (argc > 1) {
        mpm_host::InstancedRenderer r;
        const float vp[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        r.draw({}, vp);
    }
    return 0;
}
