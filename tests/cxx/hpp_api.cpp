// C++ callers of the reference's host-level names (include/ctlab_mi355x.hpp):
// raw file round trip in float and double, the helical range helpers, and
// the driver's error path (a method other than branchless), all without a GPU.
#include "ctlab_mi355x.hpp"

#include <cstdio>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    std::vector<float> a = {1.5f, -2.25f, 3.0f}, b;
    writeData_host(dir + "/f.raw", a, 3);
    readData_host(dir + "/f.raw", b, 3);
    std::vector<double> c = {0.1, 0.2}, d;
    writeData_host(dir + "/d.raw", c, 2);
    readData_host(dir + "/d.raw", d, 2);
    std::vector<float> z = {-6.f, -2.f, 2.f, 6.f};
    int s = -1, e = -1, ps = -1, pe = -1;
    getVolZIdxPair(z, 1, 3, 1.5f, 2.0f, 4, 7.5f, 1.0f, 16, s, e);
    getPrjIdxPair(z, 4, 8, 7.5f, 1.0f, 16, 1.5f, 2.0f, 4, ps, pe);
    int threw = 0;
    std::vector<float> r, init, h(2 * 2 * 2, 0.f), ang(2, 0.f), zp;
    std::vector<unsigned char> m;
    try {
        OS_SART(r, init, h, ang, zp, m, std::string(), 1, 1, false, 10.f, 20.f, 2, 2, 2, 0.f, 0.f, 0.f, 2, 2, 2, 1.f,
                1.f, 1.f, 1.f, 0.f, 0.f, PROJ_PSEUDODISTANCE, BACK_BRANCHLESS);
    } catch (const std::runtime_error&) {
        threw = ctlab_last_error() == CTLAB_ERR_MODE;
    }
    std::printf("%g %g %g %g %g %d %d %d %d %d\n", b[0], b[1], b[2], d[0], d[1], s, e, ps, pe, threw);
    return 0;
}
