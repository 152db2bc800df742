// Driver that runs CTLAB's own Ellipsoid.h (chord length through a rotated
// ellipse, and the SheppLoganPhantom table) on seeded inputs and prints inputs
// and outputs as JSON, so that ctlab_amd/phantom.py can be checked against the
// reference code itself.  Built by oracle/Makefile (target ref) from the header
// where it lies in the reference tree; output goes to oracle/_ref/ only.
// TEST INFRASTRUCTURE ONLY.
#include <cstdint>
#include <cstdio>

#include "Ellipsoid.h"

static uint64_t st = 0x0123456789abcdefull;
static float urand()   // xorshift64*, uniform [0,1)
{
    st ^= st >> 12; st ^= st << 25; st ^= st >> 27;
    return (float)((st * 2685821657736338717ull) >> 40) / 16777216.0f;
}

int main()
{
    std::printf("{\n\"shepp_logan\": [");
    CTLAB::SheppLoganPhantom ph;
    for (size_t i = 0; i < ph.ele.size(); ++i) {
        const CTLAB::Ellipsoid& e = ph.ele[i];
        std::printf("%s\n[%.9g, %.9g, %.9g, %.9g, %.9g, %.9g]", i ? "," : "", e.getCenter().x, e.getCenter().y,
                    e.getAxis().x, e.getAxis().y, e.getRotAng(), e.getIntensity());
    }
    std::printf("],\n\"chords\": [");
    const int N = 64;
    for (int i = 0; i < N; ++i) {
        const float cx = 4.0f * urand() - 2.0f, cy = 4.0f * urand() - 2.0f;
        const float a = 0.2f + 2.0f * urand(), b = 0.2f + 2.0f * urand();
        const float rot = 6.2831853f * urand() - 3.1415927f;
        const float sx = 8.0f * urand() - 4.0f, sy = 8.0f * urand() - 4.0f;
        const float ex = 8.0f * urand() - 4.0f, ey = 8.0f * urand() - 4.0f;
        CTLAB::Ellipsoid e(CTLAB::float2(cx, cy), CTLAB::float2(a, b), rot, 1.0f);
        const float len = e.intersectionLength(CTLAB::float2(sx, sy), CTLAB::float2(ex, ey));
        std::printf("%s\n[%.9g, %.9g, %.9g, %.9g, %.9g, %.9g, %.9g, %.9g, %.9g, %.9g]", i ? "," : "", cx, cy, a, b,
                    rot, sx, sy, ex, ey, len);
    }
    std::printf("]\n}\n");
    return 0;
}
