/* A C99 call site written the way CTLAB's callers call DD3Proj_gpu: it must
 * compile against include/ctlab_mi355x.h with a C compiler and link against the
 * library.  prjMode 1 is rejected before any device work (CTLAB_ERR_MODE) and the
 * output is left untouched, so this runs without a GPU. */
#include <stdio.h>
#include <stdlib.h>

#include "ctlab_mi355x.h"

int main(void)
{
    enum { DNU = 4, DNV = 2, PN = 3, XN = 4, YN = 4, ZN = 2 };
    float xds[DNU] = {-1.5f, -0.5f, 0.5f, 1.5f}, yds[DNU] = {-40.f, -40.f, -40.f, -40.f};
    float zds[DNV] = {-0.5f, 0.5f}, hangs[PN] = {0.f, 0.1f, 0.2f}, hzPos[PN] = {0.f, 0.f, 0.f};
    float vol[XN * YN * ZN], prj[PN * DNU * DNV];
    unsigned char mask[XN * YN];
    for (int i = 0; i < XN * YN * ZN; ++i) vol[i] = 1.0f;
    for (int i = 0; i < PN * DNU * DNV; ++i) prj[i] = -7.0f;
    for (int i = 0; i < XN * YN; ++i) mask[i] = 1;
    DD3Proj_gpu(0.f, 60.f, 0.f, DNU, DNV, xds, yds, zds, 0.f, 0.f, 0.f, hangs, hzPos, PN, XN, YN, ZN, vol, prj,
                1.0f, 1.0f, mask, 0, /* prjMode */ 1);
    const int err = ctlab_last_error();
    int untouched = 1;
    for (int i = 0; i < PN * DNU * DNV; ++i) untouched &= prj[i] == -7.0f;
    printf("%d %d %s\n", err, untouched, ctlab_last_error_message());
    return (err == CTLAB_ERR_MODE && untouched) ? 0 : 1;
}
