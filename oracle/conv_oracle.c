/*
 * conv_oracle.c -- FP64 CPU 3-D convolution (3x3x3, zero padding, "same"),
 * its input gradient and its weight gradient, for the learned proximal CNN.
 *
 * TEST INFRASTRUCTURE ONLY (checker of the HIP conv kernels).  The CNN has no
 * counterpart in the reference; this follows the textbook definitions:
 *   out[v][co] = b[co] + sum_{t, ci} w[t][co][ci] * in[v + off(t)][ci]
 * with voxels in (Z,X,Y) order v = (y*XN + x)*ZN + z, channels fastest,
 * tap t = (dy+1)*9 + (dx+1)*3 + (dz+1), off(t) = dz + dx*ZN + dy*XN*ZN,
 * and in[] = 0 outside the volume.
 */
#include <stdlib.h>
#include <string.h>

static int nb_ok(int y, int x, int z, int dy, int dx, int dz, int YN, int XN, int ZN)
{
    return y + dy >= 0 && y + dy < YN && x + dx >= 0 && x + dx < XN && z + dz >= 0 && z + dz < ZN;
}

/* out[nvox][cout] = conv(in[nvox][cin]; w[27][cout][cin]) + b[cout] (b may be NULL) */
void co_conv3d_fwd(const double *in, const double *w, const double *b, double *out,
                   int XN, int YN, int ZN, int cin, int cout, int unused)
{
    (void)unused;
#pragma omp parallel for collapse(2)
    for (int y = 0; y < YN; ++y)
        for (int x = 0; x < XN; ++x)
            for (int z = 0; z < ZN; ++z) {
                long v = ((long)y * XN + x) * ZN + z;
                double *o = out + v * cout;
                for (int co = 0; co < cout; ++co) o[co] = b ? b[co] : 0.0;
                for (int t = 0; t < 27; ++t) {
                    int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
                    if (!nb_ok(y, x, z, dy, dx, dz, YN, XN, ZN)) continue;
                    const double *a = in + (v + dz + (long)dx * ZN + (long)dy * XN * ZN) * cin;
                    for (int co = 0; co < cout; ++co) {
                        const double *wr = w + ((long)t * cout + co) * cin;
                        double s = 0.0;
                        for (int ci = 0; ci < cin; ++ci) s += wr[ci] * a[ci];
                        o[co] += s;
                    }
                }
            }
}

/* gin[nvox][cin] = d/d in of sum(gout * out) */
void co_conv3d_dgrad(const double *gout, const double *w, double *gin,
                     int XN, int YN, int ZN, int cin, int cout)
{
#pragma omp parallel for collapse(2)
    for (int y = 0; y < YN; ++y)
        for (int x = 0; x < XN; ++x)
            for (int z = 0; z < ZN; ++z) {
                long v = ((long)y * XN + x) * ZN + z;
                double *gi = gin + v * cin;
                for (int ci = 0; ci < cin; ++ci) gi[ci] = 0.0;
                /* in[v] feeds out[v - off(t)] through tap t */
                for (int t = 0; t < 27; ++t) {
                    int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
                    if (!nb_ok(y, x, z, -dy, -dx, -dz, YN, XN, ZN)) continue;
                    const double *g = gout + (v - dz - (long)dx * ZN - (long)dy * XN * ZN) * cout;
                    for (int co = 0; co < cout; ++co) {
                        const double *wr = w + ((long)t * cout + co) * cin;
                        for (int ci = 0; ci < cin; ++ci) gi[ci] += wr[ci] * g[co];
                    }
                }
            }
}

/* gw[27][cout][cin] and gb[cout] (gb may be NULL) */
void co_conv3d_wgrad(const double *in, const double *gout, double *gw, double *gb,
                     int XN, int YN, int ZN, int cin, int cout)
{
    memset(gw, 0, sizeof(double) * 27 * cout * cin);
    if (gb) memset(gb, 0, sizeof(double) * cout);
    for (int y = 0; y < YN; ++y)
        for (int x = 0; x < XN; ++x)
            for (int z = 0; z < ZN; ++z) {
                long v = ((long)y * XN + x) * ZN + z;
                const double *g = gout + v * cout;
                if (gb)
                    for (int co = 0; co < cout; ++co) gb[co] += g[co];
                for (int t = 0; t < 27; ++t) {
                    int dy = t / 9 - 1, dx = (t / 3) % 3 - 1, dz = t % 3 - 1;
                    if (!nb_ok(y, x, z, dy, dx, dz, YN, XN, ZN)) continue;
                    const double *a = in + (v + dz + (long)dx * ZN + (long)dy * XN * ZN) * cin;
                    for (int co = 0; co < cout; ++co)
                        for (int ci = 0; ci < cin; ++ci) gw[((long)t * cout + co) * cin + ci] += g[co] * a[ci];
                }
            }
}
