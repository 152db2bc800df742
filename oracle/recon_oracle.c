/*
 * recon_oracle.c -- FP64 CPU restatement of CTLAB's reconstruction update
 * functors, plus the TV proximal operator used by FISTA-TV.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker; the product
 * library never calls it.
 *
 *   SART_Weighting<T>::Proj / Back ... src/SARTWeighting.h:22-136 (eps 1e-7, :138)
 *   ossart_update_functor ............ src/utilities.hpp:1202-1223 (IS_ZERO, :316-320)
 *   FISTA_functor / FISTA ............ src/FastIterativeShrinkageAlgorithm.h:17-69
 *   _softThreshold_functor ........... src/utilities.hpp:1097-1117
 *   CG_update{1,2,3}_functor ......... src/utilities.hpp:1121-1156
 *   _betaupdate_functor / calculateBetaForCG  src/utilities.hpp:1228-1268, 1611-1621
 * The TV prox is not in the reference (its TV routines have no bodies,
 * src/useful.hpp:238-416); it is the fast gradient projection (FGP) of
 * Beck & Teboulle, IEEE TIP 18(11) 2009, on the isotropic TV with forward
 * differences and zero flux at the far faces, volume in (Z,X,Y) order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* SART_Weighting's epsilon is `static const float epsilon = 1.0e-7`
 * (src/SARTWeighting.h:138): the guard compares against 1e-7 rounded to float. */
static const double kSartEps = (double)1.0e-7f;

/* r = (b - Ax) / rowSum, 0 where rowSum <= eps */
void ro_sart_prj_weight(double *prj, const double *realPrj, const double *rowSum, long n)
{
    for (long i = 0; i < n; ++i) prj[i] = rowSum[i] > kSartEps ? (realPrj[i] - prj[i]) / rowSum[i] : 0.0;
}

/* recon += vol / colSum, 0 where colSum <= eps */
void ro_sart_bak_weight(const double *vol, double *recon, const double *colSum, long n)
{
    for (long i = 0; i < n; ++i) recon[i] += colSum[i] > kSartEps ? vol[i] / colSum[i] : 0.0;
}

/* img = (img + lambda*cor/weg)*msk, 0 where |weg| < 1e-9 */
void ro_ossart_update(double *img, const double *cor, const double *weg, const double *msk,
                      double lambda, long n)
{
    for (long i = 0; i < n; ++i) {
        int zero = (weg[i] < 1.0e-9) && (weg[i] > -1.0e-9);
        img[i] = zero ? 0.0 : (img[i] + lambda * cor[i] / weg[i]) * msk[i];
    }
}

/* cur = clamp(cur + (t1-1)/t2*(cur - las), minV, maxV) */
void ro_fista(double *cur, const double *las, double t1, double t2, double minV, double maxV, long n)
{
    for (long i = 0; i < n; ++i) {
        double r = cur[i] + (t1 - 1.0) / t2 * (cur[i] - las[i]);
        cur[i] = r < minV ? minV : (r > maxV ? maxV : r);
    }
}

void ro_soft_threshold(double *x, double g, long n)
{
    for (long i = 0; i < n; ++i) x[i] = x[i] < -g ? x[i] + g : (x[i] > g ? x[i] - g : 0.0);
}

/* beta = num/den with (gk, gk1, d) as in _betaupdate_functor<cases> */
double ro_cg_beta(const double *gk, const double *gk1, const double *d, long n, int cases)
{
    double num = 0.0, den = 0.0;
    for (long i = 0; i < n; ++i) {
        double a = gk[i], b = gk1[i], e = d[i];
        switch (cases) {
        case 1: num += b * (b - a); den += a * a; break;
        case 2: num += b * (b - a); den += e * (b - a); break;
        case 3: num += b * b; den += -e * a; break;
        case 4: num += b * b; den += e * (b - a); break;
        default: num += b * b; den += a * a; break;
        }
    }
    return num / den;
}

/* ---- isotropic TV prox by FGP on a (Z,X,Y)-ordered volume ---------------- */
typedef struct { int XN, YN, ZN; } dims_t;

/* L(p) = "negative divergence": p_d[i] - p_d[i - e_d], p_d = 0 outside and at the far face */
static void tv_L(const dims_t *d, const double *p, double *out)
{
    long n = (long)d->XN * d->YN * d->ZN;
    const double *px = p, *py = p + n, *pz = p + 2 * n;
    for (int y = 0; y < d->YN; ++y)
        for (int x = 0; x < d->XN; ++x)
            for (int z = 0; z < d->ZN; ++z) {
                long i = ((long)y * d->XN + x) * d->ZN + z;
                double v = px[i] + py[i] + pz[i];
                if (x > 0) v -= px[i - d->ZN];
                if (y > 0) v -= py[i - (long)d->XN * d->ZN];
                if (z > 0) v -= pz[i - 1];
                out[i] = v;
            }
}

/* L^T(x): forward differences x[i] - x[i + e_d], 0 at the far face */
static void tv_LT(const dims_t *d, const double *x, double *q)
{
    long n = (long)d->XN * d->YN * d->ZN;
    double *qx = q, *qy = q + n, *qz = q + 2 * n;
    for (int y = 0; y < d->YN; ++y)
        for (int xx = 0; xx < d->XN; ++xx)
            for (int z = 0; z < d->ZN; ++z) {
                long i = ((long)y * d->XN + xx) * d->ZN + z;
                qx[i] = xx < d->XN - 1 ? x[i] - x[i + d->ZN] : 0.0;
                qy[i] = y < d->YN - 1 ? x[i] - x[i + (long)d->XN * d->ZN] : 0.0;
                qz[i] = z < d->ZN - 1 ? x[i] - x[i + 1] : 0.0;
            }
}

int ro_tv_ndim(int XN, int YN, int ZN) { return (XN > 1) + (YN > 1) + (ZN > 1); }

/* x = argmin 0.5||x - b||^2 + lam*TV(x), iters FGP steps from p = 0 */
int ro_tv_prox(const double *b, double *x, int XN, int YN, int ZN, double lam, int iters)
{
    dims_t d = {XN, YN, ZN};
    long n = (long)XN * YN * ZN;
    if (lam <= 0) { memcpy(x, b, sizeof(double) * n); return 0; }
    double *p = (double *)calloc(3 * n, sizeof(double));
    double *pold = (double *)calloc(3 * n, sizeof(double));
    double *r = (double *)calloc(3 * n, sizeof(double));
    double *q = (double *)malloc(sizeof(double) * 3 * n);
    double *tmp = (double *)malloc(sizeof(double) * n);
    if (!p || !pold || !r || !q || !tmp) return -1;
    int nd = ro_tv_ndim(XN, YN, ZN);
    double step = 1.0 / (4.0 * (nd > 0 ? nd : 1) * lam);
    double t = 1.0;
    for (int it = 0; it < iters; ++it) {
        tv_L(&d, r, tmp);
        for (long i = 0; i < n; ++i) tmp[i] = b[i] - lam * tmp[i];
        tv_LT(&d, tmp, q);
        memcpy(pold, p, sizeof(double) * 3 * n);
        for (long i = 0; i < n; ++i) {
            double a0 = r[i] + step * q[i], a1 = r[n + i] + step * q[n + i], a2 = r[2 * n + i] + step * q[2 * n + i];
            double nrm = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
            double s = nrm > 1.0 ? 1.0 / nrm : 1.0;
            p[i] = a0 * s; p[n + i] = a1 * s; p[2 * n + i] = a2 * s;
        }
        double tn = 0.5 * (1.0 + sqrt(1.0 + 4.0 * t * t));
        double mom = (t - 1.0) / tn;
        for (long i = 0; i < 3 * n; ++i) r[i] = p[i] + mom * (p[i] - pold[i]);
        t = tn;
    }
    tv_L(&d, p, tmp);
    for (long i = 0; i < n; ++i) x[i] = b[i] - lam * tmp[i];
    free(p); free(pold); free(r); free(q); free(tmp);
    return 0;
}

/* isotropic TV value (same differences as the prox) */
double ro_tv_value(const double *x, int XN, int YN, int ZN)
{
    dims_t d = {XN, YN, ZN};
    long n = (long)XN * YN * ZN;
    double *q = (double *)malloc(sizeof(double) * 3 * n);
    tv_LT(&d, x, q);
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += sqrt(q[i] * q[i] + q[n + i] * q[n + i] + q[2 * n + i] * q[2 * n + i]);
    free(q);
    return s;
}
