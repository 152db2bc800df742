/*
 * dd_oracle.c -- FP64 CPU oracle for CTLAB's branchless distance-driven (DD)
 * projector and backprojector (cone beam, and multi-slice fan beam).
 *
 * TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, as the checker or as the
 * timed CPU baseline.  The product library (ctlab_amd) never links or calls it.
 *
 * Reference behaviour restated here (paths relative to the CTLAB tree):
 *   DD3Proj_gpu, prjMode 0 ............ src/DD3_GPU_Proj.h:16-44
 *   DD3Back_gpu, prjMode 0 ............ src/DD3_GPU_Back.h:14-44
 *   DD2_multiGPU, method 0 / 2 ........ src/multiSlices_ker.cuh:15-38
 *   DD3Boundaries ..................... src/utilities.cuh:413-433
 *   regularizeAngle ................... src/utilities.cuh:279-286
 *   per-view cos/sin/rotated source ... src/utilities.cuh:313-337
 *   (Z,X,Y) volume order .............. src/utilities.cuh:441-454
 *   integral-image bilinear read ...... src/utilities.cuh:231-254
 * The reference's .cu bodies are not in its tree, so the points its headers
 * leave open are fixed as follows (the GPU library follows the same list):
 *   vol[(y*XN + x)*ZN + z], prj[(p*DNU + u)*DNV + v], mask[y*XN + x];
 *   view p: theta = regularizeAngle(angs[p]); source (x0,y0) and detector
 *     points rotated counter-clockwise by theta; source z = z0 + zpos[p],
 *     detector row z = zds[v] + zpos[p];
 *   voxel edges: X = cx - XN*dx/2 + i*dx (same for Y with dx, Z with dz);
 *   slab orientation: y-slabs (planes y = const) when |cy - Sy| >= |cx - Sx|,
 *     otherwise x-slabs;
 *   value = sum over slabs of L * (mean of the volume over the footprint of
 *     the detector cell on the slab's centre plane), L = slab thickness *
 *     |ray| / |ray component normal to the slab|, ray = source -> cell centre;
 *   footprint u-edges: source -> detector u-boundary lines; footprint z-edges
 *     of cell (u,v): source -> (cell-u centre column, row boundary v);
 *   the backprojector is the exact adjoint (transpose) of the projector.
 * Multi-slice fan beam ("parallel" mode): ZN = DNV = slice count, slice v of
 * the detector sees only slice v of the image, L uses the in-plane ray only.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int XN, YN, ZN;
    double dx, dz;
    double cx, cy, cz;
    int DNU, DNV;
    double x0, y0, z0;
    const double *xds, *yds, *zds;   /* detector cell centres at view 0 */
    int PN;
    const double *angs, *zpos;
    const unsigned char *mask;       /* XN*YN or NULL */
    int parallel;                    /* 1: multi-slice fan beam */
} ddo_geom;

/* src/utilities.cuh:413-433: N centres -> N+1 boundaries. */
void ddo_boundaries(int nb, const double *c, double *b)
{
    if (nb >= 3) {
        b[0] = 1.5 * c[0] - 0.5 * c[1];
        for (int i = 1; i <= nb - 2; ++i) b[i] = 0.5 * c[i - 1] + 0.5 * c[i];
        b[nb - 1] = 1.5 * c[nb - 2] - 0.5 * c[nb - 3];
    } else {
        b[0] = c[0] - 0.5;
        b[1] = c[0] + 0.5;
    }
}

/* src/utilities.cuh:279-286 */
static double regularize_angle(double a)
{
    const double twopi = 6.283185307179586;
    while (a >= twopi) a -= twopi;
    while (a < 0) a += twopi;
    return a;
}

typedef struct {
    int yslab;
    int NS, NR;
    long sstride, rstride;     /* in z-columns */
    double Sx, Sy, Ss, Sr, zs;
    double Smin, Rmin, Zmin;
    double c, s;
} frame_t;

static void view_frame(const ddo_geom *g, int p, frame_t *f)
{
    double th = regularize_angle(g->angs[p]);
    f->c = cos(th);
    f->s = sin(th);
    f->Sx = g->x0 * f->c - g->y0 * f->s;
    f->Sy = g->x0 * f->s + g->y0 * f->c;
    f->zs = g->z0 + (g->zpos ? g->zpos[p] : 0.0);
    f->yslab = fabs(g->cy - f->Sy) >= fabs(g->cx - f->Sx);
    double Xmin = g->cx - g->XN * g->dx * 0.5;
    double Ymin = g->cy - g->YN * g->dx * 0.5;
    f->Zmin = g->cz - g->ZN * g->dz * 0.5;
    if (f->yslab) {
        f->NS = g->YN; f->NR = g->XN;
        f->sstride = (long)g->XN; f->rstride = 1;
        f->Ss = f->Sy; f->Sr = f->Sx; f->Smin = Ymin; f->Rmin = Xmin;
    } else {
        f->NS = g->XN; f->NR = g->YN;
        f->sstride = 1; f->rstride = (long)g->XN;
        f->Ss = f->Sx; f->Sr = f->Sy; f->Smin = Xmin; f->Rmin = Ymin;
    }
}

/* rotate a view-0 point into view p; returns (slab-axis, row-axis) coords */
static void rot_sr(const frame_t *f, double x, double y, double *ps, double *pr)
{
    double X = x * f->c - y * f->s;
    double Y = x * f->s + y * f->c;
    if (f->yslab) { *ps = Y; *pr = X; } else { *ps = X; *pr = Y; }
}

typedef struct {
    double *kap;    /* DNU+1 slopes d(row)/d(slab) of the u-boundary rays */
    double *cs;     /* DNU: cell centre, slab-axis coordinate */
    double *L;      /* DNU*DNV ray lengths through one slab */
    double *zb;     /* DNV+1 row boundaries relative to the source, world units */
} viewtab_t;

static void view_tables(const ddo_geom *g, const frame_t *f,
                        const double *xb, const double *yb, const double *zbnd,
                        viewtab_t *t)
{
    for (int u = 0; u <= g->DNU; ++u) {
        double ds, dr;
        rot_sr(f, xb[u], yb[u], &ds, &dr);
        t->kap[u] = (dr - f->Sr) / (ds - f->Ss);
    }
    for (int u = 0; u < g->DNU; ++u) {
        double ds, dr;
        rot_sr(f, g->xds[u], g->yds[u], &ds, &dr);
        t->cs[u] = ds;
        double ex = ds - f->Ss, ey = dr - f->Sr;
        for (int v = 0; v < g->DNV; ++v) {
            double ez = g->parallel ? 0.0 : (g->zds[v] - g->z0);
            t->L[(long)u * g->DNV + v] = g->dx * sqrt(ex * ex + ey * ey + ez * ez) / fabs(ex);
        }
    }
    for (int v = 0; v <= g->DNV; ++v) t->zb[v] = g->parallel ? 0.0 : (zbnd[v] - g->z0);
}

/* footprint of cell (u,v) on slab j, lattice units; returns 0 if degenerate */
typedef struct { double alo, ahi, da, blo, bhi, db; } foot_t;

static int footprint(const ddo_geom *g, const frame_t *f, const viewtab_t *t,
                     int j, int u, int v, foot_t *ft)
{
    double sj = f->Smin + (j + 0.5) * g->dx;
    double a0 = (f->Sr + (sj - f->Ss) * t->kap[u] - f->Rmin) / g->dx;
    double a1 = (f->Sr + (sj - f->Ss) * t->kap[u + 1] - f->Rmin) / g->dx;
    ft->alo = a0 < a1 ? a0 : a1;
    ft->ahi = a0 < a1 ? a1 : a0;
    ft->da = ft->ahi - ft->alo;
    if (g->parallel) {
        ft->blo = v; ft->bhi = v + 1; ft->db = 1.0;
    } else {
        double tt = (sj - f->Ss) / (t->cs[u] - f->Ss);
        double b0 = (f->zs + tt * t->zb[v] - f->Zmin) / g->dz;
        double b1 = (f->zs + tt * t->zb[v + 1] - f->Zmin) / g->dz;
        ft->blo = b0 < b1 ? b0 : b1;
        ft->bhi = b0 < b1 ? b1 : b0;
        ft->db = ft->bhi - ft->blo;
    }
    return ft->da > 0 && ft->db > 0;
}

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* bilinear read of an (NR+1) x (ZN+1) integral image at lattice point (a,b);
 * src/utilities.cuh:231-254 is the reference's FP64 form of this read. */
static void corner(int NR, int ZN, double a, double b, long *i00, double w[4])
{
    int i = (int)floor(a); if (i > NR - 1) i = NR - 1; if (i < 0) i = 0;
    int k = (int)floor(b); if (k > ZN - 1) k = ZN - 1; if (k < 0) k = 0;
    double al = a - i, be = b - k;
    *i00 = (long)i * (ZN + 1) + k;
    w[0] = (1 - al) * (1 - be);  /* (i,   k)   */
    w[1] = al * (1 - be);        /* (i+1, k)   */
    w[2] = (1 - al) * be;        /* (i,   k+1) */
    w[3] = al * be;              /* (i+1, k+1) */
}

static double sat_read(const double *S, int NR, int ZN, double a, double b)
{
    long o; double w[4];
    corner(NR, ZN, a, b, &o, w);
    long r = ZN + 1;
    return w[0] * S[o] + w[1] * S[o + r] + w[2] * S[o + 1] + w[3] * S[o + r + 1];
}

static void sat_scatter(double *S, int NR, int ZN, double a, double b, double val)
{
    long o; double w[4];
    corner(NR, ZN, a, b, &o, w);
    long r = ZN + 1;
    S[o] += w[0] * val; S[o + r] += w[1] * val; S[o + 1] += w[2] * val; S[o + r + 1] += w[3] * val;
}

static double mask_at(const ddo_geom *g, long col) { return g->mask ? (g->mask[col] ? 1.0 : 0.0) : 1.0; }

/* integral images S_j[m][b] = sum_{m'<m, k<b} f(slab j, row m', k), masked */
static double *build_sats(const ddo_geom *g, const double *vol, int yslab)
{
    int NS = yslab ? g->YN : g->XN, NR = yslab ? g->XN : g->YN, ZN = g->ZN;
    long per = (long)(NR + 1) * (ZN + 1);
    double *S = (double *)calloc((size_t)NS * per, sizeof(double));
    if (!S) return NULL;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < NS; ++j) {
        double *Sj = S + (long)j * per;
        for (int m = 0; m < NR; ++m) {
            long col = yslab ? ((long)j * g->XN + m) : ((long)m * g->XN + j);
            double mk = mask_at(g, col);
            double run = 0.0;
            for (int k = 0; k < ZN; ++k) {
                run += vol[col * ZN + k] * mk;
                Sj[(long)(m + 1) * (ZN + 1) + k + 1] = Sj[(long)m * (ZN + 1) + k + 1] + run;
            }
        }
    }
    return S;
}

static int alloc_tabs(const ddo_geom *g, viewtab_t *t)
{
    t->kap = (double *)malloc(sizeof(double) * (g->DNU + 1));
    t->cs = (double *)malloc(sizeof(double) * g->DNU);
    t->L = (double *)malloc(sizeof(double) * (size_t)g->DNU * g->DNV);
    t->zb = (double *)malloc(sizeof(double) * (g->DNV + 1));
    return t->kap && t->cs && t->L && t->zb;
}

static void free_tabs(viewtab_t *t) { free(t->kap); free(t->cs); free(t->L); free(t->zb); }

static int check_geom(const ddo_geom *g)
{
    if (g->XN <= 0 || g->YN <= 0 || g->ZN <= 0 || g->DNU <= 0 || g->DNV <= 0 || g->PN <= 0) return -1;
    if (g->parallel && g->ZN != g->DNV) return -2;
    if (!(g->dx > 0) || !(g->dz > 0)) return -3;
    return 0;
}

typedef struct { double *xb, *yb, *zb; } bnd_t;

static int make_bnd(const ddo_geom *g, bnd_t *b)
{
    b->xb = (double *)malloc(sizeof(double) * (g->DNU + 1));
    b->yb = (double *)malloc(sizeof(double) * (g->DNU + 1));
    b->zb = (double *)malloc(sizeof(double) * (g->DNV + 1));
    if (!b->xb || !b->yb || !b->zb) return -1;
    ddo_boundaries(g->DNU + 1, g->xds, b->xb);
    ddo_boundaries(g->DNU + 1, g->yds, b->yb);
    if (!g->parallel) ddo_boundaries(g->DNV + 1, g->zds, b->zb);
    else for (int v = 0; v <= g->DNV; ++v) b->zb[v] = v;
    return 0;
}

static void free_bnd(bnd_t *b) { free(b->xb); free(b->yb); free(b->zb); }

/* Forward projection over views [pb, pe).  prj holds all PN views. */
int ddo_proj_range(const ddo_geom *g, const double *vol, double *prj, int pb, int pe)
{
    int rc = check_geom(g);
    if (rc) return rc;
    bnd_t bd;
    if (make_bnd(g, &bd)) return -10;
    double *SAT[2] = {NULL, NULL};
    SAT[0] = build_sats(g, vol, 0);
    SAT[1] = build_sats(g, vol, 1);
    if (!SAT[0] || !SAT[1]) { free(SAT[0]); free(SAT[1]); free_bnd(&bd); return -10; }
    const int DNU = g->DNU, DNV = g->DNV, ZN = g->ZN;
#pragma omp parallel
    {
        viewtab_t t;
        alloc_tabs(g, &t);
#pragma omp for schedule(dynamic, 1)
        for (int p = pb; p < pe; ++p) {
            frame_t f;
            view_frame(g, p, &f);
            view_tables(g, &f, bd.xb, bd.yb, bd.zb, &t);
            const double *S = SAT[f.yslab];
            long per = (long)(f.NR + 1) * (ZN + 1);
            double *out = prj + (long)p * DNU * DNV;
            for (long e = 0; e < (long)DNU * DNV; ++e) out[e] = 0.0;
            for (int j = 0; j < f.NS; ++j) {
                const double *Sj = S + (long)j * per;
                for (int u = 0; u < DNU; ++u) {
                    for (int v = 0; v < DNV; ++v) {
                        foot_t ft;
                        if (!footprint(g, &f, &t, j, u, v, &ft)) continue;
                        double al = clampd(ft.alo, 0, f.NR), ah = clampd(ft.ahi, 0, f.NR);
                        double bl = clampd(ft.blo, 0, ZN), bh = clampd(ft.bhi, 0, ZN);
                        if (ah <= al || bh <= bl) continue;
                        double I = sat_read(Sj, f.NR, ZN, ah, bh) - sat_read(Sj, f.NR, ZN, al, bh)
                                 - sat_read(Sj, f.NR, ZN, ah, bl) + sat_read(Sj, f.NR, ZN, al, bl);
                        out[(long)u * DNV + v] += t.L[(long)u * DNV + v] * I / (ft.da * ft.db);
                    }
                }
            }
        }
        free_tabs(&t);
    }
    free(SAT[0]); free(SAT[1]); free_bnd(&bd);
    return 0;
}

int ddo_proj(const ddo_geom *g, const double *vol, double *prj)
{
    return ddo_proj_range(g, vol, prj, 0, g->PN);
}

/* Exact adjoint of ddo_proj: transpose of the integral-image reads, then the
 * transpose of the integral-image construction (2-D suffix sums). */
int ddo_back(const ddo_geom *g, const double *prj, double *vol)
{
    int rc = check_geom(g);
    if (rc) return rc;
    bnd_t bd;
    if (make_bnd(g, &bd)) return -10;
    const int DNU = g->DNU, DNV = g->DNV, ZN = g->ZN;
    long nvox = (long)g->XN * g->YN * ZN;
    for (long e = 0; e < nvox; ++e) vol[e] = 0.0;
    int *orient = (int *)malloc(sizeof(int) * g->PN);
    for (int p = 0; p < g->PN; ++p) { frame_t f; view_frame(g, p, &f); orient[p] = f.yslab; }
    for (int o = 0; o < 2; ++o) {
        int NS = o ? g->YN : g->XN, NR = o ? g->XN : g->YN;
        long per = (long)(NR + 1) * (ZN + 1);
#pragma omp parallel
        {
            viewtab_t t;
            alloc_tabs(g, &t);
            double *Sb = (double *)malloc(sizeof(double) * per);
            double *acc = (double *)malloc(sizeof(double) * ZN);
#pragma omp for schedule(dynamic, 1)
            for (int j = 0; j < NS; ++j) {
                memset(Sb, 0, sizeof(double) * per);
                for (int p = 0; p < g->PN; ++p) {
                    if (orient[p] != o) continue;
                    frame_t f;
                    view_frame(g, p, &f);
                    view_tables(g, &f, bd.xb, bd.yb, bd.zb, &t);
                    const double *in = prj + (long)p * DNU * DNV;
                    for (int u = 0; u < DNU; ++u) {
                        for (int v = 0; v < DNV; ++v) {
                            foot_t ft;
                            if (!footprint(g, &f, &t, j, u, v, &ft)) continue;
                            double al = clampd(ft.alo, 0, NR), ah = clampd(ft.ahi, 0, NR);
                            double bl = clampd(ft.blo, 0, ZN), bh = clampd(ft.bhi, 0, ZN);
                            if (ah <= al || bh <= bl) continue;
                            double w = in[(long)u * DNV + v] * t.L[(long)u * DNV + v] / (ft.da * ft.db);
                            sat_scatter(Sb, NR, ZN, ah, bh, w);
                            sat_scatter(Sb, NR, ZN, al, bh, -w);
                            sat_scatter(Sb, NR, ZN, ah, bl, -w);
                            sat_scatter(Sb, NR, ZN, al, bl, w);
                        }
                    }
                }
                /* transpose of the integral image:
                 * f_bar[m][k] = sum_{m' > m} sum_{b > k} Sb[m'][b]
                 *             = f_bar[m+1][k] + sum_{b > k} Sb[m+1][b]      */
                for (int k = 0; k < ZN; ++k) acc[k] = 0.0;
                for (int m = NR - 1; m >= 0; --m) {
                    const double *row = Sb + (long)(m + 1) * (ZN + 1);
                    double run = 0.0;
                    for (int k = ZN - 1; k >= 0; --k) {
                        run += row[k + 1];
                        acc[k] += run;
                    }
                    long col = o ? ((long)j * g->XN + m) : ((long)m * g->XN + j);
                    double mk = mask_at(g, col);
                    for (int k = 0; k < ZN; ++k) vol[col * ZN + k] += acc[k] * mk;
                }
            }
            free(Sb);
            free(acc);
            free_tabs(&t);
        }
    }
    free(orient);
    free_bnd(&bd);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Brute force: the dense system matrix A[(p*DNU+u)*DNV+v][(y*XN+x)*ZN+z] from
 * direct interval overlaps (no integral image).  Tiny inputs only.  Pins the
 * integral-image reads of ddo_proj and the transpose in ddo_back.
 * ------------------------------------------------------------------------- */
static double ovl(double a0, double a1, double b0, double b1)
{
    double lo = a0 > b0 ? a0 : b0, hi = a1 < b1 ? a1 : b1;
    return hi > lo ? hi - lo : 0.0;
}

int ddo_dense(const ddo_geom *g, double *A)
{
    int rc = check_geom(g);
    if (rc) return rc;
    bnd_t bd;
    if (make_bnd(g, &bd)) return -10;
    const int DNU = g->DNU, DNV = g->DNV, ZN = g->ZN;
    long ncol = (long)g->XN * g->YN * ZN;
    long nrow = (long)g->PN * DNU * DNV;
    memset(A, 0, sizeof(double) * (size_t)(ncol * nrow));
    viewtab_t t;
    alloc_tabs(g, &t);
    for (int p = 0; p < g->PN; ++p) {
        frame_t f;
        view_frame(g, p, &f);
        view_tables(g, &f, bd.xb, bd.yb, bd.zb, &t);
        for (int j = 0; j < f.NS; ++j)
            for (int u = 0; u < DNU; ++u)
                for (int v = 0; v < DNV; ++v) {
                    foot_t ft;
                    if (!footprint(g, &f, &t, j, u, v, &ft)) continue;
                    long row = ((long)p * DNU + u) * DNV + v;
                    double s = t.L[(long)u * DNV + v] / (ft.da * ft.db);
                    for (int m = 0; m < f.NR; ++m) {
                        double ox = ovl(ft.alo, ft.ahi, m, m + 1);
                        if (ox == 0.0) continue;
                        long col = f.yslab ? ((long)j * g->XN + m) : ((long)m * g->XN + j);
                        double mk = mask_at(g, col);
                        for (int k = 0; k < ZN; ++k) {
                            double oz = ovl(ft.blo, ft.bhi, k, k + 1);
                            if (oz == 0.0) continue;
                            A[row * ncol + col * ZN + k] += s * ox * oz * mk;
                        }
                    }
                }
    }
    free_tabs(&t);
    free_bnd(&bd);
    return 0;
}

/* per-view orientation (1 = y-slabs), exported so tests can pick views away
 * from the 45-degree ties */
void ddo_orientation(const ddo_geom *g, int *out)
{
    for (int p = 0; p < g->PN; ++p) { frame_t f; view_frame(g, p, &f); out[p] = f.yslab; }
}

/* ---------------------------------------------------------------------------
 * Conventional (overlap-merge) fan-beam DD, De Man & Basu (2004): pixel
 * boundaries of each slab and detector boundaries are mapped onto a common
 * axis through the isocentre parallel to the slabs and merged.  Restates the
 * CPU DDM_ED_proj / DDM_EA_proj / DDM_ED_bproj / DDM_EA_bproj interface
 * (src/DDM_proj.h:28-55, 194-221; src/DDM_back.h:139-209), whose bodies are
 * not in the reference tree.  Detector frames follow src/utilities.hpp:
 * 3294-3326 (arc: (S2D sin g, S2O - S2D cos g); flat: y = -O2D).  Image
 * img[y*XN + x] centred on the isocentre, proj[p*DN + u].  The backprojector
 * is the exact transpose.  Built in double and float.
 * ------------------------------------------------------------------------- */
#define DDM_T double
#define DDM_SUF _d
#include "ddm_core.inc"
#undef DDM_T
#undef DDM_SUF
#define DDM_T float
#define DDM_SUF _f
#include "ddm_core.inc"
#undef DDM_T
#undef DDM_SUF

/* OpenMP threads of the oracle's parallel loops (the CPU baseline times 1 and all) */
void ddo_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int ddo_max_threads(void) { return omp_get_max_threads(); }
