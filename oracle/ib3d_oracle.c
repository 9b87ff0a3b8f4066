/*
 * ib3d_oracle.c -- serial FP64 restatement of the IB3D hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle and the "port" CPU
 * baseline.  Nothing in the product path links or calls this file.
 *
 * Each function names the reference subroutine (file:line) it restates.
 * Expressions keep the reference's operand order; Fortran evaluates
 * a*b/c as (a*b)/c and a+b+c as (a+b)+c, which the explicit parentheses
 * below spell out.  Build with -ffp-contract=off and without fast-math.
 * `x**2` with an integer exponent is x*x.  Real exponents follow what the
 * reference build (amdflang, -fdefault-real-8) emits: `x**2.0` -> x*x,
 * `x**3.0` -> x*(x*x), `x**4.0` -> pow(x, 4.0); the dumped coefficient
 * tables of the reference match this bit for bit.
 *
 * Parity pins (see tests/): every routine here is compared bitwise with the
 * reference built from its own sources (oracle/_ref) on generated cases;
 * closed-form and invariant checks in tests/test_oracle_*.py cover the
 * remaining properties.
 */
#include "ib3d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { UU = 0, VV = 1, WW = 2, UV = 3, VU = 4, WU = 5 };

struct orc_state {
    int64_t nx, ny, nz, sx, sy, sz;
    double re, rev, deltat, eps1, dt_order;
    int64_t inor;
    /* 1-D grid arrays, 1-based (index 0 unused) */
    double *x1, *y1, *z1, *deltax, *deltay, *deltaz;
    double *xu, *yu, *zu, *xv, *yv, *zv, *xw, *yw, *zw, *xp, *yp, *zp;
    /* non-uniform convection tables: 6 distinct sets x 12 coefficients */
    double *tab[6];
    int64_t tn[6];
    double *Acx, *Acy, *Acz; /* (n,3) column-major */
    /* fields */
    double *u, *v, *w, *ut, *vt, *wt, *p, *pc, *pco, *b, *resi_u, *resi_v, *resi_w;
    int64_t *cell, *minElemcell;
    /* surface */
    int64_t ibNodes, ibElems;
    double *xnode, *ynode, *znode, *xnode1, *ynode1, *znode1;
    int64_t *ibElP1, *ibElP2, *ibElP3;
    double *xcent, *ycent, *zcent, *cosAlpha, *cosBeta, *cosGamma, *alpha3, *beta3, *gamma3;
    /* lists, (count,3) column-major like the reference */
    int64_t ibCount, fluidCount, solidCount, redCount, blackCount;
    int64_t *ibList, *fluidList, *solidList, *redList, *blackList;
    int64_t *nelp, *nelu1, *nelu2, *nelv1, *nelv2, *nelw1, *nelw2;
    double *pNormDis, *u1NormDis, *u2NormDis, *v1NormDis, *v2NormDis, *w1NormDis, *w2NormDis;
    int64_t *Elemcell, *ucell, *vcell, *wcell, *pcell;
    int stress_ready;
    /* rigid-body motion (zero for a fixed body): wall velocity used by the
     * forcing, the solid cells and the surface forces; body acceleration
     * for dp/dn at the wall */
    double wall_vel[3], wall_acc[3];
    int64_t *cell_old;
    /* per-element surface-force contributions of the last stress call
     * (7 x ibElems: viscous drag, pressure drag, viscous lift, pressure lift,
     * area, area_xz, area_yz) and the body id of every element */
    double *contrib;
    int32_t *body;
    int32_t nbody;
    char err[256];
};

#define IX(i, j, k) ((size_t)((i) - 1) + (size_t)s->sx * ((size_t)((j) - 1) + (size_t)s->sy * (size_t)((k) - 1)))
#define F(a, i, j, k) ((a)[IX(i, j, k)])
#define LST(l, cnt, n, c) ((l)[((n) - 1) + (cnt) * ((c) - 1)])
#define ACX(i, c) (s->Acx[((i) - 1) + s->nx * ((c) - 1)])
#define ACY(j, c) (s->Acy[((j) - 1) + s->ny * ((c) - 1)])
#define ACZ(k, c) (s->Acz[((k) - 1) + s->nz * ((c) - 1)])
#define TAB(set, c, i) (s->tab[set][(int64_t)(c) * s->tn[set] + (i)])

static int fail(orc_state *s, const char *msg) {
    snprintf(s->err, sizeof s->err, "%s", msg);
    return -1;
}

static void *zalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

orc_state *orc_create(int64_t nx, int64_t ny, int64_t nz) {
    if (nx < 4 || ny < 4 || nz < 4) return NULL;
    orc_state *s = (orc_state *)calloc(1, sizeof *s);
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->sx = nx + 2; s->sy = ny + 2; s->sz = nz + 2;
    size_t P = (size_t)s->sx * s->sy * s->sz;
    s->x1 = zalloc(nx + 4, 8); s->y1 = zalloc(ny + 4, 8); s->z1 = zalloc(nz + 4, 8);
    s->deltax = zalloc(nx + 4, 8); s->deltay = zalloc(ny + 4, 8); s->deltaz = zalloc(nz + 4, 8);
    s->xu = zalloc(nx + 4, 8); s->yu = zalloc(ny + 4, 8); s->zu = zalloc(nz + 4, 8);
    s->xv = zalloc(nx + 4, 8); s->yv = zalloc(ny + 4, 8); s->zv = zalloc(nz + 4, 8);
    s->xw = zalloc(nx + 4, 8); s->yw = zalloc(ny + 4, 8); s->zw = zalloc(nz + 4, 8);
    s->xp = zalloc(nx + 4, 8); s->yp = zalloc(ny + 4, 8); s->zp = zalloc(nz + 4, 8);
    int64_t nn[6] = {nx, ny, nz, nx, ny, nz};
    for (int t = 0; t < 6; t++) {
        s->tn[t] = nn[t] + 3;
        s->tab[t] = zalloc((size_t)12 * s->tn[t], 8);
    }
    s->Acx = zalloc(nx * 3, 8); s->Acy = zalloc(ny * 3, 8); s->Acz = zalloc(nz * 3, 8);
    s->u = zalloc(P, 8); s->v = zalloc(P, 8); s->w = zalloc(P, 8);
    s->ut = zalloc(P, 8); s->vt = zalloc(P, 8); s->wt = zalloc(P, 8);
    s->p = zalloc(P, 8); s->pc = zalloc(P, 8); s->pco = zalloc(P, 8); s->b = zalloc(P, 8);
    s->resi_u = zalloc(P, 8); s->resi_v = zalloc(P, 8); s->resi_w = zalloc(P, 8);
    s->cell = zalloc(P, 8); s->minElemcell = zalloc(P, 8);
    return s;
}

#define FREEP(x) do { free(x); (x) = NULL; } while (0)
static void free_lists(orc_state *s) {
    FREEP(s->ibList); FREEP(s->fluidList); FREEP(s->solidList); FREEP(s->redList); FREEP(s->blackList);
    FREEP(s->nelp); FREEP(s->nelu1); FREEP(s->nelu2); FREEP(s->nelv1); FREEP(s->nelv2);
    FREEP(s->nelw1); FREEP(s->nelw2);
    FREEP(s->pNormDis); FREEP(s->u1NormDis); FREEP(s->u2NormDis); FREEP(s->v1NormDis);
    FREEP(s->v2NormDis); FREEP(s->w1NormDis); FREEP(s->w2NormDis);
}
static void free_mesh(orc_state *s) {
    FREEP(s->xnode); FREEP(s->ynode); FREEP(s->znode);
    FREEP(s->xnode1); FREEP(s->ynode1); FREEP(s->znode1);
    FREEP(s->ibElP1); FREEP(s->ibElP2); FREEP(s->ibElP3);
    FREEP(s->xcent); FREEP(s->ycent); FREEP(s->zcent);
    FREEP(s->cosAlpha); FREEP(s->cosBeta); FREEP(s->cosGamma);
    FREEP(s->alpha3); FREEP(s->beta3); FREEP(s->gamma3);
    FREEP(s->Elemcell); FREEP(s->ucell); FREEP(s->vcell); FREEP(s->wcell); FREEP(s->pcell);
}

void orc_free(orc_state *s) {
    if (s) {
        free(s->cell_old);
        free(s->contrib);
        free(s->body);
    }
    if (!s) return;
    free_lists(s);
    free_mesh(s);
    double *d[] = {s->x1, s->y1, s->z1, s->deltax, s->deltay, s->deltaz, s->xu, s->yu, s->zu,
                   s->xv, s->yv, s->zv, s->xw, s->yw, s->zw, s->xp, s->yp, s->zp,
                   s->Acx, s->Acy, s->Acz, s->u, s->v, s->w, s->ut, s->vt, s->wt, s->p,
                   s->pc, s->pco, s->b, s->resi_u, s->resi_v, s->resi_w};
    for (size_t i = 0; i < sizeof d / sizeof d[0]; i++) free(d[i]);
    for (int t = 0; t < 6; t++) free(s->tab[t]);
    free(s->cell);
    free(s->minElemcell);
    free(s);
}

const char *orc_last_error(orc_state *s) { return s ? s->err : "null state"; }

/* subReadInput.f90:22-24: rev = 1/re, deltat = dt_order*1e-3 */
int orc_set_params(orc_state *s, double re, double dt_order, double eps1, int64_t inor) {
    s->re = re;
    s->rev = 1.0 / re;
    s->dt_order = dt_order;
    s->deltat = dt_order * 1e-3;
    s->eps1 = eps1;
    s->inor = inor;
    return 0;
}

/* subReadInput.f90:56-129: spacings, ghost nodes, staggered coordinates */
static void grid_axis(int64_t n, const double *g, double *x1, double *dx) {
    for (int64_t i = 2; i <= n + 2; i++) x1[i] = g[i - 2];
    for (int64_t i = 2; i <= n + 1; i++) dx[i] = x1[i + 1] - x1[i];
    dx[1] = dx[2];
    dx[n + 2] = dx[n + 1];
    x1[1] = x1[2] - dx[1];
    x1[n + 3] = x1[n + 2] + dx[n + 2];
}

int orc_set_grid(orc_state *s, const double *gx, const double *gy, const double *gz) {
    int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    grid_axis(nx, gx, s->x1, s->deltax);
    grid_axis(ny, gy, s->y1, s->deltay);
    grid_axis(nz, gz, s->z1, s->deltaz);
    for (int64_t i = 1; i <= nx + 3; i++) s->xu[i] = s->x1[i];
    for (int64_t i = 1; i <= ny + 3; i++) s->yv[i] = s->y1[i];
    for (int64_t i = 1; i <= nz + 3; i++) s->zw[i] = s->z1[i];
    for (int64_t i = 1; i <= ny + 2; i++) {
        s->yu[i] = 0.5 * (s->y1[i] + s->y1[i + 1]);
        s->yw[i] = s->yu[i];
        s->yp[i] = s->yu[i];
    }
    for (int64_t i = 1; i <= nx + 2; i++) {
        s->xv[i] = 0.5 * (s->x1[i] + s->x1[i + 1]);
        s->xw[i] = s->xv[i];
        s->xp[i] = s->xv[i];
    }
    for (int64_t i = 1; i <= nz + 2; i++) {
        s->zu[i] = 0.5 * (s->z1[i] + s->z1[i + 1]);
        s->zv[i] = s->zu[i];
        s->zp[i] = s->zu[i];
    }
    return 0;
}

/* subReadInput.f90:147-167 (after parsing): nodes and triangle connectivity */
int orc_set_mesh(orc_state *s, int64_t n_nodes, const double *nodes, int64_t n_elems, const int64_t *conn) {
    free_mesh(s);
    s->stress_ready = 0;
    s->ibNodes = n_nodes;
    s->ibElems = n_elems;
    s->xnode = zalloc(n_nodes + 1, 8); s->ynode = zalloc(n_nodes + 1, 8); s->znode = zalloc(n_nodes + 1, 8);
    s->xnode1 = zalloc(n_nodes + 1, 8); s->ynode1 = zalloc(n_nodes + 1, 8); s->znode1 = zalloc(n_nodes + 1, 8);
    for (int64_t n = 1; n <= n_nodes; n++) {
        s->xnode[n] = nodes[3 * (n - 1) + 0];
        s->ynode[n] = nodes[3 * (n - 1) + 1];
        s->znode[n] = nodes[3 * (n - 1) + 2];
    }
    s->ibElP1 = zalloc(n_elems + 1, 8); s->ibElP2 = zalloc(n_elems + 1, 8); s->ibElP3 = zalloc(n_elems + 1, 8);
    for (int64_t n = 1; n <= n_elems; n++) {
        int64_t a = conn[3 * (n - 1)], b = conn[3 * (n - 1) + 1], c = conn[3 * (n - 1) + 2];
        if (a < 1 || b < 1 || c < 1 || a > n_nodes || b > n_nodes || c > n_nodes)
            return fail(s, "mesh connectivity out of range");
        s->ibElP1[n] = a; s->ibElP2[n] = b; s->ibElP3[n] = c;
    }
    size_t m = (size_t)n_elems + 1;
    s->xcent = zalloc(m, 8); s->ycent = zalloc(m, 8); s->zcent = zalloc(m, 8);
    s->cosAlpha = zalloc(m, 8); s->cosBeta = zalloc(m, 8); s->cosGamma = zalloc(m, 8);
    s->alpha3 = zalloc(m, 8); s->beta3 = zalloc(m, 8); s->gamma3 = zalloc(m, 8);
    return 0;
}

/* subSearch.f90:26-35 shiftSurfaceNodesInitial */
int orc_shift_nodes_initial(orc_state *s, double xs, double ys, double zs) {
    for (int64_t i = 1; i <= s->ibNodes; i++) {
        double xr = s->xnode[i], yr = s->ynode[i], zr = s->znode[i];
        s->xnode1[i] = xr + xs;
        s->ynode1[i] = yr + ys;
        s->znode1[i] = zr + zs;
    }
    return 0;
}

/* subSearch.f90:95-125 computeSurfaceNorm */
int orc_compute_surface_norm(orc_state *s) {
    double inor = (double)s->inor;
    for (int64_t n = 1; n <= s->ibElems; n++) {
        double p1x = s->xnode1[s->ibElP1[n]], p1y = s->ynode1[s->ibElP1[n]], p1z = s->znode1[s->ibElP1[n]];
        double p2x = s->xnode1[s->ibElP2[n]], p2y = s->ynode1[s->ibElP2[n]], p2z = s->znode1[s->ibElP2[n]];
        double p3x = s->xnode1[s->ibElP3[n]], p3y = s->ynode1[s->ibElP3[n]], p3z = s->znode1[s->ibElP3[n]];
        s->xcent[n] = ((p2x + p1x) + p3x) / 3.0;
        s->ycent[n] = ((p2y + p1y) + p3y) / 3.0;
        s->zcent[n] = ((p2z + p1z) + p3z) / 3.0;
        double ca = (p2y - p1y) * (p3z - p1z) - (p3y - p1y) * (p2z - p1z);
        double cb = (p2z - p1z) * (p3x - p1x) - (p3z - p1z) * (p2x - p1x);
        double cg = (p2x - p1x) * (p3y - p1y) - (p3x - p1x) * (p2y - p1y);
        s->alpha3[n] = ca;
        s->beta3[n] = cb;
        s->gamma3[n] = cg;
        double len = sqrt(((ca * ca) + (cb * cb)) + (cg * cg));
        s->cosAlpha[n] = (ca / len) * inor;
        s->cosBeta[n] = (cb / len) * inor;
        s->cosGamma[n] = (cg / len) * inor;
    }
    return 0;
}

/* subSearch.f90:204-247: signed distances of the 8 cell corners to the plane
 * of element e, and the tag they imply (1 solid, 0 fluid, 2 intercepted). */
static int64_t corner_tag(const orc_state *s, int64_t e, double n2x, double n3x, double n2y,
                          double n3y, double n2z, double n3z) {
    double cx = s->xcent[e], cy = s->ycent[e], cz = s->zcent[e];
    double a = s->cosAlpha[e], b = s->cosBeta[e], g = s->cosGamma[e];
    double d[8];
    d[0] = ((n2x - cx) * a + (n2y - cy) * b) + (n2z - cz) * g;
    d[1] = ((n2x - cx) * a + (n3y - cy) * b) + (n2z - cz) * g;
    d[2] = ((n3x - cx) * a + (n2y - cy) * b) + (n2z - cz) * g;
    d[3] = ((n3x - cx) * a + (n3y - cy) * b) + (n2z - cz) * g;
    d[4] = ((n2x - cx) * a + (n2y - cy) * b) + (n3z - cz) * g;
    d[5] = ((n2x - cx) * a + (n3y - cy) * b) + (n3z - cz) * g;
    d[6] = ((n3x - cx) * a + (n2y - cy) * b) + (n3z - cz) * g;
    d[7] = ((n3x - cx) * a + (n3y - cy) * b) + (n3z - cz) * g;
    int all_in = 1, all_out = 1;
    for (int c = 0; c < 8; c++) {
        if (!(d[c] <= -1e-16)) all_in = 0;
        if (!(d[c] > -1e-16)) all_out = 0;
    }
    if (all_in) return 1;
    if (all_out) return 0;
    return 2;
}

/* subSearch.f90:166-198: first-index argmin of the centroid distance */
static int64_t nearest_element(const orc_state *s, double n1x, double n1y, double n1z, int64_t prev) {
    double minDis = 1e14;
    int64_t nel = prev;
    for (int64_t m = 1; m <= s->ibElems; m++) {
        double dy = n1y - s->ycent[m], dx = n1x - s->xcent[m], dz = n1z - s->zcent[m];
        double dis = sqrt(((dy * dy) + (dx * dx)) + (dz * dz));
        if (dis < minDis) {
            minDis = dis;
            nel = m;
        }
    }
    return nel;
}

static void count_tags(orc_state *s) {
    s->ibCount = s->fluidCount = s->solidCount = 0;
    for (int64_t k = 2; k <= s->nz + 1; k++)
        for (int64_t j = 2; j <= s->ny + 1; j++)
            for (int64_t i = 2; i <= s->nx + 1; i++) {
                int64_t c = F(s->cell, i, j, k);
                if (c == 1) s->solidCount++;
                else if (c == 0) s->fluidCount++;
                else if (c == 2) s->ibCount++;
            }
}

/* subSearch.f90:134-271 tagging */
int orc_tagging(orc_state *s, const int64_t win[6]) {
    size_t P = (size_t)s->sx * s->sy * s->sz;
    memset(s->cell, 0, P * 8);
    memset(s->minElemcell, 0, P * 8);
    if (win[0] < 1 || win[1] > s->nx + 1 || win[2] < 1 || win[3] > s->ny + 1 || win[4] < 1 ||
        win[5] > s->nz + 1)
        return fail(s, "search window outside the grid");
    int64_t nel = 0;
    for (int64_t k = win[4]; k <= win[5]; k++)
        for (int64_t j = win[2]; j <= win[3]; j++)
            for (int64_t i = win[0]; i <= win[1]; i++) {
                nel = nearest_element(s, s->xp[i], s->yp[j], s->zp[k], nel);
                if (nel < 1) return fail(s, "no element closer than 1e14");
                F(s->minElemcell, i, j, k) = nel;
                F(s->cell, i, j, k) = corner_tag(s, nel, s->x1[i], s->x1[i + 1], s->y1[j], s->y1[j + 1],
                                                 s->z1[k], s->z1[k + 1]);
            }
    count_tags(s);
    return 0;
}

/* subSearch.f90:41-76 shiftSurfaceNodesMovingbody with the free motion
 * replaced by a prescribed rigid displacement (dx,dy,dz) of the initially
 * placed body: xnode1 = (xnode + xShift) + dx */
int orc_shift_nodes_moving(orc_state *s, double xs, double ys, double zs, double dx, double dy, double dz) {
    for (int64_t i = 1; i <= s->ibNodes; i++) {
        s->xnode1[i] = (s->xnode[i] + xs) + dx;
        s->ynode1[i] = (s->ynode[i] + ys) + dy;
        s->znode1[i] = (s->znode[i] + zs) + dz;
    }
    return 0;
}

/* subSearch.f90:338-463 selectiveRetagging: the 6x6x6 block
 * (i1-2..i1+3, j1-2..j1+3, k1-2..k1+3) around every intercepted cell of the
 * current list is tagged again against the moved surface; every other cell
 * keeps its tag.  minElemcell is cleared first (338-340).  Block cells outside
 * the padded index range 1..n+2 are skipped (the reference would index its
 * coordinate arrays outside their bounds there).  Cells that were solid and
 * are no longer take the wall velocity (no counterpart in the reference,
 * whose body never moves). */
int orc_selective_retagging(orc_state *s) {
    size_t P = (size_t)s->sx * s->sy * s->sz;
    if (!s->cell_old) s->cell_old = zalloc(P, 8);
    memcpy(s->cell_old, s->cell, P * 8);
    memset(s->minElemcell, 0, P * 8);
    int64_t nb = s->ibCount, nel = 0;
    for (int64_t nn = 1; nn <= nb; nn++) {
        int64_t i1 = LST(s->ibList, nb, nn, 1), j1 = LST(s->ibList, nb, nn, 2), k1 = LST(s->ibList, nb, nn, 3);
        for (int64_t k = k1 - 2; k <= k1 + 3; k++)
            for (int64_t j = j1 - 2; j <= j1 + 3; j++)
                for (int64_t i = i1 - 2; i <= i1 + 3; i++) {
                    if (i < 1 || i > s->nx + 2 || j < 1 || j > s->ny + 2 || k < 1 || k > s->nz + 2) continue;
                    nel = nearest_element(s, s->xp[i], s->yp[j], s->zp[k], nel);
                    if (nel < 1) return fail(s, "no element closer than 1e14");
                    F(s->minElemcell, i, j, k) = nel;
                    F(s->cell, i, j, k) = corner_tag(s, nel, s->x1[i], s->x1[i + 1], s->y1[j], s->y1[j + 1],
                                                     s->z1[k], s->z1[k + 1]);
                }
    }
    for (size_t q = 0; q < P; q++) {
        if (s->cell_old[q] == 1 && s->cell[q] != 1) {
            s->u[q] = s->wall_vel[0];
            s->v[q] = s->wall_vel[1];
            s->w[q] = s->wall_vel[2];
        }
    }
    count_tags(s);
    return 0;
}

int orc_set_wall_motion(orc_state *s, const double vel[3], const double acc[3]) {
    for (int q = 0; q < 3; q++) {
        s->wall_vel[q] = vel[q];
        s->wall_acc[q] = acc[q];
    }
    return 0;
}

/* dp/dn at the wall from the body acceleration a: -(a . n); the reference's
 * disabled form for a transverse motion is -(dv/dt) cosBeta (subForcing.f90:20) */
static inline double wall_dpdn(const orc_state *s, double na, double nb, double ng) {
    if (s->wall_acc[0] == 0.0 && s->wall_acc[1] == 0.0 && s->wall_acc[2] == 0.0) return 0.0;
    return -(((s->wall_acc[0] * na) + (s->wall_acc[1] * nb)) + (s->wall_acc[2] * ng));
}

/* subSearch.f90:515-596 cellCount: lists in (k,j,i) order, red = mod(i+j+k,2)==1 */
int orc_cell_count(orc_state *s) {
    free_lists(s);
    count_tags(s);
    s->ibList = zalloc(s->ibCount * 3, 8);
    s->fluidList = zalloc(s->fluidCount * 3, 8);
    s->solidList = zalloc(s->solidCount * 3, 8);
    int64_t a = 0, f = 0, q = 0;
    for (int64_t k = 2; k <= s->nz + 1; k++)
        for (int64_t j = 2; j <= s->ny + 1; j++)
            for (int64_t i = 2; i <= s->nx + 1; i++) {
                int64_t c = F(s->cell, i, j, k);
                if (c == 0) {
                    f++;
                    LST(s->fluidList, s->fluidCount, f, 1) = i;
                    LST(s->fluidList, s->fluidCount, f, 2) = j;
                    LST(s->fluidList, s->fluidCount, f, 3) = k;
                } else if (c == 1) {
                    q++;
                    LST(s->solidList, s->solidCount, q, 1) = i;
                    LST(s->solidList, s->solidCount, q, 2) = j;
                    LST(s->solidList, s->solidCount, q, 3) = k;
                } else if (c == 2) {
                    a++;
                    LST(s->ibList, s->ibCount, a, 1) = i;
                    LST(s->ibList, s->ibCount, a, 2) = j;
                    LST(s->ibList, s->ibCount, a, 3) = k;
                }
            }
    s->redCount = s->blackCount = 0;
    for (int64_t n = 1; n <= s->fluidCount; n++) {
        int64_t i = LST(s->fluidList, s->fluidCount, n, 1), j = LST(s->fluidList, s->fluidCount, n, 2),
                k = LST(s->fluidList, s->fluidCount, n, 3);
        if ((i + j + k) % 2 == 1) s->redCount++; else s->blackCount++;
    }
    s->redList = zalloc(s->redCount * 3, 8);
    s->blackList = zalloc(s->blackCount * 3, 8);
    int64_t r = 0, bk = 0;
    for (int64_t n = 1; n <= s->fluidCount; n++) {
        int64_t i = LST(s->fluidList, s->fluidCount, n, 1), j = LST(s->fluidList, s->fluidCount, n, 2),
                k = LST(s->fluidList, s->fluidCount, n, 3);
        if ((i + j + k) % 2 == 1) {
            r++;
            LST(s->redList, s->redCount, r, 1) = i;
            LST(s->redList, s->redCount, r, 2) = j;
            LST(s->redList, s->redCount, r, 3) = k;
        } else {
            bk++;
            LST(s->blackList, s->blackCount, bk, 1) = i;
            LST(s->blackList, s->blackCount, bk, 2) = j;
            LST(s->blackList, s->blackCount, bk, 3) = k;
        }
    }
    return 0;
}

/* subSearch.f90:621-672 computeNormDistance */
int orc_compute_norm_distance(orc_state *s) {
    int64_t nb = s->ibCount;
    size_t m = (size_t)nb + 1;
    int64_t **ia[] = {&s->nelp, &s->nelu1, &s->nelu2, &s->nelv1, &s->nelv2, &s->nelw1, &s->nelw2};
    double **da[] = {&s->pNormDis, &s->u1NormDis, &s->u2NormDis, &s->v1NormDis, &s->v2NormDis, &s->w1NormDis,
                     &s->w2NormDis};
    for (int q = 0; q < 7; q++) {
        free(*ia[q]);
        free(*da[q]);
        *ia[q] = NULL;
        *da[q] = NULL;
    }
    s->nelp = zalloc(m, 8); s->nelu1 = zalloc(m, 8); s->nelu2 = zalloc(m, 8);
    s->nelv1 = zalloc(m, 8); s->nelv2 = zalloc(m, 8); s->nelw1 = zalloc(m, 8); s->nelw2 = zalloc(m, 8);
    s->pNormDis = zalloc(m, 8); s->u1NormDis = zalloc(m, 8); s->u2NormDis = zalloc(m, 8);
    s->v1NormDis = zalloc(m, 8); s->v2NormDis = zalloc(m, 8); s->w1NormDis = zalloc(m, 8);
    s->w2NormDis = zalloc(m, 8);
    for (int64_t n = 1; n <= nb; n++) {
        int64_t i = LST(s->ibList, nb, n, 1), j = LST(s->ibList, nb, n, 2), k = LST(s->ibList, nb, n, 3);
        double n1x = s->xp[i], n2x = s->x1[i], n3x = s->x1[i + 1];
        double n1y = s->yp[j], n2y = s->y1[j], n3y = s->y1[j + 1];
        double n1z = s->zp[k], n2z = s->z1[k], n3z = s->z1[k + 1];
        int64_t e = F(s->minElemcell, i, j, k);
        if (e < 1) return fail(s, "intercepted cell without a nearest element");
        s->nelp[n] = s->nelu1[n] = s->nelu2[n] = s->nelv1[n] = s->nelv2[n] = s->nelw1[n] = s->nelw2[n] = e;
        double cx = s->xcent[e], cy = s->ycent[e], cz = s->zcent[e];
        double a = s->cosAlpha[e], b = s->cosBeta[e], g = s->cosGamma[e];
        s->pNormDis[n] = ((n1x - cx) * a + (n1y - cy) * b) + (n1z - cz) * g;
        s->u1NormDis[n] = ((n2x - cx) * a + (n1y - cy) * b) + (n1z - cz) * g;
        s->u2NormDis[n] = ((n3x - cx) * a + (n1y - cy) * b) + (n1z - cz) * g;
        s->v1NormDis[n] = ((n1x - cx) * a + (n2y - cy) * b) + (n1z - cz) * g;
        s->v2NormDis[n] = ((n1x - cx) * a + (n3y - cy) * b) + (n1z - cz) * g;
        s->w1NormDis[n] = ((n1x - cx) * a + (n1y - cy) * b) + (n2z - cz) * g;
        s->w2NormDis[n] = ((n1x - cx) * a + (n1y - cy) * b) + (n3z - cz) * g;
    }
    return 0;
}

/* subInitialConditions.f90:9-42: zero everything, u = 1 on fluid cells */
int orc_initial_conditions(orc_state *s) {
    size_t P = (size_t)s->sx * s->sy * s->sz;
    double *z[] = {s->u, s->ut, s->v, s->vt, s->w, s->wt, s->p, s->resi_u, s->resi_v, s->resi_w};
    for (size_t a = 0; a < sizeof z / sizeof z[0]; a++) memset(z[a], 0, P * 8);
    for (int64_t n = 1; n <= s->fluidCount; n++) {
        int64_t i = LST(s->fluidList, s->fluidCount, n, 1), j = LST(s->fluidList, s->fluidCount, n, 2),
                k = LST(s->fluidList, s->fluidCount, n, 3);
        F(s->u, i, j, k) = 1.0;
    }
    return 0;
}

/* subCoefficientMatrix.f90:8-53 */
int orc_coefficient_matrix(orc_state *s) {
    int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    for (int64_t j = 2; j <= ny + 1; j++) {
        double r1 = s->yp[j] - s->yp[j - 1], r2 = s->yp[j + 1] - s->yp[j], rs = r1 + r2;
        ACY(j - 1, 1) = 2.0 / (r1 * rs);
        ACY(j - 1, 2) = -2.0 / (r1 * r2);
        ACY(j - 1, 3) = 2.0 / (r2 * rs);
    }
    for (int64_t i = 2; i <= nx + 1; i++) {
        double r1 = s->xp[i] - s->xp[i - 1], r2 = s->xp[i + 1] - s->xp[i], rs = r1 + r2;
        ACX(i - 1, 1) = 2.0 / (r1 * rs);
        ACX(i - 1, 2) = -2.0 / (r1 * r2);
        ACX(i - 1, 3) = 2.0 / (r2 * rs);
    }
    for (int64_t i = 2; i <= nz + 1; i++) {
        double r1 = s->zp[i] - s->zp[i - 1], r2 = s->zp[i + 1] - s->zp[i], rs = r1 + r2;
        ACZ(i - 1, 1) = 2.0 / (r1 * rs);
        ACZ(i - 1, 2) = -2.0 / (r1 * r2);
        ACZ(i - 1, 3) = 2.0 / (r2 * rs);
    }
    ACX(1, 2) = ACX(1, 2) + ACX(1, 1);
    ACX(1, 1) = 0.0;
    ACX(nx, 2) = ACX(nx, 2) - ACX(nx, 3);
    ACX(nx, 3) = 0.0;
    ACY(1, 2) = ACY(1, 2) + ACY(1, 1);
    ACY(1, 1) = 0.0;
    ACY(ny, 2) = ACY(ny, 2) + ACY(ny, 3);
    ACY(ny, 3) = 0.0;
    ACZ(1, 2) = ACZ(1, 2) + ACZ(1, 1);
    ACZ(1, 1) = 0.0;
    ACZ(nz, 2) = ACZ(nz, 2) + ACZ(nz, 3);
    ACZ(nz, 3) = 0.0;
    return 0;
}

/* real powers with integral exponents, evaluated as the reference build does */
static inline double P2(double x) { return x * x; }
static inline double P3(double x) { return x * (x * x); }
static inline double P4(double x) { return pow(x, 4.0); }

/* subNavierStokes.f90:36-116: the twelve coefficients for one spacing triple.
 * `**2.0`, `**3.0`, `**4.0` carry real exponents in the reference. */
static void kk_coeffs(double t1, double t2, double t3, double out[12]) {
    double f_1 = t3;
    double f_2 = 2.0 * t3 + P2(t3);
    double f_3 = (3.0 * t3 + 3.0 * P2(t3)) + P3(t3);
    double f_4 = ((4.0 * t3 + 6.0 * P2(t3)) + 4.0 * P3(t3)) + P4(t3);
    double f_5 = 1.0 / (t1 * t2);
    double f_6 = (2.0 * t1 + 1.0) / (P2(t1) * P2(t2));
    double f_7 = ((3.0 * P2(t1) + 3.0 * t1) + 1.0) / (P3(t1) * P3(t2));
    double f_8 = (((4.0 * P3(t1) + 6.0 * P2(t1)) + 4.0 * t1) + 1.0) / (P4(t1) * P4(t2));

    double s51 = -1.0 / P4(t2);
    double s52 = f_4, s53 = f_4, s54 = s51;
    double s55 = -(f_1 / P4(t2) + f_4 / t2);
    double s56 = (f_4 / P2(t2) - f_2 / P4(t2));
    double s57 = -(f_3 / P4(t2) + f_4 / P3(t2));

    double s61 = -1.0, s62 = f_8, s63 = f_8, s64 = -1.0;
    double s65 = (f_5 + f_8), s66 = (f_8 - f_6), s67 = (f_7 + f_8);

    double s71 = -1.0, s72 = (1.0 + f_4), s73 = f_4;
    double s74 = (f_4 - f_1), s75 = (f_4 - f_2), s76 = (f_4 - f_3);

    double s81 = -1.0 / P4(t2);
    double s82 = (f_8 + (1.0 / P4(t2)));
    double s83 = f_8;
    double s84 = (f_5 / P4(t2) - f_8 / t2);
    double s85 = (f_8 / P2(t2) - f_6 / P4(t2));
    double s86 = (f_7 / P4(t2) - f_8 / P3(t2));

    double s91 = -s66 * s51;
    double s92 = s56 * s61;
    double s93 = -(s54 * s66 + s56 * s62);
    double s94 = (s56 * s64 + s66 * s52);
    double s95 = (s56 * s63 - s66 * s53);
    double s96 = (s56 * s65 - s66 * s55);
    double s97 = (s56 * s67 - s66 * s57);

    double s101 = -s85 * s71;
    double s102 = -s85 * s72;
    double s103 = s75 * s81;
    double s104 = s75 * s82;
    double s105 = (s75 * s83 - s85 * s73);
    double s106 = (s75 * s84 - s85 * s74);
    double s107 = (s75 * s86 - s85 * s76);

    out[0] = (s97 * s101 - s107 * s91);
    out[1] = (s97 * s102 + s107 * s93);
    out[2] = (s95 * s107 - s105 * s97);
    out[3] = (s94 * s107 + s104 * s97);
    out[4] = (s97 * s103 - s107 * s92);
    out[5] = (s97 * s106 - s107 * s96);

    double ak_1 = ((1.0 + 2.0 * t1) * (t3 + P2(t3))) * t2;
    double ak_2 = (1.0 + t1) * (2.0 * t3 + P2(t3));
    double ak_3 = (1.0 + t1);
    double ak_4 = (1.0 + t1) * P2(1.0 + t3);
    double ak_5 = (P2(1.0 + t1) * (t3 + P2(t3))) * t2;
    double ak_6 = (P2(t1) * t2) * (t3 + P2(t3));
    double ak_7 = ((1.0 + t1) * (t3 + P2(t3))) * t2;
    out[6] = ak_3;
    out[7] = -ak_4;
    out[8] = (ak_1 + ak_2);
    out[9] = -ak_5;
    out[10] = ak_6;
    out[11] = ak_7;
}

/* subNavierStokes.f90:3-619 non_uni_coeff: face spacing (36-283) and centre
 * spacing (285-595) tables; uw=uv, vw=vu, wv=wu are copies. */
int orc_non_uni_coeff(orc_state *s) {
    const double *d[3] = {s->deltax, s->deltay, s->deltaz};
    int64_t nn[3] = {s->nx, s->ny, s->nz};
    double c[12];
    for (int a = 0; a < 3; a++) {
        const double *dl = d[a];
        for (int64_t i = 2; i <= nn[a]; i++) {
            double t1 = dl[i] / dl[i - 1], t2 = dl[i + 1] / dl[i], t3 = dl[i + 2] / dl[i + 1];
            kk_coeffs(t1, t2, t3, c);
            for (int q = 0; q < 12; q++) TAB(a, q, i) = c[q];
        }
        for (int64_t i = 2; i <= nn[a]; i++) {
            double tmp1;
            if (i == 2) tmp1 = dl[i - 1];
            else tmp1 = 0.5 * (dl[i - 1] + dl[i - 2]);
            double tmp2 = 0.5 * (dl[i] + dl[i - 1]);
            double tmp3 = 0.5 * (dl[i] + dl[i + 1]);
            double tmp4 = 0.5 * (dl[i + 1] + dl[i + 2]);
            kk_coeffs(tmp2 / tmp1, tmp3 / tmp2, tmp4 / tmp3, c);
            for (int q = 0; q < 12; q++) TAB(3 + a, q, i) = c[q];
        }
    }
    return 0;
}

/* 3rd-order Kawamura-Kuwahara term of subNavierStokes.f90:812-831:
 * vel*(ca1 f2 + ca2 f1 + ca3 f0 + ca4 fm1 + ca5 fm2)/(ca6*dA)
 *   + |vel|*(ck1 f2 + ... + ck5 fm2)/(2*ck6*dB) */
static inline double kk3(const orc_state *s, int set, int64_t i, double vel, double f2, double f1,
                         double f0, double fm1, double fm2, double dA, double dB) {
    double sa = (((TAB(set, 0, i) * f2 + TAB(set, 1, i) * f1) + TAB(set, 2, i) * f0) + TAB(set, 3, i) * fm1) +
                TAB(set, 4, i) * fm2;
    double sk = (((TAB(set, 6, i) * f2 + TAB(set, 7, i) * f1) + TAB(set, 8, i) * f0) + TAB(set, 9, i) * fm1) +
                TAB(set, 10, i) * fm2;
    return (vel * sa) / (TAB(set, 5, i) * dA) + (fabs(vel) * sk) / ((2.0 * TAB(set, 11, i)) * dB);
}

/* non-uniform 1st-order upwind term of subNavierStokes.f90:847-866:
 * (vel/den)*((-1/(r(r+1))) fm - (1-1/r) f0 + (r/(r+1)) fp)
 *   - al*|vel|*(1/(2 den))*(2/(r(r+1)) fm - 2/r f0 + 2/(r+1) fp) */
static inline double up1(double vel, double den, double r, double fm, double f0, double fp) {
    const double al = 1.0;
    double t1 = (((-1.0 / (r * (r + 1.0))) * fm - (1.0 - 1.0 / r) * f0) + (r / (r + 1.0)) * fp);
    double t2 = ((2.0 / (r * (r + 1.0))) * fm - (2.0 / r) * f0) + (2.0 / (r + 1.0)) * fp;
    return (vel / den) * t1 - ((al * fabs(vel)) * (1.0 / (2.0 * den))) * t2;
}

/* subNavierStokes.f90:622-1122 nsMomentum (explicit Euler predictor) */
int orc_ns_momentum(orc_state *s) {
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz, nf = s->fluidCount;
    const double rev = s->rev, deltat = s->deltat;
    const double *u = s->u, *v = s->v, *w = s->w, *p = s->p;
    const double *dxa = s->deltax, *dya = s->deltay, *dza = s->deltaz;
    const int64_t *cell = s->cell;
#pragma omp parallel for schedule(static)
    for (int64_t n = 1; n <= nf; n++) {
        int64_t i = LST(s->fluidList, nf, n, 1), j = LST(s->fluidList, nf, n, 2), k = LST(s->fluidList, nf, n, 3);
        double dxr = dxa[i + 1], dx = dxa[i], dxl = dxa[i - 1];
        double dye = dya[j + 1], dy = dya[j], dyw = dya[j - 1];
        double dzt = dza[k + 1], dz = dza[k], dzb = dza[k - 1];

        double u22 = F(u, i - 1, j, k) - F(u, i, j, k);
        double u4 = F(u, i, j, k) - F(u, i + 1, j, k);
        double u6 = F(u, i, j - 1, k) - F(u, i, j, k);
        double u8 = F(u, i, j, k) - F(u, i, j + 1, k);
        double u10 = F(u, i, j, k - 1) - F(u, i, j, k);
        double u12 = F(u, i, j, k) - F(u, i, j, k + 1);
        double v4 = F(v, i - 1, j, k) - F(v, i, j, k);
        double v6 = F(v, i, j, k) - F(v, i + 1, j, k);
        double v8 = F(v, i, j - 1, k) - F(v, i, j, k);
        double v10 = F(v, i, j, k) - F(v, i, j + 1, k);
        double v12 = F(v, i, j, k - 1) - F(v, i, j, k);
        double v14 = F(v, i, j, k) - F(v, i, j, k + 1);
        double w6 = F(w, i - 1, j, k) - F(w, i, j, k);
        double w8 = F(w, i, j, k) - F(w, i + 1, j, k);
        double w10 = F(w, i, j - 1, k) - F(w, i, j, k);
        double w12 = F(w, i, j, k) - F(w, i, j + 1, k);
        double w14 = F(w, i, j, k - 1) - F(w, i, j, k);
        double w16 = F(w, i, j, k) - F(w, i, j, k + 1);

        double dpdx = (F(p, i, j, k) - F(p, i + 1, j, k)) / (0.5 * (dxr + dx));
        double dpdy = (F(p, i, j, k) - F(p, i, j + 1, k)) / (0.5 * (dye + dy));
        double dpdz = (F(p, i, j, k) - F(p, i, j, k + 1)) / (0.5 * (dzt + dz));
        double dx2xr = dx + dxr, dx2xl = dx + dxl, dy2ye = dy + dye, dy2yw = dy + dyw;
        double dz2zt = dz + dzt, dz2zb = dz + dzb;

        int no_ib = F(cell, i + 1, j, k) != 2 && F(cell, i - 1, j, k) != 2 && F(cell, i, j + 1, k) != 2 &&
                    F(cell, i, j - 1, k) != 2 && F(cell, i, j, k + 1) != 2 && F(cell, i, j, k - 1) != 2;

        /* ---- U momentum (792-897) ---- */
        double vu_e = F(v, i + 1, j, k) + (dxr / dx2xr) * (F(v, i, j, k) - F(v, i + 1, j, k));
        double vu_w = F(v, i + 1, j - 1, k) + (dxr / dx2xr) * (F(v, i, j - 1, k) - F(v, i + 1, j - 1, k));
        double v_in_um = 0.5 * (vu_e + vu_w);
        double wu_n = F(w, i + 1, j, k) + (dxr / dx2xr) * (F(w, i, j, k) - F(w, i + 1, j, k));
        double wu_s = F(w, i + 1, j, k - 1) + (dxr / dx2xr) * (F(w, i, j, k - 1) - F(w, i + 1, j, k - 1));
        double w_in_um = 0.5 * (wu_n + wu_s);
        double duudx, dvudy, dwudz;
        if (i != 2 && i < nx && j != 2 && j < ny + 1 && k != 2 && k < nz + 1 && no_ib) {
            double ddy = 0.5 * (dya[j] + dya[j - 1]);
            double ddye = 0.5 * (dya[j] + dya[j + 1]);
            double ddz = 0.5 * (dza[k] + dza[k - 1]);
            double ddzr = 0.5 * (dza[k] + dza[k + 1]);
            double uc = F(u, i, j, k);
            duudx = kk3(s, UU, i, uc, F(u, i + 2, j, k), F(u, i + 1, j, k), uc, F(u, i - 1, j, k),
                        F(u, i - 2, j, k), dxa[i + 1], dxa[i]);
            dvudy = kk3(s, VU, j, v_in_um, F(u, i, j + 2, k), F(u, i, j + 1, k), uc, F(u, i, j - 1, k),
                        F(u, i, j - 2, k), ddye, ddy);
            dwudz = kk3(s, WU, k, w_in_um, F(u, i, j, k + 2), F(u, i, j, k + 1), uc, F(u, i, j, k - 1),
                        F(u, i, j, k - 2), ddzr, ddz);
        } else {
            double r1x = dxa[i] / dxa[i + 1];
            double r1yn = 0.5 * (dya[j] + dya[j - 1]), r1yd = 0.5 * (dya[j] + dya[j + 1]);
            double r1y = r1yn / r1yd;
            double r1zn = 0.5 * (dza[k] + dza[k - 1]), r1zd = 0.5 * (dza[k] + dza[k + 1]);
            double r1z = r1zn / r1zd;
            duudx = up1(F(u, i, j, k), dxa[i + 1], r1x, F(u, i - 1, j, k), F(u, i, j, k), F(u, i + 1, j, k));
            dvudy = up1(v_in_um, r1yd, r1y, F(u, i, j - 1, k), F(u, i, j, k), F(u, i, j + 1, k));
            dwudz = up1(w_in_um, r1zd, r1z, F(u, i, j, k - 1), F(u, i, j, k), F(u, i, j, k + 1));
        }
        double d2udx2 = (2.0 / dx2xr) * ((-u4 / dxr) + (u22 / dx));
        double d2udy2 = (2.0 / (0.5 * (dy2ye + dy2yw))) * ((-u8 / (0.5 * dy2ye)) + (u6 / (0.5 * dy2yw)));
        double d2udz2 = (2.0 / (0.5 * (dz2zt + dz2zb))) * ((-u12 / (0.5 * dz2zt)) + (u10 / (0.5 * dz2zb)));
        double xtt2 = rev * ((d2udx2 + d2udy2) + d2udz2);
        double residu = ((-duudx - dvudy) - dwudz) + xtt2;
        F(s->ut, i, j, k) = F(u, i, j, k) + deltat * (residu + dpdx);
        F(s->resi_u, i, j, k) = residu;

        /* ---- V momentum (901-1006) ---- */
        double uv_e = F(u, i, j + 1, k) + (dye / dy2ye) * (F(u, i, j, k) - F(u, i, j + 1, k));
        double uv_w = F(u, i - 1, j + 1, k) + (dye / dy2ye) * (F(u, i - 1, j, k) - F(u, i - 1, j + 1, k));
        double u_in_vm = 0.5 * (uv_e + uv_w);
        double wv_n = F(w, i, j + 1, k) + (dye / dy2ye) * (F(w, i, j, k) - F(w, i, j + 1, k));
        double wv_s = F(w, i, j + 1, k - 1) + (dye / dy2ye) * (F(w, i, j, k - 1) - F(w, i, j + 1, k - 1));
        double w_in_vm = 0.5 * (wv_n + wv_s);
        double duvdx, dvvdy, dwvdz;
        if (i != 2 && i < nx + 1 && j != 2 && j < ny && k != 2 && k < nz + 1 && no_ib) {
            double ddx = 0.5 * (dxa[i] + dxa[i - 1]);
            double ddxr = 0.5 * (dxa[i] + dxa[i + 1]);
            double ddz = 0.5 * (dza[k] + dza[k - 1]);
            double ddzr = 0.5 * (dza[k] + dza[k + 1]);
            double vc = F(v, i, j, k);
            duvdx = kk3(s, UV, i, u_in_vm, F(v, i + 2, j, k), F(v, i + 1, j, k), vc, F(v, i - 1, j, k),
                        F(v, i - 2, j, k), ddxr, ddx);
            dvvdy = kk3(s, VV, j, vc, F(v, i, j + 2, k), F(v, i, j + 1, k), vc, F(v, i, j - 1, k),
                        F(v, i, j - 2, k), dya[j + 1], dya[j]);
            dwvdz = kk3(s, WU, k, w_in_vm, F(v, i, j, k + 2), F(v, i, j, k + 1), vc, F(v, i, j, k - 1),
                        F(v, i, j, k - 2), ddzr, ddz);
        } else {
            double r1xn = 0.5 * (dxa[i] + dxa[i - 1]), r1xd = 0.5 * (dxa[i] + dxa[i + 1]);
            double r1x = r1xn / r1xd;
            double r1y = dya[j] / dya[j + 1];
            double r1zn = 0.5 * (dza[k] + dza[k - 1]), r1zd = 0.5 * (dza[k] + dza[k + 1]);
            double r1z = r1zn / r1zd;
            duvdx = up1(u_in_vm, r1xd, r1x, F(v, i - 1, j, k), F(v, i, j, k), F(v, i + 1, j, k));
            dvvdy = up1(F(v, i, j, k), dya[j + 1], r1y, F(v, i, j - 1, k), F(v, i, j, k), F(v, i, j + 1, k));
            dwvdz = up1(w_in_vm, r1zd, r1z, F(v, i, j, k - 1), F(v, i, j, k), F(v, i, j, k + 1));
        }
        double d2vdx2 = (2.0 / (0.5 * (dx2xr + dx2xl))) * ((-v6 / (0.5 * dx2xr)) + (v4 / (0.5 * dx2xl)));
        double d2vdy2 = (2.0 / dy2ye) * ((-v10 / dye) + (v8 / dy));
        double d2vdz2 = (2.0 / (0.5 * (dz2zt + dz2zb))) * ((-v14 / (0.5 * dz2zt)) + (v12 / (0.5 * dz2zb)));
        double ytt2 = rev * ((d2vdx2 + d2vdy2) + d2vdz2);
        double residv = ((-duvdx - dvvdy) - dwvdz) + ytt2;
        F(s->vt, i, j, k) = F(v, i, j, k) + deltat * (residv + dpdy);
        F(s->resi_v, i, j, k) = residv;

        /* ---- W momentum (1010-1116) ---- */
        double uw_e = F(u, i, j, k + 1) + (dzt / dz2zt) * (F(u, i, j, k) - F(u, i, j, k + 1));
        double uw_w = F(u, i - 1, j, k + 1) + (dzt / dz2zt) * (F(u, i - 1, j, k) - F(u, i - 1, j, k + 1));
        double u_in_wm = 0.5 * (uw_e + uw_w);
        double vw_n = F(v, i, j, k + 1) + (dzt / dz2zt) * (F(v, i, j, k) - F(v, i, j, k + 1));
        double vw_s = F(v, i, j - 1, k + 1) + (dzt / dz2zt) * (F(v, i, j - 1, k) - F(v, i, j - 1, k + 1));
        double v_in_wm = 0.5 * (vw_n + vw_s);
        double duwdx, dvwdy, dwwdz;
        if (i != 2 && i < nx + 1 && j != 2 && j < ny + 1 && k != 2 && k < nz && no_ib) {
            double ddx = 0.5 * (dxa[i] + dxa[i - 1]);
            double ddxr = 0.5 * (dxa[i] + dxa[i + 1]);
            double ddy = 0.5 * (dya[j] + dya[j - 1]);
            double ddye = 0.5 * (dya[j] + dya[j + 1]);
            double wc = F(w, i, j, k);
            duwdx = kk3(s, UV, i, u_in_wm, F(w, i + 2, j, k), F(w, i + 1, j, k), wc, F(w, i - 1, j, k),
                        F(w, i - 2, j, k), ddxr, ddx);
            dvwdy = kk3(s, VU, j, v_in_wm, F(w, i, j + 2, k), F(w, i, j + 1, k), wc, F(w, i, j - 1, k),
                        F(w, i, j - 2, k), ddye, ddy);
            dwwdz = kk3(s, WW, k, wc, F(w, i, j, k + 2), F(w, i, j, k + 1), wc, F(w, i, j, k - 1),
                        F(w, i, j, k - 2), dza[k + 1], dza[k]);
        } else {
            double r1xn = 0.5 * (dxa[i] + dxa[i - 1]), r1xd = 0.5 * (dxa[i] + dxa[i + 1]);
            double r1x = r1xn / r1xd;
            double r1yn = 0.5 * (dya[j] + dya[j - 1]), r1yd = 0.5 * (dya[j] + dya[j + 1]);
            double r1y = r1yn / r1yd;
            double r1z = dza[k] / dza[k + 1];
            duwdx = up1(u_in_wm, r1xd, r1x, F(w, i - 1, j, k), F(w, i, j, k), F(w, i + 1, j, k));
            dvwdy = up1(v_in_wm, r1yd, r1y, F(w, i, j - 1, k), F(w, i, j, k), F(w, i, j + 1, k));
            dwwdz = up1(F(w, i, j, k), dza[k + 1], r1z, F(w, i, j, k - 1), F(w, i, j, k), F(w, i, j, k + 1));
        }
        double d2wdx2 = (2.0 / (0.5 * (dx2xr + dx2xl))) * ((-w8 / (0.5 * dx2xr)) + (w6 / (0.5 * dx2xl)));
        double d2wdy2 = (2.0 / (0.5 * (dy2ye + dy2yw))) * ((-w12 / (0.5 * dy2ye)) + (w10 / (0.5 * dy2yw)));
        double d2wdz2 = (2.0 / dz2zt) * ((-w16 / dzt) + (w14 / dz));
        double ztt2 = rev * ((d2wdx2 + d2wdy2) + d2wdz2);
        double residw = ((-duwdx - dvwdy) - dwwdz) + ztt2;
        F(s->wt, i, j, k) = F(w, i, j, k) + deltat * (residw + dpdz);
        F(s->resi_w, i, j, k) = residw;
    }
    return 0;
}

/* subBoundaryConditions.f90:139-193 velocityBC (the live branch after GO TO 11) */
int orc_velocity_bc(orc_state *s) {
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    double *ut = s->ut, *vt = s->vt, *wt = s->wt;
    for (int64_t k = 2; k <= nz + 1; k++)
        for (int64_t j = 2; j <= ny + 1; j++) {
            F(ut, 1, j, k) = 1.0;
            F(vt, 1, j, k) = -F(vt, 2, j, k);
            F(wt, 1, j, k) = -F(wt, 2, j, k);
            F(ut, nx + 1, j, k) = F(ut, nx, j, k);
            F(vt, nx + 2, j, k) = F(vt, nx + 1, j, k);
            F(wt, nx + 2, j, k) = F(wt, nx + 1, j, k);
        }
    for (int64_t k = 2; k <= nz + 1; k++)
        for (int64_t i = 2; i <= nx + 1; i++) {
            F(ut, i, 1, k) = F(ut, i, 2, k);
            F(wt, i, 1, k) = F(wt, i, 2, k);
            F(vt, i, 1, k) = 0.0;
            F(ut, i, ny + 2, k) = F(ut, i, ny + 1, k);
            F(wt, i, ny + 2, k) = F(wt, i, ny + 1, k);
            F(vt, i, ny + 1, k) = 0.0;
        }
    for (int64_t j = 2; j <= ny + 1; j++)
        for (int64_t i = 2; i <= nx + 1; i++) {
            F(ut, i, j, 1) = F(ut, i, j, 2);
            F(vt, i, j, 1) = F(vt, i, j, 2);
            F(wt, i, j, 1) = 0.0;
            F(ut, i, j, nz + 2) = F(ut, i, j, nz + 1);
            F(vt, i, j, nz + 2) = F(vt, i, j, nz + 1);
            F(wt, i, j, nz + 1) = 0.0;
        }
    return 0;
}

/* subBoundaryConditions.f90:207-215 solidCellBC */
int orc_solid_cell_bc(orc_state *s) {
    int64_t ns = s->solidCount;
    for (int64_t n = 1; n <= ns; n++) {
        int64_t i = LST(s->solidList, ns, n, 1), j = LST(s->solidList, ns, n, 2), k = LST(s->solidList, ns, n, 3);
        F(s->ut, i, j, k) = s->wall_vel[0];
        F(s->vt, i, j, k) = s->wall_vel[1];
        F(s->wt, i, j, k) = s->wall_vel[2];
        F(s->p, i, j, k) = 0.0;
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Image-point interpolation shared by the forcing and stress routines
 * (subForcing.f90:156-217, subStressCal1.f90:134-193): linear grid-line scans
 * that keep the last match and leave the index untouched when nothing
 * matches, trilinear interpolation and the 3-point non-uniform derivative.
 * ------------------------------------------------------------------------- */
static void scan(const double *c, int64_t lo, int64_t hi, double pos, int64_t *idx) {
    for (int64_t l = lo; l <= hi; l++)
        if (pos >= c[l] && pos < c[l + 1]) *idx = l;
}

static inline double lin(double a, double b, double r, double d) { return a + ((b - a) * r) / d; }

static inline double d3(double h1, double h2, double f1, double f2, double fp) {
    return (((h1 * h1) * f2 - (h2 * h2) * f1) + ((h2 * h2) - (h1 * h1)) * fp) / (((h1 * h2) * (h1 + h2)) + 1e-16);
}

typedef struct {
    const double *cx, *cy, *cz;
    int ox, oy, oz; /* 1 on the staggered axis: the lower value is F(i_x1-1) */
} comp_t;

/* value at the image point and its normal derivative */
static void image_interp(const orc_state *s, const double *Fd, const comp_t *c, double px, double py,
                         double pz, int64_t ix, int64_t iy, int64_t iz, double na, double nb, double ng,
                         double *fpos, double *dfdn) {
    int64_t ax = ix - c->ox, ay = iy - c->oy, az = iz - c->oz;
    double f000 = F(Fd, ax, ay, az), f100 = F(Fd, ax + 1, ay, az);
    double f010 = F(Fd, ax, ay + 1, az), f110 = F(Fd, ax + 1, ay + 1, az);
    double f001 = F(Fd, ax, ay, az + 1), f101 = F(Fd, ax + 1, ay, az + 1);
    double f011 = F(Fd, ax, ay + 1, az + 1), f111 = F(Fd, ax + 1, ay + 1, az + 1);
    double rx = px - c->cx[ix], gx = c->cx[ix + 1] - c->cx[ix];
    double ry = py - c->cy[iy], gy = c->cy[iy + 1] - c->cy[iy];
    double rz = pz - c->cz[iz], gz = c->cz[iz + 1] - c->cz[iz];
    double x1z1 = lin(f000, f100, rx, gx), x2z1 = lin(f010, f110, rx, gx);
    double x1z2 = lin(f001, f101, rx, gx), x2z2 = lin(f011, f111, rx, gx);
    double vz1 = lin(x1z1, x2z1, ry, gy), vz2 = lin(x1z2, x2z2, ry, gy);
    double vy1 = lin(x1z1, x1z2, rz, gz), vy2 = lin(x2z1, x2z2, rz, gz);
    double z1x1 = lin(f000, f001, rz, gz), z2x1 = lin(f010, f011, rz, gz);
    double z1x2 = lin(f100, f101, rz, gz), z2x2 = lin(f110, f111, rz, gz);
    double vx1 = lin(z1x1, z2x1, ry, gy), vx2 = lin(z1x2, z2x2, ry, gy);
    double vp = lin(vx1, vx2, rx, gx);
    double h2, h1;
    h2 = fabs(c->cx[ix + 1] - px); h1 = fabs(c->cx[ix] - px);
    double ddx = d3(h1, h2, vx1, vx2, vp);
    h2 = fabs(c->cy[iy + 1] - py); h1 = fabs(c->cy[iy] - py);
    double ddy = d3(h1, h2, vy1, vy2, vp);
    h2 = fabs(c->cz[iz + 1] - pz); h1 = fabs(c->cz[iz] - pz);
    double ddz = d3(h1, h2, vz1, vz2, vp);
    *fpos = vp;
    *dfdn = (ddx * na + ddy * nb) + ddz * ng;
}

/* One ghost face of subForcing.f90:143-224 (and its five siblings): image
 * point at pt1 along the normal, quadratic u(d)=a d^2+b d+c with c = 0. */
static int ghost_face(orc_state *s, double *Fd, const comp_t *c, int64_t i, int64_t j, int64_t k,
                      double nx0, double ny0, double nz0, double sur, int64_t e, int64_t sxh, int64_t syh,
                      int64_t szh, int clamp_axis, int64_t from, int64_t to, int64_t idx[3], double *out) {
    const double surf = s->wall_vel[c->ox ? 0 : c->oy ? 1 : 2];
    double pt1 = 1.5 * sqrt(((s->deltax[i] * s->deltax[i]) + (s->deltay[j] * s->deltay[j])) +
                            (s->deltaz[k] * s->deltaz[k])) + (fabs(sur) - sur) * 0.5;
    double na = s->cosAlpha[e], nb = s->cosBeta[e], ng = s->cosGamma[e];
    double px = nx0 + pt1 * na, py = ny0 + pt1 * nb, pz = nz0 + pt1 * ng;
    scan(c->cx, 1, sxh, px, &idx[0]);
    scan(c->cy, 1, syh, py, &idx[1]);
    scan(c->cz, 1, szh, pz, &idx[2]);
    if (idx[0] < 1 || idx[1] < 1 || idx[2] < 1) return fail(s, "image point outside the grid lines");
    if (clamp_axis >= 0 && idx[clamp_axis] == from) idx[clamp_axis] = to;
    double fpos, dn;
    image_interp(s, Fd, c, px, py, pz, idx[0], idx[1], idx[2], na, nb, ng, &fpos, &dn);
    double n1 = pt1 + sur;
    double cval = surf;
    double bval = (2.0 / n1) * (fpos - surf) - dn;
    double aval = dn / n1 - (fpos - surf) / (n1 * n1);
    *out = ((aval * (sur * sur)) + bval * sur) + cval;
    return 0;
}

/* subForcing.f90:107-652 velocityForcing1 (serial order: the later list entry
 * wins a shared face, and later faces read values written earlier) */
int orc_velocity_forcing(orc_state *s) {
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz, nb = s->ibCount;
    comp_t cu = {s->xu, s->yu, s->zu, 1, 0, 0};
    comp_t cv = {s->xv, s->yv, s->zv, 0, 1, 0};
    comp_t cw = {s->xw, s->yw, s->zw, 0, 0, 1};
    int64_t idx[3] = {-1, -1, -1};
    double val;
    for (int64_t n = 1; n <= nb; n++) {
        int64_t i = LST(s->ibList, nb, n, 1), j = LST(s->ibList, nb, n, 2), k = LST(s->ibList, nb, n, 3);
        if (ghost_face(s, s->ut, &cu, i, j, k, s->xu[i + 1], s->yu[j], s->zu[k], s->u2NormDis[n], s->nelu2[n],
                       nx + 2, ny + 1, nz + 1, 0, nx + 2, nx + 1, idx, &val)) return -1;
        F(s->ut, i, j, k) = val;
        if (ghost_face(s, s->ut, &cu, i, j, k, s->xu[i], s->yu[j], s->zu[k], s->u1NormDis[n], s->nelu1[n],
                       nx + 1, ny + 1, nz + 1, 0, 1, 2, idx, &val)) return -1;
        F(s->ut, i - 1, j, k) = val;
        if (ghost_face(s, s->vt, &cv, i, j, k, s->xv[i], s->yv[j + 1], s->zv[k], s->v2NormDis[n], s->nelv2[n],
                       nx + 1, ny + 2, nz + 1, 1, ny + 2, ny + 1, idx, &val)) return -1;
        F(s->vt, i, j, k) = val;
        if (ghost_face(s, s->vt, &cv, i, j, k, s->xv[i], s->yv[j], s->zv[k], s->v1NormDis[n], s->nelv1[n],
                       nx + 1, ny + 1, nz + 1, 1, 1, 2, idx, &val)) return -1;
        F(s->vt, i, j - 1, k) = val;
        if (ghost_face(s, s->wt, &cw, i, j, k, s->xw[i], s->yw[j], s->zw[k + 1], s->w2NormDis[n], s->nelw2[n],
                       nx + 1, ny + 1, nz + 2, 2, nz + 2, nz + 1, idx, &val)) return -1;
        F(s->wt, i, j, k) = val;
        if (ghost_face(s, s->wt, &cw, i, j, k, s->xw[i], s->yw[j], s->zw[k], s->w1NormDis[n], s->nelw1[n],
                       nx + 1, ny + 1, nz + 1, 2, 1, 2, idx, &val)) return -1;
        F(s->wt, i, j, k - 1) = val;
    }
    return 0;
}

/* subForcing.f90:2-103 pressureForcing1: dp/dn = 0 at the wall */
int orc_pressure_forcing(orc_state *s) {
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz, nb = s->ibCount;
    comp_t cp = {s->xp, s->yp, s->zp, 0, 0, 0};
    int64_t idx[3] = {-1, -1, -1};
    for (int64_t n = 1; n <= nb; n++) {
        int64_t i = LST(s->ibList, nb, n, 1), j = LST(s->ibList, nb, n, 2), k = LST(s->ibList, nb, n, 3);
        double sur = s->pNormDis[n];
        double pt1 = 1.5 * sqrt(((s->deltax[i] * s->deltax[i]) + (s->deltay[j] * s->deltay[j])) +
                                (s->deltaz[k] * s->deltaz[k])) + (fabs(sur) - sur) * 0.5;
        int64_t e = s->nelp[n];
        double na = s->cosAlpha[e], nb_ = s->cosBeta[e], ng = s->cosGamma[e];
        double px = s->xp[i] + pt1 * na, py = s->yp[j] + pt1 * nb_, pz = s->zp[k] + pt1 * ng;
        scan(s->xp, 1, nx + 1, px, &idx[0]);
        scan(s->yp, 1, ny + 1, py, &idx[1]);
        scan(s->zp, 1, nz + 1, pz, &idx[2]);
        if (idx[0] < 1 || idx[1] < 1 || idx[2] < 1) return fail(s, "image point outside the grid lines");
        double fpos, dn;
        image_interp(s, s->p, &cp, px, py, pz, idx[0], idx[1], idx[2], na, nb_, ng, &fpos, &dn);
        const double dpdn = wall_dpdn(s, na, nb_, ng);
        double n1 = pt1 + sur;
        double bval = dpdn;
        double aval = (dn - dpdn) / (2.0 * n1);
        double cval = fpos - ((dn + dpdn) * n1) * 0.5;
        F(s->p, i, j, k) = ((aval * (sur * sur)) + bval * sur) + cval;
    }
    return 0;
}

/* subPcorVcor.f90:91-99 computeDiv */
int orc_compute_div(orc_state *s, double *divmax) {
    int64_t nf = s->fluidCount;
    double dm = 0.0;
#pragma omp parallel for schedule(static) reduction(max : dm)
    for (int64_t n = 1; n <= nf; n++) {
        int64_t i = LST(s->fluidList, nf, n, 1), j = LST(s->fluidList, nf, n, 2), k = LST(s->fluidList, nf, n, 3);
        double bb = ((F(s->ut, i, j, k) - F(s->ut, i - 1, j, k)) / s->deltax[i] +
                     (F(s->vt, i, j, k) - F(s->vt, i, j - 1, k)) / s->deltay[j]) +
                    (F(s->wt, i, j, k) - F(s->wt, i, j, k - 1)) / s->deltaz[k];
        F(s->b, i, j, k) = bb;
        double ab = fabs(bb);
        dm = (ab > dm || dm != dm) ? ab : dm;
    }
    *divmax = dm;
    return 0;
}

/* one colour sweep of subPcorVcor.f90:164-190 */
static void sor_colour(orc_state *s, const int64_t *lst, int64_t cnt, const double *src, double omega) {
    const double deltat = s->deltat;
#pragma omp parallel for schedule(static)
    for (int64_t n = 1; n <= cnt; n++) {
        int64_t i = LST(lst, cnt, n, 1), j = LST(lst, cnt, n, 2), k = LST(lst, cnt, n, 3);
        double g = ((((((F(s->b, i, j, k) / deltat - ACX(i - 1, 1) * F(src, i - 1, j, k)) -
                        ACX(i - 1, 3) * F(src, i + 1, j, k)) -
                       ACY(j - 1, 1) * F(src, i, j - 1, k)) -
                      ACY(j - 1, 3) * F(src, i, j + 1, k)) -
                     ACZ(k - 1, 1) * F(src, i, j, k - 1)) -
                    ACZ(k - 1, 3) * F(src, i, j, k + 1)) /
                   ((ACX(i - 1, 2) + ACY(j - 1, 2)) + ACZ(k - 1, 2));
        F(s->pc, i, j, k) = (1.0 - omega) * F(s->pco, i, j, k) + omega * g;
    }
}

/* subPcorVcor.f90:148-212 REDBLACKSOR: red reads pco, black reads pc, both
 * relax against pco; stop once max|pc-pco| < epsi over fluid cells. */
int orc_redblack_sor(orc_state *s, double epsi, int64_t *isum, double *derr2_out) {
    const double omega = 1.99;
    int64_t nf = s->fluidCount;
    int64_t it = 0;
    double derr2;
    do {
        it++;
        derr2 = 0.0;
        sor_colour(s, s->redList, s->redCount, s->pco, omega);
        sor_colour(s, s->blackList, s->blackCount, s->pc, omega);
#pragma omp parallel for schedule(static) reduction(max : derr2)
        for (int64_t n = 1; n <= nf; n++) {
            int64_t i = LST(s->fluidList, nf, n, 1), j = LST(s->fluidList, nf, n, 2), k = LST(s->fluidList, nf, n, 3);
            double d = fabs(F(s->pc, i, j, k) - F(s->pco, i, j, k));
            derr2 = (d > derr2 || derr2 != derr2) ? d : derr2;
            F(s->pco, i, j, k) = F(s->pc, i, j, k);
        }
        if (it > 100000000) return fail(s, "SOR did not converge");
    } while (derr2 >= epsi);
    *isum = it;
    *derr2_out = derr2;
    return 0;
}

/* subPcorVcor.f90:113-118 correctPressure */
int orc_correct_pressure(orc_state *s) {
    int64_t nf = s->fluidCount;
#pragma omp parallel for schedule(static)
    for (int64_t n = 1; n <= nf; n++) {
        int64_t i = LST(s->fluidList, nf, n, 1), j = LST(s->fluidList, nf, n, 2), k = LST(s->fluidList, nf, n, 3);
        F(s->p, i, j, k) = F(s->p, i, j, k) + F(s->pc, i, j, k);
    }
    return 0;
}

/* subPcorVcor.f90:132-140 correctVelocity */
int orc_correct_velocity(orc_state *s) {
    int64_t nf = s->fluidCount;
    const double dt = s->deltat;
#pragma omp parallel for schedule(static)
    for (int64_t n = 1; n <= nf; n++) {
        int64_t i = LST(s->fluidList, nf, n, 1), j = LST(s->fluidList, nf, n, 2), k = LST(s->fluidList, nf, n, 3);
        F(s->ut, i, j, k) = F(s->ut, i, j, k) -
                            (dt / (0.5 * (s->deltax[i + 1] + s->deltax[i]))) * (F(s->pc, i + 1, j, k) - F(s->pc, i, j, k));
        F(s->vt, i, j, k) = F(s->vt, i, j, k) -
                            (dt / (0.5 * (s->deltay[j + 1] + s->deltay[j]))) * (F(s->pc, i, j + 1, k) - F(s->pc, i, j, k));
        F(s->wt, i, j, k) = F(s->wt, i, j, k) -
                            (dt / (0.5 * (s->deltaz[k + 1] + s->deltaz[k]))) * (F(s->pc, i, j, k + 1) - F(s->pc, i, j, k));
    }
    return 0;
}

/* subPcorVcor.f90:2-80 poissonSolver */
int orc_poisson_solver(orc_state *s, int64_t *niter, double *divmax, double *derr_std) {
    size_t P = (size_t)s->sx * s->sy * s->sz;
    memset(s->b, 0, P * 8);
    memset(s->pc, 0, P * 8);
    memset(s->pco, 0, P * 8);
    orc_compute_div(s, divmax);
    double derr2;
    if (orc_redblack_sor(s, s->eps1, niter, &derr2)) return -1;
    orc_correct_pressure(s);
    orc_correct_velocity(s);
    orc_velocity_bc(s);
    int64_t nf = s->fluidCount;
    double m = 0.0;
    for (int64_t n = 1; n <= nf; n++) {
        int64_t i = LST(s->fluidList, nf, n, 1), j = LST(s->fluidList, nf, n, 2), k = LST(s->fluidList, nf, n, 3);
        double a = fabs(F(s->ut, i, j, k) - F(s->u, i, j, k)) / s->deltat;
        double b = fabs(F(s->vt, i, j, k) - F(s->v, i, j, k)) / s->deltat;
        double c = fabs(F(s->wt, i, j, k) - F(s->w, i, j, k)) / s->deltat;
        if (a > m) m = a;
        if (b > m) m = b;
        if (c > m) m = c;
    }
    *derr_std = m;
    memcpy(s->u, s->ut, P * 8);
    memcpy(s->v, s->vt, P * 8);
    memcpy(s->w, s->wt, P * 8);
    return 0;
}

/* subStressCal1.f90:85-450 / subStressCal2.f90:78-371: per-triangle wall
 * shear and pressure, summed in element order. */
int orc_stress(orc_state *s, int first, double out[4]) {
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz, ne = s->ibElems;
    if (first || !s->stress_ready) {
        FREEP(s->Elemcell); FREEP(s->ucell); FREEP(s->vcell); FREEP(s->wcell); FREEP(s->pcell);
        s->Elemcell = zalloc(ne * 3, 8); s->ucell = zalloc(ne * 3, 8); s->vcell = zalloc(ne * 3, 8);
        s->wcell = zalloc(ne * 3, 8); s->pcell = zalloc(ne * 3, 8);
    }
    comp_t cu = {s->xu, s->yu, s->zu, 1, 0, 0};
    comp_t cv = {s->xv, s->yv, s->zv, 0, 1, 0};
    comp_t cw = {s->xw, s->yw, s->zw, 0, 0, 1};
    comp_t cp = {s->xp, s->yp, s->zp, 0, 0, 0};
    double viscousDrag = 0.0, pressureDrag = 0.0, viscousLift = 0.0, pressureLift = 0.0;
    double area_Sx = 0.0, area_Sy = 0.0, surf_area = 0.0;
    int64_t ic = -1, jc = -1, kc = -1, idx[3] = {-1, -1, -1};
    int use_cache = !first && s->stress_ready;
    for (int64_t e = 1; e <= ne; e++) {
        if (use_cache) {
            ic = LST(s->Elemcell, ne, e, 1); jc = LST(s->Elemcell, ne, e, 2); kc = LST(s->Elemcell, ne, e, 3);
        } else {
            scan(s->x1, 2, nx + 1, s->xcent[e], &ic);
            scan(s->y1, 2, ny + 1, s->ycent[e], &jc);
            scan(s->z1, 2, nz + 1, s->zcent[e], &kc);
            if (ic < 1 || jc < 1 || kc < 1) return fail(s, "element centroid outside the grid");
            LST(s->Elemcell, ne, e, 1) = ic; LST(s->Elemcell, ne, e, 2) = jc; LST(s->Elemcell, ne, e, 3) = kc;
        }
        double xs = s->xcent[e], ys = s->ycent[e], zs = s->zcent[e];
        double dX = s->x1[ic + 1] - s->x1[ic], dY = s->y1[jc + 1] - s->y1[jc], dZ = s->z1[kc + 1] - s->z1[kc];
        double diagdis = sqrt(((dX * dX) + (dY * dY)) + (dZ * dZ));
        double normdis = diagdis;
        double na = s->cosAlpha[e], nb = s->cosBeta[e], ng = s->cosGamma[e];
        double px = xs + normdis * na, py = ys + normdis * nb, pz = zs + normdis * ng;
        const double usurf = s->wall_vel[0], vsurf = s->wall_vel[1], wsurf = s->wall_vel[2];
        const double dpdn = wall_dpdn(s, na, nb, ng);
        double fpos, dn, dudn_s, dvdn_s, dwdn_s;
        const comp_t *cs[3] = {&cu, &cv, &cw};
        int64_t *cache[3] = {s->ucell, s->vcell, s->wcell};
        double *fld[3] = {s->ut, s->vt, s->wt};
        double *dnds[3] = {&dudn_s, &dvdn_s, &dwdn_s};
        const double surfv[3] = {usurf, vsurf, wsurf};
        for (int q = 0; q < 3; q++) {
            if (use_cache) {
                idx[0] = LST(cache[q], ne, e, 1); idx[1] = LST(cache[q], ne, e, 2); idx[2] = LST(cache[q], ne, e, 3);
            } else {
                scan(cs[q]->cx, 2, nx + 1, px, &idx[0]);
                scan(cs[q]->cy, 2, ny + 1, py, &idx[1]);
                scan(cs[q]->cz, 2, nz + 1, pz, &idx[2]);
                if (idx[0] < 1 || idx[1] < 1 || idx[2] < 1) return fail(s, "probe point outside the grid");
                LST(cache[q], ne, e, 1) = idx[0]; LST(cache[q], ne, e, 2) = idx[1]; LST(cache[q], ne, e, 3) = idx[2];
            }
            image_interp(s, fld[q], cs[q], px, py, pz, idx[0], idx[1], idx[2], na, nb, ng, &fpos, &dn);
            *dnds[q] = (2.0 / normdis) * (fpos - surfv[q]) - dn;
        }
        double alen = sqrt((P2(s->alpha3[e]) + P2(s->beta3[e])) + P2(s->gamma3[e]));
        double area = alen / 2.0;
        double area_yz = 0.5 * fabs(s->alpha3[e]);
        double area_xz = 0.5 * fabs(s->beta3[e]);
        double proj = (dudn_s * na + dvdn_s * nb) + dwdn_s * ng;
        double stx1 = dudn_s - proj * na;
        double sty1 = dvdn_s - proj * nb;
        double shear_x = ((1.0 / s->re) * stx1) * area;
        double shear_y = ((1.0 / s->re) * sty1) * area;
        if (use_cache) {
            idx[0] = LST(s->pcell, ne, e, 1); idx[1] = LST(s->pcell, ne, e, 2); idx[2] = LST(s->pcell, ne, e, 3);
        } else {
            scan(s->xp, 2, nx + 1, px, &idx[0]);
            scan(s->yp, 2, ny + 1, py, &idx[1]);
            scan(s->zp, 2, nz + 1, pz, &idx[2]);
            LST(s->pcell, ne, e, 1) = idx[0]; LST(s->pcell, ne, e, 2) = idx[1]; LST(s->pcell, ne, e, 3) = idx[2];
        }
        image_interp(s, s->p, &cp, px, py, pz, idx[0], idx[1], idx[2], na, nb, ng, &fpos, &dn);
        double cval = fpos - ((dn + dpdn) * diagdis) * 0.5;
        double f_surf = cval * area;
        double f_surf_x = -f_surf * na;
        double f_surf_y = -f_surf * nb;
        if (s->contrib) {
            double *o = s->contrib;
            o[(e - 1)] = shear_x;
            o[ne + (e - 1)] = f_surf_x;
            o[2 * ne + (e - 1)] = shear_y;
            o[3 * ne + (e - 1)] = f_surf_y;
            o[4 * ne + (e - 1)] = area;
            o[5 * ne + (e - 1)] = area_xz;
            o[6 * ne + (e - 1)] = area_yz;
        }
        viscousDrag = viscousDrag + shear_x;
        pressureDrag = pressureDrag + f_surf_x;
        viscousLift = viscousLift + shear_y;
        pressureLift = pressureLift + f_surf_y;
        surf_area = surf_area + area;
        area_Sx = area_Sx + area_xz;
        area_Sy = area_Sy + area_yz;
    }
    s->stress_ready = 1;
    area_Sx = 0.5 * area_Sx;
    area_Sy = 0.5 * area_Sy;
    out[0] = 2.0 * (viscousDrag / area_Sx);
    out[1] = 2.0 * (pressureDrag / area_Sy);
    out[2] = 2.0 * (viscousLift / area_Sy);
    out[3] = 2.0 * (pressureLift / area_Sx);
    return 0;
}

/* Per-body coefficients: the sums of subStressCal1.f90:327-331 and the
 * normalisation of 441-450 restricted to the elements of each body (body ids
 * 1..nbody, one per element in mesh order); no counterpart in the reference,
 * which sums over all elements. */
int orc_set_bodies(orc_state *s, const int32_t *body, int32_t nbody) {
    free(s->body);
    free(s->contrib);
    s->body = zalloc(s->ibElems, sizeof(int32_t));
    memcpy(s->body, body, (size_t)s->ibElems * sizeof(int32_t));
    s->nbody = nbody;
    s->contrib = zalloc((size_t)7 * s->ibElems, 8);
    return 0;
}

int orc_stress_bodies(orc_state *s, double *out) {
    if (!s->contrib) return fail(s, "orc_set_bodies has not run");
    const int64_t ne = s->ibElems;
    for (int32_t k = 1; k <= s->nbody; k++) {
        double vd = 0.0, pd = 0.0, vl = 0.0, pl = 0.0, ax = 0.0, ay = 0.0;
        for (int64_t e = 0; e < ne; e++) {
            if (s->body[e] != k) continue;
            vd = vd + s->contrib[e];
            pd = pd + s->contrib[ne + e];
            vl = vl + s->contrib[2 * ne + e];
            pl = pl + s->contrib[3 * ne + e];
            ax = ax + s->contrib[5 * ne + e];
            ay = ay + s->contrib[6 * ne + e];
        }
        ax = 0.5 * ax;
        ay = 0.5 * ay;
        double *o = out + 4 * (k - 1);
        o[0] = 2.0 * (vd / ax);
        o[1] = 2.0 * (pd / ay);
        o[2] = 2.0 * (vl / ay);
        o[3] = 2.0 * (pl / ax);
    }
    return 0;
}

/* main.f90:156-174 (first step: 139-149 with stressCal1) */
int orc_step(orc_state *s, int first, int64_t *niter, double *divmax, double *derr_std, double coef[4]) {
    if (orc_ns_momentum(s)) return -1;
    orc_velocity_bc(s);
    orc_solid_cell_bc(s);
    if (orc_velocity_forcing(s)) return -1;
    orc_velocity_bc(s);
    if (orc_poisson_solver(s, niter, divmax, derr_std)) return -1;
    if (orc_pressure_forcing(s)) return -1;
    return orc_stress(s, first, coef);
}

/* ---- north-star Poisson modes ------------------------------------------ */
/* r = b/dt - A pc over fluid cells (A as in subPcorVcor.f90:170-173) */
int orc_residual_l2(orc_state *s, double *rnorm, double *bnorm) {
    int64_t nf = s->fluidCount;
    double rr = 0.0, bb = 0.0;
    const double dt = s->deltat;
    const double *pc = s->pc;
    for (int64_t n = 1; n <= nf; n++) {
        int64_t i = LST(s->fluidList, nf, n, 1), j = LST(s->fluidList, nf, n, 2), k = LST(s->fluidList, nf, n, 3);
        double bd = F(s->b, i, j, k) / dt;
        double ap = ((((((ACX(i - 1, 1) * F(pc, i - 1, j, k) + ACX(i - 1, 3) * F(pc, i + 1, j, k)) +
                         ACY(j - 1, 1) * F(pc, i, j - 1, k)) + ACY(j - 1, 3) * F(pc, i, j + 1, k)) +
                       ACZ(k - 1, 1) * F(pc, i, j, k - 1)) + ACZ(k - 1, 3) * F(pc, i, j, k + 1)) +
                     ((ACX(i - 1, 2) + ACY(j - 1, 2)) + ACZ(k - 1, 2)) * F(pc, i, j, k));
        double r = bd - ap;
        rr += r * r;
        bb += bd * bd;
    }
    *rnorm = sqrt(rr);
    *bnorm = sqrt(bb);
    return 0;
}

int orc_sor_l2(orc_state *s, double omega, double tol, int64_t max_sweeps, int64_t check,
               int64_t *sweeps, double *relres) {
    int64_t nf = s->fluidCount;
    double rn, bn;
    orc_residual_l2(s, &rn, &bn);
    int64_t it = 0;
    double rel = bn > 0 ? rn / bn : 0.0;
    if (check < 1) check = 1;
    while (rel > tol && it < max_sweeps) {
        for (int64_t q = 0; q < check && it < max_sweeps; q++) {
            for (int64_t n = 1; n <= nf; n++) {
                int64_t i = LST(s->fluidList, nf, n, 1), j = LST(s->fluidList, nf, n, 2), k = LST(s->fluidList, nf, n, 3);
                F(s->pco, i, j, k) = F(s->pc, i, j, k);
            }
            sor_colour(s, s->redList, s->redCount, s->pco, omega);
            sor_colour(s, s->blackList, s->blackCount, s->pc, omega);
            it++;
        }
        orc_residual_l2(s, &rn, &bn);
        rel = bn > 0 ? rn / bn : 0.0;
    }
    *sweeps = it;
    *relres = rel;
    return 0;
}

/* ---- accessors ---------------------------------------------------------- */
int orc_counts(orc_state *s, int64_t out[7]) {
    out[0] = s->ibCount; out[1] = s->fluidCount; out[2] = s->solidCount;
    out[3] = s->redCount; out[4] = s->blackCount; out[5] = s->ibElems; out[6] = s->ibNodes;
    return 0;
}

int orc_array(orc_state *s, const char *name, void **ptr, int64_t *count, int *type) {
    int64_t P = s->sx * s->sy * s->sz, nx = s->nx, ny = s->ny, nz = s->nz;
    int64_t nb = s->ibCount, ne = s->ibElems;
    struct { const char *n; void *p; int64_t c; int t; } tab[] = {
        {"u", s->u, P, 8}, {"v", s->v, P, 8}, {"w", s->w, P, 8}, {"ut", s->ut, P, 8},
        {"vt", s->vt, P, 8}, {"wt", s->wt, P, 8}, {"p", s->p, P, 8}, {"pc", s->pc, P, 8},
        {"pco", s->pco, P, 8}, {"b", s->b, P, 8}, {"resi_u", s->resi_u, P, 8},
        {"resi_v", s->resi_v, P, 8}, {"resi_w", s->resi_w, P, 8}, {"cell", s->cell, P, 64},
        {"minElemcell", s->minElemcell, P, 64},
        {"x1", s->x1 + 1, nx + 3, 8}, {"y1", s->y1 + 1, ny + 3, 8}, {"z1", s->z1 + 1, nz + 3, 8},
        {"deltax", s->deltax + 1, nx + 2, 8}, {"deltay", s->deltay + 1, ny + 2, 8},
        {"deltaz", s->deltaz + 1, nz + 2, 8},
        {"xu", s->xu + 1, nx + 3, 8}, {"yu", s->yu + 1, ny + 2, 8}, {"zu", s->zu + 1, nz + 2, 8},
        {"xv", s->xv + 1, nx + 2, 8}, {"yv", s->yv + 1, ny + 3, 8}, {"zv", s->zv + 1, nz + 2, 8},
        {"xw", s->xw + 1, nx + 2, 8}, {"yw", s->yw + 1, ny + 2, 8}, {"zw", s->zw + 1, nz + 3, 8},
        {"xp", s->xp + 1, nx + 2, 8}, {"yp", s->yp + 1, ny + 2, 8}, {"zp", s->zp + 1, nz + 2, 8},
        {"Acx", s->Acx, nx * 3, 8}, {"Acy", s->Acy, ny * 3, 8}, {"Acz", s->Acz, nz * 3, 8},
        {"tab_uu", s->tab[UU], 12 * s->tn[UU], 8}, {"tab_vv", s->tab[VV], 12 * s->tn[VV], 8},
        {"tab_ww", s->tab[WW], 12 * s->tn[WW], 8}, {"tab_uv", s->tab[UV], 12 * s->tn[UV], 8},
        {"tab_vu", s->tab[VU], 12 * s->tn[VU], 8}, {"tab_wu", s->tab[WU], 12 * s->tn[WU], 8},
        {"xnode1", s->xnode1 ? s->xnode1 + 1 : NULL, s->ibNodes, 8},
        {"ynode1", s->ynode1 ? s->ynode1 + 1 : NULL, s->ibNodes, 8},
        {"znode1", s->znode1 ? s->znode1 + 1 : NULL, s->ibNodes, 8},
        {"xcent", s->xcent ? s->xcent + 1 : NULL, ne, 8}, {"ycent", s->ycent ? s->ycent + 1 : NULL, ne, 8},
        {"zcent", s->zcent ? s->zcent + 1 : NULL, ne, 8},
        {"cosAlpha", s->cosAlpha ? s->cosAlpha + 1 : NULL, ne, 8},
        {"cosBeta", s->cosBeta ? s->cosBeta + 1 : NULL, ne, 8},
        {"cosGamma", s->cosGamma ? s->cosGamma + 1 : NULL, ne, 8},
        {"alpha3", s->alpha3 ? s->alpha3 + 1 : NULL, ne, 8}, {"beta3", s->beta3 ? s->beta3 + 1 : NULL, ne, 8},
        {"gamma3", s->gamma3 ? s->gamma3 + 1 : NULL, ne, 8},
        {"interceptedIndexPtr", s->ibList, nb * 3, 64}, {"fluidIndexPtr", s->fluidList, s->fluidCount * 3, 64},
        {"solidIndexPtr", s->solidList, s->solidCount * 3, 64},
        {"redCellIndexPtr", s->redList, s->redCount * 3, 64},
        {"blackCellIndexPtr", s->blackList, s->blackCount * 3, 64},
        {"nelp", s->nelp ? s->nelp + 1 : NULL, nb, 64},
        {"pNormDis", s->pNormDis ? s->pNormDis + 1 : NULL, nb, 8},
        {"u1NormDis", s->u1NormDis ? s->u1NormDis + 1 : NULL, nb, 8},
        {"u2NormDis", s->u2NormDis ? s->u2NormDis + 1 : NULL, nb, 8},
        {"v1NormDis", s->v1NormDis ? s->v1NormDis + 1 : NULL, nb, 8},
        {"v2NormDis", s->v2NormDis ? s->v2NormDis + 1 : NULL, nb, 8},
        {"w1NormDis", s->w1NormDis ? s->w1NormDis + 1 : NULL, nb, 8},
        {"w2NormDis", s->w2NormDis ? s->w2NormDis + 1 : NULL, nb, 8},
        {"Elemcell", s->Elemcell, ne * 3, 64}, {"ucell", s->ucell, ne * 3, 64},
        {"vcell", s->vcell, ne * 3, 64}, {"wcell", s->wcell, ne * 3, 64}, {"pcell", s->pcell, ne * 3, 64},
    };
    for (size_t q = 0; q < sizeof tab / sizeof tab[0]; q++) {
        if (strcmp(tab[q].n, name) == 0) {
            *ptr = tab[q].p;
            *count = tab[q].p ? tab[q].c : 0;
            *type = tab[q].t;
            return 0;
        }
    }
    return fail(s, "unknown array name");
}
