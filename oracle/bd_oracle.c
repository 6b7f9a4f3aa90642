/* bd_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the band rebuild of the free-boundary step, as the
 * GPU runs it: the reference's eikonal update of the fast march
 * (levelset/heap.cc:183-442: calc, fourway, threeway1/2, corner,
 * corner_bail, look_l/r/d/u) applied in rounds -- each round accepts every
 * trial cell within delta of the smallest tentative value and below the band
 * limit, then recomputes the trial cells next to the newly accepted ones
 * (axis neighbours at distance 1 and 2, heap.cc:466-470) -- positive side
 * first, then the negative side; afterwards the statuses follow the sign
 * pattern and cells outside the band are clamped to +-limit
 * (levelset.cc:416-421, 740-771).
 *
 * It pins the GPU march bit for bit (same rounds, same arithmetic; built
 * with -ffp-contract=off) and is itself compared with the reference's heap march,
 * which it follows statistically.  Only tests/ call it.
 */
#include "stz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int m, n, pos;
    double dx, dy, dx2, dy2, dxbydy, dx2bydy2;
    const double *phi;
    const int *st;
} eik;

static int st_code(int s) { return s & 15; }
static int in_(const eik *E, int ij) { return E->pos ? st_code(E->st[ij]) < 6 : st_code(E->st[ij]) > 1; }
static int cm(const eik *E, double a, double b) { return E->pos ? a < b : a > b; }
static double ct(const eik *E, double a, double b) { return E->pos ? a + b : a - b; }

/* heap.cc:396-442 */
static int look_l(const eik *E, int i, int ij, double *e) {
    const double *p = E->phi;
    if (i > 1 && in_(E, ij - 2) && cm(E, p[ij - 2], p[ij - 1])) { *e = 2 * p[ij - 1] - 0.5 * p[ij - 2]; return 1; }
    *e = p[ij - 1];
    return 0;
}
static int look_r(const eik *E, int i, int ij, double *e) {
    const double *p = E->phi;
    if (i < E->m - 2 && in_(E, ij + 2) && cm(E, p[ij + 2], p[ij + 1])) { *e = 2 * p[ij + 1] - 0.5 * p[ij + 2]; return 1; }
    *e = p[ij + 1];
    return 0;
}
static int look_d(const eik *E, int ij, double *e) {
    const double *p = E->phi;
    const int m = E->m;
    if (ij >= 2 * m && in_(E, ij - 2 * m) && cm(E, p[ij - 2 * m], p[ij - m])) {
        *e = 2 * p[ij - m] - 0.5 * p[ij - 2 * m];
        return 1;
    }
    *e = p[ij - m];
    return 0;
}
static int look_u(const eik *E, int ij, double *e) {
    const double *p = E->phi;
    const int m = E->m, mn = E->m * E->n;
    if (ij < mn - 2 * m && in_(E, ij + 2 * m) && cm(E, p[ij + 2 * m], p[ij + m])) {
        *e = 2 * p[ij + m] - 0.5 * p[ij + 2 * m];
        return 1;
    }
    *e = p[ij + m];
    return 0;
}

/* heap.cc:385-394 */
static double corner_bail(const eik *E, int h, int v, double a, double b) {
    const double ti = 2.0 / 3.0;
    double ht = h ? ct(E, a, E->dx) * ti : ct(E, a, E->dx);
    double vt = v ? ct(E, b, E->dy) * ti : ct(E, b, E->dy);
    return cm(E, ht, vt) ? ht : vt;
}

/* heap.cc:358-383 */
static double corner(const eik *E, int h, int v, double a, double b) {
    const double dx2 = E->dx2, dy2 = E->dy2, r = E->dx2bydy2;
    const double f2a = 2.25 * (dy2 + dx2), f2b = 2.25 * dy2 + dx2, f2c = dy2 + 2.25 * dx2, f2d = dy2 + dx2;
    const double b1 = 2.25 * dx2 > dy2 ? 1 + 2.25 * r : 1 + dy2 / (2.25 * dx2);
    const double b2 = dx2 > dy2 ? 1 + r : 1 + dy2 / dx2;
    const double b3 = dx2 > 2.25 * dy2 ? 1 + r / 2.25 : 1 + 2.25 * dy2 / dx2;
    const double i2a = dy2 / f2a, i2b = dy2 / f2b, i2c = dy2 / f2c, i2d = dy2 / f2d;
    double f;
    if (h) {
        if (v) {
            f = 1.5 * (a - b);
            f *= f;
            return f * b2 < f2a ? ct(E, a * 1.5 + b * 1.5 * r, E->dxbydy * sqrt(f2a - f)) * i2a
                                : corner_bail(E, 1, 1, a, b);
        }
        f = a - 1.5 * b;
        f *= f;
        return f * b3 < f2b ? ct(E, a * 1.5 + b * r, E->dxbydy * sqrt(f2b - f)) * i2b : corner_bail(E, 1, 0, a, b);
    }
    if (v) {
        f = 1.5 * a - b;
        f *= f;
        return f * b1 < f2c ? ct(E, a + b * 1.5 * r, E->dxbydy * sqrt(f2c - f)) * i2c : corner_bail(E, 0, 1, a, b);
    }
    f = a - b;
    f *= f;
    return f * b2 < f2d ? ct(E, a + b * r, E->dxbydy * sqrt(f2d - f)) * i2d : corner_bail(E, 0, 0, a, b);
}

static void clamp0(const eik *E, double *g) {
    if (cm(E, *g, 0)) *g = 0;
}

/* heap.cc:270-296 */
static double fourway(const eik *E, double a, double b, double c, double d) {
    double e = 2 * (c - a), f = 2 * (d - b), g, h;
    const double r = E->dx2bydy2, dx2 = E->dx2;
    if (cm(E, e, f)) {
        g = e - a;
        h = e - b;
        clamp0(E, &g);
        clamp0(E, &h);
        if (g * g + h * h * r < dx2) {
            g = 1.5 * f - c;
            h = f - b;
            clamp0(E, &g);
            clamp0(E, &h);
            if (g * g + h * h * r < dx2) return corner(E, 1, 1, c, d);
            return corner(E, 1, 0, c, b);
        }
        return corner(E, 0, 0, a, b);
    }
    g = f - a;
    h = f - b;
    clamp0(E, &g);
    clamp0(E, &h);
    if (g * g + h * h * r < dx2) {
        g = e - a;
        h = 1.5 * e - d;
        clamp0(E, &g);
        clamp0(E, &h);
        if (g * g + h * h * r < dx2) return corner(E, 1, 1, c, d);
        return corner(E, 0, 1, a, d);
    }
    return corner(E, 0, 0, a, b);
}

/* heap.cc:298-326 */
static double threeway1(const eik *E, int h1, int h2, int v, double a, double c, double b) {
    double d;
    const double r = E->dx2bydy2, dx2 = E->dx2;
    if (h1) {
        if (h2) {
            d = cm(E, a, c) ? a : c;
            return corner(E, 1, v, d, b);
        } else {
            d = 2 * (a - c);
            double e = d - c, f = (v ? 1.5 * d : d) - b;
            clamp0(E, &e);
            clamp0(E, &f);
            return e * e + f * f * r < dx2 ? corner(E, 1, v, a, b) : corner(E, 0, v, c, b);
        }
    }
    if (h2) {
        d = 2 * (c - a);
        double e = d - a, f = (v ? 1.5 * d : d) - b;
        clamp0(E, &e);
        clamp0(E, &f);
        return e * e + f * f * r < dx2 ? corner(E, 1, v, c, b) : corner(E, 0, v, a, b);
    }
    d = cm(E, a, c) ? a : c;
    return corner(E, 0, v, d, b);
}

/* heap.cc:328-356 */
static double threeway2(const eik *E, int h, int v1, int v2, double b, double a, double c) {
    double d;
    const double r = E->dx2bydy2, dx2 = E->dx2;
    if (v1) {
        if (v2) {
            d = cm(E, a, c) ? a : c;
            return corner(E, h, 1, b, d);
        } else {
            d = 2 * (a - c);
            double e = d - c, f = (h ? 1.5 * d : d) - b;
            clamp0(E, &e);
            clamp0(E, &f);
            return f * f + e * e * r < dx2 ? corner(E, h, 1, b, a) : corner(E, h, 0, b, c);
        }
    }
    if (v2) {
        d = 2 * (c - a);
        double e = d - a, f = (h ? 1.5 * d : d) - b;
        clamp0(E, &e);
        clamp0(E, &f);
        return f * f + e * e * r < dx2 ? corner(E, h, 1, b, c) : corner(E, h, 0, b, a);
    }
    d = cm(E, a, c) ? a : c;
    return corner(E, h, 0, b, d);
}

/* heap.cc:183-268 */
static double calc(const eik *E, int i, int ij) {
    const int m = E->m, mn = E->m * E->n;
    const double ti = 2.0 / 3.0, dx = E->dx, dy = E->dy;
    double a, b, c, d;
    int h1, h2, v;
    const int D = ij >= m && in_(E, ij - m), U = ij < mn - m && in_(E, ij + m);
    const int R = i < m - 1 && in_(E, ij + 1), L = i > 0 && in_(E, ij - 1);
    if (D) {
        if (U) {
            if (R) {
                if (L) {
                    if (look_d(E, ij, &a)) {
                        if (look_u(E, ij, &c)) {
                            d = cm(E, a, c) ? a : c;
                            h1 = look_l(E, i, ij, &a);
                            h2 = look_r(E, i, ij, &c);
                            return threeway1(E, h1, h2, 1, a, c, d);
                        }
                        if (look_l(E, i, ij, &b)) {
                            if (look_r(E, i, ij, &d)) {
                                b = cm(E, b, d) ? b : d;
                                return threeway2(E, 1, 1, 0, b, a, c);
                            }
                            return fourway(E, d, c, b, a);
                        }
                        if (look_r(E, i, ij, &d)) return fourway(E, b, c, d, a);
                        b = cm(E, b, d) ? b : d;
                        return threeway2(E, 0, 1, 0, b, a, c);
                    }
                    if (look_u(E, ij, &c)) {
                        if (look_l(E, i, ij, &b)) {
                            if (look_r(E, i, ij, &d)) {
                                b = cm(E, b, d) ? b : d;
                                return threeway2(E, 1, 1, 0, b, c, a);
                            }
                            return fourway(E, d, a, b, c);
                        }
                        if (look_r(E, i, ij, &d)) return fourway(E, b, a, d, c);
                        b = cm(E, b, d) ? b : d;
                        return threeway2(E, 0, 1, 0, b, c, a);
                    }
                    d = cm(E, a, c) ? a : c;
                    h1 = look_l(E, i, ij, &a);
                    h2 = look_r(E, i, ij, &c);
                    return threeway1(E, h1, h2, 0, a, c, d);
                }
                h1 = look_r(E, i, ij, &b);
                h2 = look_u(E, ij, &a);
                v = look_d(E, ij, &c);
                return threeway2(E, h1, h2, v, b, a, c);
            }
            if (L) {
                h1 = look_l(E, i, ij, &b);
                h2 = look_u(E, ij, &a);
                v = look_d(E, ij, &c);
                return threeway2(E, h1, h2, v, b, a, c);
            }
            c = look_d(E, ij, &a) ? ct(E, a, dy) * ti : ct(E, a, dy);
            d = look_u(E, ij, &b) ? ct(E, b, dy) * ti : ct(E, b, dy);
            return cm(E, c, d) ? c : d;
        }
        if (R) {
            if (L) {
                h1 = look_l(E, i, ij, &a);
                h2 = look_r(E, i, ij, &c);
                v = look_d(E, ij, &b);
                return threeway1(E, h1, h2, v, a, c, b);
            }
            h1 = look_r(E, i, ij, &a);
            v = look_d(E, ij, &b);
            return corner(E, h1, v, a, b);
        }
        if (L) {
            h1 = look_l(E, i, ij, &a);
            v = look_d(E, ij, &b);
            return corner(E, h1, v, a, b);
        }
        return look_d(E, ij, &a) ? ct(E, a, dy) * ti : ct(E, a, dy);
    }
    if (U) {
        if (R) {
            if (L) {
                h1 = look_l(E, i, ij, &a);
                h2 = look_r(E, i, ij, &c);
                v = look_u(E, ij, &b);
                return threeway1(E, h1, h2, v, a, c, b);
            }
            h1 = look_r(E, i, ij, &a);
            v = look_u(E, ij, &b);
            return corner(E, h1, v, a, b);
        }
        if (L) {
            h1 = look_l(E, i, ij, &a);
            v = look_u(E, ij, &b);
            return corner(E, h1, v, a, b);
        }
        return look_u(E, ij, &a) ? ct(E, a, dy) * ti : ct(E, a, dy);
    }
    if (R) {
        if (L) {
            c = look_l(E, i, ij, &a) ? ct(E, a, dx) * ti : ct(E, a, dx);
            d = look_r(E, i, ij, &b) ? ct(E, b, dx) * ti : ct(E, b, dx);
            return cm(E, c, d) ? c : d;
        }
        return look_r(E, i, ij, &a) ? ct(E, a, dx) * ti : ct(E, a, dx);
    }
    if (L) return look_l(E, i, ij, &a) ? ct(E, a, dx) * ti : ct(E, a, dx);
    return 0;
}

enum { NEW = 16 };

/* one side of the march in rounds; st codes as levelset.hh s[] */
static int march_side(int m, int n, double dx, double dy, double delta, double lim, double *phi, int *st, int pos) {
    const int mn = m * n;
    const int iface = pos ? 4 : 3, acc = pos ? 5 : 2, trial = pos ? 6 : 1, far = pos ? 7 : 0;
    eik E;
    E.m = m; E.n = n; E.pos = pos; E.dx = dx; E.dy = dy;
    E.dx2 = dx * dx; E.dy2 = dy * dy; E.dxbydy = dx / dy; E.dx2bydy2 = E.dxbydy * E.dxbydy;
    E.phi = phi; E.st = st;
    int *nst = (int *)malloc(sizeof(int) * mn);
    double *nphi = (double *)malloc(sizeof(double) * mn);
    int rounds = 0;
    /* seeds (read the state before any write, as a parallel launch does) */
    memcpy(nst, st, sizeof(int) * mn);
    memcpy(nphi, phi, sizeof(double) * mn);
    for (int ij = 0; ij < mn; ij++) {
        const int i = ij % m;
        if (st_code(st[ij]) != far) continue;
        if ((i > 0 && st_code(st[ij - 1]) == iface) || (i < m - 1 && st_code(st[ij + 1]) == iface) ||
            (ij >= m && st_code(st[ij - m]) == iface) || (ij < mn - m && st_code(st[ij + m]) == iface)) {
            nphi[ij] = calc(&E, i, ij);
            nst[ij] = trial;
        }
    }
    memcpy(st, nst, sizeof(int) * mn);
    memcpy(phi, nphi, sizeof(double) * mn);
    for (;;) {
        /* smallest tentative key */
        int have = 0;
        double kmin = 0;
        for (int ij = 0; ij < mn; ij++)
            if (st_code(st[ij]) == trial) {
                const double k = pos ? phi[ij] : -phi[ij];
                if (!have || k < kmin) { kmin = k; have = 1; }
            }
        /* accept (and clear last round's marks) */
        int accepted = 0;
        for (int ij = 0; ij < mn; ij++) {
            if (st[ij] & NEW) { st[ij] &= 15; continue; }
            if (!have || st_code(st[ij]) != trial) continue;
            const double k = pos ? phi[ij] : -phi[ij];
            if (k < kmin + delta && k < lim) { st[ij] = acc | NEW; accepted++; }
        }
        rounds++;
        if (!accepted) break;
        /* update the trial and far cells next to the new acceptances */
        memcpy(nst, st, sizeof(int) * mn);
        memcpy(nphi, phi, sizeof(double) * mn);
        for (int ij = 0; ij < mn; ij++) {
            const int s = st_code(st[ij]), i = ij % m;
            if (s != trial && s != far) continue;
#define ISNEW(k) ((st[(k)] & NEW) != 0)
            int go = (i > 0 && ISNEW(ij - 1)) || (i < m - 1 && ISNEW(ij + 1)) || (ij >= m && ISNEW(ij - m)) ||
                     (ij < mn - m && ISNEW(ij + m));
            if (s == trial && !go)
                go = (i > 1 && ISNEW(ij - 2)) || (i < m - 2 && ISNEW(ij + 2)) || (ij >= 2 * m && ISNEW(ij - 2 * m)) ||
                     (ij < mn - 2 * m && ISNEW(ij + 2 * m));
#undef ISNEW
            if (!go) continue;
            nphi[ij] = calc(&E, i, ij);
            if (s == far) nst[ij] = trial;
        }
        memcpy(st, nst, sizeof(int) * mn);
        memcpy(phi, nphi, sizeof(double) * mn);
    }
    for (int ij = 0; ij < mn; ij++) st[ij] &= 15;
    free(nst);
    free(nphi);
    return rounds;
}

int orc_bd_remarch(int m, int n, double dx, double dy, double delta, double *phi, int *st) {
    const int mn = m * n;
    const double lim = 5 * dx;
    /* the interface cells stay, the rest of the band is forgotten
     * (levelset.hh:169-200) */
    for (int ij = 0; ij < mn; ij++) {
        const int s = st_code(st[ij]);
        if (s != 3 && s != 4) st[ij] = s < 4 ? 0 : 7;
    }
    int rounds = march_side(m, n, dx, dy, delta, lim, phi, st, 1);
    rounds += march_side(m, n, dx, dy, delta, lim, phi, st, 0);
    /* statuses from the sign pattern, clean (levelset.cc:416-421, 740-771) */
    int *out = (int *)malloc(sizeof(int) * mn);
    for (int ij = 0; ij < mn; ij++) {
        const int s = st_code(st[ij]), i = ij % m;
        const int band = s >= 2 && s <= 5;
        if (!band) {
            phi[ij] = s <= 1 ? -lim : lim;
            out[ij] = s;
            continue;
        }
        const int p = phi[ij] >= 0;
        int opp = 0;
        const int nb[4] = {i > 0 ? ij - 1 : -1, i < m - 1 ? ij + 1 : -1, ij >= m ? ij - m : -1,
                           ij < mn - m ? ij + m : -1};
        for (int q = 0; q < 4; q++) {
            if (nb[q] < 0) continue;
            const int t = st_code(st[nb[q]]);
            const int side = (t >= 2 && t <= 5) ? phi[nb[q]] >= 0 : t >= 4;
            if (side != p) opp = 1;
        }
        out[ij] = p ? (opp ? 4 : 5) : (opp ? 3 : 2);
    }
    memcpy(st, out, sizeof(int) * mn);
    free(out);
    return rounds;
}
