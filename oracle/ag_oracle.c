/*
 * ag_oracle.c -- FP64 CPU restatement of autogrid4's grid-map path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the MI355X engine.  Only tests,
 * the smoke check and the benchmark's CPU-baseline leg may load it.
 *
 * Parity pins (see tests/test_oracle.py):
 *   - the compiled reference binary (oracle/_ref/autogrid4) on the
 *     reference's own GPFs and on seeded receptors: every .map value line
 *     must be byte-identical;
 *   - PyAutoDock maps held by the reference's tests (Tests/py_hsg1_sm.*.map,
 *     copied as fixtures) at the reference's tolerance with the sign fixed;
 *   - closed forms: eps(r) against the Mehler-Solmajer formula, the LJ
 *     coefficient known answer of Tests/hsg1_sm_nbpcoeffs.gpf, a one-atom
 *     receptor's radial e map, min-window smoothing against brute force;
 *   - orc_vina_tables: whole rows against the Vina scoring function's
 *     closed forms in FP64 (tests/test_vina.py);
 *   - orc_analytic_maps: one-atom E(r) against the brute-force smoothing
 *     definition and the reference's table, and whole map sets against the
 *     table path within per-point Lipschitz bounds
 *     (tests/test_analytic_host.py).
 */
#include "ag_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define APPROX_ZERO 1.0e-6  /* autocomm.h:72 */
#define EINTCLAMP 100000.0  /* autocomm.h:89 */
#define NEINT 1024          /* constants.h:160 */
#define NDIEL 8192          /* constants.h:171 */
#define A_DIV 128           /* constants.h:151 */
#define SOFTNBC 8.0         /* constants.h:150 */
#define NBC 8.0             /* constants.h:148 (USE_8A_NBCUTOFF); the tree's Loop-2 radius, main.cpp:352 */
#define MAXNB 4             /* AG_MAX_NBONDS, autogrid.h:59 */

static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }

/* main.cpp:2040-2266 */
int orc_vdw_tables(const ag_setup *s, const ag_mapset *ms, float *out) {
    const int T = s->n_rec_types, M = s->n_atom_maps;
    const int i_smooth = (int)(s->r_smooth * A_DIV / 2.);
    float e[NEINT];
    double sm[NEINT];
    memset(out, 0, sizeof(float) * (size_t)T * M * NEINT);
    for (int ia = 0; ia < M; ia++) {
        if (ms->maps[ia].is_covalent) continue;
        for (int i = 0; i < T; i++) {
            const ag_pair *p = &s->pairs[(size_t)ia * T + i];
            double cA = p->cA, cB = p->cB;
            const int xA = p->xA, xB = p->xB;
            if (cA == 0 && cB == 0) {
                double tmpconst;
                cA = (tmpconst = p->eps / (double)(xA - xB)) * pow(p->Rij, (double)xA) * (double)xB;
                cB = tmpconst * pow(p->Rij, (double)xB) * (double)xA;
            }
            if (isnan(cA) || isnan(cB) || xA == 0 || xB == 0) return -1;
            for (int k = 1; k < NEINT; k++) {
                double r = ((double)k) * (1. / A_DIV);
                double rA = pow(r, (double)xA);
                double rB = pow(r, (double)xB);
                double v = cA / rA - cB / rB;
                e[k] = (float)((EINTCLAMP < v) ? EINTCLAMP : v);
            }
            e[0] = (float)EINTCLAMP;
            e[NEINT - 1] = 0.f;
            if (i_smooth > 0) {
                for (int k = 0; k < NEINT; k++) {
                    sm[k] = 100000.;
                    int j0 = k - i_smooth > 0 ? k - i_smooth : 0;
                    int j1 = k + i_smooth + 1 < NEINT ? k + i_smooth + 1 : NEINT;
                    for (int j = j0; j < j1; j++) sm[k] = (sm[k] < e[j]) ? sm[k] : e[j];
                }
                for (int k = 0; k < NEINT; k++) e[k] = (float)sm[k];
            }
            memcpy(out + ((size_t)i * M + ia) * NEINT, e, sizeof e);
        }
    }
    return 0;
}

/* main.cpp:2272-2276 */
void orc_sol_fn(double coeff_desolv, double *out) {
    double sigma = 3.6;
    out[0] = 0.0;
    for (int k = 1; k < NDIEL; k++) {
        double r = ((double)k) * (1. / A_DIV);
        out[k] = coeff_desolv * exp(-(r * r) / (2. * (sigma * sigma)));
    }
}

/* distdepdiel.cpp:30-57; the denominator is long double as in the reference */
double orc_ddd(double distance) {
    double lambda = 0.003627L, epsilon0 = 78.4L, A = -8.5525L, rk = 7.7839L;
    double B = epsilon0 - A;
    double lambda_B = -lambda * B;
    double epsilon = A + B / (1.0L + rk * exp(lambda_B * distance));
    if (epsilon < APPROX_ZERO) epsilon = 1.0L;
    return epsilon;
}

/* main.cpp:1612-1628 */
void orc_r_eps(double *out) {
    double eps[NDIEL];
    eps[0] = 1.0;
    for (int k = 1; k < NDIEL; k++) eps[k] = orc_ddd(((double)k) * (1. / A_DIV));
    for (int k = 0; k < NDIEL; k++) out[k] = 332.0 / eps[k];
}

/* --- bond vectors ------------------------------------------------------ */

static double len3(const double *v) { return hypot(v[0], hypot(v[1], v[2])); } /* vect3len */

static double vnorm(double *v) { /* vect_normalize, main.cpp:3606 */
    double rd = len3(v), inv;
    if (rd < APPROX_ZERO) inv = 1. / APPROX_ZERO;
    else inv = 1. / rd;
    for (int i = 0; i < 3; i++) v[i] *= inv;
    return rd;
}

static double vsubnorm(double *res, const double *a, const double *b) { /* main.cpp:3620 */
    double rd2 = 0., rd, inv;
    for (int i = 0; i < 3; i++) {
        res[i] = a[i] - b[i];
        rd2 += res[i] * res[i];
    }
    if (rd2 < APPROX_ZERO) rd = APPROX_ZERO;
    else rd = sqrt(rd2);
    inv = 1. / rd;
    for (int i = 0; i < 3; i++) res[i] *= inv;
    return rd;
}

static int dist_le(const double *a, const double *b, double lim, double *dc, double *d) { /* main.cpp:121 */
    if ((dc[0] = a[0] - b[0]) > lim || dc[0] < -lim) return 0;
    if ((dc[1] = a[1] - b[1]) > lim || dc[1] < -lim) return 0;
    if ((dc[2] = a[2] - b[2]) > lim || dc[2] < -lim) return 0;
    *d = len3(dc);
    return *d <= lim;
}

int orc_bond_vectors(const ag_setup *s, int n, const double *coord, const int32_t *type,
                     const int32_t *hbond, int32_t *rexp, double *rvector, double *rvector2,
                     uint8_t *disorder) {
    const int hydrogen = s->t_hydrogen, nonHB_hydrogen = s->t_nonhb_hydrogen, carbon = s->t_carbon;
    const int arom_carbon = s->t_arom_carbon, oxygen = s->t_oxygen, nitrogen = s->t_nitrogen;
    const int nonHB_nitrogen = s->t_nonhb_nitrogen, sulphur = s->t_sulphur;
    int *nb = calloc((size_t)n + 1, sizeof(int));
    int *bd = calloc((size_t)n * MAXNB + 1, sizeof(int));
    int rc = 0;
    memset(rexp, 0, sizeof(int32_t) * (size_t)n);
    memset(rvector, 0, sizeof(double) * 3 * (size_t)n);
    memset(rvector2, 0, sizeof(double) * 3 * (size_t)n);
    memset(disorder, 0, (size_t)n);
#define C(a) (coord + 3 * (size_t)(a))
#define ADDBOND(i, j)                                                   \
    do {                                                                \
        if (nb[i] >= MAXNB || nb[j] >= MAXNB) { rc = -1; goto done; }   \
        bd[(size_t)(i) * MAXNB + nb[i]++] = (j);                        \
        bd[(size_t)(j) * MAXNB + nb[j]++] = (i);                        \
    } while (0)
    for (int ia = 0; ia < n; ia++) {
        int from = ia - 20 > 0 ? ia - 20 : 0;
        int to = ia + 20 < n - 1 ? ia + 20 : n - 1;
        double *rv = rvector + 3 * (size_t)ia, *rv2 = rvector2 + 3 * (size_t)ia;
        if (hbond[ia] == 2) { /* D1, main.cpp:2320-2414 */
            for (int ib = from; ib <= to; ib++) {
                double d[3], rd;
                if (ib == ia) continue;
                if (!dist_le(C(ia), C(ib), 1.378, d, &rd)) continue;
                for (int i = 0; i < 3; i++) rv[i] = d[i];
                vnorm(rv);
                if (type[ib] == nitrogen) rexp[ia] = 2;
                else if (type[ib] == oxygen || type[ib] == sulphur) rexp[ia] = 4;
                else rexp[ia] = 2;
                if (s->disorder_h) ADDBOND(ia, ib);
            }
        } else if (hbond[ia] == 5) { /* A2, main.cpp:2416-2594 */
            int nbond = 0, i1 = 0, i2 = 0;
            for (int ib = from; ib <= to; ib++) {
                double dc[3], rd;
                if (ib == ia) continue;
                if (!dist_le(C(ia), C(ib), (type[ib] == hydrogen || type[ib] == nonHB_hydrogen) ? 1.3 : 1.9, dc, &rd))
                    continue;
                if (nbond == 2) {
                } else if (nbond == 1) {
                    nbond = 2;
                    i2 = ib;
                } else if (nbond == 0) {
                    nbond = 1;
                    i1 = ib;
                }
            }
            if (nbond == 1) {
                vsubnorm(rv, C(ia), C(i1));
                for (i2 = from; i2 <= to; i2++) {
                    double rdcx, d[3];
                    if (i2 == i1 || i2 == ia) continue;
                    if (!dist_le(C(i1), C(i2), type[i2] == hydrogen ? 1.3 : 1.7, d, &rdcx)) continue;
                    vnorm(d);
                    rv2[0] = rv[1] * d[2] - rv[2] * d[1];
                    rv2[1] = rv[2] * d[0] - rv[0] * d[2];
                    rv2[2] = rv[0] * d[1] - rv[1] * d[0];
                    vnorm(rv2);
                }
            } else if (nbond == 2) {
                if (s->disorder_h && (type[i1] == hydrogen || type[i2] == hydrogen) && type[i1] != type[i2]) {
                    int ib;
                    if (type[i1] == carbon || type[i1] == arom_carbon) ib = i1;
                    else if (type[i2] == carbon || type[i2] == arom_carbon) ib = i2;
                    else ib = i1;
                    ADDBOND(ia, ib);
                    vsubnorm(rv, C(ia), C(ib));
                } else {
                    double rdot = 0.;
                    vsubnorm(rv2, C(i2), C(i1));
                    for (int i = 0; i < 3; i++) rdot += (C(ia)[i] - C(i1)[i]) * rv2[i];
                    for (int i = 0; i < 3; i++) rv[i] = C(ia)[i] - ((rdot * rv2[i]) + C(i1)[i]);
                    vnorm(rv);
                }
            }
        } else if (hbond[ia] == 4) { /* A1, main.cpp:2596-2657 */
            int nbond = 0;
            double bondv[3][3], bondl[3], mean[3] = {0., 0., 0.};
            for (int ib = from; ib <= to; ib++) {
                if (ib == ia) continue;
                if (!dist_le(C(ia), C(ib), (type[ib] == hydrogen || type[ib] == nonHB_hydrogen) ? 1.3 : 1.7,
                             bondv[nbond], &bondl[nbond]))
                    continue;
                if (++nbond == 3) break;
            }
            for (int b = 0; b < nbond; b++)
                for (int i = 0; i < 3; i++) mean[i] += (bondv[b][i]) / nbond;
            vnorm(mean);
            for (int i = 0; i < 3; i++) rv[i] = -1 * mean[i];
        }
    }
    if (s->disorder_h) { /* main.cpp:2684-2736 */
        for (int ia = 0; ia < n; ia++) {
            int hcount = 0;
            for (int b = 0; b < MAXNB && b < nb[ia]; b++)
                if (type[bd[(size_t)ia * MAXNB + b]] == hydrogen) hcount++;
            if ((type[ia] == nitrogen && hcount == 2) || (type[ia] == nonHB_nitrogen && hcount == 3) ||
                (type[ia] == oxygen && hcount == 1) || (type[ia] == sulphur && hcount == 1)) {
                disorder[ia] = 1;
                for (int b = 0; b < MAXNB && b < nb[ia]; b++) {
                    int ba = bd[(size_t)ia * MAXNB + b];
                    if (type[ba] == hydrogen) disorder[ba] = 1;
                }
            }
        }
    }
done:
#undef ADDBOND
#undef C
    free(nb);
    free(bd);
    return rc;
}

/* --- per-point loop, main.cpp:2805-3431 --------------------------------- */

/* Radius query of a -DUSE_BHTREE build (bhtree.cpp:320-379): the atoms
 * within `cutoff` of x by the tree's FP32 test -- each component of
 * x - atom within +-cutoff, then d0*d0 + d1*d1 + d2*d2 <= cutoff*cutoff,
 * every operation rounded to float -- with the FP32 distance sqrt(D).  The
 * tree prunes by the same bounds, so this visits the atoms it returns (in
 * index order; the tree's order only permutes an FP64 sum).  Stores up to
 * `cap` (index, distance) pairs, returns the count. */
static int bh_query(const ag_receptor *rec, const float x[3], float cutoff, int *idx, float *dist, int cap) {
    const float C = cutoff * cutoff;
    int n = 0;
    for (int ia = 0; ia < rec->n_atoms; ia++) {
        const double *a = rec->coord + 3 * (size_t)ia;
        float d[3];
        int out = 0;
        for (int i = 0; i < 3 && !out; i++) {
            d[i] = x[i] - (float)a[i];
            out = d[i] > cutoff || d[i] < -cutoff;
        }
        if (out) continue;
        const float D0 = d[0] * d[0];
        const float D1 = D0 + d[1] * d[1];
        const float D = D1 + d[2] * d[2];
        if (D > C) continue;
        if (n < cap) {
            if (idx) idx[n] = ia;
            if (dist) dist[n] = sqrtf(D);
        }
        n++;
    }
    return n;
}

int orc_maps(const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms, const ag_tables *tab,
             double *maps, double *fmap, int32_t *cmap, int z0, int z1) {
    const int N = rec->n_atoms, M = ms->n_atom_maps;
    const int n1x = grid->n1[0], n1y = grid->n1[1];
    const long long P = (long long)grid->n1[0] * grid->n1[1] * grid->n1[2];
    const double spacing = grid->spacing;
    const int elecPE = ms->elec_index, dsolvPE = ms->dsolv_index;
    const int hydrogen = ms->hydrogen_type;
    const double PI_halved = 3.14159265358979323846 / 2.;
    const double ln_half = log(0.5);
    int nh12 = 0;
    int *h12 = malloc(sizeof(int) * ((size_t)N + 1));
    for (int i = 0; i < N; i++)
        if (rec->hbond[i] == 1 || rec->hbond[i] == 2) h12[nh12++] = i;
    if (M > 0 && nh12 == 0) {
        free(h12);
        return -1;
    }
    const float *evdw = tab->e_vdw_hb;
#define EV(t, m, k) ((double)evdw[((size_t)(t) * M + (m)) * NEINT + (k)])
#pragma omp parallel for schedule(dynamic, 1)
    for (int iz = z0; iz < z1; iz++) {
        double c[3];
        int *bh_idx = ms->bhtree ? malloc(sizeof(int) * ((size_t)N + 1)) : NULL;
        float *bh_dist = ms->bhtree ? malloc(sizeof(float) * ((size_t)N + 1)) : NULL;
        double hbondmin[AG_MAX_MAPS], hbondmax[AG_MAX_MAPS];
        int hbondflag[AG_MAX_MAPS];
        c[2] = ((double)(iz - grid->ne[2])) * spacing;
        for (int iy = 0; iy < n1y; iy++) {
            c[1] = ((double)(iy - grid->ne[1])) * spacing;
            for (int ix = 0; ix < n1x; ix++) {
                const long long mapi = ix + (long long)n1x * (iy + (long long)n1y * iz);
                double d[3];
                c[0] = ((double)(ix - grid->ne[0])) * spacing;
                /* covalent maps (main.cpp:2866-2879) */
                for (int j = 0; j < M; j++)
                    if (ms->maps[j].is_covalent) {
                        double rcov;
                        for (int ii = 0; ii < 3; ii++) d[ii] = ms->cov_pos[ii] - c[ii];
                        rcov = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / ms->cov_half_width;
                        if (rcov < APPROX_ZERO) rcov = APPROX_ZERO;
                        maps[(size_t)j * P + mapi] = ms->cov_barrier * (1. - exp(ln_half * rcov * rcov));
                    } else {
                        maps[(size_t)j * P + mapi] = 0.;
                    }
                if (elecPE >= 0) maps[(size_t)elecPE * P + mapi] = 0.;
                if (dsolvPE >= 0) maps[(size_t)dsolvPE * P + mapi] = 0.;
                if (fmap) fmap[mapi] = 1.0e12;
                if (cmap) cmap[mapi] = 0;
                for (int m = 0; m < M; m++) {
                    hbondmin[m] = 999999.;
                    hbondmax[m] = -999999.;
                    hbondflag[m] = 0;
                }
                /* -DUSE_BHTREE build (main.cpp:2893-2928): the constriction count
                 * by the tree, and unless map_receptor_interior, a collision
                 * within 1.5 A sets e = 0 and fmap = 1.5 and skips Loop 1 */
                const int bh = ms->bhtree;
                int is_collision = 0;
                if (bh) {
                    const float fcc[3] = {(float)c[0], (float)c[1], (float)c[2]};
                    if (cmap) cmap[mapi] = bh_query(rec, fcc, (float)ms->constriction_cutoff, NULL, NULL, 0);
                    if (bh == 1 && (elecPE >= 0 || fmap) && bh_query(rec, fcc, 1.5f, NULL, NULL, 0) > 0) {
                        is_collision = 1;
                        if (elecPE >= 0) maps[(size_t)elecPE * P + mapi] = 0.;
                        if (fmap) fmap[mapi] = 1.5;
                    }
                }
                /* Loop 1 (main.cpp:2940-2982): all atoms, hypot distance */
                if (!is_collision && (elecPE >= 0 || fmap || (cmap && !bh))) {
                    double e = 0;
                    for (int ia = 0; ia < N; ia++) {
                        const double *a = rec->coord + 3 * (size_t)ia;
                        double r = hypot((a[0] - c[0]), hypot((a[1] - c[1]), (a[2] - c[2])));
                        if (r < APPROX_ZERO) r = APPROX_ZERO;
                        double inv_rmax = 1. / dmax(r, 0.5);
                        int indx_r = (int)(r * A_DIV);
                        if (indx_r > NDIEL - 1) indx_r = NDIEL - 1;
                        if (fmap) fmap[mapi] = dmin(r, fmap[mapi]);
                        if (cmap && !bh && r <= ms->constriction_cutoff) cmap[mapi]++;
                        if (elecPE >= 0) {
                            if (ms->dddiel) e += rec->charge[ia] * inv_rmax * tab->r_eps[indx_r] * ms->coeff_estat;
                            else e += rec->charge[ia] * inv_rmax * ms->inv_diel_cal * ms->coeff_estat;
                        }
                    }
                    if (elecPE >= 0) maps[(size_t)elecPE * P + mapi] = e;
                }
                /* closest donor H (main.cpp:2985-3015) */
                double rminH = 999999.;
                int closestH = -1;
                if (M > 0)
                    for (int inh = 0; inh < nh12; inh++) {
                        int ia = h12[inh];
                        int brk = 0;
                        double dd[3];
                        for (int i = 0; i < 3; i++) {
                            dd[i] = rec->coord[3 * (size_t)ia + i] - c[i];
                            if (fabs(dd[i]) > rminH) { brk = 1; break; }
                        }
                        if (brk) continue;
                        double r = sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
                        if (r < rminH) {
                            rminH = r;
                            closestH = ia;
                        }
                    }
                if (!(M > 0 || dsolvPE >= 0)) continue;
                /* Loop 2 (main.cpp:3017-3404); with the tree, over the atoms
                 * its 8 A query returns, at their FP32 distance (main.cpp:3019-3028) */
                int nbh = N;
                if (bh) {
                    const float fcc[3] = {(float)c[0], (float)c[1], (float)c[2]};
                    nbh = bh_query(rec, fcc, (float)NBC, bh_idx, bh_dist, N);
                }
                for (int ib = 0; ib < nbh; ib++) {
                    const int ia = bh ? bh_idx[ib] : ib;
                    const double *a = rec->coord + 3 * (size_t)ia;
                    const double *rv = rec->rvector + 3 * (size_t)ia;
                    const double *rv2 = rec->rvector2 + 3 * (size_t)ia;
                    const int hb = rec->hbond[ia], at = rec->type[ia];
                    const int dis = rec->disorder[ia] != 0;
                    double dnorm[3], racc, rdon, Hramp;
                    for (int i = 0; i < 3; i++) d[i] = a[i] - c[i];
                    double r = bh ? (double)bh_dist[ib] : sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    if (r > SOFTNBC) continue;
                    if (ms->disorder_h && dis && at == hydrogen) continue;
                    if (r < APPROX_ZERO) r = APPROX_ZERO;
                    double inv_r = 1. / r;
                    for (int i = 0; i < 3; i++) dnorm[i] = d[i] * inv_r;
                    int indx_r = (int)(r * A_DIV);
                    if (indx_r > NDIEL - 1) indx_r = NDIEL - 1;
                    int indx_n = (int)(r * A_DIV);
                    if (indx_n > NEINT - 1) indx_n = NEINT - 1;
                    racc = 1.;
                    rdon = 1.;
                    Hramp = 1.;
                    if (hb == 2) { /* D1 */
                        double cos_theta = 0.;
                        for (int i = 0; i < 3; i++) cos_theta -= dnorm[i] * rv[i];
                        if (cos_theta <= 0.) {
                            racc = 0.;
                        } else {
                            switch (rec->rexp[ia]) {
                                double tmp;
                                case 1:
                                default: racc = cos_theta; break;
                                case 2: racc = cos_theta * cos_theta; break;
                                case 4:
                                    tmp = cos_theta * cos_theta;
                                    racc = tmp * tmp;
                                    break;
                            }
                            if (ia == closestH || closestH < 0) {
                                /* closestH < 0 only without atom maps, where Hramp is unused */
                                Hramp = 1.;
                            } else {
                                const double *rh = rec->rvector + 3 * (size_t)closestH;
                                double ct = 0.;
                                for (int i = 0; i < 3; i++) ct += rh[i] * rv[i];
                                ct = dmin(ct, 1.0);
                                ct = dmax(ct, -1.0);
                                double theta = acos(ct);
                                Hramp = 0.5 - 0.5 * cos(theta * 120. / 90.);
                            }
                        }
                    } else if (hb == 4) { /* A1 */
                        double cos_theta = 0.;
                        for (int i = 0; i < 3; i++) cos_theta -= dnorm[i] * rv[i];
                        if (cos_theta <= 0.) rdon = 0.;
                        else rdon = cos_theta * cos_theta;
                    } else if (hb == 5 && !dis) { /* A2 */
                        double t0, ti, cos_theta = 0.;
                        for (int i = 0; i < 3; i++) cos_theta -= dnorm[i] * rv[i];
                        t0 = 0.;
                        for (int i = 0; i < 3; i++) t0 += dnorm[i] * rv2[i];
                        if (t0 > 1.) t0 = 1.;
                        else if (t0 < -1.) t0 = -1.;
                        t0 = PI_halved - acos(t0);
                        double cross[3];
                        cross[0] = dnorm[1] * rv2[2] - dnorm[2] * rv2[1];
                        cross[1] = dnorm[2] * rv2[0] - dnorm[0] * rv2[2];
                        cross[2] = dnorm[0] * rv2[1] - dnorm[1] * rv2[0];
                        double rd2 = cross[0] * cross[0] + cross[1] * cross[1] + cross[2] * cross[2];
                        if (rd2 < APPROX_ZERO) rd2 = APPROX_ZERO;
                        double inv_rd = 1. / sqrt(rd2);
                        ti = 0.;
                        for (int i = 0; i < 3; i++) ti += cross[i] * inv_rd * rv[i];
                        rdon = 0.;
                        if (cos_theta >= 0.) {
                            if (ti > 1.) ti = 1.;
                            else if (ti < -1.) ti = -1.;
                            ti = acos(ti) - PI_halved;
                            if (ti < 0.) ti = -ti;
                            rdon = (0.9 + 0.1 * sin(ti + ti)) * cos(t0);
                        } else if (cos_theta >= -0.34202) {
                            rdon = 562.25 * pow(0.116978 - cos_theta * cos_theta, 3.) * cos(t0);
                        }
                    } else if (dis && hb == 5) { /* disordered hydroxyl */
                        double cos_theta = 0.;
                        for (int i = 0; i < 3; i++) cos_theta -= dnorm[i] * rv[i];
                        if (cos_theta > 1.) cos_theta = 1.;
                        else if (cos_theta < -1.) cos_theta = -1.;
                        double theta = acos(cos_theta);
                        racc = 0.;
                        rdon = 0.;
                        if (theta <= 1.24791 + PI_halved) {
                            rdon = pow(cos(theta - 1.24791), 4.);
                            racc = rdon;
                        }
                    }
                    for (int m = 0; m < M; m++) {
                        const ag_map_desc *md = &ms->maps[m];
                        if (md->is_covalent) continue;
                        double *E = &maps[(size_t)m * P + mapi];
                        if (md->is_hbonder) {
                            double rsph = EV(at, m, indx_n) / 100.;
                            rsph = dmax(rsph, 0.);
                            rsph = dmin(rsph, 1.);
                            const int mh = md->hbond;
                            if ((mh == 3 || mh == 5 || mh == 6) && (hb == 1 || hb == 2 || hb == 6)) {
                                if (!dis) {
                                    *E += EV(at, m, indx_n) * Hramp * (racc + (1. - racc) * rsph);
                                } else {
                                    int indx_h = (int)((r - 1.10) * A_DIV);
                                    indx_h = indx_h > 0 ? indx_h : 0;
                                    indx_h = indx_h < NEINT - 1 ? indx_h : NEINT - 1;
                                    *E += EV(hydrogen, m, indx_h) * Hramp * (racc + (1. - racc) * rsph);
                                }
                            } else if ((mh == 4 || mh == 6) && (hb == 1 || hb == 2 || hb == 6)) {
                                hbondmin[m] = dmin(hbondmin[m], EV(at, m, indx_n) * (racc + (1. - racc) * rsph));
                                hbondmax[m] = dmax(hbondmax[m], EV(at, m, indx_n) * (racc + (1. - racc) * rsph));
                                hbondflag[m] = 1;
                            } else if ((mh == 1 || mh == 2 || mh == 6) && (hb > 2 || hb == 6)) {
                                if (dis) {
                                    int indx_h = (int)((r - 1.10) * A_DIV);
                                    indx_h = indx_h > 0 ? indx_h : 0;
                                    indx_h = indx_h < NEINT - 1 ? indx_h : NEINT - 1;
                                    *E += EV(hydrogen, m, indx_h) * Hramp * (rdon + (1. - rdon) * rsph);
                                } else {
                                    double t = EV(at, m, indx_n) * (rdon + (1. - rdon) * rsph);
                                    hbondmin[m] = dmin(hbondmin[m], t);
                                    hbondmax[m] = dmax(hbondmax[m], t);
                                    hbondflag[m] = 1;
                                }
                            } else {
                                *E += EV(at, m, indx_n);
                            }
                        } else {
                            *E += EV(at, m, indx_n);
                        }
                        int md2 = m;
                        if (ms->separate_desolvation) md2 = ++m;
                        maps[(size_t)md2 * P + mapi] +=
                            ms->maps[md2].solpar_probe * rec->vol[ia] * tab->sol_fn[indx_r] +
                            (rec->solpar[ia] + ms->solpar_q * fabs(rec->charge[ia])) * ms->maps[md2].vol_probe *
                                tab->sol_fn[indx_r];
                    }
                    if (dsolvPE >= 0)
                        maps[(size_t)dsolvPE * P + mapi] += ms->solpar_q * rec->vol[ia] * tab->sol_fn[indx_r];
                }
                /* min/max fix-up (main.cpp:3420-3431) */
                for (int m = 0; m < M; m++)
                    if (hbondflag[m]) maps[(size_t)m * P + mapi] += (hbondmin[m] + dmax(hbondmax[m], 0));
            }
        }
        free(bh_idx);
        free(bh_dist);
    }
#undef EV
    free(h12);
    return 0;
}

/* --- Vina potential (main.cpp:1975-2127, the use_vina_potential branch) --- */

/* xs_vdw radii of read_parameter_library.cc:322-358 (ADVINA): float values,
 * exact name match, 1.8 (double) otherwise. */
int orc_xs_radius(const char *type, double *radius) {
    static const char *names[] = {"C", "A", "N", "NA", "NS", "O", "OA", "OS", "S", "SA", "P", "F", "Mg"};
    static const float vals[] = {1.9f, 1.9f, 1.8f, 1.8f, 1.8f, 1.7f, 1.7f, 1.7f, 2.0f, 2.0f, 2.1f, 1.5f, 1.2f};
    for (int i = 0; i < 13; i++)
        if (strcmp(type, names[i]) == 0) {
            *radius = vals[i];
            return 1;
        }
    *radius = 1.8;
    return 0;
}

int orc_vina_tables(const ag_setup *s, const ag_mapset *ms, const double *xs_map, const double *xs_rec,
                    float *out) {
    const int T = s->n_rec_types, M = s->n_atom_maps;
    /* the reference's weights are float locals (main.cpp:1977-1981) */
    float wt_gauss1 = -0.035579, wt_gauss2 = -0.005156, wt_repulsion = 0.840245, wt_hydrogen = -0.587439,
          wt_hydrophobic = -0.035069;
    memset(out, 0, sizeof(float) * (size_t)T * M * NEINT);
    for (int ia = 0; ia < M; ia++) {
        if (ms->maps[ia].is_covalent) continue;
        for (int i = 0; i < T; i++) {
            float *e = out + ((size_t)i * M + ia) * NEINT;
            double Rij = xs_map[ia];        /* ligand radius, double */
            float map_Rij = (float)xs_rec[i]; /* receptor radius through a float */
            int hbonder = s->pairs[(size_t)ia * T + i].hbonder;
            /* gridmap[i].hbonder[ia] -- map i, type ia; 0 outside the arrays */
            int hbonder_x = (i < M && ia < T) ? s->pairs[(size_t)i * T + ia].hbonder : 0;
            for (int indx_r = 1; indx_r < NEINT; indx_r++) {
                double r = (double)indx_r / A_DIV;
                double rddist = r - (map_Rij + Rij);
                double delta_e = wt_gauss1 * exp(-pow(((rddist) / 0.5), 2)) +
                                 wt_gauss2 * exp(-pow(((rddist - 3.) / 2.), 2));
                e[indx_r] += delta_e;
                if (rddist < 0) {
                    delta_e = wt_repulsion * pow(rddist, 2);
                    e[indx_r] += delta_e;
                }
                if (hbonder > 0) {
                    if (rddist <= 0.7) {
                        delta_e = 1 * wt_hydrogen;
                        e[indx_r] += delta_e;
                    }
                    if ((-0.7 < rddist) && (rddist <= 0.)) {
                        delta_e = (rddist / 0.7) * wt_hydrogen;
                        e[indx_r] -= delta_e;
                    }
                }
                if ((hbonder == 0) && (hbonder_x == 0)) {
                    delta_e = 0.;
                    if (rddist < 0.5) {
                        delta_e = 1 * wt_hydrophobic;
                    } else if (rddist < 1.5) {
                        delta_e = (0.5 - rddist) * wt_hydrophobic;
                    }
                    e[indx_r] += delta_e;
                }
            }
        }
    }
    return 0;
}

/* Vina map set: covalent maps and Loop 1's r_min / count as in orc_maps; per
 * atom map the table value at min(floor(128 r), 1023) of every atom within
 * 8 A (sqrt distance, disordered H skipped under disorder_h); e and d maps
 * zero. */
int orc_maps_vina(const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms, const ag_tables *tab,
                  double *maps, double *fmap, int32_t *cmap, int z0, int z1) {
    if (ms->bhtree) return -3;   /* the tree build's Vina branch is dead code as well */
    const int N = rec->n_atoms, M = ms->n_atom_maps;
    const int n1x = grid->n1[0], n1y = grid->n1[1];
    const long long P = (long long)grid->n1[0] * grid->n1[1] * grid->n1[2];
    const double spacing = grid->spacing;
    const double ln_half = log(0.5);
    const float *evdw = tab->e_vdw_hb;
#pragma omp parallel for schedule(dynamic, 1)
    for (int iz = z0; iz < z1; iz++) {
        double c[3], d[3];
        c[2] = ((double)(iz - grid->ne[2])) * spacing;
        for (int iy = 0; iy < n1y; iy++) {
            c[1] = ((double)(iy - grid->ne[1])) * spacing;
            for (int ix = 0; ix < n1x; ix++) {
                const long long mapi = ix + (long long)n1x * (iy + (long long)n1y * iz);
                c[0] = ((double)(ix - grid->ne[0])) * spacing;
                for (int j = 0; j < ms->n_maps; j++) maps[(size_t)j * P + mapi] = 0.;
                for (int j = 0; j < M; j++)
                    if (ms->maps[j].is_covalent) {
                        double rcov;
                        for (int ii = 0; ii < 3; ii++) d[ii] = ms->cov_pos[ii] - c[ii];
                        rcov = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / ms->cov_half_width;
                        if (rcov < APPROX_ZERO) rcov = APPROX_ZERO;
                        maps[(size_t)j * P + mapi] = ms->cov_barrier * (1. - exp(ln_half * rcov * rcov));
                    }
                if (fmap || cmap) {
                    if (fmap) fmap[mapi] = 1.0e12;
                    if (cmap) cmap[mapi] = 0;
                    for (int ia = 0; ia < N; ia++) {
                        const double *a = rec->coord + 3 * (size_t)ia;
                        double r = hypot((a[0] - c[0]), hypot((a[1] - c[1]), (a[2] - c[2])));
                        if (r < APPROX_ZERO) r = APPROX_ZERO;
                        if (fmap) fmap[mapi] = dmin(r, fmap[mapi]);
                        if (cmap && r <= ms->constriction_cutoff) cmap[mapi]++;
                    }
                }
                for (int ia = 0; ia < N; ia++) {
                    const double *a = rec->coord + 3 * (size_t)ia;
                    const int at = rec->type[ia];
                    for (int i = 0; i < 3; i++) d[i] = a[i] - c[i];
                    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    if (r > SOFTNBC) continue;
                    if (ms->disorder_h && rec->disorder[ia] && at == ms->hydrogen_type) continue;
                    if (r < APPROX_ZERO) r = APPROX_ZERO;
                    int indx_n = (int)(r * A_DIV);
                    if (indx_n > NEINT - 1) indx_n = NEINT - 1;
                    for (int m = 0; m < M; m++)
                        if (!ms->maps[m].is_covalent)
                            maps[(size_t)m * P + mapi] += (double)evdw[((size_t)at * M + m) * NEINT + indx_n];
                }
            }
        }
    }
    return 0;
}

/* --- Analytic model (the engine's differentiable mode) -------------------
 *
 * The table path's functions evaluated at the exact distance instead of
 * through 1/128 A bins:
 *   E_mt(r)  = min(EINTCLAMP, Phi(s(r))), Phi the 12-6 / 12-10 pair
 *              potential of main.cpp:2135-2190 (cA, cB from Rij, eps, or as
 *              given by nbp_coeffs), and s(r) the continuous form of the
 *              min-window smoothing of main.cpp:2200-2216: for a potential
 *              with one minimum at r*, min over [r - w, r + w] of Phi is
 *              Phi(r + w) left of the well, Phi(r*) inside it and
 *              Phi(r - w) right of it; w = floor(r_smooth * 64) / 128;
 *   sol(r)   = coeff_desolv exp(-r^2 / (2 * 3.6^2))    (main.cpp:2272-2276);
 *   eps(r)   the Mehler-Solmajer dielectric at r      (distdepdiel.cpp:30-57);
 * and the hard 8 A cutoff replaced by the switch S(r) = 1 - u^3 (10 - 15 u
 * + 6 u^2), u = clamp((r - 7.5) / 0.5, 0, 1), multiplying every 8 A term.
 * The per-point assembly (closest H, racc / rdon / Hramp, the H-bond case
 * split with its min / max, desolvation, d map) is main.cpp:2985-3431 with
 * those functions in place of the tables.  The e map is
 * sum_i q_i coeff_estat (332 / eps(r)) / max(r, 0.5) over all atoms
 * (main.cpp:2940-2982).  disorder_h is not part of the analytic model.
 *
 * Parity: on a one-atom receptor E(r) equals the brute-force minimum over
 * the continuous window of the textbook potential, and lies within the
 * Lipschitz bound of the reference's table at floor(128 r); whole map sets
 * lie within per-point Lipschitz bounds of the table path (orc_maps). */

static double an_phi(const ag_pair *pp, double s) {
    const int a = pp->xA, b = pp->xB;
    if (pp->cA == 0 && pp->cB == 0) {
        const double rho = pp->Rij / s;
        return pp->eps / (double)(a - b) * ((double)b * pow(rho, a) - (double)a * pow(rho, b));
    }
    return pp->cA / pow(s, a) - pp->cB / pow(s, b);
}

static double an_rstar(const ag_pair *pp) {
    const int a = pp->xA, b = pp->xB;
    if (pp->cA == 0 && pp->cB == 0) return pp->Rij;
    return pow((double)a * pp->cA / ((double)b * pp->cB), 1.0 / (double)(a - b));
}

static double an_energy(const ag_pair *pp, double r, double w) {
    const double rs = an_rstar(pp);
    double s;
    if (r < rs - w) s = r + w;
    else if (r > rs + w) s = r - w;
    else s = rs;
    return dmin(EINTCLAMP, an_phi(pp, s));
}

static double an_switch(double r) {
    double u = (r - 7.5) / 0.5;
    if (u <= 0) return 1.0;
    if (u >= 1) return 0.0;
    return 1.0 - u * u * u * (10.0 - 15.0 * u + 6.0 * u * u);
}

static double an_reps(double r, int dddiel, double inv_diel_cal) {
    if (!dddiel) return inv_diel_cal;
    const double lambda = 0.003627, epsilon0 = 78.4, A = -8.5525, rk = 7.7839;
    const double B = epsilon0 - A;
    double eps = A + B / (1.0 + rk * exp(-lambda * B * r));
    if (eps < APPROX_ZERO) eps = 1.0;
    return 332.0 / eps;
}

int orc_analytic_maps(const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms, const ag_pair *pairs,
                      int T, double r_smooth, double coeff_desolv, int npts, const int64_t *points, double *out) {
    const int N = rec->n_atoms, M = ms->n_atom_maps;
    const int n1x = grid->n1[0], n1y = grid->n1[1];
    const double spacing = grid->spacing;
    const double w = floor(r_smooth * A_DIV / 2.) / A_DIV;
    const double ln_half = log(0.5);
    const double PI_halved = 3.14159265358979323846 / 2.;
    const int elecPE = ms->elec_index, dsolvPE = ms->dsolv_index;
    if (ms->disorder_h) return -2;
#pragma omp parallel for schedule(dynamic, 64)
    for (int k = 0; k < npts; k++) {
        const int64_t mapi = points[k];
        const int ix = (int)(mapi % n1x), iy = (int)((mapi / n1x) % n1y), iz = (int)(mapi / ((int64_t)n1x * n1y));
        double c[3] = {(ix - grid->ne[0]) * spacing, (iy - grid->ne[1]) * spacing, (iz - grid->ne[2]) * spacing};
        double en[AG_MAX_MAPS], hmin[AG_MAX_MAPS], hmax[AG_MAX_MAPS];
        int hflag[AG_MAX_MAPS];
        for (int m = 0; m < ms->n_maps; m++) en[m] = 0.0;
        for (int m = 0; m < M; m++) {
            hmin[m] = 999999.;
            hmax[m] = -999999.;
            hflag[m] = 0;
            if (ms->maps[m].is_covalent) {
                double d[3], rcov;
                for (int ii = 0; ii < 3; ii++) d[ii] = ms->cov_pos[ii] - c[ii];
                rcov = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / ms->cov_half_width;
                if (rcov < APPROX_ZERO) rcov = APPROX_ZERO;
                en[m] = ms->cov_barrier * (1. - exp(ln_half * rcov * rcov));
            }
        }
        /* e map: all atoms */
        if (elecPE >= 0) {
            double e = 0.0;
            for (int ia = 0; ia < N; ia++) {
                const double *a = rec->coord + 3 * (size_t)ia;
                const double dx = a[0] - c[0], dy = a[1] - c[1], dz = a[2] - c[2];
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                e += rec->charge[ia] * ms->coeff_estat * an_reps(r, ms->dddiel, ms->inv_diel_cal) / dmax(r, 0.5);
            }
            en[elecPE] = e;
        }
        /* closest donor H, FP64, lowest index on ties */
        double rminH = 999999.;
        int closestH = -1;
        for (int ia = 0; ia < N; ia++) {
            if (rec->hbond[ia] != 1 && rec->hbond[ia] != 2) continue;
            const double *a = rec->coord + 3 * (size_t)ia;
            const double dx = a[0] - c[0], dy = a[1] - c[1], dz = a[2] - c[2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r < rminH) {
                rminH = r;
                closestH = ia;
            }
        }
        for (int ia = 0; ia < N; ia++) {
            const double *a = rec->coord + 3 * (size_t)ia;
            const double *rv = rec->rvector + 3 * (size_t)ia;
            const double *rv2 = rec->rvector2 + 3 * (size_t)ia;
            const int hb = rec->hbond[ia], at = rec->type[ia];
            double d[3] = {a[0] - c[0], a[1] - c[1], a[2] - c[2]};
            double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (r >= 8.0) continue;
            if (r < APPROX_ZERO) r = APPROX_ZERO;
            const double S = an_switch(r);
            double dnorm[3];
            for (int i = 0; i < 3; i++) dnorm[i] = d[i] / r;
            double racc = 1., rdon = 1., Hramp = 1.;
            if (hb == 2) { /* D1 */
                double cos_theta = 0.;
                for (int i = 0; i < 3; i++) cos_theta -= dnorm[i] * rv[i];
                if (cos_theta <= 0.) {
                    racc = 0.;
                } else {
                    switch (rec->rexp[ia]) {
                        case 2: racc = cos_theta * cos_theta; break;
                        case 4: { double t = cos_theta * cos_theta; racc = t * t; } break;
                        default: racc = cos_theta; break;
                    }
                    if (ia != closestH && closestH >= 0) {
                        double ct = 0.;
                        for (int i = 0; i < 3; i++) ct += rec->rvector[3 * (size_t)closestH + i] * rv[i];
                        ct = dmin(ct, 1.0);
                        ct = dmax(ct, -1.0);
                        Hramp = 0.5 - 0.5 * cos(acos(ct) * 120. / 90.);
                    }
                }
            } else if (hb == 4) { /* A1 */
                double cos_theta = 0.;
                for (int i = 0; i < 3; i++) cos_theta -= dnorm[i] * rv[i];
                rdon = cos_theta <= 0. ? 0. : cos_theta * cos_theta;
            } else if (hb == 5) { /* A2, ordered (disorder_h off) */
                double cos_theta = 0.;
                for (int i = 0; i < 3; i++) cos_theta -= dnorm[i] * rv[i];
                double t0 = 0.;
                for (int i = 0; i < 3; i++) t0 += dnorm[i] * rv2[i];
                t0 = dmin(dmax(t0, -1.0), 1.0);
                t0 = PI_halved - acos(t0);
                rdon = 0.;
                if (cos_theta >= 0.) {
                    double cr[3] = {dnorm[1] * rv2[2] - dnorm[2] * rv2[1], dnorm[2] * rv2[0] - dnorm[0] * rv2[2],
                                    dnorm[0] * rv2[1] - dnorm[1] * rv2[0]};
                    double rd2 = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
                    if (rd2 < APPROX_ZERO) rd2 = APPROX_ZERO;
                    const double inv = 1. / sqrt(rd2);
                    double ti = 0.;
                    for (int i = 0; i < 3; i++) ti += cr[i] * inv * rv[i];
                    ti = dmin(dmax(ti, -1.0), 1.0);
                    ti = fabs(acos(ti) - PI_halved);
                    rdon = (0.9 + 0.1 * sin(ti + ti)) * cos(t0);
                } else if (cos_theta >= -0.34202) {
                    const double u = 0.116978 - cos_theta * cos_theta;
                    rdon = 562.25 * u * u * u * cos(t0);
                }
            }
            for (int m = 0; m < M; m++) {
                if (ms->maps[m].is_covalent) continue;
                int md = m; /* map receiving the desolvation term */
                if (ms->separate_desolvation) {
                    md = m + 1;
                }
                const ag_pair *pp = pairs + (size_t)m * T + at;
                const double E = an_energy(pp, r, w) * S;
                const double rsph = dmax(0.0, dmin(1.0, E / 100.));
                const int mh = ms->maps[m].hbond;
                if (ms->maps[m].is_hbonder) {
                    if ((mh == 3 || mh == 5 || mh == 6) && (hb == 1 || hb == 2 || hb == 6)) {
                        en[m] += E * Hramp * (racc + (1. - racc) * rsph);
                    } else if ((mh == 4 || mh == 6) && (hb == 1 || hb == 2 || hb == 6)) {
                        const double v = E * (racc + (1. - racc) * rsph);
                        hmin[m] = dmin(hmin[m], v);
                        hmax[m] = dmax(hmax[m], v);
                        hflag[m] = 1;
                    } else if ((mh == 1 || mh == 2 || mh == 6) && (hb > 2 || hb == 6)) {
                        const double v = E * (rdon + (1. - rdon) * rsph);
                        hmin[m] = dmin(hmin[m], v);
                        hmax[m] = dmax(hmax[m], v);
                        hflag[m] = 1;
                    } else {
                        en[m] += E;
                    }
                } else {
                    en[m] += E;
                }
                const double sol = coeff_desolv * exp(-(r * r) / (2. * 3.6 * 3.6)) * S;
                en[md] += ms->maps[md].solpar_probe * rec->vol[ia] * sol +
                          (rec->solpar[ia] + ms->solpar_q * fabs(rec->charge[ia])) * ms->maps[md].vol_probe * sol;
                if (ms->separate_desolvation) m++;
            }
            if (dsolvPE >= 0)
                en[dsolvPE] += ms->solpar_q * rec->vol[ia] * coeff_desolv * exp(-(r * r) / (2. * 3.6 * 3.6)) * S;
        }
        for (int m = 0; m < M; m++)
            if (hflag[m]) en[m] += hmin[m] + dmax(hmax[m], 0.);
        for (int m = 0; m < ms->n_maps; m++) out[(size_t)m * npts + k] = en[m];
    }
    return 0;
}

/* main.cpp:3486-3497 with HAVE_ROUND unset (main.cpp:85-89) */
int orc_format_value(double e, char *out) {
    if (fabs(e) < 0.0001) {
        strcpy(out, "0\n");
        return 2;
    }
    return sprintf(out, "%.3f\n", (float)(floor(e * 1000.0 + 0.5) / 1000.0));
}

/* Floating-grid value (main.cpp:3502-3507): "%.3f" of (float)round3dp(r_min). */
int orc_format_rmin(double r, char *out) {
    return sprintf(out, "%.3f\n", (float)(floor(r * 1000.0 + 0.5) / 1000.0));
}

/* ------------------------------------------------------------------------
 * W_A: the number of (point, atom) pairs Loop 2 keeps, i.e. atoms with
 * r <= SOFTNBC by the reference's FP64 distance (main.cpp:3036-3049), over
 * z planes [z0, z1).  Atoms are binned into cells of SOFTNBC + 0.1 A so
 * each point tests the 27 cells around its own; the test itself is the
 * oracle's Loop-2 distance.  Returns the count, or -1 on a bad grid.
 * ---------------------------------------------------------------------- */
long long orc_count_pairs(const ag_receptor *rec, const ag_grid *grid, int z0, int z1) {
    const int N = rec->n_atoms;
    const int n1x = grid->n1[0], n1y = grid->n1[1];
    const double h = grid->spacing, cell = SOFTNBC + 0.1;
    if (N <= 0) return 0;
    if (z0 < 0 || z1 > grid->n1[2] || z0 >= z1) return -1;
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = -grid->ne[k] * h - cell;
        hi[k] = (grid->n1[k] - 1 - grid->ne[k]) * h + cell;
    }
    int nc[3];
    for (int k = 0; k < 3; k++) nc[k] = (int)((hi[k] - lo[k]) / cell) + 1;
    const long long ncell = (long long)nc[0] * nc[1] * nc[2];
    int *start = (int *)calloc((size_t)ncell + 1, sizeof(int));
    int *cellof = (int *)malloc(sizeof(int) * (size_t)N);
    int *order = (int *)malloc(sizeof(int) * (size_t)N);
    if (!start || !cellof || !order) {
        free(start); free(cellof); free(order);
        return -1;
    }
    for (int i = 0; i < N; i++) {
        int ci[3], out = 0;
        for (int k = 0; k < 3; k++) {
            const double x = rec->coord[3 * (size_t)i + k];
            if (x < lo[k] || x >= hi[k]) out = 1;
            ci[k] = out ? 0 : (int)((x - lo[k]) / cell);
        }
        cellof[i] = out ? -1 : ci[0] + nc[0] * (ci[1] + nc[1] * ci[2]);
        if (!out) start[cellof[i] + 1]++;
    }
    for (long long c = 0; c < ncell; c++) start[c + 1] += start[c];
    {
        int *fill = (int *)calloc((size_t)ncell, sizeof(int));
        for (int i = 0; i < N; i++)
            if (cellof[i] >= 0) order[start[cellof[i]] + fill[cellof[i]]++] = i;
        free(fill);
    }
    long long total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int iz = z0; iz < z1; iz++) {
        double c[3];
        c[2] = (iz - grid->ne[2]) * h;
        for (int iy = 0; iy < n1y; iy++) {
            c[1] = (iy - grid->ne[1]) * h;
            for (int ix = 0; ix < n1x; ix++) {
                c[0] = (ix - grid->ne[0]) * h;
                int ci[3];
                for (int k = 0; k < 3; k++) ci[k] = (int)((c[k] - lo[k]) / cell);
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const int x = ci[0] + dx, y = ci[1] + dy, z = ci[2] + dz;
                            if (x < 0 || y < 0 || z < 0 || x >= nc[0] || y >= nc[1] || z >= nc[2]) continue;
                            const long long cc = x + (long long)nc[0] * (y + (long long)nc[1] * z);
                            for (int s = start[cc]; s < start[cc + 1]; s++) {
                                const double *a = rec->coord + 3 * (size_t)order[s];
                                const double r = hypot((a[0] - c[0]), hypot((a[1] - c[1]), (a[2] - c[2])));
                                if (!(r > SOFTNBC)) total++;
                            }
                        }
            }
        }
    }
    free(start);
    free(cellof);
    free(order);
    return total;
}
