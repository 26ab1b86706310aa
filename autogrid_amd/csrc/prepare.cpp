// Host-side precompute of the grid-map path: energy tables and H-bond
// geometry.  These run once per receptor in milliseconds and feed the device
// kernels; they follow the reference's arithmetic term by term so that the
// tables the kernels read are the reference's tables.

#include <cmath>
#include <cstring>

#include "ag_internal.h"

namespace ag {

// e_vdW_Hb[type][map][0..1023] (main.cpp:2040-2266): 12-6 / 12-10 pair
// potential clamped at EINTCLAMP, E[0] = EINTCLAMP, E[1023] = 0, then the
// min-window smoothing of half-width floor(r_smooth*128/2) bins.  The table
// element type is float as in the reference (typedefs.h:82); the arithmetic
// before storing is double.
void build_vdw_tables(const ag_setup &s, const ag_mapset &ms, float *out) {
    const int T = s.n_rec_types, M = s.n_atom_maps;
    std::vector<float> e(AG_NEINT);
    std::vector<double> smooth(AG_NEINT);
    memset(out, 0, sizeof(float) * (size_t)T * M * AG_NEINT);
    const int i_smooth = (int)(s.r_smooth * AG_A_DIV / 2.);
    for (int m = 0; m < M; m++) {
        if (ms.maps[m].is_covalent) continue;
        for (int t = 0; t < T; t++) {
            const ag_pair &p = s.pairs[(size_t)m * T + t];
            double cA = p.cA, cB = p.cB;
            const int xA = p.xA, xB = p.xB;
            if (cA == 0 && cB == 0) {
                double tmpconst = p.eps / (double)(xA - xB);
                cA = tmpconst * pow(p.Rij, (double)xA) * (double)xB;
                cB = tmpconst * pow(p.Rij, (double)xB) * (double)xA;
            }
            if (std::isnan(cA))
                fail(AG_ERR_BAD_INPUT, "Van der Waals coefficient cA is not a number.  AutoGrid must exit.");
            if (std::isnan(cB))
                fail(AG_ERR_BAD_INPUT, "Van der Waals coefficient cB is not a number.  AutoGrid must exit.");
            if (xA == 0) fail(AG_ERR_BAD_INPUT, "Van der Waals exponent xA is 0.  AutoGrid must exit.");
            if (xB == 0) fail(AG_ERR_BAD_INPUT, "Van der Waals exponent xB is 0.  AutoGrid must exit.");
            for (int k = 1; k < AG_NEINT; k++) {
                const double r = (double)k * (1. / AG_A_DIV);
                const double v = cA / pow(r, (double)xA) - cB / pow(r, (double)xB);
                e[k] = (float)(kEintClamp < v ? kEintClamp : v);
            }
            e[0] = (float)kEintClamp;
            e[AG_NEINT - 1] = 0.f;
            if (i_smooth > 0) {
                for (int k = 0; k < AG_NEINT; k++) {
                    double lo = 100000.;
                    const int j0 = k - i_smooth < 0 ? 0 : k - i_smooth;
                    const int j1 = k + i_smooth + 1 > AG_NEINT ? AG_NEINT : k + i_smooth + 1;
                    for (int j = j0; j < j1; j++) lo = lo < e[j] ? lo : (double)e[j];
                    smooth[k] = lo;
                }
                for (int k = 0; k < AG_NEINT; k++) e[k] = (float)smooth[k];
            }
            memcpy(out + ((size_t)t * M + m) * AG_NEINT, e.data(), sizeof(float) * AG_NEINT);
        }
    }
}

// sol_fn[k] = coeff_desolv * exp(-r^2 / (2 sigma^2)), sigma = 3.6, sol_fn[0]
// left at 0 (main.cpp:2272-2276).
void build_sol_fn(double coeff_desolv, double *out) {
    const double sigma = 3.6;
    out[0] = 0.0;
    for (int k = 1; k < AG_NDIEL; k++) {
        const double r = (double)k * (1. / AG_A_DIV);
        out[k] = coeff_desolv * exp(-(r * r) / (2. * (sigma * sigma)));
    }
}

// Mehler-Solmajer distance-dependent dielectric (distdepdiel.cpp:30-57); the
// reference evaluates the denominator in long double (1.0L literal).
static double ddd_mehler_solmajer(double distance) {
    const double lambda = 0.003627L, epsilon0 = 78.4L, A = -8.5525L, rk = 7.7839L;
    const double B = epsilon0 - A;
    const double lambda_B = -lambda * B;
    double epsilon = A + B / (1.0L + rk * exp(lambda_B * distance));
    if (epsilon < kApproxZero) epsilon = 1.0L;
    return epsilon;
}

// r_epsilon_fn[k] = 332 / eps(k/128), eps[0] = 1 (main.cpp:1612-1628).
void build_r_eps(double *out) {
    out[0] = 332.0 / 1.0;
    for (int k = 1; k < AG_NDIEL; k++) out[k] = 332.0 / ddd_mehler_solmajer((double)k * (1. / AG_A_DIV));
}

// --- bond vectors (main.cpp:2297-2736) -------------------------------------

static inline double len3(const double v[3]) { return hypot(v[0], hypot(v[1], v[2])); }

// vect_normalize (main.cpp:3606-3617)
static double normalize(double v[3]) {
    double rd = len3(v);
    double inv = rd < kApproxZero ? 1. / kApproxZero : 1. / rd;
    for (int i = 0; i < 3; i++) v[i] *= inv;
    return rd;
}

// vect_sub_normalize (main.cpp:3620-3633)
static double sub_normalize(double res[3], const double a[3], const double b[3]) {
    double rd2 = 0.;
    for (int i = 0; i < 3; i++) {
        res[i] = a[i] - b[i];
        rd2 += res[i] * res[i];
    }
    double rd = rd2 < kApproxZero ? kApproxZero : sqrt(rd2);
    double inv = 1. / rd;
    for (int i = 0; i < 3; i++) res[i] *= inv;
    return rd;
}

// distance_le (main.cpp:121-129): componentwise early-out, then the hypot
// length; dc = a - b.
static bool distance_le(const double *a, const double *b, double limit, double dc[3], double &d) {
    dc[0] = a[0] - b[0];
    if (dc[0] > limit || dc[0] < -limit) return false;
    dc[1] = a[1] - b[1];
    if (dc[1] > limit || dc[1] < -limit) return false;
    dc[2] = a[2] - b[2];
    if (dc[2] > limit || dc[2] < -limit) return false;
    d = len3(dc);
    return d <= limit;
}

// Atoms of the index window [from, to] (other than `self`) whose squared
// distance from atom `c` could pass a distance_le test of at most `lim`: a
// branch-free pass over the window, in index order.  The squared distance of
// any atom left out exceeds lim^2 (1 + 1e-9), so its hypot length exceeds
// lim and distance_le would reject it; the callers run distance_le on the
// rest in the same order, so they find what a scan of the whole window finds.
static int window_candidates(const double *coord, int c, int self, int from, int to, double lim, int *out) {
    const double cx = coord[3 * (size_t)c], cy = coord[3 * (size_t)c + 1], cz = coord[3 * (size_t)c + 2];
    const double l2 = lim * lim * (1.0 + 1e-9);
    int k = 0;
    for (int ib = from; ib <= to; ib++) {
        const double dx = coord[3 * (size_t)ib] - cx, dy = coord[3 * (size_t)ib + 1] - cy,
                     dz = coord[3 * (size_t)ib + 2] - cz;
        out[k] = ib;
        k += (dx * dx + dy * dy + dz * dz <= l2) & (ib != self);
    }
    return k;
}

struct Bonds {
    std::vector<int> n;
    std::vector<int> b;  // [atom][kMaxNbonds]
    // the list is only kept for disorder_h; without it no bond is added
    explicit Bonds(int natoms) : n(natoms, 0), b((size_t)natoms * kMaxNbonds, -1) {}
    Bonds() = default;
    void add(int i, int j) {  // bondmanager.cpp:44-93
        if (n[i] >= kMaxNbonds || n[j] >= kMaxNbonds) fail(AG_ERR_BAD_INPUT, "could not add bond");
        b[(size_t)i * kMaxNbonds + n[i]++] = j;
        b[(size_t)j * kMaxNbonds + n[j]++] = i;
    }
};

void build_bond_vectors(const ag_setup &s, int n, const double *coord, const int32_t *type,
                        const int32_t *hbond, int32_t *rexp, double *rvector, double *rvector2,
                        uint8_t *disorder, FILE *log, BondRecipe *recipes) {
    if (n <= 0) return;
    if (recipes)
        for (int i = 0; i < n; i++) recipes[i] = BondRecipe{};
    const int hydrogen = s.t_hydrogen, nonHB_hydrogen = s.t_nonhb_hydrogen;
    const int carbon = s.t_carbon, arom_carbon = s.t_arom_carbon, oxygen = s.t_oxygen;
    const int nitrogen = s.t_nitrogen, nonHB_nitrogen = s.t_nonhb_nitrogen, sulphur = s.t_sulphur;
    const bool disorder_h = s.disorder_h != 0;
    for (int i = 0; i < n; i++) rexp[i] = 0;
    memset(rvector, 0, sizeof(double) * 3 * (size_t)n);
    memset(rvector2, 0, sizeof(double) * 3 * (size_t)n);
    memset(disorder, 0, (size_t)n);
    Bonds bonds = disorder_h ? Bonds(n) : Bonds();
    auto C = [&](int a) { return coord + 3 * (size_t)a; };
    for (int ia = 0; ia < n; ia++) {
        const int from = ia - 20 > 0 ? ia - 20 : 0;
        const int to = ia + 20 < n - 1 ? ia + 20 : n - 1;
        double *rv = rvector + 3 * (size_t)ia, *rv2 = rvector2 + 3 * (size_t)ia;
        int cand[41];
        if (hbond[ia] == AG_HB_D1) {
            // Donor hydrogen: every atom within 1.378 A in the window sets the
            // bond vector (the last one wins) and rexp.
            const int nc = window_candidates(coord, ia, ia, from, to, 1.378, cand);
            for (int q = 0; q < nc; q++) {
                const int ib = cand[q];
                double d[3], rd;
                if (!distance_le(C(ia), C(ib), 1.378, d, rd)) continue;
                for (int i = 0; i < 3; i++) rv[i] = d[i];
                normalize(rv);
                if (recipes) recipes[ia] = BondRecipe{BondRecipe::kDiff, ib, -1, -1, 0};
                if (type[ib] == nitrogen) rexp[ia] = 2;
                else if (type[ib] == oxygen || type[ib] == sulphur) rexp[ia] = 4;
                else rexp[ia] = 2;
                if (disorder_h) bonds.add(ia, ib);
            }
        } else if (hbond[ia] == AG_HB_A2) {
            int nbond = 0, i1 = -1, i2 = -1;
            const int nc = window_candidates(coord, ia, ia, from, to, 1.9, cand);
            for (int q = 0; q < nc; q++) {
                const int ib = cand[q];
                double dc[3], rd;
                const double lim = (type[ib] == hydrogen || type[ib] == nonHB_hydrogen) ? 1.3 : 1.9;
                if (!distance_le(C(ia), C(ib), lim, dc, rd)) continue;
                if (nbond == 2) {
                    if (log)
                        fprintf(log, "\nWARNING: Found an H-bonding O atom with three bonded atoms, atom serial number %d\n", ia + 1);
                } else if (nbond == 1) {
                    nbond = 2;
                    i2 = ib;
                } else {
                    nbond = 1;
                    i1 = ib;
                }
            }
            if (nbond == 0 && log)
                fprintf(log, "\nWARNING: Oxygen atom found with no bonded atoms, atom serial number %d, atom_type %d\n",
                        ia + 1, type[ia]);
            if (nbond == 1) {
                // Carbonyl O=C-X: rvector = O - C, rvector2 = (O-C) x (C-X).
                sub_normalize(rv, C(ia), C(i1));
                if (recipes) recipes[ia] = BondRecipe{BondRecipe::kCarbonyl, i1, -1, -1, 0};
                int cj[41];
                const int ncj = window_candidates(coord, i1, i1, from, to, 1.7, cj);
                for (int q = 0; q < ncj; q++) {
                    const int j = cj[q];
                    if (j == ia) continue;
                    double d[3], rdcx;
                    if (!distance_le(C(i1), C(j), type[j] == hydrogen ? 1.3 : 1.7, d, rdcx)) continue;
                    if (recipes) recipes[ia].b = j;
                    normalize(d);
                    rv2[0] = rv[1] * d[2] - rv[2] * d[1];
                    rv2[1] = rv[2] * d[0] - rv[0] * d[2];
                    rv2[2] = rv[0] * d[1] - rv[1] * d[0];
                    normalize(rv2);
                }
            } else if (nbond == 2) {
                if (disorder_h && (type[i1] == hydrogen || type[i2] == hydrogen) && type[i1] != type[i2]) {
                    int ib;
                    if (type[i1] == carbon || type[i1] == arom_carbon) ib = i1;
                    else if (type[i2] == carbon || type[i2] == arom_carbon) ib = i2;
                    else ib = i1;
                    bonds.add(ia, ib);
                    sub_normalize(rv, C(ia), C(ib));
                    if (recipes) recipes[ia] = BondRecipe{BondRecipe::kDiff, ib, -1, -1, 0};
                } else {
                    // Hydroxyl / ether X1-O-X2: lone-pair plane along X1->X2,
                    // rvector from the X1-X2 line to O.
                    sub_normalize(rv2, C(i2), C(i1));
                    if (recipes) recipes[ia] = BondRecipe{BondRecipe::kEther, i1, i2, -1, 0};
                    double rdot = 0.;
                    for (int i = 0; i < 3; i++) rdot += (C(ia)[i] - C(i1)[i]) * rv2[i];
                    for (int i = 0; i < 3; i++) rv[i] = C(ia)[i] - ((rdot * rv2[i]) + C(i1)[i]);
                    normalize(rv);
                }
            }
        } else if (hbond[ia] == AG_HB_A1) {
            // Directional N acceptor: up to three bonded atoms; rvector is the
            // negated normalised mean of the (N - X) vectors.
            int nbond = 0;
            double bondv[3][3];
            int bonded[3] = {-1, -1, -1};
            const int nc = window_candidates(coord, ia, ia, from, to, 1.7, cand);
            for (int q = 0; q < nc; q++) {
                const int ib = cand[q];
                double rd;
                const double lim = (type[ib] == hydrogen || type[ib] == nonHB_hydrogen) ? 1.3 : 1.7;
                if (!distance_le(C(ia), C(ib), lim, bondv[nbond], rd)) continue;
                bonded[nbond] = ib;
                if (++nbond == 3) break;
            }
            if (recipes && nbond > 0) recipes[ia] = BondRecipe{BondRecipe::kA1, bonded[0], bonded[1], bonded[2], nbond};
            if (nbond == 0 && log)
                fprintf(log, "\nWARNING: Nitrogen atom found with no bonded atoms, atom serial number %d\n", ia + 1);
            double mean[3] = {0, 0, 0};
            for (int b = 0; b < nbond; b++)
                for (int i = 0; i < 3; i++) mean[i] += bondv[b][i] / nbond;
            normalize(mean);
            for (int i = 0; i < 3; i++) rv[i] = -1 * mean[i];
        }
    }
    if (disorder_h) {
        // main.cpp:2684-2736
        for (int ia = 0; ia < n; ia++) {
            int hcount = 0;
            for (int b = 0; b < kMaxNbonds && b < bonds.n[ia]; b++)
                if (type[bonds.b[(size_t)ia * kMaxNbonds + b]] == hydrogen) hcount++;
            if ((type[ia] == nitrogen && hcount == 2) || (type[ia] == nonHB_nitrogen && hcount == 3) ||
                (type[ia] == oxygen && hcount == 1) || (type[ia] == sulphur && hcount == 1)) {
                disorder[ia] = 1;
                for (int b = 0; b < kMaxNbonds && b < bonds.n[ia]; b++) {
                    int ba = bonds.b[(size_t)ia * kMaxNbonds + b];
                    if (type[ba] == hydrogen) disorder[ba] = 1;
                }
            }
        }
    }
}

void job_prepare(Job &job) {
    const int T = (int)job.rec_types.size();
    const int M = job.num_atom_maps;
    job.pairs.assign((size_t)M * T, ag_pair{});
    for (int m = 0; m < M; m++)
        for (int t = 0; t < T; t++) job.pairs[(size_t)m * T + t] = job.maps[m].pair[t];
    job.map_desc.assign(M, ag_map_desc{});
    for (int m = 0; m < M; m++) {
        ag_map_desc &d = job.map_desc[m];
        d.is_covalent = job.maps[m].is_covalent;
        d.is_hbonder = job.maps[m].is_hbonder;
        d.hbond = job.maps[m].hbond;
        d.solpar_probe = job.maps[m].solpar_probe;
        d.vol_probe = job.maps[m].vol_probe;
    }
    ag_receptor rec;
    ag_grid grid;
    ag_mapset ms;
    ag_tables tab;
    ag_setup setup;
    // views need the vectors sized first
    job.e_vdw_hb.assign((size_t)T * M * AG_NEINT, 0.f);
    job.sol_fn.assign(AG_NDIEL, 0.0);
    job.r_eps.assign(AG_NDIEL, 0.0);
    job.rexp.assign(job.n_atoms, 0);
    job.rvector.assign(3 * (size_t)job.n_atoms, 0.0);
    job.rvector2.assign(3 * (size_t)job.n_atoms, 0.0);
    job.disorder.assign(job.n_atoms, 0);
    job.prepared = true;
    job_views(job, &rec, &grid, &ms, &tab, &setup);
    build_vdw_tables(setup, ms, job.e_vdw_hb.data());
    build_sol_fn(job.lib.coeff_desolv, job.sol_fn.data());
    if (job.dddiel) build_r_eps(job.r_eps.data());
    build_bond_vectors(setup, job.n_atoms, job.coord.data(), job.type.data(), job.hbond.data(),
                       job.rexp.data(), job.rvector.data(), job.rvector2.data(), job.disorder.data(),
                       job.log);
}

}  // namespace ag

namespace ag {

// xs_vdw radius of a ligand or receptor type: the ADVINA table of
// read_parameter_library.cc:320-358 (exact type-name match, float values);
// types outside it get the 1.8 A default there (a double literal).
bool xs_radius_lookup(const char *type, double *radius) {
    static const struct {
        const char *name;
        float value;
    } table[] = {{"C", 1.9f},  {"A", 1.9f},  {"N", 1.8f},  {"NA", 1.8f}, {"NS", 1.8f},
                 {"O", 1.7f},  {"OA", 1.7f}, {"OS", 1.7f}, {"S", 2.0f},  {"SA", 2.0f},
                 {"P", 2.1f},  {"F", 1.5f},  {"Mg", 1.2f}};
    for (const auto &e : table)
        if (strcmp(type, e.name) == 0) {
            *radius = (double)e.value;
            return true;
        }
    *radius = 1.8;
    return false;
}

// Vina-potential tables (the use_vina_potential branch of main.cpp:2075-2127,
// never enabled in the shipped reference).  For map m, receptor type t and
// bin k = 1..1023, with rdd = k/128 - (xs_rec + xs_map):
//   gauss1 + gauss2 terms, repulsion rdd^2 for rdd < 0,
//   H-bond pairs (hbonder[m][t] > 0): wt_h for rdd <= 0.7, then minus
//     (rdd/0.7) wt_h for -0.7 < rdd <= 0,
//   hydrophobic (hbonder[m][t] == 0 and hbonder[t][m] == 0, the latter
//     read as 0 where t or m is outside the map or type range): wt_p below
//     0.5, (0.5 - rdd) wt_p below 1.5.
// Each term is added to the float table entry in that order, as the
// reference's += on a float table does; bin 0 stays 0; no smoothing and no
// clamp.  The weights are the reference's float constants; the receptor
// radius passes through a float (map_Rij), the ligand one stays double.
void build_vina_tables(const ag_setup &s, const ag_mapset &ms, const double *xs_map, const double *xs_rec,
                       float *out) {
    const int T = s.n_rec_types, M = s.n_atom_maps;
    const float wt_gauss1 = -0.035579f, wt_gauss2 = -0.005156f, wt_repulsion = 0.840245f,
                wt_hydrogen = -0.587439f, wt_hydrophobic = -0.035069f;
    memset(out, 0, sizeof(float) * (size_t)T * M * AG_NEINT);
    for (int m = 0; m < M; m++) {
        if (ms.maps[m].is_covalent) continue;
        for (int t = 0; t < T; t++) {
            const double lig_r = xs_map[m];
            const float rec_r = (float)xs_rec[t];
            const int hb = s.pairs[(size_t)m * T + t].hbonder;
            const int hb_cross = (t < M && m < T) ? s.pairs[(size_t)t * T + m].hbonder : 0;
            float *e = out + ((size_t)t * M + m) * AG_NEINT;
            for (int k = 1; k < AG_NEINT; k++) {
                const double r = (double)k / AG_A_DIV;
                const double rdd = r - ((double)rec_r + lig_r);
                double de = wt_gauss1 * exp(-pow(rdd / 0.5, 2)) + wt_gauss2 * exp(-pow((rdd - 3.) / 2., 2));
                e[k] += de;
                if (rdd < 0) {
                    de = wt_repulsion * pow(rdd, 2);
                    e[k] += de;
                }
                if (hb > 0) {
                    if (rdd <= 0.7) {
                        de = 1 * wt_hydrogen;
                        e[k] += de;
                    }
                    if (-0.7 < rdd && rdd <= 0.) {
                        de = (rdd / 0.7) * wt_hydrogen;
                        e[k] -= de;
                    }
                }
                if (hb == 0 && hb_cross == 0) {
                    de = 0.;
                    if (rdd < 0.5) de = 1 * wt_hydrophobic;
                    else if (rdd < 1.5) de = (0.5 - rdd) * wt_hydrophobic;
                    e[k] += de;
                }
            }
        }
    }
}

void job_prepare_vina(Job &job) {
    if (job.separate_desolvation)
        fail(AG_ERR_UNSUPPORTED, "Vina-potential maps do not combine with separate_desolvation_maps");
    if (!job.prepared) job_prepare(job);
    const int T = (int)job.rec_types.size();
    const int M = job.num_atom_maps;
    std::vector<double> xs_map(M), xs_rec(T);
    for (int m = 0; m < M; m++)
        if (!xs_radius_lookup(job.maps[m].type.c_str(), &xs_map[m]))
            job.logf("%s: WARNING: atom type %s not found in the xs_vdw_lookup table.\n\n", job.program.c_str(),
                     job.maps[m].type.c_str());
    for (int t = 0; t < T; t++)
        if (!xs_radius_lookup(job.rec_types[t].c_str(), &xs_rec[t]))
            job.logf("%s: WARNING: atom type %s not found in the xs_vdw_lookup table.\n\n", job.program.c_str(),
                     job.rec_types[t].c_str());
    ag_receptor rec;
    ag_grid grid;
    ag_mapset ms;
    ag_tables tab;
    ag_setup setup;
    job_views(job, &rec, &grid, &ms, &tab, &setup);
    build_vina_tables(setup, ms, xs_map.data(), xs_rec.data(), job.e_vdw_hb.data());
    job.vina = true;
    job.logf("\n Using Vina potential for calculation.\n\n");
}

}  // namespace ag

namespace ag {

// --- gradients through the bond vectors ------------------------------------

static void norm_vjp(const double v[3], const double g[3], double out[3]) {
    // u = v / |v|: dL/dv = (g - (g.u) u) / |v|
    const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len < kApproxZero) {
        for (int k = 0; k < 3; k++) out[k] = 0.0;
        return;
    }
    const double u[3] = {v[0] / len, v[1] / len, v[2] / len};
    const double d = g[0] * u[0] + g[1] * u[1] + g[2] * u[2];
    for (int k = 0; k < 3; k++) out[k] = (g[k] - d * u[k]) / len;
}

static void unit(const double v[3], double u[3]) {
    const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; k++) u[k] = len < kApproxZero ? 0.0 : v[k] / len;
}

// Adds into g_coord the gradient that flows from dL/drvector and
// dL/drvector2 through the bond-vector construction of each atom (the
// geometry of its recipe is held fixed: which atoms are bonded does not
// change under an infinitesimal move).
void bond_vectors_vjp(int n, const double *coord, const BondRecipe *rec, const double *g_rv,
                      const double *g_rv2, double *g_coord) {
    auto X = [&](int i) { return coord + 3 * (size_t)i; };
    auto G = [&](int i) { return g_coord + 3 * (size_t)i; };
    for (int i = 0; i < n; i++) {
        const BondRecipe &r = rec[i];
        const double *gr = g_rv + 3 * (size_t)i, *gr2 = g_rv2 + 3 * (size_t)i;
        switch (r.kind) {
            case BondRecipe::kDiff: {   // rv = N(x_i - x_a)
                double v[3], gv[3];
                for (int k = 0; k < 3; k++) v[k] = X(i)[k] - X(r.a)[k];
                norm_vjp(v, gr, gv);
                for (int k = 0; k < 3; k++) {
                    G(i)[k] += gv[k];
                    G(r.a)[k] -= gv[k];
                }
            } break;
            case BondRecipe::kCarbonyl: {   // rv = N(x_i - x_a); rv2 = N(rv x N(x_a - x_b))
                double v[3], u[3];
                for (int k = 0; k < 3; k++) v[k] = X(i)[k] - X(r.a)[k];
                unit(v, u);
                double gu[3] = {gr[0], gr[1], gr[2]};
                if (r.b >= 0) {
                    double tv[3], t[3];
                    for (int k = 0; k < 3; k++) tv[k] = X(r.a)[k] - X(r.b)[k];
                    unit(tv, t);
                    const double c[3] = {u[1] * t[2] - u[2] * t[1], u[2] * t[0] - u[0] * t[2], u[0] * t[1] - u[1] * t[0]};
                    double gc[3];
                    norm_vjp(c, gr2, gc);
                    // c = u x t: du = t x gc, dt = gc x u
                    gu[0] += t[1] * gc[2] - t[2] * gc[1];
                    gu[1] += t[2] * gc[0] - t[0] * gc[2];
                    gu[2] += t[0] * gc[1] - t[1] * gc[0];
                    const double gt[3] = {gc[1] * u[2] - gc[2] * u[1], gc[2] * u[0] - gc[0] * u[2],
                                          gc[0] * u[1] - gc[1] * u[0]};
                    double gtv[3];
                    norm_vjp(tv, gt, gtv);
                    for (int k = 0; k < 3; k++) {
                        G(r.a)[k] += gtv[k];
                        G(r.b)[k] -= gtv[k];
                    }
                }
                double gv[3];
                norm_vjp(v, gu, gv);
                for (int k = 0; k < 3; k++) {
                    G(i)[k] += gv[k];
                    G(r.a)[k] -= gv[k];
                }
            } break;
            case BondRecipe::kEther: {   // rv2 = N(x_b - x_a); rv = N(o - ((o - a).rv2 rv2 + a))
                const double *o = X(i), *a = X(r.a), *b = X(r.b);
                double w[3], rv2[3];
                for (int k = 0; k < 3; k++) w[k] = b[k] - a[k];
                unit(w, rv2);
                double oa[3], rdot = 0.0;
                for (int k = 0; k < 3; k++) {
                    oa[k] = o[k] - a[k];
                    rdot += oa[k] * rv2[k];
                }
                double v[3];
                for (int k = 0; k < 3; k++) v[k] = o[k] - (rdot * rv2[k] + a[k]);
                double gv[3];
                norm_vjp(v, gr, gv);
                double go[3], ga[3], grv2[3] = {gr2[0], gr2[1], gr2[2]};
                double grdot = 0.0;
                for (int k = 0; k < 3; k++) {
                    go[k] = gv[k];
                    ga[k] = -gv[k];
                    grdot -= gv[k] * rv2[k];
                    grv2[k] -= rdot * gv[k];
                }
                for (int k = 0; k < 3; k++) {
                    go[k] += grdot * rv2[k];
                    ga[k] -= grdot * rv2[k];
                    grv2[k] += grdot * oa[k];
                }
                double gw[3];
                norm_vjp(w, grv2, gw);
                for (int k = 0; k < 3; k++) {
                    G(i)[k] += go[k];
                    G(r.a)[k] += ga[k] - gw[k];
                    G(r.b)[k] += gw[k];
                }
            } break;
            case BondRecipe::kA1: {   // rv = -N(sum_k (x_i - x_k) / nb)
                const int idx[3] = {r.a, r.b, r.c};
                double m[3] = {0, 0, 0};
                for (int q = 0; q < r.nb; q++)
                    for (int k = 0; k < 3; k++) m[k] += (X(i)[k] - X(idx[q])[k]) / r.nb;
                const double ng[3] = {-gr[0], -gr[1], -gr[2]};
                double gm[3];
                norm_vjp(m, ng, gm);
                for (int k = 0; k < 3; k++) {
                    G(i)[k] += gm[k];
                    for (int q = 0; q < r.nb; q++) G(idx[q])[k] -= gm[k] / r.nb;
                }
            } break;
            default:
                break;
        }
    }
}

}  // namespace ag
