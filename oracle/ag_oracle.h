/*
 * ag_oracle.h -- FP64 CPU restatement of autogrid4's grid-map path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker of the MI355X
 * engine: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it.  The engine itself never links or calls it.
 *
 * Every function restates one part of the reference (ccsb-scripps/AutoGrid,
 * main.cpp and friends) in plain C99 double arithmetic, term for term and in
 * the reference's summation order, so that on x86-64 (no FMA contraction)
 * it reproduces the reference's doubles exactly.  It reads the seam structs
 * of include/autogrid_amd.h.
 */
#ifndef AG_ORACLE_H
#define AG_ORACLE_H
#include "../include/autogrid_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* e_vdW_Hb tables [n_rec_types][n_atom_maps][1024] (main.cpp:2040-2266).
 * Returns 0, or -1 when the reference would stop (NaN coefficient, zero
 * exponent). */
int orc_vdw_tables(const ag_setup *s, const ag_mapset *ms, float *out);
/* sol_fn[8192] (main.cpp:2272-2276). */
void orc_sol_fn(double coeff_desolv, double *out);
/* r_epsilon_fn[8192] = 332/eps (main.cpp:1612-1628, distdepdiel.cpp:30-57). */
void orc_r_eps(double *out);
/* Mehler-Solmajer epsilon(r) (distdepdiel.cpp:30-57). */
double orc_ddd(double r);
/* rexp, rvector, rvector2, disorder (main.cpp:2297-2736). Returns 0, or -1
 * when the reference stops ("could not add bond"). */
int orc_bond_vectors(const ag_setup *s, int n, const double *coord, const int32_t *type,
                     const int32_t *hbond, int32_t *rexp, double *rvector,
                     double *rvector2, uint8_t *disorder);
/* Per-point loop (main.cpp:2805-3431) over z planes [z0, z1): maps is
 * [n_maps][P] double.  fmap (r_min) and cmap (count) may be NULL.
 * Returns 0, or -1 when no donor H exists and atom maps are requested. */
int orc_maps(const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
             const ag_tables *tab, double *maps, double *fmap, int32_t *cmap,
             int z0, int z1);
/* W_A: (point, atom) pairs with r <= 8 A by Loop 2's FP64 distance over z
 * planes [z0, z1) (main.cpp:3036-3049); -1 on a bad plane range. */
long long orc_count_pairs(const ag_receptor *rec, const ag_grid *grid, int z0, int z1);
/* xs_vdw radius of a type (read_parameter_library.cc:322-358); 1 if listed. */
int orc_xs_radius(const char *type, double *radius);
/* Vina tables [n_rec_types][n_atom_maps][1024] (main.cpp:2075-2127). */
int orc_vina_tables(const ag_setup *s, const ag_mapset *ms, const double *xs_map,
                    const double *xs_rec, float *out);
/* Vina map set over z planes [z0, z1) (main.cpp:2805-3431 with
 * use_vina_potential on and no directional factors). */
int orc_maps_vina(const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                  const ag_tables *tab, double *maps, double *fmap, int32_t *cmap,
                  int z0, int z1);
/* Analytic model (the engine's differentiable mode) in FP64 at the given
 * point indices: out [n_maps][npts].  pairs [n_atom_maps][T] hold the
 * effective (weighted) eps; coeff_estat comes from the map set.  Returns -2
 * when disorder_h is set (not part of the analytic model). */
int orc_analytic_maps(const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                      const ag_pair *pairs, int T, double r_smooth, double coeff_desolv,
                      int npts, const int64_t *points, double *out);
/* One map value as the reference writes it (main.cpp:3486-3497). */
int orc_format_value(double e, char *out);
/* One floating-grid value (main.cpp:3502-3507). */
int orc_format_rmin(double r, char *out);

#ifdef __cplusplus
}
#endif
#endif
