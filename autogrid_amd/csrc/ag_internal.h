// Internal host-side types of the MI355X AutoGrid4 engine.
#pragma once

#include <cstdio>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>
#include <stdexcept>

#include "autogrid_amd.h"

namespace ag {

// Error carrying an ag_status; thrown inside the library, turned into a status
// plus ag_last_error() text at the C boundary.
struct Error : std::runtime_error {
    ag_status status;
    Error(ag_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] void fail(ag_status s, const std::string &msg);
void set_last_error(const std::string &msg);

// Constants of the 8 A build.
constexpr double kApproxZero = 1.0e-6;   // APPROX_ZERO, autocomm.h:72
constexpr double kEintClamp = 100000.0;  // EINTCLAMP, autocomm.h:89
constexpr double kBig = 1.0e12;          // BIG, autocomm.h:73
constexpr double kSoftNbc = 8.0;         // SOFTNBC, constants.h:150
constexpr double kPrecision = 0.0001;    // PRECISION, autogrid.h:68
constexpr int kMaxNbonds = 4;            // AG_MAX_NBONDS, autogrid.h:59

// One atom_par entry of the parameter library (ParameterEntry, parameters.h:39-61).
struct ParamEntry {
    char type[8] = {0};
    double Rij = 0, epsij_unweighted = 0, vol = 0, solpar = 0;
    double Rij_hb = 0, epsij_hb_unweighted = 0;
    int hbond = 0;
    int rec_index = -1, map_index = -1, bond_index = 0;
};

// Parameter dictionary keyed on the first two characters of the type name,
// case-sensitive (atom_parameter_manager.cpp:37-55).
struct ParamLibrary {
    std::vector<int> slot;             // 65536 entries, index into entries or -1
    std::vector<ParamEntry> entries;
    double coeff_vdW = 0, coeff_hbond = 0, coeff_estat = 0, coeff_desolv = 0,
           coeff_tors = 0;
    ParamLibrary();
    static unsigned key(const char *t);
    ParamEntry *find(const char *t);
    void enter(const ParamEntry &e);
    // Apply one parameter-library line (read_parameter_library.cc:84-199).
    void apply_line(const char *line, ParamEntry &scratch);
    void load_builtin_ad41_bound();
    void load_file(const std::string &path);
};

struct MapInfo {
    std::string type;         // ligand type, e.g. "OA"
    std::string filename;
    FILE *fp = nullptr;
    bool is_covalent = false, is_hbonder = false;
    int hbond = 0;
    double solpar_probe = 0, vol_probe = 0;
    double Rij = 0, epsij = 0, Rij_hb = 0, epsij_hb = 0;
    ag_pair pair[AG_MAX_REC_TYPES];
};

// Everything main() holds once the GPF loop ends, plus the prepared state.
struct Job {
    // bookkeeping / files
    std::string program = "autogrid4";
    std::string gpf_path, receptor_filename, fld_filename, xyz_filename;
    std::string fmap_filename, cmap_filename;
    FILE *log = nullptr;
    bool own_log = false;
    bool make_files = true;
    bool bhtree = false;                // semantics of a -DUSE_BHTREE build
    bool map_receptor_interior = false; // main.cpp:580
    bool binary_maps = false;           // raw FP32 side-output next to the ASCII maps
    bool headers_written = false;       // the 6 header lines of every map file
    FILE *fld_fp = nullptr, *xyz_fp = nullptr, *fmap_fp = nullptr, *cmap_fp = nullptr;
    bool xyz_open = false;       // gridfld seen and .xyz not yet written
    int outlev = 1;

    ParamLibrary lib;

    // receptor
    std::vector<double> coord;         // [n][3], centred when gridcenter seen
    std::vector<double> charge, vol, solpar;
    std::vector<int32_t> type, hbond;
    int n_atoms = 0;
    std::vector<std::string> rec_types;   // receptor_types[] in rec_index order
    bool has_receptor_types_in_gpf = false;
    int receptor_types_gpf_ct = 0;
    double cmean[3] = {0, 0, 0};

    // grid
    int nelements[3] = {0, 0, 0};
    int n1[3] = {0, 0, 0};
    int ne[3] = {0, 0, 0};
    bool npts_set = false;
    double spacing = 0.375;
    double center[3] = {0, 0, 0};

    // maps
    std::vector<MapInfo> maps;         // num_atom_maps + 2 slots as in the reference
    int num_atom_maps = 0, num_maps = 0;
    int elecPE = -1, dsolvPE = -1;
    bool ligand_types_seen = false;
    bool separate_desolvation = false;
    bool disorder_h = false;
    bool dddiel = false;
    double inv_diel_cal = 0;
    double r_smooth = 0.5;
    double solpar_q = 0.01097;
    double cov_pos[3] = {0, 0, 0};
    double cov_half_width = 1.0, cov_barrier = 1000.0;
    bool floating_grid = false, constriction_grid = false;
    double constriction_cutoff = 7.5;
    int num_distance_maps = 0;

    // prepared state (host)
    bool prepared = false;
    std::vector<ag_pair> pairs;        // [num_atom_maps][n_rec_types]
    std::vector<float> e_vdw_hb;       // [n_rec_types][num_atom_maps][1024]
    std::vector<double> sol_fn, r_eps; // [8192]
    std::vector<int32_t> rexp;
    std::vector<double> rvector, rvector2;
    std::vector<uint8_t> disorder;
    std::vector<ag_map_desc> map_desc;
    bool vina = false;                 // e_vdw_hb holds Vina-potential tables
    int t_hydrogen = -1, t_nonhb_hydrogen = -1, t_carbon = -1, t_arom_carbon = -1;
    int t_oxygen = -1, t_sulphur = -1, t_nitrogen = -1, t_nonhb_nitrogen = -1;

    ~Job();
    int64_t points() const { return (int64_t)n1[0] * n1[1] * n1[2]; }
    int rec_index(const char *t) { ParamEntry *p = lib.find(t); return p ? p->rec_index : -1; }
    int map_index(const char *t) { ParamEntry *p = lib.find(t); return p ? p->map_index : -1; }
    void logf(const char *fmt, ...) __attribute__((format(printf, 2, 3)));
};

// front end
void job_parse_gpf(Job &job, const std::string &gpf_path);
void job_write_fld(Job &job);
void job_write_map_headers(Job &job);
// prepare
void job_prepare(Job &job);
void job_views(const Job &job, ag_receptor *rec, ag_grid *grid, ag_mapset *ms,
               ag_tables *tab, ag_setup *setup);
void build_vdw_tables(const ag_setup &s, const ag_mapset &ms, float *out);
bool xs_radius_lookup(const char *type, double *radius);
void build_vina_tables(const ag_setup &s, const ag_mapset &ms, const double *xs_map, const double *xs_rec,
                       float *out);
void job_prepare_vina(Job &job);
void build_sol_fn(double coeff_desolv, double *out);
void build_r_eps(double *out);
// How one atom's bond vectors were built from atom positions (kind and the
// atoms involved), so gradients can flow back through them.
struct BondRecipe {
    enum Kind : int { kNone = 0, kDiff = 1, kCarbonyl = 2, kEther = 3, kA1 = 4 };
    int kind = kNone;
    int a = -1, b = -1, c = -1;   // atoms of the construction (see bond_vectors_vjp)
    int nb = 0;                   // bonded atoms of an A1 nitrogen
};
void build_bond_vectors(const ag_setup &s, int n, const double *coord,
                        const int32_t *type, const int32_t *hbond,
                        int32_t *rexp, double *rvector, double *rvector2,
                        uint8_t *disorder, FILE *log, BondRecipe *recipes = nullptr);
void bond_vectors_vjp(int n, const double *coord, const BondRecipe *rec, const double *g_rv,
                      const double *g_rv2, double *g_coord);
// writer
int format_map_value(float v, char *out);     // returns bytes incl. '\n'
// Formats and writes every map file; wait_ready(j), when given, blocks until
// map j (n_maps: fmap, n_maps + 1: cmap) is on the host.
struct MapExtrema {
    std::vector<double> emin, emax;
};
MapExtrema write_all_maps(Job &job, const float *maps, const float *fmap,
                          const int32_t *cmap, int n_threads,
                          const std::function<void(int)> &wait_ready = nullptr);
void log_map_summary(Job &job, const MapExtrema &ext);
void log_plane_progress(Job &job, double seconds);

}  // namespace ag

// Device side (ag_device.hip).
struct ag_ctx;
struct ag_plan;
struct ag_analytic;
struct ag_scorer;
namespace agdev {
ag_status context_create(ag_ctx **out, int device);
void context_destroy(ag_ctx *ctx);
void context_timing(ag_ctx *ctx, float *ms);
// Pinned host memory on ctx's device; null when the allocation fails.
void *host_alloc_pinned(ag_ctx *ctx, size_t bytes);
void host_free_pinned(void *p);
ag_plan *plan_create(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                     const ag_tables *tab, int potential = AG_POTENTIAL_AD4);
void plan_destroy(ag_plan *p);
ag_plan *batch_create(ag_ctx *c, int n, const ag_receptor *recs, const ag_grid *grids, const ag_mapset *ms,
                      const ag_tables *tab, int potential, const ag_tune *tune);
void plan_tune(const ag_plan *p, ag_tune *out);
void plan_layout(const ag_plan *p, long long *map_off, long long *pt_off, long long *map_floats, long long *points);
int plan_count(const ag_plan *p);
void maps_batch(ag_ctx *c, int n, const ag_receptor *recs, const ag_grid *grids, const ag_mapset *ms,
                const ag_tables *tab, float *const *maps_out, float *const *fmap_out, int32_t *const *cmap_out,
                int potential, const ag_tune *tune);
void plan_execute(ag_plan *p, float *d_maps, float *d_fmap, int32_t *d_cmap);
void plan_execute_host(ag_plan *p, float *d_maps, float *d_fmap, int32_t *d_cmap, float *const *maps_out,
                       float *const *fmap_out, int32_t *const *cmap_out);
void plan_audit(ag_plan *p, long long out[4]);
void maps_plain_streamed(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                         const ag_tables *tab, float *maps_out, float *fmap, int32_t *cmap,
                         const std::function<void(int)> &on_ready, int potential = AG_POTENTIAL_AD4);
// ag_maps_plain with device outputs, every device buffer drawn from the
// device's scratch array (no allocation).
void maps_plain_noalloc(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                        const ag_tables *tab, float *maps, float *fmap, int32_t *cmap);
void maps_plain(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid,
                const ag_mapset *ms, const ag_tables *tab, float *maps,
                float *fmap, int32_t *cmap, int outputs_on_device, int potential = AG_POTENTIAL_AD4);
// analytic mode (ag_analytic.hip)
ag_analytic *analytic_create(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                             const ag_setup *setup, const double w0[4]);
void analytic_destroy(ag_analytic *a);
void analytic_set(ag_analytic *a, const double *coord, const double *charge, const double *rij, const double *eps0,
                  const double *w);
void analytic_forward(ag_analytic *a, float *d_maps);
void analytic_adjoint(ag_analytic *a, const float *d_dmaps, double *g_coord, double *g_charge, double *g_rij,
                      double *g_eps0, double g_w[4]);
void analytic_info(const ag_analytic *a, int *n_atoms, int *n_maps, int64_t *points, int *n_pairs);
// ligand scoring (ag_score.hip)
ag_scorer *scorer_create(ag_ctx *c, const ag_grid *grid, int n_maps, int elec_index, int dsolv_index, int n_poses,
                         const int32_t *pose_start, const double *xyz, const int32_t *map_index,
                         const double *charge, double outside);
void scorer_set_positions(ag_scorer *s, const double *xyz);
void scorer_score(ag_scorer *s, const float *d_maps, float *d_scores, float *d_grad);
void scorer_adjoint(ag_scorer *s, const float *d_dscore, float *d_dmaps);
void scorer_destroy(ag_scorer *s);
}  // namespace agdev
