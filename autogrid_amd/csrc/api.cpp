// C-ABI of the MI355X AutoGrid4 engine (include/autogrid_amd.h).
//
// Every entry point catches the library's exceptions and turns them into an
// ag_status plus a thread-local message; none of them calls exit().

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

#include "ag_internal.h"

struct ag_job {
    ag::Job job;
};

namespace ag {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

void job_views(const Job &job, ag_receptor *rec, ag_grid *grid, ag_mapset *ms, ag_tables *tab,
               ag_setup *setup) {
    if (rec) {
        rec->n_atoms = job.n_atoms;
        rec->coord = job.coord.data();
        rec->charge = job.charge.data();
        rec->vol = job.vol.data();
        rec->solpar = job.solpar.data();
        rec->type = job.type.data();
        rec->hbond = job.hbond.data();
        rec->rexp = job.prepared ? job.rexp.data() : nullptr;
        rec->rvector = job.prepared ? job.rvector.data() : nullptr;
        rec->rvector2 = job.prepared ? job.rvector2.data() : nullptr;
        rec->disorder = job.prepared ? job.disorder.data() : nullptr;
    }
    if (grid) {
        for (int d = 0; d < 3; d++) {
            grid->n1[d] = job.n1[d];
            grid->ne[d] = job.ne[d];
        }
        grid->spacing = job.spacing;
    }
    if (ms) {
        memset(ms, 0, sizeof *ms);
        ms->n_atom_maps = job.num_atom_maps;
        ms->maps = job.prepared ? job.map_desc.data() : nullptr;
        ms->elec_index = job.elecPE;
        ms->dsolv_index = job.dsolvPE;
        ms->n_maps = job.num_maps;
        ms->separate_desolvation = job.separate_desolvation;
        ms->disorder_h = job.disorder_h;
        ms->hydrogen_type = job.t_hydrogen;
        ms->dddiel = job.dddiel;
        ms->floating_grid = job.floating_grid;
        ms->constriction_grid = job.constriction_grid;
        ms->bhtree = job.bhtree ? (job.map_receptor_interior ? 2 : 1) : 0;
        ms->inv_diel_cal = job.inv_diel_cal;
        ms->coeff_estat = job.lib.coeff_estat;
        ms->solpar_q = job.solpar_q;
        for (int d = 0; d < 3; d++) ms->cov_pos[d] = job.cov_pos[d];
        ms->cov_half_width = job.cov_half_width;
        ms->cov_barrier = job.cov_barrier;
        ms->constriction_cutoff = job.constriction_cutoff;
    }
    if (tab) {
        memset(tab, 0, sizeof *tab);
        tab->n_rec_types = (int32_t)job.rec_types.size();
        tab->e_vdw_hb = job.prepared ? job.e_vdw_hb.data() : nullptr;
        tab->sol_fn = job.prepared ? job.sol_fn.data() : nullptr;
        tab->r_eps = job.prepared ? job.r_eps.data() : nullptr;
    }
    if (setup) {
        memset(setup, 0, sizeof *setup);
        setup->n_rec_types = (int32_t)job.rec_types.size();
        setup->n_atom_maps = job.num_atom_maps;
        setup->pairs = job.prepared ? job.pairs.data() : nullptr;
        setup->r_smooth = job.r_smooth;
        setup->coeff_desolv = job.lib.coeff_desolv;
        setup->dddiel = job.dddiel;
        setup->disorder_h = job.disorder_h;
        setup->t_hydrogen = job.t_hydrogen;
        setup->t_nonhb_hydrogen = job.t_nonhb_hydrogen;
        setup->t_carbon = job.t_carbon;
        setup->t_arom_carbon = job.t_arom_carbon;
        setup->t_oxygen = job.t_oxygen;
        setup->t_sulphur = job.t_sulphur;
        setup->t_nitrogen = job.t_nitrogen;
        setup->t_nonhb_nitrogen = job.t_nonhb_nitrogen;
    }
}

}  // namespace ag

template <typename F>
static ag_status guarded(F &&f) {
    try {
        f();
        return AG_OK;
    } catch (const ag::Error &e) {
        ag::set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        ag::set_last_error("host allocation failed");
        return AG_ERR_OOM;
    } catch (const std::exception &e) {
        ag::set_last_error(e.what());
        return AG_ERR_BAD_INPUT;
    }
}

// print_error(FATAL_ERROR) text in the log (main.cpp:133-177).
static void log_fatal(ag::Job &job, const std::string &msg) {
    char out[2048];
    snprintf(out, sizeof out, "\n%s: %s:  %s\n", job.program.c_str(), "ERROR", msg.c_str());
    if (job.log) {
        fprintf(job.log, "%s\n", out);
        fprintf(job.log, "\n%s: Unsuccessful Completion.\n", job.program.c_str());
        fflush(job.log);
    }
    fprintf(stderr, "%s\n", out);
}

extern "C" {

int ag_abi_version(void) { return AG_ABI_VERSION; }

const char *ag_last_error(void) { return ag::g_last_error.c_str(); }

ag_status ag_context_create(ag_ctx **out, int device) {
    if (!out) {
        ag::set_last_error("null output pointer");
        return AG_ERR_BAD_INPUT;
    }
    *out = nullptr;
    return guarded([&] { agdev::context_create(out, device); });
}

ag_status ag_context_destroy(ag_ctx *ctx) {
    return guarded([&] { agdev::context_destroy(ctx); });
}

ag_status ag_job_open(const char *gpf_path, const char *glg_path, const char *program_name, uint32_t flags,
                      ag_job **out) {
    if (!out) {
        ag::set_last_error("null output pointer");
        return AG_ERR_BAD_INPUT;
    }
    *out = nullptr;
    ag_job *j = new (std::nothrow) ag_job();
    if (!j) {
        ag::set_last_error("host allocation failed");
        return AG_ERR_OOM;
    }
    ag::Job &job = j->job;
    job.make_files = !(flags & AG_JOB_NO_FILES);
    job.bhtree = (flags & AG_JOB_BHTREE) != 0;
    job.binary_maps = (flags & AG_JOB_BINARY_MAPS) != 0;
    if (program_name) job.program = program_name;
    ag_status st = guarded([&] {
        if (glg_path) {
            job.log = fopen(glg_path, "w");
            if (!job.log) ag::fail(AG_ERR_IO, std::string("Sorry, I can't create the log file \"") + glg_path + "\"");
            job.own_log = true;
        }
        if (!gpf_path) ag::fail(AG_ERR_BAD_INPUT, "no grid parameter file given");
        ag::job_parse_gpf(job, gpf_path);
    });
    if (st != AG_OK) {
        log_fatal(job, ag::g_last_error);
        delete j;
        return st;
    }
    *out = j;
    return AG_OK;
}

ag_status ag_job_prepare(ag_job *j) {
    if (!j) {
        ag::set_last_error("null job");
        return AG_ERR_BAD_INPUT;
    }
    ag_status st = guarded([&] { ag::job_prepare(j->job); });
    if (st != AG_OK) log_fatal(j->job, ag::g_last_error);
    return st;
}

ag_status ag_job_views(const ag_job *j, ag_receptor *rec, ag_grid *grid, ag_mapset *ms, ag_tables *tab,
                       ag_setup *setup) {
    if (!j) {
        ag::set_last_error("null job");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { ag::job_views(j->job, rec, grid, ms, tab, setup); });
}

int64_t ag_job_points(const ag_job *j) { return j ? j->job.points() : 0; }

void ag_job_close(ag_job *j) { delete j; }

ag_status ag_maps_plain(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                        const ag_tables *tab, float *maps, float *fmap, int32_t *cmap, int outputs_on_device) {
    if (!ctx || !rec || !grid || !ms || !tab || !maps) {
        ag::set_last_error("null argument to ag_maps_plain");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { agdev::maps_plain(ctx, rec, grid, ms, tab, maps, fmap, cmap, outputs_on_device); });
}

ag_status ag_maps_plain_noalloc(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                                const ag_tables *tab, float *maps, float *fmap, int32_t *cmap) {
    if (!ctx || !rec || !grid || !ms || !tab || !maps) {
        ag::set_last_error("null argument to ag_maps_plain_noalloc");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { agdev::maps_plain_noalloc(ctx, rec, grid, ms, tab, maps, fmap, cmap); });
}

ag_status ag_plan_create(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                         const ag_tables *tab, ag_plan **out) {
    if (!ctx || !rec || !grid || !ms || !tab || !out) {
        ag::set_last_error("null argument to ag_plan_create");
        return AG_ERR_BAD_INPUT;
    }
    *out = nullptr;
    return guarded([&] { *out = agdev::plan_create(ctx, rec, grid, ms, tab); });
}

ag_status ag_plan_create_potential(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                                   const ag_tables *tab, int potential, ag_plan **out) {
    if (!ctx || !rec || !grid || !ms || !tab || !out) {
        ag::set_last_error("null argument to ag_plan_create_potential");
        return AG_ERR_BAD_INPUT;
    }
    *out = nullptr;
    return guarded([&] { *out = agdev::plan_create(ctx, rec, grid, ms, tab, potential); });
}

int ag_xs_radius(const char *type, double *radius) {
    double r = 1.8;
    const int found = type ? (int)ag::xs_radius_lookup(type, &r) : 0;
    if (radius) *radius = r;
    return found;
}

ag_status ag_vina_tables(const ag_setup *setup, const ag_mapset *ms, const double *xs_map, const double *xs_rec,
                         float *out) {
    if (!setup || !ms || !xs_map || !xs_rec || !out || !setup->pairs || (setup->n_atom_maps > 0 && !ms->maps)) {
        ag::set_last_error("null argument to ag_vina_tables");
        return AG_ERR_BAD_INPUT;
    }
    if (setup->n_atom_maps != ms->n_atom_maps) {
        ag::set_last_error("setup and map set disagree on the number of atom maps");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { ag::build_vina_tables(*setup, *ms, xs_map, xs_rec, out); });
}

ag_status ag_job_prepare_vina(ag_job *j) {
    if (!j) {
        ag::set_last_error("null job");
        return AG_ERR_BAD_INPUT;
    }
    ag_status st = guarded([&] { ag::job_prepare_vina(j->job); });
    if (st != AG_OK) log_fatal(j->job, ag::g_last_error);
    return st;
}

ag_status ag_maps_vina(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                       const ag_tables *tab, float *maps, float *fmap, int32_t *cmap, int outputs_on_device) {
    if (!ctx || !rec || !grid || !ms || !tab || !maps) {
        ag::set_last_error("null argument to ag_maps_vina");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] {
        agdev::maps_plain(ctx, rec, grid, ms, tab, maps, fmap, cmap, outputs_on_device, AG_POTENTIAL_VINA);
    });
}

ag_status ag_build_tables(const ag_setup *setup, const ag_mapset *ms, float *e_vdw_hb, double *sol_fn,
                          double *r_eps) {
    if (!setup || !ms || !e_vdw_hb || !sol_fn || (setup->dddiel && !r_eps) ||
        (setup->n_atom_maps > 0 && (!setup->pairs || !ms->maps)) || setup->n_atom_maps != ms->n_atom_maps) {
        ag::set_last_error("bad arguments to ag_build_tables");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] {
        ag::build_vdw_tables(*setup, *ms, e_vdw_hb);
        ag::build_sol_fn(setup->coeff_desolv, sol_fn);
        if (setup->dddiel) ag::build_r_eps(r_eps);
    });
}

ag_status ag_bond_vectors(const ag_setup *setup, int n, const double *coord, const int32_t *type,
                          const int32_t *hbond, int32_t *rexp, double *rvector, double *rvector2,
                          uint8_t *disorder) {
    if (!setup || n < 0 || (n > 0 && (!coord || !type || !hbond || !rexp || !rvector || !rvector2 || !disorder))) {
        ag::set_last_error("bad arguments to ag_bond_vectors");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] {
        for (int i = 0; i < n; i++) rexp[i] = 0;
        for (int i = 0; i < 3 * n; i++) rvector[i] = rvector2[i] = 0.0;
        for (int i = 0; i < n; i++) disorder[i] = 0;
        ag::build_bond_vectors(*setup, n, coord, type, hbond, rexp, rvector, rvector2, disorder, nullptr);
    });
}

ag_status ag_bond_vectors_batch(const ag_setup *setup, int n, const ag_receptor *recs, const ag_bond_out *outs,
                                int n_threads) {
    if (!setup || n < 0 || (n > 0 && (!recs || !outs))) {
        ag::set_last_error("bad arguments to ag_bond_vectors_batch");
        return AG_ERR_BAD_INPUT;
    }
    for (int i = 0; i < n; i++) {
        const ag_receptor &r = recs[i];
        const ag_bond_out &o = outs[i];
        if (r.n_atoms < 0 || (r.n_atoms > 0 && (!r.coord || !r.type || !r.hbond || !o.rexp || !o.rvector ||
                                                !o.rvector2 || !o.disorder))) {
            ag::set_last_error("bad receptor or output in ag_bond_vectors_batch");
            return AG_ERR_BAD_INPUT;
        }
    }
    return guarded([&] {
        int nt = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
        nt = std::max(1, std::min(nt, n));
        // largest conformations first, so the threads finish together
        std::vector<int> order(n);
        for (int i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return recs[a].n_atoms > recs[b].n_atoms; });
        std::atomic<int> next{0};
        std::vector<std::string> errors(nt);
        std::vector<ag_status> codes(nt, AG_OK);
        auto work = [&](int t) {
            for (;;) {
                const int k = next.fetch_add(1);
                if (k >= n) return;
                const int i = order[k];
                try {
                    ag::build_bond_vectors(*setup, recs[i].n_atoms, recs[i].coord, recs[i].type, recs[i].hbond,
                                           outs[i].rexp, outs[i].rvector, outs[i].rvector2, outs[i].disorder,
                                           nullptr);
                } catch (const ag::Error &e) {
                    codes[t] = e.status;
                    errors[t] = e.what();
                    return;
                }
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(work, t);
        work(0);
        for (auto &x : th) x.join();
        for (int t = 0; t < nt; t++)
            if (codes[t] != AG_OK) ag::fail(codes[t], errors[t]);
    });
}

ag_status ag_bond_vectors_vjp(const ag_setup *setup, int n, const double *coord, const int32_t *type,
                              const int32_t *hbond, const double *g_rv, const double *g_rv2, double *g_coord) {
    if (!setup || n < 0 || (n > 0 && (!coord || !type || !hbond || !g_rv || !g_rv2 || !g_coord))) {
        ag::set_last_error("bad arguments to ag_bond_vectors_vjp");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] {
        std::vector<int32_t> rexp(n);
        std::vector<double> rv(3 * (size_t)n), rv2(3 * (size_t)n);
        std::vector<uint8_t> dis(n);
        std::vector<ag::BondRecipe> rec(n);
        ag::build_bond_vectors(*setup, n, coord, type, hbond, rexp.data(), rv.data(), rv2.data(), dis.data(),
                               nullptr, rec.data());
        ag::bond_vectors_vjp(n, coord, rec.data(), g_rv, g_rv2, g_coord);
    });
}

ag_status ag_batch_create(ag_ctx *ctx, int n, const ag_receptor *recs, const ag_grid *grids, const ag_mapset *ms,
                          const ag_tables *tab, int potential, const ag_tune *tune, ag_plan **out) {
    if (!ctx || n < 1 || !recs || !grids || !ms || !tab || !out) {
        ag::set_last_error("bad arguments to ag_batch_create");
        return AG_ERR_BAD_INPUT;
    }
    *out = nullptr;
    return guarded([&] { *out = agdev::batch_create(ctx, n, recs, grids, ms, tab, potential, tune); });
}

ag_status ag_plan_layout(const ag_plan *plan, int64_t *map_offsets, int64_t *point_offsets, int64_t *map_floats,
                         int64_t *points) {
    if (!plan) {
        ag::set_last_error("null plan");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] {
        const int n = agdev::plan_count(plan);
        std::vector<long long> mo(n), po(n);
        long long mf = 0, pt = 0;
        agdev::plan_layout(plan, mo.data(), po.data(), &mf, &pt);
        for (int i = 0; i < n; i++) {
            if (map_offsets) map_offsets[i] = mo[i];
            if (point_offsets) point_offsets[i] = po[i];
        }
        if (map_floats) *map_floats = mf;
        if (points) *points = pt;
    });
}

int ag_plan_conformations(const ag_plan *plan) { return plan ? agdev::plan_count(plan) : 0; }

ag_status ag_plan_tune(const ag_plan *plan, ag_tune *out) {
    if (!plan || !out) {
        ag::set_last_error("null argument to ag_plan_tune");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { agdev::plan_tune(plan, out); });
}

ag_status ag_maps_batch(ag_ctx *ctx, int n, const ag_receptor *recs, const ag_grid *grids, const ag_mapset *ms,
                        const ag_tables *tab, float *const *maps_out, float *const *fmap_out,
                        int32_t *const *cmap_out, int potential, const ag_tune *tune) {
    if (!ctx || n < 1 || !recs || !grids || !ms || !tab || !maps_out) {
        ag::set_last_error("bad arguments to ag_maps_batch");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] {
        agdev::maps_batch(ctx, n, recs, grids, ms, tab, maps_out, fmap_out, cmap_out, potential, tune);
    });
}

ag_status ag_analytic_create(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                             const ag_setup *setup, const ag_weights *w, ag_analytic **out) {
    if (!ctx || !rec || !grid || !ms || !setup || !w || !out || (rec->n_atoms > 0 && (!rec->coord || !rec->charge ||
        !rec->vol || !rec->solpar || !rec->type || !rec->hbond)) || (ms->n_atom_maps > 0 && (!ms->maps || !setup->pairs))) {
        ag::set_last_error("bad arguments to ag_analytic_create");
        return AG_ERR_BAD_INPUT;
    }
    *out = nullptr;
    const double w0[4] = {w->vdw, w->hbond, w->estat, w->desolv};
    return guarded([&] { *out = agdev::analytic_create(ctx, rec, grid, ms, setup, w0); });
}

ag_status ag_analytic_set(ag_analytic *a, const double *coord, const double *charge, const double *rij,
                          const double *eps, const ag_weights *w) {
    if (!a) {
        ag::set_last_error("null analytic plan");
        return AG_ERR_BAD_INPUT;
    }
    double wv[4];
    if (w) {
        wv[0] = w->vdw;
        wv[1] = w->hbond;
        wv[2] = w->estat;
        wv[3] = w->desolv;
    }
    return guarded([&] { agdev::analytic_set(a, coord, charge, rij, eps, w ? wv : nullptr); });
}

ag_status ag_analytic_forward(ag_analytic *a, float *maps_dev) {
    if (!a || !maps_dev) {
        ag::set_last_error("null argument to ag_analytic_forward");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { agdev::analytic_forward(a, maps_dev); });
}

ag_status ag_analytic_adjoint(ag_analytic *a, const float *dmaps_dev, double *g_coord, double *g_charge,
                              double *g_rij, double *g_eps, ag_weights *g_w) {
    if (!a || !dmaps_dev) {
        ag::set_last_error("null argument to ag_analytic_adjoint");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] {
        double gw[4];
        agdev::analytic_adjoint(a, dmaps_dev, g_coord, g_charge, g_rij, g_eps, gw);
        if (g_w) {
            g_w->vdw = gw[0];
            g_w->hbond = gw[1];
            g_w->estat = gw[2];
            g_w->desolv = gw[3];
        }
    });
}

ag_status ag_analytic_destroy(ag_analytic *a) {
    return guarded([&] { agdev::analytic_destroy(a); });
}

ag_status ag_scorer_create(ag_ctx *ctx, const ag_grid *grid, int n_maps, int elec_index, int dsolv_index, int n_poses,
                           const int32_t *pose_start, const double *xyz, const int32_t *map_index,
                           const double *charge, double outside, ag_scorer **out) {
    if (!ctx || !grid || !pose_start || !out) {
        ag::set_last_error("null argument to ag_scorer_create");
        return AG_ERR_BAD_INPUT;
    }
    *out = nullptr;
    return guarded([&] {
        *out = agdev::scorer_create(ctx, grid, n_maps, elec_index, dsolv_index, n_poses, pose_start, xyz, map_index,
                                    charge, outside);
    });
}

ag_status ag_scorer_set_positions(ag_scorer *s, const double *xyz) {
    if (!s) {
        ag::set_last_error("null scorer");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { agdev::scorer_set_positions(s, xyz); });
}

ag_status ag_scorer_score(ag_scorer *s, const float *maps_dev, float *scores_dev, float *grad_dev) {
    if (!s || !maps_dev || !scores_dev) {
        ag::set_last_error("null argument to ag_scorer_score");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { agdev::scorer_score(s, maps_dev, scores_dev, grad_dev); });
}

ag_status ag_scorer_adjoint(ag_scorer *s, const float *dscore_dev, float *dmaps_dev) {
    if (!s || !dscore_dev || !dmaps_dev) {
        ag::set_last_error("null argument to ag_scorer_adjoint");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { agdev::scorer_adjoint(s, dscore_dev, dmaps_dev); });
}

ag_status ag_scorer_destroy(ag_scorer *s) {
    return guarded([&] { agdev::scorer_destroy(s); });
}

ag_status ag_job_weights(const ag_job *j, ag_weights *out) {
    if (!j || !out) {
        ag::set_last_error("null argument to ag_job_weights");
        return AG_ERR_BAD_INPUT;
    }
    out->vdw = j->job.lib.coeff_vdW;
    out->hbond = j->job.lib.coeff_hbond;
    out->estat = j->job.lib.coeff_estat;
    out->desolv = j->job.lib.coeff_desolv;
    return AG_OK;
}

ag_status ag_plan_execute(ag_plan *plan, float *maps_dev, float *fmap_dev, int32_t *cmap_dev) {
    if (!plan || !maps_dev) {
        ag::set_last_error("null argument to ag_plan_execute");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { agdev::plan_execute(plan, maps_dev, fmap_dev, cmap_dev); });
}

ag_status ag_plan_execute_host(ag_plan *plan, float *maps_dev, float *fmap_dev, int32_t *cmap_dev,
                               float *const *maps_out, float *const *fmap_out, int32_t *const *cmap_out) {
    if (!plan || !maps_dev || !maps_out) {
        ag::set_last_error("null argument to ag_plan_execute_host");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] {
        agdev::plan_execute_host(plan, maps_dev, fmap_dev, cmap_dev, maps_out, fmap_out, cmap_out);
    });
}

ag_status ag_plan_audit(ag_plan *plan, ag_audit *out) {
    if (!plan || !out) {
        ag::set_last_error("null argument to ag_plan_audit");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] {
        long long v[4];
        agdev::plan_audit(plan, v);
        out->far_pairs = v[0];
        out->far_bin_flips = v[1];
        out->near_unguarded_flips = v[2];
        out->max_flips_per_point = v[3];
    });
}

ag_status ag_plan_destroy(ag_plan *plan) {
    return guarded([&] { agdev::plan_destroy(plan); });
}

ag_status ag_context_timing(ag_ctx *ctx, float ms[4]) {
    if (!ctx || !ms) {
        ag::set_last_error("null argument to ag_context_timing");
        return AG_ERR_BAD_INPUT;
    }
    return guarded([&] { agdev::context_timing(ctx, ms); });
}

int64_t ag_format_map_values(const float *values, int64_t n, char *buf, int64_t cap) {
    int64_t used = 0;
    char tmp[64];
    for (int64_t i = 0; i < n; i++) {
        int k = ag::format_map_value(values[i], tmp);
        if (used + k > cap) return -1;
        memcpy(buf + used, tmp, (size_t)k);
        used += k;
    }
    return used;
}

ag_status ag_job_write_maps(ag_job *j, const float *maps, const float *fmap, const int32_t *cmap,
                            int n_threads) {
    if (!j || !maps) {
        ag::set_last_error("null argument to ag_job_write_maps");
        return AG_ERR_BAD_INPUT;
    }
    ag_status st = guarded([&] {
        if (!j->job.prepared) ag::fail(AG_ERR_BAD_INPUT, "ag_job_write_maps needs a prepared job");
        // the .fld and the map headers, unless already written
        ag::job_write_fld(j->job);
        ag::job_write_map_headers(j->job);
        ag::log_map_summary(j->job, ag::write_all_maps(j->job, maps, fmap, cmap, n_threads > 0 ? n_threads : 1));
        j->job.logf("\n%s: Successful Completion.\n", j->job.program.c_str());
        if (j->job.log) fflush(j->job.log);
    });
    if (st != AG_OK) log_fatal(j->job, ag::g_last_error);
    return st;
}

ag_status ag_run_gpf(ag_ctx *ctx, const char *gpf_path, const char *glg_path, const char *program_name) {
    return ag_run_gpf_flags(ctx, gpf_path, glg_path, program_name, 0);
}

ag_status ag_run_gpf_flags(ag_ctx *ctx, const char *gpf_path, const char *glg_path, const char *program_name,
                           uint32_t flags) {
    // With no context given, device 0 is opened on a thread of its own while
    // the inputs are read and the tables built.  Its outcome is looked at
    // only after every input error has had its chance to surface, so the
    // errors reported are the same as when opening it afterwards.
    ag_ctx *own = nullptr;
    ag_status dev_status = AG_OK;
    std::string dev_error;
    struct Own {
        ag_ctx *&c;
        ~Own() { if (c) agdev::context_destroy(c); }
    } own_guard{own};
    std::thread opener;
    struct Join {   // declared after own_guard: the thread ends before the context goes
        std::thread &t;
        ~Join() { if (t.joinable()) t.join(); }
    } opener_join{opener};
    if (!ctx)
        opener = std::thread([&] {
            try {
                agdev::context_create(&own, 0);
            } catch (const ag::Error &e) {
                dev_status = e.status;
                dev_error = e.what();
            } catch (const std::exception &e) {
                dev_status = AG_ERR_DEVICE;
                dev_error = e.what();
            }
        });
    ag_job *j = nullptr;
    ag_status st = ag_job_open(gpf_path, glg_path, program_name, flags & (AG_JOB_BHTREE | AG_JOB_BINARY_MAPS), &j);
    if (st != AG_OK) return st;
    ag::Job &job = j->job;
    st = guarded([&] {
        ag::job_write_fld(job);
        ag::job_prepare(job);
        job.logf("\n\nCalculating Pairwise Interaction Energies\n");
        job.logf("Beginning grid calculations.\n");
        job.logf("\nCalculating %d grid%s over %lld element%s, around %d receptor atom%s.\n\n",
                 job.num_maps + job.num_distance_maps, (job.num_maps + job.num_distance_maps) == 1 ? "" : "s",
                 (long long)job.points(), job.points() == 1 ? "" : "s", job.n_atoms, job.n_atoms == 1 ? "" : "s");
        ag::job_write_map_headers(job);
        ag_receptor rec;
        ag_grid grid;
        ag_mapset ms;
        ag_tables tab;
        ag::job_views(job, &rec, &grid, &ms, &tab, nullptr);
        if (!ctx) {
            opener.join();
            if (dev_status != AG_OK) ag::fail(dev_status, dev_error);
            ctx = own;
        }
        // The maps come back into pinned host memory (device-to-host copies
        // at full link speed, no zero fill); pageable memory if that fails.
        const size_t n_values = (size_t)job.num_maps * job.points();
        struct HostMaps {
            float *pinned = nullptr;
            std::vector<float> pageable;
            ~HostMaps() { agdev::host_free_pinned(pinned); }
        } host;
        host.pinned = (float *)agdev::host_alloc_pinned(ctx, sizeof(float) * std::max<size_t>(n_values, 1));
        if (!host.pinned) host.pageable.resize(n_values);
        float *const maps = host.pinned ? host.pinned : host.pageable.data();
        std::vector<float> fmap(job.floating_grid ? job.points() : 0);
        std::vector<int32_t> cmap(job.constriction_grid ? job.points() : 0);
        // The device computes the map set; a copy thread brings the maps to
        // the host one by one while writer threads format and write those
        // already there.
        const int nready = job.num_maps + 2;
        std::vector<char> ready(nready, 0);
        std::mutex mu;
        std::condition_variable cv;
        std::string copy_error;
        ag_status copy_status = AG_OK;
        auto t0 = std::chrono::steady_clock::now();
        std::thread copier([&] {
            auto mark = [&](int k) {
                std::lock_guard<std::mutex> lk(mu);
                ready[k] = 1;
                cv.notify_all();
            };
            try {
                agdev::maps_plain_streamed(ctx, &rec, &grid, &ms, &tab, maps,
                                           job.floating_grid ? fmap.data() : nullptr,
                                           job.constriction_grid ? cmap.data() : nullptr, mark);
            } catch (const ag::Error &e) {
                copy_status = e.status;
                copy_error = e.what();
            } catch (const std::exception &e) {
                copy_status = AG_ERR_DEVICE;
                copy_error = e.what();
            }
            std::lock_guard<std::mutex> lk(mu);
            for (auto &r : ready) r = r ? r : 2;   // 2: never arrives
            cv.notify_all();
        });
        auto wait_ready = [&](int k) {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return ready[k] != 0; });
            if (ready[k] == 2) ag::fail(copy_status != AG_OK ? copy_status : AG_ERR_DEVICE, copy_error);
        };
        unsigned hw = std::thread::hardware_concurrency();
        ag::MapExtrema ext;
        try {
            ext = ag::write_all_maps(job, maps, fmap.data(), cmap.data(), (int)std::min(16u, hw ? hw : 1u),
                                     wait_ready);
        } catch (...) {
            copier.join();
            if (copy_status != AG_OK) ag::fail(copy_status, copy_error);
            throw;
        }
        copier.join();
        if (copy_status != AG_OK) ag::fail(copy_status, copy_error);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        ag::log_plane_progress(job, secs);
        ag::log_map_summary(job, ext);
        job.logf("\n%s: Successful Completion.\n", job.program.c_str());
        if (job.log) fflush(job.log);
    });
    if (st != AG_OK) log_fatal(job, ag::g_last_error);
    ag_job_close(j);
    return st;
}

}  // extern "C"
