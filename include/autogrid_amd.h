/*
 * autogrid_amd.h -- C-ABI of the MI355X AutoGrid4 grid-map engine.
 *
 * The reference autogrid4 has no library interface: its hot path is inline in
 * main() (main.cpp:2805-3464) and its only boundary is the command line plus
 * files (setflags.cpp:96-134).  This header is the in-process seam placed at
 * the point where the reference has finished reading the GPF and building its
 * tables and bond vectors (main.cpp:2785), plus the front-end calls that take
 * a GPF to that point and the writer that turns computed maps into the
 * reference's .map files.
 *
 * Conventions
 *   - Plain C types only; every array is a caller-visible pointer plus a size.
 *   - Functions return ag_status; the message of the last failure on the
 *     calling thread is returned by ag_last_error().  Nothing calls exit().
 *   - Map layout everywhere: maps[m][iz][iy][ix], x fastest, the index of a
 *     point is ix + n1x*(iy + n1y*iz) (main.cpp:322).
 *   - Atom coordinates are in the grid frame: PDBQT coordinate minus grid
 *     centre (main.cpp:1135-1139).  Grid point (ix,iy,iz) sits at
 *     ((ix-ne_x)*spacing, (iy-ne_y)*spacing, (iz-ne_z)*spacing)
 *     (main.cpp:2815-2859).
 */
#ifndef AUTOGRID_AMD_H
#define AUTOGRID_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AG_ABI_VERSION 3

/* Table sizes of the 8 A build (constants.h:148-171, USE_8A_NBCUTOFF). */
#define AG_A_DIV 128           /* bins per Angstrom                         */
#define AG_NEINT 1024          /* vdW/H-bond table length: 8 A * 128        */
#define AG_NDIEL 8192          /* dielectric / desolvation table length     */
#define AG_MAX_REC_TYPES 20    /* NUM_RECEPTOR_TYPES, autogrid.h:95         */
#define AG_MAX_MAPS 128        /* MAX_MAPS, autocomm.h:94                   */
#define AG_MAX_ATOMS 100000    /* AG_MAX_ATOMS, autogrid.h:58               */
#define AG_MAX_GRID_PTS 1025   /* MAX_GRID_PTS, autocomm.h:87               */

typedef enum ag_status {
    AG_OK = 0,
    AG_ERR_BAD_INPUT = 1,   /* malformed GPF/PDBQT/parameters, bad sizes   */
    AG_ERR_NO_DONOR_H = 2,  /* affinity maps requested, no DS/D1 atom      */
    AG_ERR_OOM = 3,         /* host or device allocation failed            */
    AG_ERR_DEVICE = 4,      /* HIP error, no device, kernel fault          */
    AG_ERR_IO = 5,          /* file could not be opened or written         */
    AG_ERR_UNSUPPORTED = 6  /* request outside what this build implements  */
} ag_status;

/* Pair potentials of the map path: AutoDock 4 (the reference's only live
 * path) and the Vina potential of the reference's disabled
 * use_vina_potential branch (main.cpp:263, 1975-2127). */
enum { AG_POTENTIAL_AD4 = 0, AG_POTENTIAL_VINA = 1 };

/* H-bond classes of parameters.h:36-37 plus the AD code the hot loop tests. */
enum { AG_HB_NON = 0, AG_HB_DS = 1, AG_HB_D1 = 2, AG_HB_AS = 3,
       AG_HB_A1 = 4, AG_HB_A2 = 5, AG_HB_AD = 6 };

/* ------------------------------------------------------------------------ */
/* Prepared state: the in-process seam (main.cpp:2785).                      */
/* ------------------------------------------------------------------------ */

/* Receptor atoms.  Replaces the static arrays of main.cpp:359-380. */
typedef struct ag_receptor {
    int32_t        n_atoms;
    const double  *coord;     /* [n][3] grid frame                         */
    const double  *charge;    /* [n] e                                     */
    const double  *vol;       /* [n] atomic fragmental volume              */
    const double  *solpar;    /* [n] atomic solvation parameter            */
    const int32_t *type;      /* [n] receptor type index (rec_index)       */
    const int32_t *hbond;     /* [n] AG_HB_*                               */
    const int32_t *rexp;      /* [n] donor exponent (main.cpp:2360-2390)   */
    const double  *rvector;   /* [n][3] bond vector (main.cpp:2297-2664)   */
    const double  *rvector2;  /* [n][3] lone-pair plane vector             */
    const uint8_t *disorder;  /* [n] disordered hydroxyl flags             */
} ag_receptor;

/* Grid.  n1 = nelements+1 points per axis, ne = nelements/2 (main.cpp:1086). */
typedef struct ag_grid {
    int32_t n1[3];
    int32_t ne[3];
    double  spacing;
} ag_grid;

/* One ligand-type map (MapObject, main.cpp:289-315). */
typedef struct ag_map_desc {
    int32_t is_covalent;
    int32_t is_hbonder;
    int32_t hbond;            /* AG_HB_* of the probe type                 */
    int32_t reserved;
    double  solpar_probe;
    double  vol_probe;
} ag_map_desc;

/* The map set and the scalars the hot loop reads. */
typedef struct ag_mapset {
    int32_t n_atom_maps;      /* ligand maps incl. covalent and separate
                                 desolvation maps (num_atom_maps)           */
    const ag_map_desc *maps;  /* [n_atom_maps]                             */
    int32_t elec_index;       /* output slot of the e map, or -1 (elecPE)  */
    int32_t dsolv_index;      /* output slot of the d map, or -1 (dsolvPE) */
    int32_t n_maps;           /* n_atom_maps + (e) + (d)  (num_maps)       */
    int32_t separate_desolvation;
    int32_t disorder_h;
    int32_t hydrogen_type;    /* rec index of HD, -1 if none (main.cpp:2280) */
    int32_t dddiel;           /* 1: distance-dependent dielectric          */
    int32_t floating_grid;    /* 1: also produce r_min (fmap)              */
    int32_t constriction_grid;/* 1: also produce the atom count (cmap)     */
    int32_t bhtree;           /* semantics of a -DUSE_BHTREE build (main.cpp:2893-2928, 3019-3042):
                                 0: the default build's (no tree);
                                 1: Loop 2 and the cmap count select atoms by the tree's FP32
                                    distance (bhtree.cpp:341-362) and take their bins from it;
                                    a point within 1.5 A of an atom (BH_collision_dist,
                                    main.cpp:346) gets e = 0 and fmap = 1.5;
                                 2: as 1 with map_receptor_interior on: no collision test */
    double  inv_diel_cal;     /* 332/diel for a constant dielectric        */
    double  coeff_estat;
    double  solpar_q;         /* qasp, charge term of desolvation          */
    double  cov_pos[3];       /* covalent attachment point, grid frame     */
    double  cov_half_width;
    double  cov_barrier;
    double  constriction_cutoff;
} ag_mapset;

/* Lookup tables (EnergyTables, structs.h:257-268). */
typedef struct ag_tables {
    int32_t       n_rec_types;
    int32_t       reserved;
    const float  *e_vdw_hb;   /* [n_rec_types][n_atom_maps][AG_NEINT]      */
    const double *sol_fn;     /* [AG_NDIEL] desolvation Gaussian           */
    const double *r_eps;      /* [AG_NDIEL] 332/eps(r) (dddiel only)       */
} ag_tables;

/* One (map, receptor type) Lennard-Jones pair after GPF processing
 * (main.cpp:1315-1357, nbp_r_eps / nbp_coeffs main.cpp:1741-1806).          */
typedef struct ag_pair {
    double  cA, cB;           /* both 0: derive from Rij, eps              */
    double  Rij, eps;         /* eps already weighted                      */
    int32_t xA, xB;
    int32_t hbonder;
    int32_t reserved;
} ag_pair;

/* What the table and bond-vector builders read (the state at main.cpp:1819). */
typedef struct ag_setup {
    int32_t        n_rec_types;
    int32_t        n_atom_maps;
    const ag_pair *pairs;     /* [n_atom_maps][n_rec_types]                */
    double         r_smooth;
    double         coeff_desolv;
    int32_t        dddiel;
    int32_t        disorder_h;
    /* receptor type indices used by the bond-vector pass, -1 if absent
       (main.cpp:2280-2288) */
    int32_t t_hydrogen, t_nonhb_hydrogen, t_carbon, t_arom_carbon;
    int32_t t_oxygen, t_sulphur, t_nitrogen, t_nonhb_nitrogen;
} ag_setup;

/* ------------------------------------------------------------------------ */
/* Context: one GPU, its compute, second compute and copy streams, and     */
/* device scratch reused across calls.                                      */
/* ------------------------------------------------------------------------ */
typedef struct ag_ctx ag_ctx;

/* Create a context on HIP device `device`.  Fails with AG_ERR_DEVICE when no
 * device is usable; there is no CPU fallback. */
ag_status ag_context_create(ag_ctx **out, int device);
ag_status ag_context_destroy(ag_ctx *ctx);
const char *ag_last_error(void);
int ag_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Front end: GPF + PDBQT + parameter library -> job.                       */
/* ------------------------------------------------------------------------ */
typedef struct ag_job ag_job;

/* Flags of ag_job_open. */
#define AG_JOB_NO_FILES 1u    /* parse only: open no map/.fld/.xyz files   */
#define AG_JOB_BHTREE   2u    /* the semantics of a -DUSE_BHTREE build: the
                                 ag_mapset bhtree mode, map_receptor_interior
                                 off accepted (main.cpp:1643-1658), and the
                                 log's USE_BHTREE line reading yes          */
#define AG_JOB_BINARY_MAPS 4u /* also write each map as raw FP32 (<map>.bin:
                                 little-endian, x fastest, no header; counts
                                 as float) and <name>.maps.bin.fld, an AVS
                                 field over them (filetype=binary); the
                                 reference writes ASCII only (main.cpp:1917-1920,
                                 3481-3497)                                 */

/* Read a GPF with the keyword semantics of main.cpp:778-1819 (order
 * dependence included), read the receptor PDBQT (main.cpp:823-1047) and any
 * parameter_file, and, unless AG_JOB_NO_FILES, create the output files at
 * the points the reference does.  `glg_path` may be NULL (log discarded);
 * `program_name` is what the log and .fld print as the program (argv[0]). */
ag_status ag_job_open(const char *gpf_path, const char *glg_path,
                      const char *program_name, uint32_t flags, ag_job **out);
/* Build e_vdW_Hb / sol_fn / r_eps (main.cpp:1607-1639, 2040-2276) and the
 * bond vectors and disorder flags (main.cpp:2297-2736) on the host. */
ag_status ag_job_prepare(ag_job *job);
ag_status ag_job_views(const ag_job *job, ag_receptor *rec, ag_grid *grid,
                       ag_mapset *mapset, ag_tables *tables, ag_setup *setup);
/* Points per map (n1x*n1y*n1z). */
int64_t ag_job_points(const ag_job *job);
void ag_job_close(ag_job *job);

/* Whole drop-in: open + prepare + compute on ctx + write every output file +
 * log "Successful Completion" (main.cpp:3565).  On failure the log ends with
 * the reference's "Unsuccessful Completion" line (main.cpp:173-176).  With
 * ctx NULL a context on device 0 is created while the inputs are read; its
 * failure is reported only after the inputs have been checked.  An
 * empty gpf_path reads the GPF from standard input (setflags.cpp:78). */
ag_status ag_run_gpf(ag_ctx *ctx, const char *gpf_path, const char *glg_path,
                     const char *program_name);
/* ag_run_gpf with ag_job_open flags (AG_JOB_BHTREE, AG_JOB_BINARY_MAPS). */
ag_status ag_run_gpf_flags(ag_ctx *ctx, const char *gpf_path, const char *glg_path,
                           const char *program_name, uint32_t flags);

/* ------------------------------------------------------------------------ */
/* Hot path.                                                                 */
/* ------------------------------------------------------------------------ */

/* Compute every map of the set for one receptor on one grid
 * (main.cpp:2805-3431): covalent maps, the electrostatic map (all atoms, no
 * cutoff), closest donor-H, directional H-bond terms, vdW/H-bond table
 * lookups within 8 A, desolvation, the d map and the H-bond min/max fix-up.
 *
 * maps  : [n_maps][P] float; slot order as in the GPF (atom maps, then e and
 *         d at elec_index / dsolv_index).
 * fmap  : [P] float, r_min of the floating grid, or NULL.
 * cmap  : [P] int32, atoms within constriction_cutoff (<= 8 A), or NULL.
 * fmap / cmap must be given exactly when the map set asks for them.
 * Output pointers are device pointers when `outputs_on_device` is nonzero,
 * else host pointers (the result is copied back before return).
 * Inputs are host pointers; they are staged to the device on ctx's stream. */
ag_status ag_maps_plain(ag_ctx *ctx, const ag_receptor *rec,
                        const ag_grid *grid, const ag_mapset *mapset,
                        const ag_tables *tables, float *maps, float *fmap,
                        int32_t *cmap, int outputs_on_device);

/* Allocation-free map set.  No counterpart in the reference: a variant of
 * ag_maps_plain with device outputs (outputs_on_device = 1) for callers
 * that cannot allocate device or pinned memory -- serving loops and
 * harnesses that forbid hipMalloc.  Every device buffer of the map set (the
 * receptor arrays, cell list, tables, far-kernel atoms and status flag)
 * comes from one device array of AG_NOALLOC_SCRATCH_BYTES that each GPU
 * holds from the library's load until the process ends; the call makes no
 * hipMalloc, hipFree or pinned-memory allocation.  A map set that needs
 * more than the array holds returns AG_ERR_UNSUPPORTED with the limit in
 * ag_last_error() and leaves the outputs untouched.  Inputs are host
 * pointers as for ag_maps_plain; maps_dev / fmap_dev / cmap_dev are device
 * pointers.  The call returns after the outputs are written.  All calls on
 * one GPU share its array, so calls that overlap on one GPU (from several
 * threads or contexts) race; the array keeps the last call's contents. */
#define AG_NOALLOC_SCRATCH_BYTES (32u << 20)
ag_status ag_maps_plain_noalloc(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid,
                                const ag_mapset *mapset, const ag_tables *tables, float *maps_dev,
                                float *fmap_dev, int32_t *cmap_dev);

/* Device-resident form of the same computation: ag_plan_create uploads the
 * receptor, tables and map set to ctx's device once; each ag_plan_execute
 * then builds the cell list and runs the near-field and far-field kernels
 * into maps_dev, a device buffer of [n_maps][P] floats, and -- when the map
 * set asks for them -- fmap_dev ([P] float) and cmap_dev ([P] int32), else
 * NULL.  A plan belongs to its context and must be destroyed before it.
 * The same calls execute and destroy batches (ag_batch_create), whose
 * buffers hold every conformation (ag_plan_layout). */
typedef struct ag_plan ag_plan;
ag_status ag_plan_create(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid,
                         const ag_mapset *mapset, const ag_tables *tables, ag_plan **out);
ag_status ag_plan_execute(ag_plan *plan, float *maps_dev, float *fmap_dev, int32_t *cmap_dev);
/* ag_plan_execute, then conformation i's outputs copied into host memory:
 * maps_out[i] ([n_maps][P_i] floats) and, when the map set asks for them,
 * fmap_out[i] / cmap_out[i] ([P_i]).  The device buffers are the work space
 * and hold the same values afterwards.  Every map but e is final once the
 * near field is done, so those copies run on a copy stream while the
 * far-field kernel computes (behind an event recorded after a batch's
 * near-field graph); a batch's far field runs in two graphs, and the e
 * maps of the first follow it, the rest the end of the map set.  With pinned host outputs
 * (hipHostMalloc) the copies overlap the kernels.  Returns after the
 * outputs are written.  It replaces the reference's hand-over of the
 * finished gridmap[j].energy arrays to its map writer (main.cpp:3474-3497),
 * for a plan or for every conformation of a batch. */
ag_status ag_plan_execute_host(ag_plan *plan, float *maps_dev, float *fmap_dev, int32_t *cmap_dev,
                               float *const *maps_out, float *const *fmap_out, int32_t *const *cmap_out);
ag_status ag_plan_destroy(ag_plan *plan);

/* Table-bin audit of a plan (a diagnostic, outside the map computation):
 * compares, for every (point, atom) pair, the bin floor(128 r) the device
 * path uses with the one the reference's FP64 distance gives.
 *   far_pairs            pairs handled by the far-field kernel (r > 8 A)
 *   far_bin_flips        of those, pairs whose FP32 bin differs from FP64
 *   near_unguarded_flips near pairs whose FP32 bin differs from FP64 outside
 *                        the FP64 recompute band (0 unless the band is too
 *                        narrow)
 *   max_flips_per_point  largest number of far flips at one grid point */
typedef struct ag_audit {
    int64_t far_pairs;
    int64_t far_bin_flips;
    int64_t near_unguarded_flips;
    int64_t max_flips_per_point;
} ag_audit;
ag_status ag_plan_audit(ag_plan *plan, ag_audit *out);

/* Milliseconds of the last execute on ctx, from HIP events recorded on ctx's
 * stream: [0] cell list, [1] near-field kernel(s), [2] far-field kernel,
 * [3] total. */
ag_status ag_context_timing(ag_ctx *ctx, float ms[4]);

/* ------------------------------------------------------------------------ */
/* Vina potential.                                                           */
/* ------------------------------------------------------------------------ */
/* The reference builds Vina tables only in a branch that a const flag keeps
 * off (main.cpp:263 `use_vina_potential = FALSE`; the GPF keyword is fatal,
 * main.cpp:1466-1470), with the per-point loop's electrostatic, desolvation
 * and d-map terms gated off (main.cpp:2968, 3379, 3391) and the bond-vector
 * pass skipped (main.cpp:2295-2662).  This build defines the Vina map set
 * as: the tables of that branch, summed over the atoms within 8 A of each
 * point at bin min(floor(128 r), 1023) with no directional factor (the
 * bond vectors the directional terms need are never built in that mode);
 * covalent maps, fmap and cmap as in the AD4 path; e and d maps all zero. */

/* xs_vdw radius of a type (read_parameter_library.cc:320-358, the ADVINA
 * table).  Returns 1 when the type is in the table, 0 when the 1.8 A
 * default applies; *radius receives the value either way. */
int ag_xs_radius(const char *type, double *radius);

/* Vina tables [n_rec_types][n_atom_maps][AG_NEINT] float (main.cpp:2075-
 * 2127): for bin k = 1..1023 and rdd = k/128 - (xs_rec + xs_map),
 *   -0.035579 exp(-(rdd/0.5)^2) - 0.005156 exp(-((rdd-3)/2)^2),
 *   + 0.840245 rdd^2 for rdd < 0,
 *   H-bond pairs (pairs[m][t].hbonder > 0): + -0.587439 for rdd <= 0.7,
 *     then - (rdd/0.7)(-0.587439) for -0.7 < rdd <= 0,
 *   other pairs (also pairs[t][m].hbonder == 0 where that pair exists):
 *     + -0.035069 for rdd < 0.5, (0.5 - rdd)(-0.035069) for rdd < 1.5;
 * each term added to the float entry in that order; bin 0 is 0; covalent
 * maps' rows are 0.  xs_map [n_atom_maps], xs_rec [n_rec_types]. */
ag_status ag_vina_tables(const ag_setup *setup, const ag_mapset *mapset, const double *xs_map,
                         const double *xs_rec, float *out);

/* ag_job_prepare, then replace the job's e_vdW_Hb tables by the Vina tables
 * of its map and receptor types; ag_job_views then returns those tables. */
ag_status ag_job_prepare_vina(ag_job *job);

/* Vina map set for one receptor on one grid; arguments as ag_maps_plain,
 * `tables->e_vdw_hb` holding Vina tables. */
ag_status ag_maps_vina(ag_ctx *ctx, const ag_receptor *rec,
                       const ag_grid *grid, const ag_mapset *mapset,
                       const ag_tables *tables, float *maps, float *fmap,
                       int32_t *cmap, int outputs_on_device);

/* Device-resident plan of either potential (ag_plan_create is the AD4
 * one). */
ag_status ag_plan_create_potential(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid,
                                   const ag_mapset *mapset, const ag_tables *tables, int potential,
                                   ag_plan **out);

/* ------------------------------------------------------------------------ */
/* Host precompute (the reference's main.cpp:1607-1639, 2040-2736).         */
/* ------------------------------------------------------------------------ */
/* e_vdW_Hb [n_rec_types][n_atom_maps][AG_NEINT] float (main.cpp:2040-2266),
 * sol_fn [AG_NDIEL] (main.cpp:2272-2276) and, when setup->dddiel,
 * r_eps [AG_NDIEL] = 332/eps(r) (main.cpp:1607-1639); r_eps may be NULL
 * otherwise. */
ag_status ag_build_tables(const ag_setup *setup, const ag_mapset *mapset, float *e_vdw_hb,
                          double *sol_fn, double *r_eps);
/* rexp, rvector, rvector2 and disorder flags of one receptor
 * (main.cpp:2297-2736); arrays as in ag_receptor. */
ag_status ag_bond_vectors(const ag_setup *setup, int n_atoms, const double *coord,
                          const int32_t *type, const int32_t *hbond, int32_t *rexp,
                          double *rvector, double *rvector2, uint8_t *disorder);

/* Bond vectors of n receptors on n_threads host threads (<= 0: all cores);
 * receptor i reads recs[i].n_atoms / coord / type / hbond and writes
 * outs[i]. */
typedef struct ag_bond_out {
    int32_t *rexp;
    double  *rvector;
    double  *rvector2;
    uint8_t *disorder;
} ag_bond_out;
ag_status ag_bond_vectors_batch(const ag_setup *setup, int n, const ag_receptor *recs,
                                const ag_bond_out *outs, int n_threads);

/* Gradient through the bond vectors: adds to g_coord [n][3] the gradient
 * of sum(g_rvector . rvector + g_rvector2 . rvector2) with respect to the
 * coordinates, the bonding found at coord held fixed. */
ag_status ag_bond_vectors_vjp(const ag_setup *setup, int n_atoms, const double *coord,
                              const int32_t *type, const int32_t *hbond,
                              const double *g_rvector, const double *g_rvector2,
                              double *g_coord);

/* ------------------------------------------------------------------------ */
/* Batches: many receptor conformations in one map set.                     */
/* ------------------------------------------------------------------------ */
/* The reference computes one receptor per process run; a batch computes n
 * receptors (conformations: their own atoms, coordinates and box, sharing
 * the map set and tables) in one replay of captured hipGraphs: the cell
 * list and one near-field launch over the bricks of every conformation,
 * then a persistent far-field kernel, in two parts, whose workgroups pull
 * the tiles of their conformations from a work counter.  Outputs are
 * concatenated: conformation i's maps start at map_offsets[i] floats
 * ([n_maps][P_i]), its fmap / cmap values at point_offsets[i]. */
typedef struct ag_tune {
    double  cell_size;        /* cell-list edge in A, 2..8                  */
    int32_t cube;             /* 1: near-kernel waves on 4x4x4 point cubes,
                                 0: on 8x8x1 slabs                          */
    int32_t reserved;
    double  tune_ms;          /* out (ag_plan_tune): autotuner time         */
} ag_tune;

/* tune NULL: the autotuner times the candidate layouts on a sample of the
 * batch (up to 8 conformations spread over its sizes) and keeps the
 * fastest.  Results are bitwise repeatable for a fixed layout. */
ag_status ag_batch_create(ag_ctx *ctx, int n, const ag_receptor *recs, const ag_grid *grids,
                          const ag_mapset *mapset, const ag_tables *tables, int potential,
                          const ag_tune *tune, ag_plan **out);
/* Output layout of a plan or batch: offsets may be NULL. */
ag_status ag_plan_layout(const ag_plan *plan, int64_t *map_offsets, int64_t *point_offsets,
                         int64_t *map_floats, int64_t *points);
int ag_plan_conformations(const ag_plan *plan);
ag_status ag_plan_tune(const ag_plan *plan, ag_tune *out);
/* Batch with host outputs: create, execute once, copy conformation i's maps
 * to maps_out[i] ([n_maps][P_i] floats; fmap_out[i] / cmap_out[i] when the
 * map set asks for them, else those arrays may be NULL), destroy.  When all
 * outputs are pinned host memory (hipHostMalloc), the conformations run in
 * chunks launched directly on two streams, and each chunk's copies overlap
 * the later chunks' kernels. */
ag_status ag_maps_batch(ag_ctx *ctx, int n, const ag_receptor *recs, const ag_grid *grids,
                        const ag_mapset *mapset, const ag_tables *tables,
                        float *const *maps_out, float *const *fmap_out,
                        int32_t *const *cmap_out, int potential, const ag_tune *tune);

/* ------------------------------------------------------------------------ */
/* Analytic (differentiable) mode.                                           */
/* ------------------------------------------------------------------------ */
/* The map set evaluated from the functional forms behind the reference's
 * tables (main.cpp:2040-2276) at the exact distance, with the reference's
 * per-point assembly (main.cpp:2805-3431):
 *   E_mt(r) = min(EINTCLAMP, Phi(s(r))), Phi = eps/(a-b) (b (R/s)^a - a (R/s)^b)
 *             (or cA/s^a - cB/s^b for nbp_coeffs pairs), s(r) the continuous
 *             form of the min-window smoothing: r + w left of the minimum r*,
 *             r* within w of it, r - w right of it, w = floor(64 r_smooth)/128;
 *   sol(r)  = coeff_desolv exp(-r^2 / 25.92);
 *   e       = sum over all atoms q coeff_estat (332/eps(r)) / max(r, 0.5);
 * every 8 A term multiplied by S(r) = 1 - u^3 (10 - 15u + 6u^2),
 * u = clamp((r - 7.5)/0.5, 0, 1), instead of the hard cutoff.  Bond vectors
 * are rebuilt from the coordinates (main.cpp:2297-2664) and differentiated
 * with the bonding held fixed.  disorder_h is not part of this mode.
 *
 * Parameters: coordinates [N][3], charges [N], per (map, receptor type)
 * Rij and unweighted eps [n_atom_maps][n_rec_types], and the four AD4
 * weights; the effective eps of a pair is the vdW (12-6) or H-bond (12-10)
 * weight times its unweighted eps.  The adjoint gathers, for each atom, over
 * the grid points within 8 A (and over all points for the e map) with
 * fixed-order sums: no atomics, bitwise repeatable. */
typedef struct ag_weights {
    double vdw, hbond, estat, desolv;   /* coeff_vdW, coeff_hbond, coeff_estat, coeff_desolv */
} ag_weights;
typedef struct ag_analytic ag_analytic;

/* `weights` are the coefficients the setup's pair eps were weighted with
 * (the parameter library's, ag_job_weights); rec's bond-vector arrays are
 * not read. */
ag_status ag_analytic_create(ag_ctx *ctx, const ag_receptor *rec, const ag_grid *grid,
                             const ag_mapset *mapset, const ag_setup *setup,
                             const ag_weights *weights, ag_analytic **out);
/* Replace any of the parameters (NULL keeps the current value). */
ag_status ag_analytic_set(ag_analytic *a, const double *coord, const double *charge,
                          const double *rij, const double *eps, const ag_weights *weights);
/* Maps into maps_dev ([n_maps][P] float, device). */
ag_status ag_analytic_forward(ag_analytic *a, float *maps_dev);
/* Given dL/dmaps ([n_maps][P] float, device) at the last forward: dL/dcoord
 * [N][3], dL/dcharge [N], dL/dRij and dL/deps (unweighted)
 * [n_atom_maps][n_rec_types] (0 for nbp_coeffs pairs) and dL/dweights.  Any
 * output may be NULL. */
ag_status ag_analytic_adjoint(ag_analytic *a, const float *dmaps_dev, double *g_coord,
                              double *g_charge, double *g_rij, double *g_eps,
                              ag_weights *g_weights);
ag_status ag_analytic_destroy(ag_analytic *a);
/* The parameter library's AD4 weights of a job (read_parameter_library.cc
 * FE_coeff_* lines). */
ag_status ag_job_weights(const ag_job *job, ag_weights *out);

/* Ligand scoring, the consumer side of the maps.  AutoDock reads AutoGrid
 * maps by trilinear interpolation: the maps are [n_maps][n1z][n1y][n1x]
 * floats, x fastest (main.cpp:322), on the grid the .fld file describes
 * (main.cpp:1903-1956), point (i, j, k) at ((i - ne_x) s, (j - ne_y) s,
 * (k - ne_z) s) in the grid frame of ag_receptor's coordinates.  A ligand
 * atom of map index t (its type's affinity map, -1 for none) and charge q
 * at x scores
 *     E(x) = sum_c w_c(x) (M_t[c] + q M_e[c] + |q| M_d[c])
 * over the 8 points c of its cell, w_c the trilinear weights; an atom whose
 * cell is not on the grid scores `outside` with zero gradient.  A pose
 * scores the sum of its atoms.  Pose p holds atoms [pose_start[p],
 * pose_start[p+1]) of xyz [n][3] (grid frame, A), map_index [n], charge [n]
 * (host arrays).  The reference has no such consumer; this is the step
 * downstream of its output. */
typedef struct ag_scorer ag_scorer;
ag_status ag_scorer_create(ag_ctx *ctx, const ag_grid *grid, int n_maps, int elec_index,
                           int dsolv_index, int n_poses, const int32_t *pose_start,
                           const double *xyz, const int32_t *map_index,
                           const double *charge, double outside, ag_scorer **out);
/* New positions for the same atoms. */
ag_status ag_scorer_set_positions(ag_scorer *s, const double *xyz);
/* scores_dev [n_poses] float (device); grad_dev [n][3] float (device,
 * dE/dx of each atom) or NULL.  Returns when the results are written. */
ag_status ag_scorer_score(ag_scorer *s, const float *maps_dev, float *scores_dev,
                          float *grad_dev);
/* dmaps_dev [n_maps][P] float (device) := S^T dscore_dev ([n_poses] float,
 * device), S the scores' matrix (d scores / d maps): a fixed-order sum per
 * map entry, bitwise repeatable. */
ag_status ag_scorer_adjoint(ag_scorer *s, const float *dscore_dev, float *dmaps_dev);
ag_status ag_scorer_destroy(ag_scorer *s);

/* Map output: the value text of main.cpp:3486-3497 -- "0" when |e| < 1e-4,
 * else "%.3f" of (float)(floor(e*1000+0.5)/1000) -- for n values into buf.
 * Returns the number of bytes written (each value ends in '\n'), or -1 if
 * cap is too small.  Needs at most 16*n bytes for finite |e| < 1e12. */
int64_t ag_format_map_values(const float *values, int64_t n, char *buf,
                             int64_t cap);

/* Write a prepared job's outputs (opened with files): the .fld
 * (main.cpp:1903-1956) and the map headers (main.cpp:2752-2776) unless
 * already written, then the per-job maps from a computed [n_maps][P] host
 * array into the .map files (+ fmap/cmap; with AG_JOB_BINARY_MAPS the raw
 * FP32 side-output too), then the min/max summary and the completion line
 * of the log.  Formatting runs on `n_threads` host threads. */
ag_status ag_job_write_maps(ag_job *job, const float *maps, const float *fmap,
                            const int32_t *cmap, int n_threads);

#ifdef __cplusplus
}
#endif
#endif /* AUTOGRID_AMD_H */
