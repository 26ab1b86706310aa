// GPF / PDBQT / parameter-library front end of the MI355X AutoGrid4 engine.
//
// Follows the keyword semantics of the reference's GPF loop
// (main.cpp:778-1819, gpfparser.cpp:54-173), including its order dependence,
// its FATAL conditions and the points at which it creates output files.
// FATAL conditions become ag::Error(AG_ERR_BAD_INPUT / AG_ERR_IO) instead of
// exit(1); the caller logs them the way print_error() does (main.cpp:133-177).

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <strings.h>

#include "ag_internal.h"

namespace ag {

void fail(ag_status s, const std::string &msg) { throw Error(s, msg); }

Job::~Job() {
    for (auto &m : maps)
        if (m.fp) fclose(m.fp);
    if (fld_fp) fclose(fld_fp);
    if (xyz_fp) fclose(xyz_fp);
    if (fmap_fp) fclose(fmap_fp);
    if (cmap_fp) fclose(cmap_fp);
    if (log && own_log) fclose(log);
}

void Job::logf(const char *fmt, ...) {
    if (!log) return;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(log, fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------------------
// Parameter library
// ---------------------------------------------------------------------------

ParamLibrary::ParamLibrary() : slot(65536, -1) {}

// Hash of atom_parameter_manager.cpp:37-45: first two characters only.
unsigned ParamLibrary::key(const char *t) {
    size_t n = strlen(t);
    if (n == 0) return 0;
    if (n == 1) return (unsigned char)t[0];
    return (unsigned char)t[0] + 256u * (unsigned char)t[1];
}

ParamEntry *ParamLibrary::find(const char *t) {
    int s = slot[key(t)];
    return s < 0 ? nullptr : &entries[s];
}

// apm_enter replaces the whole entry (rec_index/map_index included).
void ParamLibrary::enter(const ParamEntry &e) {
    unsigned k = key(e.type);
    if (slot[k] < 0) {
        slot[k] = (int)entries.size();
        entries.push_back(e);
    } else {
        entries[slot[k]] = e;
    }
}

// parse_param_line.cc:58-113: the keyword is the first blank-terminated word,
// lower-cased, and matches a table entry when it is a case-insensitive prefix
// of it (strncasecmp over the word's length).
enum { PAR_NONE = -1, PAR_NULL, PAR_COMMENT, PAR_VDW, PAR_HBOND, PAR_ESTAT,
       PAR_DESOLV, PAR_TORS, PAR_UNBOUND, PAR_ATOM_PAR };

static int parse_param_keyword(const char *line) {
    char c[512];
    int j = 0;
    for (; line[j] && line[j] != ' ' && line[j] != '\t' && line[j] != '\n' && j < 511; j++)
        c[j] = (char)tolower((unsigned char)line[j]);
    c[j] = '\0';
    if (c[0] == '\n' || c[0] == '\0') return PAR_NULL;
    if (c[0] == '#') return PAR_COMMENT;
    static const struct { const char *lex; int tok; } table[] = {
        {"FE_coeff_vdW", PAR_VDW},     {"FE_coeff_hbond", PAR_HBOND},
        {"FE_coeff_estat", PAR_ESTAT}, {"FE_coeff_desolv", PAR_DESOLV},
        {"FE_coeff_tors", PAR_TORS},   {"FE_unbound_model", PAR_UNBOUND},
        {"atom_par", PAR_ATOM_PAR}};
    for (auto &t : table)
        if (strncasecmp(t.lex, c, (size_t)j) == 0) return t.tok;
    return PAR_NONE;
}

void ParamLibrary::apply_line(const char *line, ParamEntry &p) {
    int tok = parse_param_keyword(line);
    double *coeff = nullptr;
    const char *what = "";
    switch (tok) {
        case PAR_VDW: coeff = &coeff_vdW; what = "van der Waals term"; break;
        case PAR_HBOND: coeff = &coeff_hbond; what = "H-bonding term"; break;
        case PAR_ESTAT: coeff = &coeff_estat; what = "electrostatic term"; break;
        case PAR_DESOLV: coeff = &coeff_desolv; what = "desolvation term"; break;
        case PAR_TORS: coeff = &coeff_tors; what = "torsional term"; break;
        case PAR_UNBOUND:
            fail(AG_ERR_BAD_INPUT,
                 "the unbound model cannot be specified in the parameter library file.\n"
                 "Use the DPF parameter 'unbound_model' instead.");
        case PAR_ATOM_PAR: {
            // read_parameter_library.cc:126-164; fields absent from the line
            // keep the previous entry's values, as sscanf leaves them.
            int hb = p.hbond;
            char tname[64];
            memcpy(tname, p.type, sizeof p.type);
            tname[sizeof p.type] = '\0';
            int nf = sscanf(line, "%*s %63s %lf %lf %lf %lf %lf %lf %d %d %d %d", tname,
                            &p.Rij, &p.epsij_unweighted, &p.vol, &p.solpar, &p.Rij_hb,
                            &p.epsij_hb_unweighted, &hb, &p.rec_index, &p.map_index,
                            &p.bond_index);
            if (nf < 2) return;
            memset(p.type, 0, sizeof p.type);
            memcpy(p.type, tname, std::min(strlen(tname), sizeof p.type - 1));
            p.hbond = (hb >= 0 && hb <= 5) ? hb : 0;
            enter(p);
            return;
        }
        default: return;
    }
    if (sscanf(line, "%*s %lf", coeff) < 1)
        fail(AG_ERR_BAD_INPUT,
             std::string("Must supply a ") + what + " coefficient as a floating point number.");
}

// The AutoDock 4.1 "bound" parameter set that autogrid4 loads before reading
// the GPF (read_parameter_library.cc:202-383, data ad4_shared/AD4.1_bound.dat:46-122).
void ParamLibrary::load_builtin_ad41_bound() {
    static const char *const lines[] = {
        "FE_coeff_vdW    0.1662", "FE_coeff_hbond  0.1209", "FE_coeff_estat  0.1406",
        "FE_coeff_desolv 0.1322", "FE_coeff_tors   0.2983",
        "atom_par H  2.00 0.020  0.0000  0.00051 0.0 0.0 0 -1 -1 3",
        "atom_par HD 2.00 0.020  0.0000  0.00051 0.0 0.0 2 -1 -1 3",
        "atom_par HS 2.00 0.020  0.0000  0.00051 0.0 0.0 1 -1 -1 3",
        "atom_par C  4.00 0.150 33.5103 -0.00143 0.0 0.0 0 -1 -1 0",
        "atom_par A  4.00 0.150 33.5103 -0.00052 0.0 0.0 0 -1 -1 0",
        "atom_par N  3.50 0.160 22.4493 -0.00162 0.0 0.0 0 -1 -1 1",
        "atom_par NA 3.50 0.160 22.4493 -0.00162 1.9 5.0 4 -1 -1 1",
        "atom_par NS 3.50 0.160 22.4493 -0.00162 1.9 5.0 3 -1 -1 1",
        "atom_par OA 3.20 0.200 17.1573 -0.00251 1.9 5.0 5 -1 -1 2",
        "atom_par OS 3.20 0.200 17.1573 -0.00251 1.9 5.0 3 -1 -1 2",
        "atom_par F  3.09 0.080 15.4480 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par Mg 1.30 0.875  1.5600 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par MG 1.30 0.875  1.5600 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par P  4.20 0.200 38.7924 -0.00110 0.0 0.0 0 -1 -1 5",
        "atom_par SA 4.00 0.200 33.5103 -0.00214 2.5 1.0 5 -1 -1 6",
        "atom_par S  4.00 0.200 33.5103 -0.00214 0.0 0.0 0 -1 -1 6",
        "atom_par Cl 4.09 0.276 35.8235 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par CL 4.09 0.276 35.8235 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par Ca 1.98 0.550  2.7700 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par CA 1.98 0.550  2.7700 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par Mn 1.30 0.875  2.1400 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par MN 1.30 0.875  2.1400 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par Fe 1.30 0.010  1.8400 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par FE 1.30 0.010  1.8400 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par Zn 1.48 0.550  1.7000 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par ZN 1.48 0.550  1.7000 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par Br 4.33 0.389 42.5661 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par BR 4.33 0.389 42.5661 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par I  4.72 0.550 55.0585 -0.00110 0.0 0.0 0 -1 -1 4",
        "atom_par Z  4.00 0.150 33.5103 -0.00143 0.0 0.0 0 -1 -1 0",
        "atom_par G  4.00 0.150 33.5103 -0.00143 0.0 0.0 0 -1 -1 0",
        "atom_par GA 4.00 0.150 33.5103 -0.00052 0.0 0.0 0 -1 -1 0",
        "atom_par J  4.00 0.150 33.5103 -0.00143 0.0 0.0 0 -1 -1 0",
        "atom_par Q  4.00 0.150 33.5103 -0.00143 0.0 0.0 0 -1 -1 0",
        "atom_par Si 4.10 0.200 35.8235 -0.00143 0.0 0.0 0 -1 -1 6",
        "atom_par B  3.84 0.155 29.6478 -0.00152 0.0 0.0 0 -1 -1 0",
    };
    ParamEntry scratch;
    for (const char *l : lines) apply_line(l, scratch);
}

void ParamLibrary::load_file(const std::string &path) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) fail(AG_ERR_IO, "Sorry, I can't find or open " + path);
    char line[256];
    ParamEntry scratch;
    try {
        while (fgets(line, sizeof line, f)) apply_line(line, scratch);
    } catch (...) {
        fclose(f);
        throw;
    }
    fclose(f);
}

// ---------------------------------------------------------------------------
// GPF
// ---------------------------------------------------------------------------

enum {
    GPF_UNKNOWN = -1, GPF_NULL = 0, GPF_COMMENT, GPF_RECEPTOR, GPF_GRIDFLD, GPF_NPTS,
    GPF_SPACING, GPF_GRIDCENTER, GPF_LIGAND_TYPES, GPF_MAP, GPF_NBP_COEFFS,
    GPF_NBP_R_EPS, GPF_ELECMAP, GPF_DIEL, GPF_FMAP, GPF_DISORDER, GPF_SMOOTH,
    GPF_SOL_PAR, GPF_CONSTANT, GPF_COVALENTMAP, GPF_RECEPTOR_TYPES, GPF_PARAM_FILE,
    GPF_DSOLVMAP, GPF_QASP, GPF_USE_VINA_POTENTIAL, GPF_OUTLEV,
    GPF_MAP_RECEPTOR_INTERIOR, GPF_CMAP, GPF_CONSTRICTION_DISTANCE_CUTOFF,
    GPF_SEPARATE_DESOLVATION_MAPS, GPF_COEFF_VDW, GPF_COEFF_HBOND, GPF_COEFF_ESTAT,
    GPF_COEFF_DESOLV, GPF_COEFF_TORS
};

// gpfparser.cpp:54-173: first whitespace-delimited word, case-insensitive.
static int gpf_keyword(const char *line) {
    const char *t = line;
    while (isascii((unsigned char)*t) && isspace((unsigned char)*t)) t++;
    char c[256];
    int i = 0;
    for (; isascii((unsigned char)t[i]) && !isspace((unsigned char)t[i]) && t[i] && i < 255; i++)
        c[i] = t[i];
    c[i] = '\0';
    if (c[0] == '\n' || c[0] == '\0') return GPF_NULL;
    if (c[0] == '#') return GPF_COMMENT;
    static const struct { const char *kw; int tok; } table[] = {
        {"receptor_types", GPF_RECEPTOR_TYPES}, {"receptor", GPF_RECEPTOR},
        {"gridfld", GPF_GRIDFLD}, {"npts", GPF_NPTS}, {"spacing", GPF_SPACING},
        {"gridcenter", GPF_GRIDCENTER}, {"types", GPF_LIGAND_TYPES},
        {"ligand_types", GPF_LIGAND_TYPES}, {"map", GPF_MAP}, {"elecmap", GPF_ELECMAP},
        {"dsolvmap", GPF_DSOLVMAP}, {"desolvmap", GPF_DSOLVMAP},
        {"covalentmap", GPF_COVALENTMAP}, {"coeff_vdw", GPF_COEFF_VDW},
        {"coeff_hbond", GPF_COEFF_HBOND}, {"coeff_estat", GPF_COEFF_ESTAT},
        {"coeff_desolv", GPF_COEFF_DESOLV}, {"coeff_tors", GPF_COEFF_TORS},
        {"nbp_coeffs", GPF_NBP_COEFFS}, {"nbp_r_eps", GPF_NBP_R_EPS},
        {"dielectric", GPF_DIEL}, {"qasp", GPF_QASP}, {"fmap", GPF_FMAP},
        {"disorder_h", GPF_DISORDER}, {"smooth", GPF_SMOOTH}, {"sol_par", GPF_SOL_PAR},
        {"constant", GPF_CONSTANT}, {"parameter_file", GPF_PARAM_FILE},
        {"use_vina_potential", GPF_USE_VINA_POTENTIAL}, {"outlev", GPF_OUTLEV},
        {"map_receptor_interior", GPF_MAP_RECEPTOR_INTERIOR}, {"cmap", GPF_CMAP},
        {"constriction_map", GPF_CMAP},
        {"constriction_distance_cutoff", GPF_CONSTRICTION_DISTANCE_CUTOFF},
        {"separate_desolvation_maps", GPF_SEPARATE_DESOLVATION_MAPS}};
    for (auto &e : table)
        if (strcasecmp(c, e.kw) == 0) return e.tok;
    return GPF_UNKNOWN;
}

// parsetypes.cpp:50-99: whitespace-separated words after the keyword.
static std::vector<std::string> parse_types(const char *line, int maxwords) {
    std::vector<std::string> out;
    const char *p = line;
    bool keyword = true;
    while (true) {
        while (isspace((unsigned char)*p)) p++;
        if (!*p) break;
        const char *b = p;
        while (*p && !isspace((unsigned char)*p)) p++;
        if (keyword) keyword = false;
        else out.emplace_back(b, p);
        if ((int)out.size() >= maxwords) break;
    }
    return out;
}

// The first blank-delimited word after the keyword (sscanf "%*s %s").
static std::string second_word(const char *line) {
    char buf[256] = {0};
    if (sscanf(line, "%*s %255s", buf) < 1) return "";
    return buf;
}

// check_size.cpp:59-83.
static int check_size(Job &job, int n, char axis) {
    char msg[256];
    if (n < 0) {
        snprintf(msg, sizeof msg, "\n%s: Error! Negative number of %c-grid elements!  Aborting.\n\n",
                 job.program.c_str(), axis);
        fail(AG_ERR_BAD_INPUT, msg);
    }
    if (n == 0) {
        snprintf(msg, sizeof msg, "\n%s: Error!! 0 %c-grid elements!\n\n", job.program.c_str(), axis);
        fail(AG_ERR_BAD_INPUT, msg);
    }
    if (n > AG_MAX_GRID_PTS) {
        snprintf(msg, sizeof msg, "\n%s: Error! Maximum number of %c-grid elements allowed is %d.\n",
                 job.program.c_str(), axis, AG_MAX_GRID_PTS);
        fail(AG_ERR_BAD_INPUT, msg);
    }
    int even = (n / 2) * 2;
    if (even != n)
        job.logf("%s: Number of grid elements must be even; %c-elements changed to: %d\n",
                 job.program.c_str(), axis, even);
    return even;
}

static FILE *open_out(Job &job, const std::string &name, const std::string &what) {
    if (!job.make_files) return nullptr;
    FILE *f = fopen(name.c_str(), "w");
    if (!f) fail(AG_ERR_IO, what);
    return f;
}

// PDBQT receptor reader (main.cpp:823-1047).
static void read_receptor(Job &job, const std::string &fname) {
    FILE *f = fopen(fname.c_str(), "r");
    if (!f) fail(AG_ERR_IO, "can't find or open receptor PDBQT file \"" + fname + "\".\n");
    job.logf("\nReceptor Input File :\t%s\n\nReceptor Atom Type Assignments:\n\n", fname.c_str());
    job.coord.clear(); job.charge.clear(); job.vol.clear(); job.solpar.clear();
    job.type.clear(); job.hbond.clear();
    job.n_atoms = 0;
    double q_max = -kBig, q_min = kBig, q_tot = 0;
    double cmax[3] = {-kBig, -kBig, -kBig}, cmin[3] = {kBig, kBig, kBig}, csum[3] = {0, 0, 0};
    std::vector<int> type_count(AG_MAX_REC_TYPES, 0);
    char line[256];
    char autogrid_type[64] = {0};
    int lc = 0;
    struct Closer { FILE *f; ~Closer() { fclose(f); } } closer{f};
    while (fgets(line, sizeof line, f)) {
        lc++;
        if (strncmp(line, "ATOM  ", 6) && strncmp(line, "HETATM", 6) && strncmp(line, "CHAR", 4))
            continue;
        char msg[512];
        if (job.n_atoms >= AG_MAX_ATOMS) {
            snprintf(msg, sizeof msg, "Too many atoms in receptor PDBQT file %s line %d;",
                     fname.c_str(), lc);
            fail(AG_ERR_BAD_INPUT, msg);
        }
        if (strlen(line) < 78) {
            snprintf(msg, sizeof msg,
                     "ATOM/HETATM line is too short in receptor PDBQT file %s line %d;",
                     fname.c_str(), lc);
            fail(AG_ERR_BAD_INPUT, msg);
        }
        double xyz[3];
        int got = 0;
        for (int k = 0; k < 3; k++) {
            char field[9];
            memcpy(field, line + 30 + 8 * k, 8);
            field[8] = '\0';
            got += sscanf(field, "%lf", &xyz[k]);
        }
        if (got != 3) {
            snprintf(msg, sizeof msg, "ATOM/HETATM line bad x,y,z in receptor PDBQT file %s line %d;",
                     fname.c_str(), lc);
            fail(AG_ERR_BAD_INPUT, msg);
        }
        double q = 0.0;
        {
            char field[7];
            memcpy(field, line + 70, 6);
            field[6] = '\0';
            sscanf(field, "%lf", &q);
        }
        {
            char field[3];
            memcpy(field, line + 77, 2);
            field[2] = '\0';
            sscanf(field, "%63s", autogrid_type);  // a blank field keeps the previous type
        }
        ParamEntry *p = job.lib.find(autogrid_type);
        if (!p) {
            snprintf(msg, sizeof msg,
                     "\n\nreceptor file contains unknown type: '%s line %d'\nadd parameters for it to the parameter library first\n",
                     autogrid_type, lc);
            fail(AG_ERR_BAD_INPUT, msg);
        }
        if (p->rec_index < 0) {
            if ((int)job.rec_types.size() >= AG_MAX_REC_TYPES)
                fail(AG_ERR_BAD_INPUT, "too many receptor atom types");
            p->rec_index = (int)job.rec_types.size();
            job.rec_types.emplace_back(p->type);
        }
        for (int k = 0; k < 3; k++) job.coord.push_back(xyz[k]);
        job.charge.push_back(q);
        job.type.push_back(p->rec_index);
        job.solpar.push_back(p->solpar);
        job.vol.push_back(p->vol);
        job.hbond.push_back(p->hbond);
        type_count[p->rec_index]++;
        q_max = std::max(q_max, q);
        q_min = std::min(q_min, q);
        for (int k = 0; k < 3; k++) {
            cmax[k] = std::max(cmax[k], xyz[k]);
            cmin[k] = std::min(cmin[k], xyz[k]);
            csum[k] += xyz[k];
        }
        q_tot += q;
        job.n_atoms++;
    }
    if (job.has_receptor_types_in_gpf && (int)job.rec_types.size() != job.receptor_types_gpf_ct) {
        char msg[512];
        snprintf(msg, sizeof msg,
                 "The number of atom types found in the receptor PDBQT (%d) does not match the number specified by the \"receptor_types\" command (%d) in the GPF!\n\n",
                 (int)job.rec_types.size(), job.receptor_types_gpf_ct);
        fail(AG_ERR_BAD_INPUT, msg);
    }
    job.logf("\nMaximum partial atomic charge found = %+.3lf e\n", q_max);
    job.logf("Minimum partial atomic charge found = %+.3lf e\n\n", q_min);
    if (q_max == 0. && q_min == 0.)
        fail(AG_ERR_BAD_INPUT, "No partial atomic charges were found in the receptor PDBQT file " +
                                   fname + "!\n\n");
    job.logf("Atom\tAtom\tNumber of this Type\n");
    job.logf("Type\t ID \t in Receptor\n");
    job.logf("____\t____\t___________________\n");
    for (size_t t = 0; t < job.rec_types.size(); t++)
        if (type_count[t]) job.logf(" %d\t %s\t\t%6d\n", (int)t, job.rec_types[t].c_str(), type_count[t]);
    job.logf("\nTotal number of atoms :\t\t%d atoms \n", job.n_atoms);
    job.logf("Total charge :\t\t\t%.2lf e\n", q_tot);
    job.logf("\nMaximum coordinates :\t\t(%.3lf, %.3lf, %.3lf)\n", cmax[0], cmax[1], cmax[2]);
    job.logf("Minimum coordinates :\t\t(%.3lf, %.3lf, %.3lf)\n\n\n", cmin[0], cmin[1], cmin[2]);
    for (int k = 0; k < 3; k++) job.cmean[k] = csum[k] / (double)job.n_atoms;
}

// Per-map and per-(map, receptor type) parameters set when ligand_types is
// read (main.cpp:1289-1357); coeff_vdW / coeff_hbond are applied here, with
// the values in force at this line.
static void setup_map_pairs(Job &job) {
    const double cv = job.lib.coeff_vdW, ch = job.lib.coeff_hbond;
    for (int i = 0; i < job.num_atom_maps; i++) {
        MapInfo &m = job.maps[i];
        ParamEntry *lp = job.lib.find(m.type.c_str());
        m.is_covalent = (m.type == "Z");
        m.is_hbonder = false;
        m.solpar_probe = lp->solpar;
        m.vol_probe = lp->vol;
        m.Rij = lp->Rij;
        m.epsij = lp->epsij_unweighted * cv;
        m.hbond = lp->hbond;
        m.Rij_hb = lp->Rij_hb;
        m.epsij_hb = lp->epsij_hb_unweighted * ch;
        if (m.hbond > 0) m.is_hbonder = true;
        for (int j = 0; j < AG_MAX_REC_TYPES; j++) m.pair[j] = ag_pair{};
        for (size_t j = 0; j < job.rec_types.size(); j++) {
            ParamEntry *rp = job.lib.find(job.rec_types[j].c_str());
            ag_pair &pr = m.pair[j];
            pr.Rij = (m.Rij + rp->Rij) / 2.;
            pr.eps = sqrt(m.epsij * rp->epsij_unweighted * cv);
            pr.xA = 12;
            pr.xB = 6;
            pr.hbonder = 0;
            const int mh = m.hbond, rh = rp->hbond;
            if ((mh > 2 || mh == 6) && (rh == 1 || rh == 2 || rh == 6)) {
                pr.xB = 10;
                pr.hbonder = 1;
                m.is_hbonder = true;
                pr.Rij = m.Rij_hb;
                pr.eps = m.epsij_hb;
            } else if ((mh == 1 || mh == 2 || mh == 6) && rh > 2) {
                pr.xB = 10;
                pr.hbonder = 1;
                m.is_hbonder = true;
                pr.Rij = rp->Rij_hb;
                pr.eps = rp->epsij_hb_unweighted * ch;
            }
        }
    }
}

void job_parse_gpf(Job &job, const std::string &gpf_path) {
    // An empty path reads the GPF from standard input (setflags.cpp:78).
    FILE *gpf = gpf_path.empty() ? stdin : fopen(gpf_path.c_str(), "r");
    if (!gpf) fail(AG_ERR_IO, "Sorry, I can't find or open Grid Parameter File \"" + gpf_path + "\"");
    struct Closer { FILE *f; ~Closer() { if (f != stdin) fclose(f); } } closer{gpf};
    job.gpf_path = gpf_path;
    job.lib.load_builtin_ad41_bound();
    char raw[256];
    while (fgets(raw, sizeof raw, gpf)) {
        char line[256];
        strcpy(line, raw);
        int kw = gpf_keyword(line);
        job.logf("GPF> %s", raw);
        if (kw == GPF_UNKNOWN) {
            job.logf("\n%s: WARNING:  Unrecognized keyword in grid parameter file.\n\n\n",
                     job.program.c_str());
            continue;
        }
        if (kw != GPF_NULL && kw != GPF_COMMENT) {
            char *hash = strchr(line, '#');
            if (hash) *hash = '\0';
        }
        char msg[1024];
        switch (kw) {
            case GPF_NULL:
            case GPF_COMMENT:
            case GPF_CONSTANT:
                break;

            case GPF_RECEPTOR: {
                std::string fn = second_word(line);
                if (fn.empty()) fail(AG_ERR_BAD_INPUT, "No name specified for receptor filename");
                job.receptor_filename = fn;
                read_receptor(job, fn);
                break;
            }

            case GPF_GRIDFLD: {
                std::string fn = second_word(line);
                if (fn.empty()) fail(AG_ERR_BAD_INPUT, "No name specified for AVS_fld_filename");
                if (fn.find(".fld") == std::string::npos)
                    fail(AG_ERR_BAD_INPUT, "Grid data file needs the extension \".fld\" for AVS input\n\n");
                job.fld_filename = fn;
                job.xyz_filename = fn;
                job.xyz_filename.replace(fn.find("fld"), 3, "xyz");
                if (job.make_files) {
                    job.fld_fp = open_out(job, fn, "can't create grid dimensions data file " + fn + "\n");
                    job.xyz_fp = open_out(job, job.xyz_filename,
                                          "can't create grid extrema data file " + job.xyz_filename + "\n");
                }
                job.xyz_open = true;
                job.logf("\nCreating (AVS-readable) grid maps file : %s\n", fn.c_str());
                job.logf("\nCreating (AVS-readable) grid-coordinates extrema file : %s\n\n",
                         job.xyz_filename.c_str());
                break;
            }

            case GPF_NPTS: {
                sscanf(line, "%*s %d %d %d", &job.nelements[0], &job.nelements[1], &job.nelements[2]);
                const char axes[] = "xyz";
                for (int i = 0; i < 3; i++) {
                    job.nelements[i] = check_size(job, job.nelements[i], axes[i]);
                    job.ne[i] = job.nelements[i] / 2;
                    job.n1[i] = job.nelements[i] + 1;
                }
                job.npts_set = true;
                job.logf("\nNumber of grid points in x-direction:\t%d\n", job.n1[0]);
                job.logf("Number of grid points in y-direction:\t%d\n", job.n1[1]);
                job.logf("Number of grid points in z-direction:\t%d\n\n", job.n1[2]);
                break;
            }

            case GPF_SPACING:
                sscanf(line, "%*s %lf", &job.spacing);
                job.logf("Grid Spacing :\t\t\t%.3lf Angstrom\n\n", job.spacing);
                break;

            case GPF_GRIDCENTER: {
                // main.cpp:1116-1165: needs an open .fld and .xyz, centres the
                // atoms read so far and writes the .xyz extents.
                if (job.fld_filename.empty() || !job.xyz_open)
                    fail(AG_ERR_BAD_INPUT,
                         "You need to set the \"gridfld\" file before setting the grid center\".\n");
                std::string tok = second_word(line);
                if (tok == "auto") {
                    for (int i = 0; i < 3; i++) job.center[i] = job.cmean[i];
                    job.logf("Grid maps will be centered on the center of mass.\n");
                } else {
                    sscanf(line, "%*s %lf %lf %lf", &job.center[0], &job.center[1], &job.center[2]);
                    job.logf("\nGrid maps will be centered on user-defined coordinates:\n\n\t\t(%.3lf, %.3lf, %.3lf)\n",
                             job.center[0], job.center[1], job.center[2]);
                }
                for (int a = 0; a < job.n_atoms; a++)
                    for (int i = 0; i < 3; i++) job.coord[3 * a + i] -= job.center[i];
                double cgmin[3], cgmax[3];
                for (int i = 0; i < 3; i++) {
                    double cext = job.spacing * (double)job.ne[i];
                    cgmax[i] = job.center[i] + cext;
                    cgmin[i] = job.center[i] - cext;
                }
                if (job.xyz_fp) {
                    for (int i = 0; i < 3; i++) fprintf(job.xyz_fp, "%.3lf %.3lf\n", cgmin[i], cgmax[i]);
                    fclose(job.xyz_fp);
                    job.xyz_fp = nullptr;
                }
                job.logf("\nMaximum coordinates :\t\t(%.3lf, %.3lf, %.3lf)\n", cgmax[0], cgmax[1], cgmax[2]);
                job.logf("Minimum coordinates :\t\t(%.3lf, %.3lf, %.3lf)\n\n", cgmin[0], cgmin[1], cgmin[2]);
                // A second gridcenter line finds the .xyz closed (main.cpp:1164).
                job.xyz_open = false;
                break;
            }

            case GPF_LIGAND_TYPES: {
                if (job.n1[0] == 0)
                    fail(AG_ERR_BAD_INPUT,
                         "You need to set the \"npts\" before setting the ligand types\".\n");
                std::vector<std::string> types = parse_types(line, AG_MAX_MAPS - 2);
                std::vector<std::string> lig;
                for (int i = 0; i < job.num_atom_maps; i++) lig.push_back(job.maps[i].type);
                for (auto &t : types) {
                    lig.push_back(t);
                    if (t != "Z" && job.separate_desolvation) lig.push_back(t);
                }
                if ((int)lig.size() + 2 > AG_MAX_MAPS)
                    fail(AG_ERR_BAD_INPUT, "Too many ligand atom types");
                job.num_atom_maps = (int)lig.size();
                for (int i = 0; i < job.num_atom_maps; i++) {
                    ParamEntry *p = job.lib.find(lig[i].c_str());
                    if (!p) {
                        snprintf(msg, sizeof msg,
                                 "unknown ligand atom type %s\nadd parameters for it to the parameter library first!\n",
                                 lig[i].c_str());
                        fail(AG_ERR_BAD_INPUT, msg);
                    }
                    p->map_index = i;
                }
                for (auto &m : job.maps)
                    if (m.fp) { fclose(m.fp); m.fp = nullptr; }
                job.maps.assign(job.num_atom_maps + 2, MapInfo{});
                for (int i = 0; i < job.num_atom_maps; i++) job.maps[i].type = lig[i];
                job.ligand_types_seen = true;
                job.logf("Allocated space for %d gridmap objects\n", job.num_atom_maps + 2);
                setup_map_pairs(job);
                break;
            }

            case GPF_RECEPTOR_TYPES: {
                std::vector<std::string> types = parse_types(line, AG_MAX_MAPS - 2);
                if ((int)types.size() > AG_MAX_REC_TYPES)
                    fail(AG_ERR_BAD_INPUT, "too many receptor atom types");
                job.receptor_types_gpf_ct = (int)types.size();
                job.has_receptor_types_in_gpf = true;
                job.rec_types = types;
                for (size_t i = 0; i < types.size(); i++) {
                    ParamEntry *p = job.lib.find(types[i].c_str());
                    if (!p) {
                        snprintf(msg, sizeof msg,
                                 "Unknown receptor type: \"%s\"\n -- Add parameters for it to the parameter library first!\n",
                                 types[i].c_str());
                        fail(AG_ERR_BAD_INPUT, msg);
                    }
                    p->rec_index = (int)i;
                }
                break;
            }

            case GPF_SOL_PAR: {
                char t[64] = {0};
                double v = 0, s = 0;
                sscanf(line, "%*s %63s %lf %lf", t, &v, &s);
                ParamEntry *p = job.lib.find(t);
                if (p) {
                    p->vol = v;
                    p->solpar = s;
                    if (p->map_index >= 0 && p->map_index < (int)job.maps.size())
                        job.maps[p->map_index].solpar_probe = s;
                } else {
                    job.logf("%s key not found\n", t);
                }
                break;
            }

            case GPF_USE_VINA_POTENTIAL:
                job.logf("\n Using Vina potential for calculation.\n\n");
                fail(AG_ERR_BAD_INPUT, "Vina potential not implemented");

            case GPF_MAP: {
                if (!job.ligand_types_seen)
                    fail(AG_ERR_BAD_INPUT, "\"map\" keyword appeared before \"ligand_types\"");
                if (job.num_maps > job.num_atom_maps + (job.elecPE > -1) + (job.dsolvPE > -1)) {
                    snprintf(msg, sizeof msg,
                             "Too many \"map\" keywords (%d);  the \"ligand_types\" command declares only %d atom types.\nRemove a \"map\" keyword from the GPF.\n",
                             job.num_maps, job.num_atom_maps);
                    fail(AG_ERR_BAD_INPUT, msg);
                }
                if (job.num_maps >= (int)job.maps.size())
                    fail(AG_ERR_BAD_INPUT, "Too many map requests for the declared ligand types");
                std::string fn = second_word(line);
                if (fn.empty()) fail(AG_ERR_BAD_INPUT, "No name specified for map file");
                job.maps[job.num_maps].filename = fn;
                job.maps[job.num_maps].fp =
                    open_out(job, fn, "Cannot open grid map \"" + fn + "\" for writing.");
                job.logf("\nOutput Grid Map %d:   %s\n\n", job.num_maps + 1, fn.c_str());
                job.num_maps++;
                break;
            }

            case GPF_ELECMAP:
            case GPF_DSOLVMAP: {
                const bool elec = kw == GPF_ELECMAP;
                int &slot = elec ? job.elecPE : job.dsolvPE;
                if (slot >= 0)
                    fail(AG_ERR_BAD_INPUT, elec ? "Duplicate \"elecmap\" request"
                                                : "Duplicate \"dsolvmap\" request");
                if (!job.ligand_types_seen)
                    fail(AG_ERR_BAD_INPUT, "map request appeared before \"ligand_types\"");
                if (job.num_maps < job.num_atom_maps)
                    fail(AG_ERR_BAD_INPUT,
                         elec ? "\"elecmap\" request appeared before all atom map requests had appeared."
                              : "\"desolvmap\" request appeared before all atom map requests had appeared.");
                if (job.num_maps >= (int)job.maps.size())
                    fail(AG_ERR_BAD_INPUT, "Too many map requests for the declared ligand types");
                slot = job.num_maps;
                std::string fn = second_word(line);
                if (fn.empty())
                    fail(AG_ERR_BAD_INPUT, elec ? "No name specified for electrostatic map"
                                                : "No name specified for desolvation map");
                job.maps[slot].filename = fn;
                job.maps[slot].fp = open_out(
                    job, fn,
                    std::string("can't open file \"") + fn +
                        (elec ? "\" for writing electrostatic map.\n" : "\" for writing desolvation map.\n"));
                job.logf(elec ? "\nOutput Electrostatic Potential Energy Grid Map: %s\n\n"
                              : "\nOutput Desolvation Free Energy Grid Map: %s\n\n",
                         fn.c_str());
                job.num_maps++;
                break;
            }

            case GPF_COVALENTMAP:
                sscanf(line, "%*s %lf %lf %lf %lf %lf", &job.cov_half_width, &job.cov_barrier,
                       &job.cov_pos[0], &job.cov_pos[1], &job.cov_pos[2]);
                for (int i = 0; i < 3; i++) job.cov_pos[i] -= job.center[i];
                break;

            case GPF_DISORDER:
                job.disorder_h = true;
                job.logf("\nHydroxyls will be disordered \n\n");
                break;

            case GPF_SMOOTH:
                sscanf(line, "%*s %lf", &job.r_smooth);
                job.logf("\nPotentials will be smoothed by: %.3lf Angstrom\n\n", job.r_smooth);
                break;

            case GPF_CONSTRICTION_DISTANCE_CUTOFF:
                sscanf(line, "%*s %lf", &job.constriction_cutoff);
                break;

            case GPF_SEPARATE_DESOLVATION_MAPS:
                if (job.num_atom_maps > 0)
                    fail(AG_ERR_BAD_INPUT,
                         "\"separate_desolvation_maps\" must be specified before \"ligand_types\"");
                job.separate_desolvation = true;
                break;

            case GPF_OUTLEV:
                sscanf(line, "%*s %d", &job.outlev);
                job.logf("\nOutput level: %d\n\n", job.outlev);
                break;

            case GPF_QASP:
                sscanf(line, "%*s %lf", &job.solpar_q);
                job.logf("\nCharge component of the atomic solvation parameter: %.3lf\n\n", job.solpar_q);
                break;

            case GPF_DIEL: {
                // main.cpp:1606-1639
                double diel = 0;
                sscanf(line, "%*s %lf", &diel);
                if (diel < 0.) {
                    job.dddiel = true;
                    job.logf("\nUsing *distance-dependent* dielectric function of Mehler and Solmajer, Prot.Eng.4, 903-910.\n\n");
                } else {
                    job.dddiel = false;
                    if (diel <= kApproxZero) diel = 40.;
                    job.logf("Using a *constant* dielectric of:  %.2f\n", diel);
                    job.inv_diel_cal = 332.0 / diel;
                }
                break;
            }

            case GPF_MAP_RECEPTOR_INTERIOR: {   // main.cpp:1643-1658
                char tok[64] = "true";
                sscanf(line, "%*s %63s", tok);
                if (strcasecmp(tok, "true") == 0 || strcasecmp(tok, "on") == 0) {
                    job.map_receptor_interior = true;
                    job.logf("Mapping receptor interior instead of reporting collision if nearer than %.2f to "
                             "any atom.", 1.5);
                } else if (job.bhtree) {
                    job.map_receptor_interior = false;
                    job.logf("Reporting collision if nearer than %.2f to any atom.", 1.5);
                } else {
                    fail(AG_ERR_BAD_INPUT, "'map_receptor_interior off'  is supported only with USE_BHTREE.");
                }
                break;
            }

            case GPF_FMAP:
            case GPF_CMAP: {
                const bool fm = kw == GPF_FMAP;
                bool &flag = fm ? job.floating_grid : job.constriction_grid;
                if (flag)
                    fail(AG_ERR_BAD_INPUT, fm ? "Muliple requests for floating map are not allowed."
                                              : "Muliple requests for constriction map are not allowed.");
                std::string fn = second_word(line);
                if (fn.empty())
                    fail(AG_ERR_BAD_INPUT, fm ? "No name specified for floating map"
                                              : "No name specified for constriction map");
                FILE *f = open_out(job, fn,
                                   std::string("can't open file \"") + fn +
                                       (fm ? "\" for writing floating map.\n"
                                           : "\" for writing constriction map.\n"));
                if (fm) { job.fmap_filename = fn; job.fmap_fp = f; }
                else { job.cmap_filename = fn; job.cmap_fp = f; }
                flag = true;
                job.num_distance_maps++;
                break;
            }

            case GPF_COEFF_VDW: sscanf(line, "%*s %lf", &job.lib.coeff_vdW); break;
            case GPF_COEFF_HBOND: sscanf(line, "%*s %lf", &job.lib.coeff_hbond); break;
            case GPF_COEFF_ESTAT: sscanf(line, "%*s %lf", &job.lib.coeff_estat); break;
            case GPF_COEFF_DESOLV: sscanf(line, "%*s %lf", &job.lib.coeff_desolv); break;
            case GPF_COEFF_TORS: sscanf(line, "%*s %lf", &job.lib.coeff_tors); break;

            case GPF_PARAM_FILE: {
                std::string fn = second_word(line);
                if (fn.empty()) fail(AG_ERR_BAD_INPUT, "No name specified for FN_parameter_library");
                job.lib.load_file(fn);
                break;
            }

            case GPF_NBP_COEFFS:
            case GPF_NBP_R_EPS: {
                // main.cpp:1741-1806
                double Rij = 0, epsij = 0, cA, cB;
                int xA = 0, xB = 0;
                char p0[256] = {0}, p1[256] = {0};
                int nf = sscanf(line, "%*s %lf %lf %d %d %255s %255s", &Rij, &epsij, &xA, &xB, p0, p1);
                if (nf != 6) fail(AG_ERR_BAD_INPUT, "syntax error, not 6 values in NBP_R_EPS line");
                if (kw == GPF_NBP_R_EPS) {
                    epsij *= job.lib.coeff_vdW;
                    if (xA == xB) fail(AG_ERR_BAD_INPUT, "exponents xA and xB cannot be equal.\n");
                    double tmpconst = epsij / (double)(xA - xB);
                    cA = tmpconst * pow(Rij, (double)xA) * (double)xB;
                    cB = tmpconst * pow(Rij, (double)xB) * (double)xA;
                } else {
                    cA = Rij;
                    cB = epsij;
                }
                const char *ps[2] = {p0, p1};
                for (int i = 0; i < 2; i++) {
                    int lt = job.map_index(ps[i % 2]);
                    int rt = job.rec_index(ps[(i + 1) % 2]);
                    if (lt >= 0 && rt >= 0 && lt < (int)job.maps.size() && rt < AG_MAX_REC_TYPES) {
                        ag_pair &pr = job.maps[lt].pair[rt];
                        pr.cA = cA;
                        pr.cB = cB;
                        pr.Rij = Rij;
                        pr.eps = epsij;
                        pr.xA = xA;
                        pr.xB = xB;
                    }
                }
                break;
            }

            default:
                break;
        }
    }

    // End-of-GPF checks (main.cpp:1852-1894).
    if (job.num_atom_maps > job.num_maps - (job.elecPE >= 0) - (job.dsolvPE >= 0)) {
        job.logf("Too few \"map\" keywords ;  the \"ligand_types\" command declares %d atom types.\nAdd a \"map\" keyword from the GPF.\n",
                 job.num_atom_maps);
        fail(AG_ERR_BAD_INPUT, "Too few \"map\" keywords found for the number of ligand atom types.\n");
    }
    if (job.elecPE < 0) {
        job.logf("\n%s: INFORMATION:  The electrostatic map file is not defined in the GPF.\n\n\n", job.program.c_str());
        job.logf("\n%s: INFORMATION:  No electrostatic map file requested.\n\n\n", job.program.c_str());
    }
    if (job.dsolvPE < 0) {
        job.logf("\n%s: INFORMATION:  The desolvation map file is not defined in the GPF.\n\n\n", job.program.c_str());
        job.logf("\n%s: INFORMATION:  No desolvation map file requested.\n\n\n", job.program.c_str());
    }
    job.logf("\n>>> Closing the grid parameter file (GPF)... <<<\n\n");
    if (job.fld_filename.empty()) {
        job.logf("Missing \"gridfld\" keyword;\nAdd a \"gridfld <file.fld>\" keyword to the GPF after the \"npts\" keyword.\n");
        fail(AG_ERR_BAD_INPUT, "Missing \"gridfld\" keyword.");
    }
    if (!job.npts_set) fail(AG_ERR_BAD_INPUT, "Missing \"npts\" keyword.");
    if (job.num_maps == 0 && !job.floating_grid && !job.constriction_grid)
        fail(AG_ERR_BAD_INPUT, "No maps were requested in the GPF.");

    // Receptor type indices for the bond-vector pass (main.cpp:2280-2288).
    job.t_hydrogen = job.rec_index("HD");
    job.t_nonhb_hydrogen = job.rec_index("H");
    job.t_carbon = job.rec_index("C");
    job.t_arom_carbon = job.rec_index("A");
    job.t_oxygen = job.rec_index("OA");
    job.t_nitrogen = job.rec_index("NA");
    job.t_nonhb_nitrogen = job.rec_index("N");
    job.t_sulphur = job.rec_index("SA");
}

// AVS .fld header (main.cpp:1897-1956).
// The .fld text (main.cpp:1903-1956).  With `binary`, the variables point
// at the raw FP32 side-output files (<map>.bin, little-endian, x fastest,
// no header) instead of the ASCII maps.
static void write_fld_text(Job &job, FILE *f, bool binary) {
    const char *gpf = job.gpf_path.c_str();
    auto var = [&](int k, const std::string &fn) {
        if (binary) fprintf(f, "variable %d file=%s.bin filetype=binary skip=0\n", k, fn.c_str());
        else fprintf(f, "variable %d file=%s filetype=ascii skip=6\n", k, fn.c_str());
    };
    fprintf(f, "# AVS field file\n#\n");
    fprintf(f, "# AutoDock Atomic Affinity and Electrostatic Grids\n#\n");
    fprintf(f, "# Created by %s %s.\n#\n", job.program.c_str(), "4.2.7.x");
    fprintf(f, "#SPACING %.3f\n", (float)job.spacing);
    fprintf(f, "#NELEMENTS %d %d %d\n", job.nelements[0], job.nelements[1], job.nelements[2]);
    fprintf(f, "#CENTER %.3lf %.3lf %.3lf\n", job.center[0], job.center[1], job.center[2]);
    fprintf(f, "#MACROMOLECULE %s\n", job.receptor_filename.c_str());
    fprintf(f, "#GRID_PARAMETER_FILE %s\n#\n", gpf);
    fprintf(f, "ndim=3\t\t\t# number of dimensions in the field\n");
    fprintf(f, "dim1=%d\t\t\t# number of x-elements\n", job.n1[0]);
    fprintf(f, "dim2=%d\t\t\t# number of y-elements\n", job.n1[1]);
    fprintf(f, "dim3=%d\t\t\t# number of z-elements\n", job.n1[2]);
    fprintf(f, "nspace=3\t\t# number of physical coordinates per point\n");
    fprintf(f, "veclen=%d\t\t# number of affinity values at each point\n", job.num_maps + job.num_distance_maps);
    fprintf(f, "data=float\t\t# data type (byte, integer, float, double)\n");
    fprintf(f, "field=uniform\t\t# field type (uniform, rectilinear, irregular)\n");
    for (int i = 0; i < 3; i++)
        fprintf(f, "coord %d file=%s filetype=ascii offset=%d\n", i + 1, job.xyz_filename.c_str(), i * 2);
    for (int i = 0; i < job.num_atom_maps; i++) {
        fprintf(f, "label=%s-affinity\t# component label for variable %d\n", job.maps[i].type.c_str(), i + 1);
        if (job.separate_desolvation && !job.maps[i].is_covalent) {
            i++;
            fprintf(f, "label=%s-desolv\t# component label for variable %d\n", job.maps[i].type.c_str(), i + 1);
        }
    }
    if (job.elecPE >= 0)
        fprintf(f, "label=Electrostatics\t# component label for variable %d\n", job.elecPE + 1);
    if (job.dsolvPE >= 0)
        fprintf(f, "label=Desolvation\t# component label for variable %d\n", job.dsolvPE + 1);
    const int base = job.num_atom_maps + (job.elecPE >= 0) + (job.dsolvPE >= 0);
    if (job.floating_grid)
        fprintf(f, "label=Floating_Grid\t# component label for variable %d\n", base + 1);
    if (job.constriction_grid)
        fprintf(f, "label=Constriction_Grid\t# component label for variable %d\n",
                base + (job.floating_grid ? 1 : 0) + 1);
    fprintf(f, "#\n# location of affinity grid files and how to read them\n#\n");
    for (int i = 0; i < job.num_atom_maps; i++) var(i + 1, job.maps[i].filename);
    if (job.elecPE >= 0) var(job.num_atom_maps + 1, job.maps[job.elecPE].filename);
    if (job.dsolvPE >= 0) var(job.num_atom_maps + ((job.elecPE >= 0) ? 2 : 1), job.maps[job.dsolvPE].filename);
    if (job.floating_grid) var(job.num_maps + 1, job.fmap_filename);
    if (job.constriction_grid) var(job.num_maps + (job.floating_grid ? 1 : 0) + 1, job.cmap_filename);
}

static std::string binary_fld_name(const Job &job) {
    const std::string &fn = job.fld_filename;
    if (fn.size() > 4 && fn.compare(fn.size() - 4, 4, ".fld") == 0) return fn.substr(0, fn.size() - 4) + ".bin.fld";
    return fn + ".bin.fld";
}

void job_write_fld(Job &job) {
    FILE *f = job.fld_fp;
    if (!f) return;
    write_fld_text(job, f, false);
    fclose(f);
    job.fld_fp = nullptr;
    if (job.binary_maps && job.make_files) {
        const std::string bfn = binary_fld_name(job);
        FILE *b = fopen(bfn.c_str(), "w");
        if (!b) fail(AG_ERR_IO, "can't create the binary grid data file " + bfn + "\n");
        write_fld_text(job, b, true);
        if (fclose(b) != 0) fail(AG_ERR_IO, "Problems writing " + bfn + "\n");
    }
}

// Six header lines of every map file (main.cpp:2752-2776).
static void write_header(Job &job, FILE *f) {
    fprintf(f, "GRID_PARAMETER_FILE %s\n", job.gpf_path.c_str());
    fprintf(f, "GRID_DATA_FILE %s\n", job.fld_filename.c_str());
    fprintf(f, "MACROMOLECULE %s\n", job.receptor_filename.c_str());
    fprintf(f, "SPACING %.3lf\n", job.spacing);
    fprintf(f, "NELEMENTS %d %d %d\n", job.nelements[0], job.nelements[1], job.nelements[2]);
    fprintf(f, "CENTER %.3lf %.3lf %.3lf\n", job.center[0], job.center[1], job.center[2]);
}

void job_write_map_headers(Job &job) {
    if (job.headers_written) return;
    job.headers_written = true;
    for (int k = 0; k < job.num_maps; k++)
        if (job.maps[k].fp) write_header(job, job.maps[k].fp);
    if (job.fmap_fp) write_header(job, job.fmap_fp);
    if (job.cmap_fp) write_header(job, job.cmap_fp);
}

}  // namespace ag
