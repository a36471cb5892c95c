// deck.cpp -- hstruct deck parser and supercell setup.
//
// Same keys and token rules as the reference's main() (app/src/main.cpp:75-937):
// each line is split on single spaces; a recognised key consumes the next
// tokens ('=' then the value) and scanning continues on the same line.
// MAXSIZE also sets the flake radius max_R in Angstrom (main.cpp:95-104);
// MATERIAL snapshots the current MINSIZE/MAXSIZE/max_R/BOUNDARY_CONDITION
// (main.cpp:324-337); ANGLE is read in degrees (main.cpp:345-352).
#include <cmath>
#include <cstdlib>
#include <sstream>

#include "host.h"

namespace leo2d {

Params::Params()
{
    // Job_params::Job_params (job_params.cpp:18-156), entries on this path
    s["job_name"] = "LEO2D_JOB";
    i["solver_type"] = 0;
    i["observable_type"] = 0;
    i["cond_transform"] = 1;     // job_params.cpp:122
    i["cond_poly_par"] = 0;      // job_params.cpp:47
    i["solver_space"] = 0;
    i["nShifts"] = 1;
    i["num_shift_sheets"] = 1;
    i["num_lc_points"] = 2;
    dm["lc_points"] = {{0.0, 0.0}, {1.0, 1.0}};
    i["num_sheets"] = 1;
    i["boundary_condition"] = 0;
    i["sc_search_size"] = 1;
    i["global_shifts_on"] = 0;
    i["strain_type"] = 0;
    i["uniform_strain"] = 0;
    i["h_supercell_modify"] = 0;     // job_params.cpp:55-56
    i["hex_supercell_modify"] = 0;
    i["matrix_save"] = 0;
    i["matrix_pos_save"] = 0;
    i["verbose_save"] = 1;
    i["diagonalize"] = 0;
    i["chiral_on"] = 0;
    i["k_sampling"] = 0;
    i["k_type"] = 0;
    i["kpm_trace"] = 0;
    i["kpm_trace_samps"] = -1;
    d["energy_rescale"] = 20.0;
    d["energy_shift"] = 0.0;
    i["num_target_sheets"] = 1;
    i["num_targets"] = 0;
    i["poly_order"] = 20;
    i["dos_transform"] = 1;
    i["magOn"] = 0;
    i["elecOn"] = 0;
    d["B"] = 0.0;
    d["E"] = 0.0;
    i["ballistic_transport"] = 0;
    i["mat_soc_on"] = 0;
    i["mat_from_file"] = 0;
}

long long Params::geti(const std::string& k) const
{
    auto it = i.find(k);
    if (it == i.end()) throw DeckError("missing integer parameter " + k);
    return it->second;
}

double Params::getd(const std::string& k) const
{
    auto it = d.find(k);
    if (it == d.end()) throw DeckError("missing double parameter " + k);
    return it->second;
}

namespace {

struct Tokens {
    std::vector<std::string> t;
    size_t p = 0;
    std::string next() { return p < t.size() ? t[p++] : std::string(); }
    bool more() const { return p < t.size(); }
};

Tokens split_spaces(const std::string& line)
{
    Tokens tk;
    std::istringstream in(line);
    std::string s;
    while (std::getline(in, s, ' ')) tk.t.push_back(s);
    return tk;
}

int material_of(const std::string& name)
{
    if (name == "Graphene") return MAT_GRAPHENE;
    if (name == "StrainedGraphene") return MAT_STRAINED_GRAPHENE;
    if (name == "GrapheneSimple") return MAT_GRAPHENE_SIMPLE;
    if (name == "GrapheneSimpleStrain") return MAT_GRAPHENE_SIMPLE_STRAIN;
    throw DeckError("Failed to find material! Aborting...");
}

}  // namespace

Deck parse_deck(const std::string& text)
{
    Deck deck;
    Params& o = deck.opts;
    std::vector<int> mins, maxs;
    double max_R = 0.0;
    int bc = 0;
    int num_sheets = 0;
    int cur = -1;
    double height = 0.0, angle = 0.0;
    std::istringstream lines(text);
    std::string line;
    while (std::getline(lines, line)) {
        Tokens tk = split_spaces(line);
        while (tk.more()) {
            const std::string key = tk.next();
            auto val = [&]() { tk.next(); return tk.next(); };
            if (key == "JOB_NAME") o.s["job_name"] = val();
            else if (key == "MAXSIZE") {
                int v = std::atoi(val().c_str());
                maxs.insert(maxs.end(), {v, v, v});
                max_R = (double)v;
            } else if (key == "MINSIZE") {
                int v = std::atoi(val().c_str());
                mins.insert(mins.end(), {v, v, v});
            } else if (key == "MAX_R") max_R = std::atof(val().c_str());
            else if (key == "BOUNDARY_CONDITION") {
                bc = std::atoi(val().c_str());
                o.i["boundary_condition"] = bc;
            } else if (key == "SUPERCELL_GRID") {
                tk.next();
                o.i["x_supercell"] = std::atoi(tk.next().c_str());
                o.i["y_supercell"] = std::atoi(tk.next().c_str());
                o.i["supercell_type"] = 2;
            } else if (key == "SUPERCELL_SEARCH_SIZE") o.i["sc_search_size"] = std::atoi(val().c_str());
            else if (key == "K_SAMPLING") o.i["k_sampling"] = std::atoi(val().c_str());
            else if (key == "K_TYPE") o.i["k_type"] = std::atoi(val().c_str());
            else if (key == "K_GRID") {
                tk.next();
                o.i["num_k1"] = std::atoi(tk.next().c_str());
                o.i["num_k2"] = std::atoi(tk.next().c_str());
            } else if (key == "NUM_SHEETS") {
                num_sheets = std::atoi(val().c_str());
                if (num_sheets < 0 || num_sheets > LEO2D_MAX_SHEETS) throw DeckError("NUM_SHEETS out of range");
                deck.sheets.assign(num_sheets, SheetData());
                deck.heights.assign(num_sheets, 0.0);
                deck.angles.assign(num_sheets, 0.0);
                deck.has_material.assign(num_sheets, 0);
                o.i["num_sheets"] = num_sheets;
            } else if (key == "SUPERCELL_GROUPS") {
                // main.cpp:131-155: the sheets (1-based) of group 1, "/", then those of group 2,
                // until NUM_SHEETS sheets are placed
                tk.next();
                std::vector<int> groups((size_t)std::max(num_sheets, 0), 0);
                int placed = 0;
                for (int grp = 0; grp < 2; ++grp) {
                    while (placed < num_sheets) {
                        const std::string v = tk.next();
                        if (v == "/") break;
                        const int sh = std::atoi(v.c_str());
                        if (v.empty() || sh < 1 || sh > num_sheets) throw DeckError("SUPERCELL_GROUPS names no sheet '" + v + "'");
                        groups[(size_t)sh - 1] = grp;
                        ++placed;
                    }
                }
                o.iv["sc_groups"] = groups;
            } else if (key == "CHIRAL_ON") o.i["chiral_on"] = std::atoi(val().c_str());
            else if (key == "BALLISTIC_TRANSPORT") o.i["ballistic_transport"] = std::atoi(val().c_str());
            else if (key == "MAT_SOC_ON") o.i["mat_soc_on"] = std::atoi(val().c_str());
            else if (key == "MAT_FROM_FILE") o.i["mat_from_file"] = std::atoi(val().c_str());
            else if (key == "MLMC") o.i["mlmc"] = std::atoi(val().c_str());
            else if (key == "GSFE_Z_ON") o.i["gsfe_z_on"] = std::atoi(val().c_str());
            else if (key == "TRILAYER_SUPERCELL") {   // main.cpp:294-302: the type changes only for the value 1
                if (std::atoi(val().c_str()) == 1) o.i["supercell_type"] = 3;
            }
            else if (key == "H_SUPERCELL_MODIFY") o.i["h_supercell_modify"] = std::atoi(val().c_str());
            else if (key == "HEX_SUPERCELL_MODIFY") o.i["hex_supercell_modify"] = std::atoi(val().c_str());
            else if (key == "SUPERCELL_M_N") {
                tk.next();
                o.i["m_supercell"] = std::atoi(tk.next().c_str());
                o.i["n_supercell"] = std::atoi(tk.next().c_str());
                o.i["supercell_type"] = 1;
            } else if (key == "START_SHEET") {
                cur = std::atoi(tk.next().c_str()) - 1;
                if (cur + 1 > num_sheets) {
                    throw DeckError("Cannot take in sheet " + std::to_string(cur + 1) + " if NUM_SHEETS = " +
                                    std::to_string(num_sheets));
                }
            } else if (key == "END_SHEET") {
                if (cur == std::atoi(tk.next().c_str()) - 1 && cur >= 0) {
                    deck.heights[cur] = height;
                    deck.angles[cur] = angle;
                    cur = -1;
                }
            } else if (key == "MATERIAL") {
                const std::string name = val();
                int m = material_of(name);
                if (cur < 0 || cur >= num_sheets) throw DeckError("MATERIAL outside a START_SHEET/END_SHEET block");
                if (mins.size() < 2 || maxs.size() < 2) throw DeckError("MATERIAL before MINSIZE/MAXSIZE");
                SheetData& sd = deck.sheets[cur];
                sd.mat = m;
                sd.min_shape[0] = mins[0];
                sd.min_shape[1] = mins[1];
                sd.max_shape[0] = maxs[0];
                sd.max_shape[1] = maxs[1];
                sd.max_R = max_R;
                sd.boundary_condition = bc;
                deck.has_material[cur] = 1;
            } else if (key == "HEIGHT") height = std::atof(val().c_str());
            else if (key == "ANGLE") angle = (2 * M_PI * std::atof(val().c_str())) / (360.0);
            else if (key == "NSHIFTS") o.i["nShifts"] = std::atoi(val().c_str());
            else if (key == "ENERGY_RESCALE") o.d["energy_rescale"] = std::atof(val().c_str());
            else if (key == "ENERGY_SHIFT") o.d["energy_shift"] = std::atof(val().c_str());
            else if (key == "POLY_ORDER") o.i["poly_order"] = std::atoi(val().c_str());
            else if (key == "SOLVER_TYPE") {
                const std::string v = val();
                if (v == "SQ") o.i["solver_type"] = 1;
                else if (v == "LC") o.i["solver_type"] = 2;
                else if (v == "MLMC_VAC") o.i["solver_type"] = 3;
                else if (v == "VD_FILE") o.i["solver_type"] = 4;
                else if (v == "STRAIN") o.i["solver_type"] = 5;
                else if (v == "FIXED_SHIFT") o.i["solver_type"] = 6;
            } else if (key == "NUM_LC_POINTS") o.i["num_lc_points"] = std::atoi(val().c_str());
            else if (key == "LC_POINTS") {
                int npts = (int)o.geti("num_lc_points");
                std::vector<std::vector<double>> pts(npts, std::vector<double>(2, 0.0));
                tk.next();
                for (int c = 0; c < npts; ++c) {
                    pts[c][0] = std::atof(tk.next().c_str());
                    pts[c][1] = std::atof(tk.next().c_str());
                }
                o.dm["lc_points"] = pts;
            } else if (key == "OBSERVABLE_TYPE") {
                const std::string v = val();
                if (v == "DOS") o.i["observable_type"] = 0;
                else if (v == "COND") o.i["observable_type"] = 1;
            } else if (key == "DOS_TRANSFORM") o.i["dos_transform"] = std::atoi(val().c_str());
            else if (key == "COND_TRANSFORM") o.i["cond_transform"] = std::atoi(val().c_str());
            else if (key == "COND_POLY_PAR") o.i["cond_poly_par"] = std::atoi(val().c_str());
            else if (key == "COND_POLY_PAR_SCALE") o.i["cond_poly_par_scale"] = std::atoi(val().c_str());
            else if (key == "SOLVER_SPACE") {
                const std::string v = val();
                if (!v.empty() && v[0] == 'R') o.i["solver_space"] = 0;
                else if (!v.empty() && v[0] == 'M') o.i["solver_space"] = 1;
            } else if (key == "STRAIN_TYPE") {
                // main.cpp:595-614: the value's first letter
                const std::string v = val();
                const std::string codes = "NSCRFPD";
                const size_t c = v.empty() ? std::string::npos : codes.find(v[0]);
                if (c != std::string::npos) o.i["strain_type"] = (long long)c;
            } else if (key == "STRAIN_FILE") o.s["strain_file"] = val();
            else if (key == "UNIFORM_STRAIN") o.i["uniform_strain"] = std::atoi(val().c_str());
            else if (key == "UNIFORM_STRAIN_MAT") {
                tk.next();
                std::vector<std::vector<double>> u(2, std::vector<double>(2, 0.0));
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) u[a][b] = std::atof(tk.next().c_str());
                o.dm["u_ij"] = u;
            } else if (key == "HETEROSTRAIN") {
                // uniaxial strain eps along x with Poisson ratio nu on the last sheet
                tk.next();
                const double eps = std::atof(tk.next().c_str());
                const double nu = std::atof(tk.next().c_str());
                o.dm["heterostrain"] = {{eps, nu}};
            } else if (key == "MATRIX_SAVE") o.i["matrix_save"] = std::atoi(val().c_str());
            else if (key == "MATRIX_POS_SAVE") o.i["matrix_pos_save"] = std::atoi(val().c_str());
            else if (key == "VERBOSE_SAVE") o.i["verbose_save"] = std::atoi(val().c_str());
            else if (key == "DIAGONALIZE") o.i["diagonalize"] = std::atoi(val().c_str());
            else if (key == "KPM_TRACE") o.i["kpm_trace"] = std::atoi(val().c_str());
            else if (key == "KPM_TRACE_SAMPS") o.i["kpm_trace_samps"] = std::atoi(val().c_str());
            else if (key == "USE_B_FIELD") o.i["magOn"] = std::atoi(val().c_str());
            else if (key == "B_FIELD") o.d["B"] = std::atof(val().c_str());
            else if (key == "USE_E_FIELD") o.i["elecOn"] = std::atoi(val().c_str());
            else if (key == "E_FIELD") o.d["E"] = std::atof(val().c_str());
            else if (key == "NUM_TARGET_SHEETS") o.i["num_target_sheets"] = std::atoi(val().c_str());
            else if (key == "TARGET_SHEETS") {
                tk.next();
                std::vector<int> v;
                for (long long c = 0; c < o.geti("num_target_sheets"); ++c) v.push_back(std::atoi(tk.next().c_str()) - 1);
                o.iv["target_sheets"] = v;
            } else if (key == "NUM_SHIFT_SHEETS") o.i["num_shift_sheets"] = std::atoi(val().c_str());
            else if (key == "SHIFT_SHEETS") {
                tk.next();
                std::vector<int> v;
                for (long long c = 0; c < o.geti("num_shift_sheets"); ++c) v.push_back(std::atoi(tk.next().c_str()) - 1);
                o.iv["shift_sheets"] = v;
            } else if (key == "GLOBAL_SHIFTS_ON") o.i["global_shifts_on"] = std::atoi(val().c_str());
            else if (key == "GLOBAL_SHIFTS") {
                std::vector<std::vector<double>> gs;
                for (long long c = 0; c < o.geti("num_sheets"); ++c) {
                    tk.next();
                    std::vector<double> v(3);
                    for (int q = 0; q < 3; ++q) v[q] = std::atof(tk.next().c_str());
                    gs.push_back(v);
                }
                o.dm["global_shifts"] = gs;
            }
        }
    }
    if (o.geti("poly_order") % 4 != 0) {
        throw DeckError("poly_order = " + std::to_string(o.geti("poly_order")) +
                        " is NOT divisible 4 (needed for KPM iterative method)");
    }
    if (o.geti("poly_order") < 4) throw DeckError("POLY_ORDER must be at least 4");
    for (int s = 0; s < num_sheets; ++s)
        if (!deck.has_material[s]) throw DeckError("sheet " + std::to_string(s + 1) + " has no MATERIAL");
    if (num_sheets == 0) throw DeckError("NUM_SHEETS missing");
    return deck;
}

void setup_supercell(Deck& deck)
{
    Params& o = deck.opts;
    if (o.geti("boundary_condition") != 1) return;
    // Graphene::lattice (graphene.h:78-80); the supercell is built in each
    // sheet's own frame (locality.cpp:240-257)
    // Materials::lattice(sdata[0].mat): one unit cell for every sheet (locality.cpp:188-196, 318-330)
    const double a = 2.4768;
    const double uc[2][2] = {{1. * a, 0.}, {.5 * a, lattice_s32(deck.sheets[0].mat) * a}};
    if (!o.i.count("supercell_type")) throw DeckError("BOUNDARY_CONDITION = 1 needs SUPERCELL_M_N or SUPERCELL_GRID");
    const long long type = o.geti("supercell_type");
    const int ns = (int)deck.sheets.size();
    if (type == 1) {
        const int M = (int)o.geti("m_supercell"), N = (int)o.geti("n_supercell");
        double theta = std::acos((N * N + 4 * N * M + M * M) / (2.0 * (N * N + N * M + M * M)));
        if (M < N) theta = -theta;
        // H_SUPERCELL_MODIFY: the second group turned by a further 180 degrees, its basis with it
        // (locality.cpp:178-185)
        int h_fac = 1;
        if (o.geti("h_supercell_modify") == 1) {
            theta = theta + M_PI;
            h_fac = -1;
        }
        // sc_groups = {0, 1} for one or two sheets, else the deck's SUPERCELL_GROUPS (locality.cpp:150-170)
        std::vector<int> groups = {0, 1};
        if (ns >= 3) {
            auto it = o.iv.find("sc_groups");
            if (it == o.iv.end() || (int)it->second.size() < ns)
                throw DeckError("SUPERCELL_M_N with more than two sheets needs SUPERCELL_GROUPS");
            groups = it->second;
        }
        for (int s = 0; s < ns; ++s) {
            const int group = groups[s];
            int A[2][2];
            if (group == 0) {
                deck.angles[s] = 0;
                A[0][0] = N; A[0][1] = M; A[1][0] = -M; A[1][1] = M + N;
            } else {
                deck.angles[s] = theta;
                A[0][0] = h_fac * M; A[0][1] = h_fac * N; A[1][0] = -h_fac * N; A[1][1] = h_fac * (M + N);
            }
            // HEX_SUPERCELL_MODIFY: first components of both supercell vectors take the second's too
            // (locality.cpp:222-233)
            if (o.geti("hex_supercell_modify") == 1) {
                A[0][0] = A[0][0] + A[0][1];
                A[1][0] = A[1][0] + A[1][1];
            }
            SheetData& sd = deck.sheets[s];
            for (int r = 0; r < 2; ++r) {
                sd.supercell[r][0] = A[r][0] * uc[0][0] + A[r][1] * uc[1][0];
                sd.supercell[r][1] = A[r][0] * uc[0][1] + A[r][1] * uc[1][1];
                sd.stride[r][0] = A[r][0];
                sd.stride[r][1] = A[r][1];
            }
            if (s == 0) o.dm["supercell"] = {{sd.supercell[0][0], sd.supercell[0][1]}, {sd.supercell[1][0], sd.supercell[1][1]}};
        }
    } else if (type == 2) {
        const int X = (int)o.geti("x_supercell"), Y = (int)o.geti("y_supercell");
        const double sc[2][2] = {{X * uc[0][0], X * uc[0][1]}, {Y * uc[1][0], Y * uc[1][1]}};
        for (int s = 0; s < ns; ++s) {
            const double th = deck.angles[s];
            SheetData& sd = deck.sheets[s];
            for (int r = 0; r < 2; ++r) {
                sd.supercell[r][0] = sc[r][0] * std::cos(th) + sc[r][1] * std::sin(th);
                sd.supercell[r][1] = -sc[r][0] * std::sin(th) + sc[r][1] * std::cos(th);
            }
            sd.stride[0][0] = X; sd.stride[0][1] = 0; sd.stride[1][0] = 0; sd.stride[1][1] = Y;
        }
        o.dm["supercell"] = {{sc[0][0], sc[0][1]}, {sc[1][0], sc[1][1]}};
    } else {
        throw DeckError("supercell type " + std::to_string(type) + " is not on this path");
    }
}

}  // namespace leo2d
