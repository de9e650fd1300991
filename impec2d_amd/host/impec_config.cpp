// impec_config.cpp -- libconfig-syntax parser and the SimParams mapping of
// SimParams::SimParams (SimParams.cpp:45-463); see impec_config.hpp.
#include "impec_config.hpp"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace impec_host {

// ---------------------------------------------------------------------------
// values
// ---------------------------------------------------------------------------
const CfgValue *CfgValue::member(const std::string &name) const {
    if (kind != GROUP) return nullptr;
    for (const auto &m : members)
        if (m.first == name) return &m.second;
    return nullptr;
}

const CfgValue *CfgValue::path(const std::string &dotted) const {
    const CfgValue *v = this;
    size_t a = 0;
    while (v && a <= dotted.size()) {
        const size_t b = dotted.find('.', a);
        const std::string part = dotted.substr(a, b == std::string::npos ? std::string::npos : b - a);
        v = v->member(part);
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return v;
}

bool CfgValue::get(const std::string &name, bool &out) const {
    const CfgValue *v = member(name);
    if (!v || v->kind != BOOL) return false;
    out = v->b;
    return true;
}
bool CfgValue::get(const std::string &name, int &out) const {
    const CfgValue *v = member(name);
    if (!v || v->kind != INT) return false;
    out = (int)v->i;
    return true;
}
bool CfgValue::get(const std::string &name, double &out) const {
    const CfgValue *v = member(name);
    if (!v || (v->kind != FLOAT && v->kind != INT)) return false;
    out = v->kind == FLOAT ? v->f : (double)v->i;
    return true;
}
bool CfgValue::get(const std::string &name, std::string &out) const {
    const CfgValue *v = member(name);
    if (!v || v->kind != STRING) return false;
    out = v->s;
    return true;
}

// ---------------------------------------------------------------------------
// parser: recursive descent over a character stream
// ---------------------------------------------------------------------------
namespace {

struct Parser {
    const std::string &t;
    size_t pos = 0;
    int line = 1;
    std::string err;

    explicit Parser(const std::string &text) : t(text) {}

    bool fail(const std::string &msg) {
        if (err.empty()) err = "line " + std::to_string(line) + ": " + msg;
        return false;
    }
    int peek() const { return pos < t.size() ? (unsigned char)t[pos] : -1; }
    int get() {
        if (pos >= t.size()) return -1;
        const int c = (unsigned char)t[pos++];
        if (c == '\n') line++;
        return c;
    }
    // whitespace and comments
    bool skip() {
        for (;;) {
            const int c = peek();
            if (c == -1) return true;
            if (std::isspace(c)) {
                get();
            } else if (c == '#' || (c == '/' && pos + 1 < t.size() && t[pos + 1] == '/')) {
                while (peek() != -1 && peek() != '\n') get();
            } else if (c == '/' && pos + 1 < t.size() && t[pos + 1] == '*') {
                get();
                get();
                for (;;) {
                    const int d = get();
                    if (d == -1) return fail("unterminated comment");
                    if (d == '*' && peek() == '/') {
                        get();
                        break;
                    }
                }
            } else {
                return true;
            }
        }
    }
    bool name(std::string &out) {
        const int c = peek();
        if (!(std::isalpha(c) || c == '*')) return fail("setting name expected");
        out.clear();
        while (std::isalnum(peek()) || peek() == '_' || peek() == '-' || peek() == '*') out.push_back((char)get());
        return true;
    }
    bool string_lit(std::string &out) {
        out.clear();
        for (;;) {   // adjacent literals concatenate
            if (peek() != '"') break;
            get();
            for (;;) {
                int c = get();
                if (c == -1 || c == '\n') return fail("unterminated string");
                if (c == '"') break;
                if (c == '\\') {
                    c = get();
                    if (c == 'n') c = '\n';
                    else if (c == 't') c = '\t';
                    else if (c == 'r') c = '\r';
                    else if (c == 'f') c = '\f';
                    else if (c != '\\' && c != '"') return fail("unknown escape in string");
                }
                out.push_back((char)c);
            }
            if (!skip()) return false;
        }
        return true;
    }
    bool scalar(CfgValue &v) {
        const int c = peek();
        if (c == '"') {
            v.kind = CfgValue::STRING;
            return string_lit(v.s);
        }
        // word: boolean or number
        size_t b = pos;
        while (pos < t.size() && (std::isalnum((unsigned char)t[pos]) || t[pos] == '.' || t[pos] == '+' ||
                                  t[pos] == '-' || t[pos] == '_'))
            pos++;
        std::string w = t.substr(b, pos - b);
        if (w.empty()) return fail("value expected");
        std::string lw;
        for (char ch : w) lw.push_back((char)std::tolower((unsigned char)ch));
        if (lw == "true" || lw == "false") {
            v.kind = CfgValue::BOOL;
            v.b = lw == "true";
            return true;
        }
        std::string num = w;
        while (!num.empty() && (num.back() == 'L' || num.back() == 'l')) num.pop_back();   // 64-bit suffix
        const bool hex = num.size() > 2 && num[0] == '0' && (num[1] == 'x' || num[1] == 'X');
        const bool isfloat = !hex && num.find_first_of(".eE") != std::string::npos;
        char *end = nullptr;
        errno = 0;
        if (isfloat) {
            v.kind = CfgValue::FLOAT;
            v.f = std::strtod(num.c_str(), &end);
        } else {
            v.kind = CfgValue::INT;
            v.i = std::strtoll(num.c_str(), &end, hex ? 16 : 10);
        }
        if (end == num.c_str() || *end != '\0' || errno == ERANGE) return fail("bad value '" + w + "'");
        return true;
    }
    bool value(CfgValue &v) {
        if (!skip()) return false;
        const int c = peek();
        if (c == '{') {
            get();
            v.kind = CfgValue::GROUP;
            if (!settings(v, '}')) return false;
            get();
            return true;
        }
        if (c == '[' || c == '(') {
            const int close = c == '[' ? ']' : ')';
            get();
            v.kind = c == '[' ? CfgValue::ARRAY : CfgValue::LIST;
            if (!skip()) return false;
            if (peek() == close) {
                get();
                return true;
            }
            for (;;) {
                CfgValue e;
                if (!value(e)) return false;
                if (v.kind == CfgValue::ARRAY && (e.kind == CfgValue::GROUP || e.kind == CfgValue::ARRAY ||
                                                  e.kind == CfgValue::LIST))
                    return fail("arrays hold scalars only");
                v.items.push_back(e);
                if (!skip()) return false;
                const int d = get();
                if (d == close) break;
                if (d != ',') return fail(std::string("',' or '") + (char)close + "' expected");
            }
            if (v.kind == CfgValue::ARRAY)
                for (const CfgValue &e : v.items)
                    if (e.kind != v.items[0].kind && !(e.kind == CfgValue::INT && v.items[0].kind == CfgValue::FLOAT) &&
                        !(e.kind == CfgValue::FLOAT && v.items[0].kind == CfgValue::INT))
                        return fail("array elements of different types");
            return true;
        }
        return scalar(v);
    }
    // settings until `close` (a '}' or the end of the text, -1)
    bool settings(CfgValue &g, int close) {
        for (;;) {
            if (!skip()) return false;
            const int c = peek();
            if (c == close) return true;
            if (c == -1) return fail("'}' expected before the end of the file");
            std::string n;
            if (!name(n)) return false;
            if (!skip()) return false;
            const int sep = get();
            if (sep != '=' && sep != ':') return fail("'=' or ':' expected after '" + n + "'");
            CfgValue v;
            if (!value(v)) return false;
            if (g.member(n)) return fail("duplicate setting '" + n + "'");
            g.members.emplace_back(n, v);
            if (!skip()) return false;
            if (peek() == ';' || peek() == ',') get();
        }
    }
};

}  // namespace

bool ParseConfig(const std::string &text, CfgValue &root, std::string &err) {
    Parser p(text);
    root = CfgValue();
    root.kind = CfgValue::GROUP;
    if (!p.settings(root, -1)) {
        err = p.err;
        return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// SimParams::SimParams (SimParams.cpp:45-463): the same required settings and
// the same string options; every failure the reference reports with exit(1)
// is returned as an error
// ---------------------------------------------------------------------------
bool ReadConfigText(const std::string &text, SimParams &sp, std::string &err) {
    CfgValue root;
    if (!ParseConfig(text, root, err)) return false;
    const CfgValue *gen = root.path("application.general"), *ell = root.path("application.elliptic");
    const CfgValue *hyp = root.path("application.hyperbolic"), *wel = root.path("application.wells");
    const CfgValue *par = root.path("application.parabolic");
    if (!gen || !ell || !hyp || !wel || !par) {
        err = "missing section (application.general / elliptic / hyperbolic / wells / parabolic)";
        return false;
    }
    auto need = [&](bool ok, const char *what) {
        if (!ok && err.empty()) err = std::string("Could not find one or more items in section <<") + what + ">>!";
        return ok;
    };
    err.clear();
    int gh = 2;
    bool gnuplot = false, levelset = false;
    std::string sstype, ssync;
    if (!need(gen->get("x_cells", sp.N1) && gen->get("y_cells", sp.N2) && gen->get("num_ghostcells", gh) &&
                  gen->get("x_lb", sp.x_lb) && gen->get("x_ub", sp.x_ub) && gen->get("y_ub", sp.y_ub) &&
                  gen->get("y_lb", sp.y_lb) && gen->get("starttime", sp.tstart) && gen->get("endtime", sp.tfinal) &&
                  gen->get("I", sp.I) && gen->get("log_to_file", sp.log_to_file) &&
                  gen->get("results_dir", sp.results_dir) &&
                  gen->get("check_mass_conservation", sp.check_mass_conservation) &&
                  gen->get("check_noflow_boundary", sp.check_noflow_boundary) && gen->get("is_test", sp.is_test) &&
                  gen->get("source_type", sstype) && gen->get("gnuplot", gnuplot) &&
                  gen->get("check_divergence", sp.check_divergence) && gen->get("levelset", levelset) &&
                  gen->get("syncmode", ssync),
              "general"))
        return false;
    (void)gnuplot;
    sp.c_bounds = gh;
    if (ssync != "sync" && ssync != "async") {
        err = "Config::Synchronization: protocoll is not implemented!";
        return false;
    }
    if (sp.check_mass_conservation && !gen->get("max_mass_frac_error", sp.max_mass_frac_error)) {
        err = "Could not find one or more options in section <<general>>! You need to specify <<max_mass_frac_error>>";
        return false;
    }
    if (sp.log_to_file &&
        !(gen->get("log_skip_steps", sp.log_skip_steps) && gen->get("time_matching", sp.time_matching) &&
          gen->get("log_mass_conservation", sp.log_mass_conservation) && gen->get("log_dir", sp.log_dir))) {
        err = "Could not find one or more logging-options in section <<general>>!";
        return false;
    }
    if (sstype == "parallel") sp.stype = PARALLEL;
    else if (sstype == "diagonal") sp.stype = DIAGONAL;
    else if (sstype == "horizontal") {
        err = "source_type \"horizontal\" is not supported by this solver (whole-column pins, not a Neumann problem)";
        return false;
    } else {
        err = "Config::General: source type is not valid!";
        return false;
    }
    std::string smethod, sbtype, savg;
    double bval = 0.;
    if (!need(ell->get("ell_method", smethod) && ell->get("bound_type", sbtype) && ell->get("eps", sp.eps) &&
                  ell->get("mu_res", sp.mu_res) && ell->get("mu_inj", sp.mu_inj) &&
                  ell->get("const_bound_val", bval) && ell->get("average_procedure", savg),
              "elliptic"))
        return false;
    if (sbtype != "neumann") {
        err = "Config::Elliptic: Boundary type is not supported at the moment!";
        return false;
    }
    if (smethod != "direct" && smethod != "cg" && smethod != "mg" && smethod != "cg_cuda") {
        err = "Config::Elliptic: Method is unknown!";
        return false;
    }
    if (savg == "harmonic") sp.avgproc = HARMONIC;
    else if (savg == "mean_concentration") sp.avgproc = MEANCONC;
    else {
        err = "Config::Elliptic: Averaging procedure is unknown!";
        return false;
    }
    std::string shm, ssplit;
    if (!need(hyp->get("hyper_method", shm) && hyp->get("porosity", sp.porosity) && hyp->get("cfl", sp.cfl) &&
                  hyp->get("inj_fluid_conc", sp.inj_fluid_conc) && hyp->get("nu", sp.nu) &&
                  hyp->get("diffusion", sp.diffusion) && hyp->get("alpha_l", sp.alpha_l) &&
                  hyp->get("alpha_t", sp.alpha_t) && hyp->get("dt_diff_limiter", sp.dt_diff_limit) &&
                  hyp->get("dimensional_splitting", ssplit),
              "hyperbolic"))
        return false;
    if (shm != "waf_split" && shm != "waf" && shm != "ctu" && shm != "muscl" && shm != "upwind_split") {
        err = "Config::Hyperbolic: Method is unknown!";
        return false;
    }
    sp.hmethod = MUSCL;   // Impec2DSolver::Initialize always builds a MUSCLSplitSolver (Impec2DSolver.cpp:131)
    if (ssplit == "alternating_sweeps") sp.dim_splitting = ALTERNATING_SWEEPS;
    else if (ssplit == "x_sweep_first") sp.dim_splitting = X_SWEEP_FIRST;
    else if (ssplit == "y_sweep_first") sp.dim_splitting = Y_SWEEP_FIRST;
    else if (ssplit == "strang") sp.dim_splitting = STRANG;
    else if (ssplit == "parallel") sp.dim_splitting = SYMMETRIC_PARALLEL;
    else {
        err = "Config::Hyperbolic: Dimenstional Splitting specification is unknown!";
        return false;
    }
    const CfgValue *cl = gen->member("log_counter_list"), *tl = gen->member("log_time_list");
    if (!cl || !tl || (cl->kind != CfgValue::ARRAY && cl->kind != CfgValue::LIST) ||
        (tl->kind != CfgValue::ARRAY && tl->kind != CfgValue::LIST)) {
        err = "log_counter_list and log_time_list must be arrays";
        return false;
    }
    sp.counters.clear();
    for (const CfgValue &e : cl->items) {
        if (e.kind != CfgValue::INT) {
            err = "log_counter_list holds integers";
            return false;
        }
        sp.counters.push_back((int)e.i);
    }
    sp.times.clear();
    for (const CfgValue &e : tl->items) {
        if (e.kind != CfgValue::FLOAT && e.kind != CfgValue::INT) {
            err = "log_time_list holds numbers";
            return false;
        }
        sp.times.push_back(e.kind == CfgValue::FLOAT ? e.f : (double)e.i);
    }
    if (sp.is_test) {   // SimParams.cpp:252-316
        const CfgValue *ts = root.path("application.test");
        std::string comp, sprofile;
        int ihypnorm = 0;
        if (!(ts && ts->get("component", comp) && ts->get("testno", sp.testno) &&
              ts->get("num_result_file", sp.num_result_file) && ts->get("ex_result_file", sp.ex_result_file) &&
              ts->get("norm_file", sp.norm_file) && ts->get("initial_x_cells", sp.init_x_cells) &&
              ts->get("initial_y_cells", sp.init_y_cells) && ts->get("num_u_file", sp.num_u_file) &&
              ts->get("ex_u_file", sp.ex_u_file) && ts->get("num_v_file", sp.num_v_file) &&
              ts->get("ex_v_file", sp.ex_v_file) && ts->get("levels", sp.levels) &&
              ts->get("profile", sprofile) && ts->get("hyp_norm", ihypnorm))) {
            err = "Could not find one or more items in section <<test>>!";
            return false;
        }
        if (comp == "elliptic") sp.test_component = ELLIPTIC;
        else if (comp == "hyperbolic") sp.test_component = HYPERBOLIC;
        else {
            err = "TestComponent> Config::General: source type is not valid!";
            return false;
        }
        if (sprofile == "sine") sp.hyp_test = SINE;
        else if (sprofile == "tophat") sp.hyp_test = TOPHAT;
        else if (sprofile == "gaussbell") sp.hyp_test = GAUSS;
        else if (sprofile == "horizontal_sine") sp.hyp_test = HSINE;
        else {
            err = "TestComponent> Config:: Test profile is not valid!";
            return false;
        }
        if (ihypnorm == 1) sp.hyp_norm = L1;
        else if (ihypnorm == 2) sp.hyp_norm = L2;
        else {
            err = "TestComponent> Config:: Norm is not supported!";
            return false;
        }
        if (sp.init_x_cells < 2 || sp.init_y_cells < 2 || sp.levels < 1) {
            err = "TestComponent> Config:: initial_x_cells, initial_y_cells and levels must be positive";
            return false;
        }
    }
    std::string sti;
    double ti_eps = 0.;
    if (!(hyp->get("time_integration", sti) && hyp->get("ti_eps", ti_eps))) {
        err = "Could not find property <<time_integration>> or <<ti_eps>>!";
        return false;
    }
    std::string swp, smode;
    double qf = 0., qr = 0., ww = 0., qh = 0.;
    int fcw = 0, frw = 0;
    if (!need(wel->get("well_profile", swp) && wel->get("well_rc", sp.well_rc) && wel->get("q_freq", qf) &&
                  wel->get("q_rand_max", qr) && wel->get("front_const_width", fcw) && wel->get("well_width", ww) &&
                  wel->get("front_rand_width", frw) && wel->get("q_high_multiple", qh) && wel->get("mode", smode),
              "wells"))
        return false;
    if (smode != "pulsing" && smode != "linear" && smode != "constant") {
        err = "Unknown injection mode: " + smode;
        return false;
    }
    if (swp == "circle") sp.well_profile = CIRCULAR;
    else if (swp == "gauss") sp.well_profile = GAUSSIAN;
    else if (swp == "one_cell") sp.well_profile = ONECELL;
    else {
        err = "Unknown profile <<wells>>!";
        return false;
    }
    std::string spm;
    if (!need(par->get("para_method", spm) && par->get("eps", sp.para_eps), "parabolic")) return false;
    if (spm == "crank-nicolson") sp.ptype = CRANKNICOLSON;
    else if (spm == "explicit") sp.ptype = EXPLICIT;
    else {
        err = "Unknown parabolic solver <<" + spm + ">>!";
        return false;
    }
    if (sp.ptype == CRANKNICOLSON && smethod != "mg") {
        err = "SimParams::InvalidOptions> The Crank-Nicolson scheme requires a Multigrid-Solver. Change ell_method to MG!";
        return false;
    }
    if (sp.diffusion && sp.ptype != CRANKNICOLSON) {
        err = "explicit dispersion time stepping is not supported by this solver (use crank-nicolson)";
        return false;
    }
    sp.Finalize();
    if (sp.stype == PARALLEL && (sp.dx != sp.dy || sp.N1 != sp.N2)) {
        err = "ERROR: This resolution is not supported for the parallel configuration. The module ParallelProjector "
              "assumes square grid blocks and equal resolution in both dimensions.";
        return false;
    }
    return true;
}

bool ReadConfigFile(const std::string &path, SimParams &sp, std::string &err) {
    std::ifstream f(path.c_str());
    if (!f.is_open()) {
        err = "I/O error while reading file.";
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return ReadConfigText(ss.str(), sp, err);
}

}  // namespace impec_host
