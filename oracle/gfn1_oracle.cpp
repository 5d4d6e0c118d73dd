// GFN1-xTB oracle: a plain FP64 CPU restatement of the reference's non-periodic
// SCC energy + analytic gradient path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// (or the timed CPU baseline).  Nothing in the product links or calls it.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference repository root).  The linear algebra (symmetric eigensolver,
// GEMM) is this file's own textbook code: Householder tridiagonalisation plus
// implicit symmetric QR with Wilkinson shifts (Golub & Van Loan, Alg. 8.3.1 /
// 8.3.3); the reference calls faer 0.24 for these, whose bit-level results are
// not reproducible here (parity at that boundary is pinned only through
// invariants and LAPACK comparisons in tests/).
//
// Built by oracle/Makefile into oracle/_build/libgfn1_oracle.so (C ABI below).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double PI = 3.14159265358979323846;
// Legacy model-side eV->Eh (basis.rs:18); not the CODATA unit constant.
const double EV2H = 1.0 / 27.21138505;
const double KB = 3.1668115634556e-6;          // constants.rs:26
const double AA2AU = 1.8897261246257702;        // system.rs:8

// ---------------------------------------------------------------- data tables
// data_tables.rs:14-21 xTB atomic radii (Angstrom), index = Z.
const double ATOMIC_RADII_AA[87] = {
    0.0, 0.32, 0.37, 1.3, 0.99, 0.84, 0.75, 0.71, 0.64, 0.6, 0.62, 1.6, 1.4, 1.24, 1.14, 1.09,
    1.04, 1.0, 1.01, 2.0, 1.74, 1.59, 1.48, 1.44, 1.3, 1.29, 1.24, 1.18, 1.17, 1.22, 1.2, 1.23,
    1.2, 1.2, 1.18, 1.17, 1.16, 2.15, 1.9, 1.76, 1.64, 1.56, 1.46, 1.38, 1.36, 1.34, 1.3, 1.36,
    1.4, 1.42, 1.4, 1.4, 1.37, 1.36, 1.36, 2.38, 2.06, 1.94, 1.84, 1.9, 1.88, 1.86, 1.85, 1.83,
    1.82, 1.81, 1.8, 1.79, 1.77, 1.77, 1.78, 1.74, 1.64, 1.58, 1.5, 1.41, 1.36, 1.32, 1.3, 1.3,
    1.32, 1.44, 1.45, 1.5, 1.42, 1.48, 1.46};
// data_tables.rs:29-37 Pyykko-Atsumi 2009 covalent radii (mctc-lib), Angstrom.
const double COV_RADII_AA[87] = {
    0.0, 0.32, 0.46, 1.20, 0.94, 0.77, 0.75, 0.71, 0.63, 0.64, 0.67, 1.40, 1.25, 1.13, 1.04, 1.10,
    1.02, 0.99, 0.96, 1.76, 1.54, 1.33, 1.22, 1.21, 1.10, 1.07, 1.04, 1.00, 0.99, 1.01, 1.09, 1.12,
    1.09, 1.15, 1.10, 1.14, 1.17, 1.89, 1.67, 1.47, 1.39, 1.32, 1.24, 1.15, 1.13, 1.13, 1.08, 1.15,
    1.23, 1.28, 1.26, 1.26, 1.23, 1.32, 1.31, 2.09, 1.76, 1.62, 1.47, 1.58, 1.57, 1.56, 1.55, 1.51,
    1.52, 1.51, 1.50, 1.49, 1.49, 1.48, 1.53, 1.46, 1.37, 1.31, 1.23, 1.18, 1.16, 1.11, 1.12, 1.13,
    1.32, 1.30, 1.30, 1.36, 1.31, 1.38, 1.42};
// data_tables.rs:60-69 Pauling electronegativities (Z <= 86).
const double PAULING_EN[87] = {
    0.0, 2.20, 3.00, 0.98, 1.57, 2.04, 2.55, 3.04, 3.44, 3.98, 4.50, 0.93, 1.31, 1.61, 1.90, 2.19,
    2.58, 3.16, 3.50, 0.82, 1.00, 1.36, 1.54, 1.63, 1.66, 1.55, 1.83, 1.88, 1.91, 1.90, 1.65, 1.81,
    2.01, 2.18, 2.55, 2.96, 3.00, 0.82, 0.95, 1.22, 1.33, 1.60, 2.16, 1.90, 2.20, 2.28, 2.20, 1.93,
    1.69, 1.78, 1.96, 2.05, 2.10, 2.66, 2.60, 0.79, 0.89, 1.10, 1.12, 1.13, 1.14, 1.15, 1.17, 1.18,
    1.20, 1.21, 1.22, 1.23, 1.24, 1.25, 1.26, 1.27, 1.30, 1.50, 2.36, 1.90, 2.20, 2.20, 2.28, 2.54,
    2.00, 1.62, 2.33, 2.02, 2.00, 2.20, 2.20};

// GFN1 reference occupations per angular momentum (s, p, d) (basis.rs:443-531).
double ref_occ(int z, int l) {
    if (z < 1 || z > 86 || l > 2) return 0.0;
    static double tab[87][3];
    static bool init = false;
    if (!init) {
        for (int i = 0; i <= 86; ++i) tab[i][0] = tab[i][1] = tab[i][2] = 0.0;
        // main-group valence s/p fillings by period position; d counts for TMs
        auto set = [&](int zz, double s, double p, double d) { tab[zz][0] = s; tab[zz][1] = p; tab[zz][2] = d; };
        set(1, 1, 0, 0); set(2, 2, 0, 0);
        const int pblock_start[5] = {3, 11, 19, 37, 55};
        (void)pblock_start;
        // period 2 and 3: alkali, alkaline earth, then p-block 1..6
        for (int base : {3, 11}) {
            set(base, 1, 0, 0); set(base + 1, 2, 0, 0);
            for (int k = 1; k <= 6; ++k) set(base + 1 + k, 2, k, 0);
        }
        // periods 4, 5: s-block, d-block (d = 1..9 then Zn/Cd d = 0), p-block
        for (int base : {19, 37}) {
            set(base, 1, 0, 0); set(base + 1, 2, 0, 0);
            for (int k = 1; k <= 9; ++k) set(base + 1 + k, 2, 0, k);
            set(base + 11, 2, 0, 0);
            for (int k = 1; k <= 6; ++k) set(base + 11 + k, 2, k, 0);
        }
        // period 6: Cs, Ba, La..Lu (d = 1, f-in-core), Hf..Au (d = 2..9), Hg, Tl..Rn
        set(55, 1, 0, 0); set(56, 2, 0, 0);
        for (int zz = 57; zz <= 71; ++zz) set(zz, 2, 0, 1);
        for (int k = 2; k <= 9; ++k) set(70 + k, 2, 0, k);
        set(80, 2, 0, 0);
        for (int k = 1; k <= 6; ++k) set(80 + k, 2, k, 0);
        init = true;
    }
    return tab[z][l];
}

// basis.rs:296-302: element class for the CN shift of the self energies.
int gfn1_kind(int z) {
    if (z == 1 || z == 2 || (z >= 6 && z <= 10) || (z >= 14 && z <= 18) || (z >= 32 && z <= 36) ||
        (z >= 50 && z <= 54) || (z >= 82 && z <= 86))
        return 1;
    if (z >= 21 && z <= 24) return 2;
    return 0;
}

struct Err {
    int code;
    std::string msg;
};

// ------------------------------------------------------------------ params
struct ShellParam {
    int n, l;
    double level_ev, slater, poly, lpar;
};
struct ElementParam {
    bool present = false;
    std::vector<ShellParam> shells;
    double gam = 0, gam3 = 0, repa = 0, repb = 0, cxb = 0;
};
struct Params {
    std::map<std::string, double> glob;
    std::map<std::pair<int, int>, double> pair;
    ElementParam el[87];
    double global(const std::string& k, double def) const {
        auto it = glob.find(k);
        return it == glob.end() ? def : it->second;
    }
    // params.rs:361-368 and :694-711
    double pair_scaling(int a, int b) const {
        auto key = a <= b ? std::make_pair(a, b) : std::make_pair(b, a);
        auto it = pair.find(key);
        if (it != pair.end()) return it->second;
        auto drow = [](int z) -> double {
            if (z >= 21 && z < 30) return 1.1;
            if (z >= 39 && z < 48) return 1.2;
            if (z >= 57 && z < 80) return 1.2;
            return -1.0;
        };
        double da = drow(a), db = drow(b);
        if (da > 0 && db > 0) return 0.5 * (da + db);
        return 1.0;
    }
    // params.rs:369-376
    double k_shell(int l) const {
        if (l == 0) return global("ks", 1.85);
        if (l == 1) return global("kp", 2.25);
        return global("kd", 2.00);
    }
    // params.rs:153-160
    double shell_hardness(int z, int l) const {
        double lpar = 0.0;
        for (auto& s : el[z].shells)
            if (s.l == l) { lpar = s.lpar; break; }
        return el[z].gam * (1.0 + 0.1 * lpar);
    }
};

std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

// params.rs:272-341 (parser) and :826-872 (element records)
bool parse_params(const std::string& path, Params& p, Err& err) {
    std::ifstream f(path);
    if (!f) { err = {1, "cannot open parameter file " + path}; return false; }
    enum { NONE, INFO, GLOB, PAIR, ELEM } sec = NONE;
    int z = 0;
    std::map<std::string, std::vector<double>> raw;
    std::string ao;
    auto finish = [&]() {
        if (sec != ELEM || z <= 0 || z > 86) return;
        ElementParam& e = p.el[z];
        e.present = true;
        e.shells.clear();
        std::vector<std::pair<int, int>> labels;
        for (size_t i = 0; i < ao.size();) {
            if (std::isspace((unsigned char)ao[i])) { ++i; continue; }
            int n = 0;
            while (i < ao.size() && std::isdigit((unsigned char)ao[i])) n = n * 10 + (ao[i++] - '0');
            char lc = (char)std::tolower((unsigned char)ao[i++]);
            int l = lc == 's' ? 0 : lc == 'p' ? 1 : lc == 'd' ? 2 : lc == 'f' ? 3 : 4;
            labels.push_back({n, l});
        }
        auto get = [&](const std::string& k, double d) {
            auto it = raw.find(k);
            return (it == raw.end() || it->second.empty()) ? d : it->second[0];
        };
        const char* suf = "SPDFG";
        for (size_t i = 0; i < labels.size(); ++i) {
            ShellParam sp;
            sp.n = labels[i].first;
            sp.l = labels[i].second;
            sp.level_ev = raw["LEV"][i];
            sp.slater = raw["EXP"][i];
            sp.poly = get(std::string("POLY") + suf[sp.l], 0.0);
            sp.lpar = get(std::string("LPAR") + suf[sp.l], 0.0);
            e.shells.push_back(sp);
        }
        e.gam = get("GAM", 0.0);
        e.gam3 = get("GAM3", 0.0);
        e.repa = get("REPA", 0.0);
        e.repb = get("REPB", 0.0);
        e.cxb = get("CXB", 0.0);
        raw.clear();
        ao.clear();
    };
    std::string line;
    while (std::getline(f, line)) {
        std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        if (t.rfind("$info", 0) == 0) { finish(); sec = INFO; continue; }
        if (t.rfind("$globpar", 0) == 0) { finish(); sec = GLOB; continue; }
        if (t.rfind("$pairpar", 0) == 0) { finish(); sec = PAIR; continue; }
        if (t.rfind("$Z=", 0) == 0) {
            finish();
            z = std::atoi(t.c_str() + 3);
            sec = ELEM;
            continue;
        }
        if (t.rfind("$end", 0) == 0) { finish(); sec = NONE; continue; }
        std::istringstream ss(t);
        if (sec == GLOB) {
            std::string k; double v; ss >> k >> v; p.glob[lower(k)] = v;
        } else if (sec == PAIR) {
            int a, b; double v; ss >> a >> b >> v;
            p.pair[a <= b ? std::make_pair(a, b) : std::make_pair(b, a)] = v;
        } else if (sec == ELEM) {
            if (t.rfind("ao=", 0) == 0) { ao = trim(t.substr(3)); continue; }
            size_t eq = t.find('=');
            if (eq == std::string::npos) continue;
            std::string k = trim(t.substr(0, eq));
            for (auto& c : k) c = (char)std::toupper((unsigned char)c);
            std::istringstream vs(t.substr(eq + 1));
            std::vector<double> vals; double v;
            while (vs >> v) vals.push_back(v);
            raw[k] = vals;
        }
    }
    finish();
    return true;
}

// ------------------------------------------------------------------ STO-nG
struct StoTable {
    std::map<std::pair<int, int>, std::vector<double>> rec;  // (type, ng) -> alpha..., coeff...
    std::vector<double> s6, p6;
};

bool load_sto(const std::string& path, StoTable& t, Err& err) {
    std::ifstream f(path);
    if (!f) { err = {1, "cannot open " + path}; return false; }
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string tag; int ng;
        ss >> tag >> ng;
        std::vector<double> v; double x;
        while (ss >> x) v.push_back(x);
        if (tag == "6s") t.s6 = v;
        else if (tag == "6p") t.p6 = v;
        else t.rec[{std::atoi(tag.c_str()), ng}] = v;
    }
    return true;
}

struct Prim {
    double a, c;
};

// sto.rs:790-881 slater_to_gauss (normalised primitives).
bool slater_to_gauss(const StoTable& t, int ng, int n, int l, double zeta, std::vector<Prim>& out, Err& err) {
    if (ng < 1 || ng > 6) { err = {2, "unsupported STO-nG"}; return false; }
    if ((n > 5 || n <= l) && !(n == 6 && ng == 6 && l <= 1)) { err = {2, "unsupported STO-nG shell"}; return false; }
    const std::vector<double>* v;
    if (n == 6 && ng == 6) {
        v = l == 0 ? &t.s6 : &t.p6;
    } else {
        int ityp = l == 0 ? n : l == 1 ? 4 + n : l == 2 ? 7 + n : l == 3 ? 9 + n : 10 + n;
        auto it = t.rec.find({ityp, ng});
        if (it == t.rec.end()) { err = {2, "no STO table entry"}; return false; }
        v = &it->second;
    }
    static const double dfact[8] = {1, 1, 3, 15, 105, 945, 10395, 135135};
    out.clear();
    for (int i = 0; i < ng; ++i) {
        double a = (*v)[i] * zeta * zeta;
        double c = (*v)[ng + i];
        c *= std::pow(2.0 / PI * a, 0.75) * std::pow(std::sqrt(4.0 * a), l) / std::sqrt(dfact[l]);
        out.push_back({a, c});
    }
    return true;
}

// basis.rs:318-344: Gram-Schmidt of a same-centre same-l shell against the first one.
void orthogonalize(const std::vector<Prim>& ref, std::vector<Prim>& sh) {
    double ov = 0.0;
    for (auto& a : ref)
        for (auto& b : sh) {
            double e = a.a + b.a;
            ov += a.c * b.c * std::pow(std::sqrt(PI / e), 3);
        }
    for (auto& a : ref) sh.push_back({a.a, -ov * a.c});
    double nrm = 0.0;
    for (auto& a : sh)
        for (auto& b : sh) {
            double e = a.a + b.a;
            nrm += a.c * b.c * std::pow(std::sqrt(PI / e), 3);
        }
    double s = 1.0 / std::sqrt(nrm);
    for (auto& a : sh) a.c *= s;
}

// ------------------------------------------------------------------ basis
struct Comp {
    int px, py, pz;
    double c;
};
struct Shell {
    int atom, z, l, nao, first_ao;
    bool valence;
    double hdiag, kcn, poly, ref_occ, hardness, gam3, rad;
    std::vector<Prim> prims;
};
struct AO {
    int atom, shell;
    std::vector<Comp> comps;
};
struct Basis {
    std::vector<Shell> sh;
    std::vector<AO> ao;
    double nref = 0.0;
};

// basis.rs:382-441 real spherical components (s; px py pz; d x2-y2, z2, xy, xz, yz)
std::vector<std::vector<Comp>> spherical(int l) {
    double s3 = std::sqrt(3.0), h3 = s3 / 2.0;
    if (l == 0) return {{{0, 0, 0, 1.0}}};
    if (l == 1) return {{{1, 0, 0, 1.0}}, {{0, 1, 0, 1.0}}, {{0, 0, 1, 1.0}}};
    return {{{2, 0, 0, h3}, {0, 2, 0, -h3}},
            {{2, 0, 0, 0.5}, {0, 2, 0, 0.5}, {0, 0, 2, -1.0}},
            {{1, 1, 0, s3}},
            {{1, 0, 1, s3}},
            {{0, 1, 1, s3}}};
}

// basis.rs:251-279
int nprim_rule(int z, int n, int l) {
    if (z <= 2) {
        if (l == 0) return n == 1 ? 4 : 3;
        if (l == 1) return 3;
        return 4;
    }
    if (l <= 1) return 6;
    return 4;
}

struct Ctx {
    Params p;
    StoTable sto;
    // D3
    int d3_max_elem = 0, d3_max_ref = 0;
    std::vector<int> d3_nref;
    std::vector<double> d3_refcn, d3_r4r2, d3_c6;
};

// basis.rs:121-233 BasisSet::build
bool build_basis(const Ctx& ctx, int nat, const int* z, int nprim_opt, Basis& b, Err& err) {
    b = Basis();
    for (int a = 0; a < nat; ++a) {
        int zz = z[a];
        if (zz < 1 || zz > 86 || !ctx.p.el[zz].present) { err = {3, "missing element"}; return false; }
        const ElementParam& e = ctx.p.el[zz];
        bool seen[5] = {false, false, false, false, false};
        std::vector<Prim> same_l_ref[5];
        bool has_ref[5] = {false, false, false, false, false};
        for (auto& sp : e.shells) {
            if (sp.l > 2) { err = {2, "only s, p, d shells are implemented"}; return false; }
            Shell s;
            s.atom = a;
            s.z = zz;
            s.l = sp.l;
            s.valence = !seen[sp.l];
            seen[sp.l] = true;
            int ng = nprim_opt == 0 ? nprim_rule(zz, sp.n, sp.l) : nprim_opt;
            if (!slater_to_gauss(ctx.sto, ng, sp.n, sp.l, sp.slater, s.prims, err)) return false;
            if (has_ref[sp.l]) {
                orthogonalize(same_l_ref[sp.l], s.prims);
            } else {
                same_l_ref[sp.l] = s.prims;
                has_ref[sp.l] = true;
            }
            s.ref_occ = s.valence ? ref_occ(zz, sp.l) : 0.0;
            s.hdiag = sp.level_ev * EV2H;
            // basis.rs:281-293 CN shift
            int kind = gfn1_kind(zz);
            double fac = 0.0;
            if (kind == 1 || kind == 2) {
                if (sp.l == 0) fac = ctx.p.global("cns", 0.6);
                else if (sp.l == 1) fac = ctx.p.global("cnp", -0.3);
                else if (sp.l == 2) fac = kind == 1 ? ctx.p.global("cnd1", -0.5) : ctx.p.global("cnd2", 0.5);
            }
            s.kcn = -sp.level_ev * EV2H * fac * 0.01;
            s.poly = sp.poly * 0.01;
            s.hardness = ctx.p.shell_hardness(zz, sp.l);
            s.gam3 = e.gam3 * 0.1;
            s.rad = ATOMIC_RADII_AA[zz] * AA2AU;
            s.first_ao = (int)b.ao.size();
            auto comps = spherical(sp.l);
            s.nao = (int)comps.size();
            for (auto& c : comps) b.ao.push_back({a, (int)b.sh.size(), c});
            b.sh.push_back(s);
            b.nref += s.ref_occ;
        }
    }
    return true;
}

// ------------------------------------------------------------------ integrals
const double PRIM_EPS = 1e-13;  // integrals.rs:17

double binom(int n, int k) {
    double r = 1.0;
    for (int i = 0; i < k; ++i) r = r * (n - i) / (i + 1);
    return r;
}

// integrals.rs:952-976 moment_1d_no_kab (binomial expansion about the product centre)
double moment1d(int ia, int jb, double alpha, double beta, double a, double b, double p, const double* g) {
    double P = (alpha * a + beta * b) / p;
    double pa = P - a, pb = P - b;
    double sum = 0.0;
    for (int u = 0; u <= ia; ++u) {
        double ca = binom(ia, u) * std::pow(pa, ia - u);
        for (int v = 0; v <= jb; ++v) {
            double cb = binom(jb, v) * std::pow(pb, jb - v);
            if ((u + v) % 2 == 0) sum += ca * cb * g[(u + v) / 2];
        }
    }
    return sum;
}

struct V3 {
    double x, y, z;
};

double comp_abs(const AO& a) {
    double s = 0;
    for (auto& c : a.comps) s += std::fabs(c.c);
    return s;
}
int degree(const AO& a) { return a.comps[0].px + a.comps[0].py + a.comps[0].pz; }

// integrals.rs:30-45 primitive_pair_bound
double prim_bound(double pref, double cs, double al, double be, double r, double r2, int order) {
    double p = al + be;
    double kab = std::exp(-al * be / p * r2);
    double spread = std::sqrt(1.0 / p);
    double gpref = (PI / p) * std::sqrt(PI / p);
    double poly = std::pow(1.0 + r + spread, order);
    return std::fabs(pref) * cs * kab * gpref * poly;
}

// integrals.rs:290-360 contracted_pair, overlap component only.
double overlap_pair(const AO& a, const AO& b, const std::vector<Prim>& pa_, const std::vector<Prim>& pb_,
                    V3 ca, V3 cb) {
    double dx = ca.x - cb.x, dy = ca.y - cb.y, dz = ca.z - cb.z;
    double r2 = dx * dx + dy * dy + dz * dz, r = std::sqrt(r2);
    double cs = comp_abs(a) * comp_abs(b);
    int order = degree(a) + degree(b) + 2;
    double s = 0.0;
    for (auto& pa : pa_)
        for (auto& pb : pb_) {
            double pref = pa.c * pb.c, al = pa.a, be = pb.a, p = al + be;
            if (prim_bound(pref, cs, al, be, r, r2, order) < PRIM_EPS) continue;
            double kab = std::exp(-al * be / p * r2);
            double g[8];
            g[0] = std::sqrt(PI / p);
            for (int m = 1; m < 8; ++m) g[m] = g[m - 1] * (m - 0.5) / p;
            for (auto& A : a.comps)
                for (auto& B : b.comps) {
                    double cf = pref * A.c * B.c * kab;
                    s += cf * moment1d(A.px, B.px, al, be, ca.x, cb.x, p, g) *
                         moment1d(A.py, B.py, al, be, ca.y, cb.y, p, g) *
                         moment1d(A.pz, B.pz, al, be, ca.z, cb.z, p, g);
                }
        }
    return s;
}

// integrals.rs:481-537 + :782-868: overlap and its bra/ket centre derivatives.
void overlap_pair_deriv(const AO& a, const AO& b, const std::vector<Prim>& pa_, const std::vector<Prim>& pb_,
                        V3 ca, V3 cb, double& s, double da[3], double db[3]) {
    double dx = ca.x - cb.x, dy = ca.y - cb.y, dz = ca.z - cb.z;
    double r2 = dx * dx + dy * dy + dz * dz, r = std::sqrt(r2);
    double cs = comp_abs(a) * comp_abs(b);
    int order = degree(a) + degree(b) + 3;
    s = 0.0;
    for (int k = 0; k < 3; ++k) da[k] = db[k] = 0.0;
    for (auto& pa : pa_)
        for (auto& pb : pb_) {
            double pref = pa.c * pb.c, al = pa.a, be = pb.a, p = al + be;
            if (prim_bound(pref, cs, al, be, r, r2, order) < PRIM_EPS) continue;
            double kab = std::exp(-al * be / p * r2);
            double g[10];
            g[0] = std::sqrt(PI / p);
            for (int m = 1; m < 10; ++m) g[m] = g[m - 1] * (m - 0.5) / p;
            for (auto& A : a.comps)
                for (auto& B : b.comps) {
                    double cf = pref * A.c * B.c * kab;
                    int ia[3] = {A.px, A.py, A.pz}, jb[3] = {B.px, B.py, B.pz};
                    double ac[3] = {ca.x, ca.y, ca.z}, bc[3] = {cb.x, cb.y, cb.z};
                    double m0[3], dA[3], dB[3];
                    for (int k = 0; k < 3; ++k) {
                        m0[k] = moment1d(ia[k], jb[k], al, be, ac[k], bc[k], p, g);
                        dA[k] = 2.0 * al * moment1d(ia[k] + 1, jb[k], al, be, ac[k], bc[k], p, g) -
                                (ia[k] > 0 ? ia[k] * moment1d(ia[k] - 1, jb[k], al, be, ac[k], bc[k], p, g) : 0.0);
                        dB[k] = 2.0 * be * moment1d(ia[k], jb[k] + 1, al, be, ac[k], bc[k], p, g) -
                                (jb[k] > 0 ? jb[k] * moment1d(ia[k], jb[k] - 1, al, be, ac[k], bc[k], p, g) : 0.0);
                    }
                    s += cf * m0[0] * m0[1] * m0[2];
                    da[0] += cf * dA[0] * m0[1] * m0[2];
                    da[1] += cf * m0[0] * dA[1] * m0[2];
                    da[2] += cf * m0[0] * m0[1] * dA[2];
                    db[0] += cf * dB[0] * m0[1] * m0[2];
                    db[1] += cf * m0[0] * dB[1] * m0[2];
                    db[2] += cf * m0[0] * m0[1] * dB[2];
                }
        }
}

// ------------------------------------------------------------------ geometry terms
struct Geo {
    int nat;
    const int* z;
    const double* x;  // 3*nat bohr
    V3 pos(int i) const { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
    double r2(int i, int j) const {
        double dx = x[3 * j] - x[3 * i], dy = x[3 * j + 1] - x[3 * i + 1], dz = x[3 * j + 2] - x[3 * i + 2];
        return dx * dx + dy * dy + dz * dz;
    }
};
const double PAIR_EPS2 = 1e-24;  // pairlist.rs:16 (DIST_EPS^2)

struct CnPair {
    int i, j;
    double rij[3];  // r_i - r_j
    double dcn_dr;
};

// coordination.rs:51-106 exponential counting CN with pair derivatives.
void coordination(const Geo& g, double cutoff, std::vector<double>& cn, std::vector<CnPair>* pairs) {
    cn.assign(g.nat, 0.0);
    if (pairs) pairs->clear();
    const double kcn = 16.0;
    double c2 = cutoff * cutoff;
    for (int i = 0; i < g.nat; ++i)
        for (int j = i + 1; j < g.nat; ++j) {
            double r2 = g.r2(i, j);
            if (r2 <= PAIR_EPS2 || r2 > c2) continue;
            double r = std::sqrt(r2);
            double rc = (COV_RADII_AA[g.z[i]] + COV_RADII_AA[g.z[j]]) * (4.0 / 3.0) * AA2AU;
            if (rc <= 1e-12) continue;
            double raw = -kcn * (rc / r - 1.0);
            double arg = std::min(80.0, std::max(-80.0, raw));
            double e = std::exp(arg);
            double val = 1.0 / (1.0 + e);
            double dv = (raw < -80.0 || raw > 80.0) ? 0.0 : -kcn * rc * e / (r * r * (1.0 + e) * (1.0 + e));
            cn[i] += val;
            cn[j] += val;
            if (pairs) {
                CnPair cp;
                cp.i = i; cp.j = j;
                cp.rij[0] = g.x[3 * i] - g.x[3 * j];
                cp.rij[1] = g.x[3 * i + 1] - g.x[3 * j + 1];
                cp.rij[2] = g.x[3 * i + 2] - g.x[3 * j + 2];
                cp.dcn_dr = dv;
                pairs->push_back(cp);
            }
        }
}

// CN chain rule: g += sum_pairs (dE/dCN_i + dE/dCN_j) dCN/dr r_ij / r (gradient.rs:475-489)
void cn_chain(const std::vector<CnPair>& pairs, const std::vector<double>& dedcn, double* grad, double eps) {
    for (auto& cp : pairs) {
        double r = std::sqrt(cp.rij[0] * cp.rij[0] + cp.rij[1] * cp.rij[1] + cp.rij[2] * cp.rij[2]);
        if (r <= eps) continue;
        double pref = (dedcn[cp.i] + dedcn[cp.j]) * cp.dcn_dr / r;
        for (int k = 0; k < 3; ++k) {
            grad[3 * cp.i + k] += cp.rij[k] * pref;
            grad[3 * cp.j + k] -= cp.rij[k] * pref;
        }
    }
}

// repulsion.rs:39-59, :139-159
double repulsion(const Ctx& ctx, const Geo& g, double* grad) {
    double e = 0.0, c2 = 25.0 * 25.0;
    double kexp = ctx.p.global("kexp", 1.5);
    double kexpl = ctx.p.global("kexplight", kexp);
    for (int i = 0; i < g.nat; ++i)
        for (int j = i + 1; j < g.nat; ++j) {
            double r2 = g.r2(i, j);
            if (r2 <= PAIR_EPS2 || r2 > c2) continue;
            double r = std::sqrt(r2);
            const ElementParam &A = ctx.p.el[g.z[i]], &B = ctx.p.el[g.z[j]];
            double alpha = std::sqrt(A.repa * B.repa);
            double zeff = A.repb * B.repb;
            double k = (g.z[i] <= 2 && g.z[j] <= 2) ? kexpl : kexp;
            double rk = std::pow(r, k);
            double en = zeff * std::exp(-alpha * rk) / r;
            double pref = (alpha * rk * k + 1.0) * en / r2;
            e += en;
            if (grad)
                for (int c = 0; c < 3; ++c) {
                    double d = (g.x[3 * j + c] - g.x[3 * i + c]) * pref;
                    grad[3 * i + c] += d;
                    grad[3 * j + c] -= d;
                }
        }
    return e;
}

bool load_d3(const std::string& path, Ctx& ctx, Err& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { err = {1, "cannot open " + path}; return false; }
    int32_t hdr[3];
    f.read((char*)hdr, sizeof(hdr));
    if (hdr[0] != 0x44334331) { err = {1, "bad D3 file"}; return false; }
    int me = hdr[1], mr = hdr[2];
    ctx.d3_max_elem = me;
    ctx.d3_max_ref = mr;
    ctx.d3_nref.resize(me);
    f.read((char*)ctx.d3_nref.data(), me * 4);
    ctx.d3_refcn.resize((size_t)me * mr);
    f.read((char*)ctx.d3_refcn.data(), ctx.d3_refcn.size() * 8);
    std::vector<double> raw(me);
    f.read((char*)raw.data(), me * 8);
    ctx.d3_r4r2.resize(me);
    // dispersion.rs:3195-3203: sqrt(0.5 * r4r2 * sqrt(Z))
    for (int i = 0; i < me; ++i) ctx.d3_r4r2[i] = std::sqrt(0.5 * raw[i] * std::sqrt((double)(i + 1)));
    size_t npair = (size_t)me * (me + 1) / 2;
    ctx.d3_c6.resize(npair * mr * mr);
    f.read((char*)ctx.d3_c6.data(), ctx.d3_c6.size() * 8);
    if (!f) { err = {1, "truncated D3 file"}; return false; }
    return true;
}

// dispersion.rs:3157-3174 D3Reference::c6
double d3_c6ref(const Ctx& c, int iref, int jref, int zi, int zj) {
    int a, b, ia, ib;
    if (zi > zj) { a = zi; b = zj; ia = iref; ib = jref; }
    else { a = zj; b = zi; ia = jref; ib = iref; }
    size_t pair = (size_t)b + (size_t)a * (a - 1) / 2 - 1;
    return c.d3_c6[pair * c.d3_max_ref * c.d3_max_ref + ib * c.d3_max_ref + ia];
}

// dispersion.rs:1232-1360 D3(BJ) two-body energy and gradient (s9 = 0 path).
double dispersion(const Ctx& ctx, const Geo& g, double* grad) {
    const int nat = g.nat, mr = ctx.d3_max_ref;
    double s6 = ctx.p.global("s6", 1.0), s8 = ctx.p.global("s8", 0), a1 = ctx.p.global("a1", 0),
           a2 = ctx.p.global("a2", 0);
    std::vector<double> cn;
    std::vector<CnPair> cnp;
    coordination(g, 25.0, cn, &cnp);
    // reference_weights (dispersion.rs:1629-1692)
    const double WF = 4.0, wf2 = 2.0 * WF;
    std::vector<double> gw((size_t)nat * mr, 0.0), dgw((size_t)nat * mr, 0.0);
    for (int i = 0; i < nat; ++i) {
        int zi = g.z[i], nref = ctx.d3_nref[zi - 1];
        double mx = -INFINITY;
        std::vector<double> logit(nref), cr(nref), raw(nref);
        for (int r = 0; r < nref; ++r) {
            cr[r] = ctx.d3_refcn[(size_t)(zi - 1) * mr + r];
            logit[r] = -WF * (cn[i] - cr[r]) * (cn[i] - cr[r]);
            mx = std::max(mx, logit[r]);
        }
        double norm = 0, dnorm = 0;
        for (int r = 0; r < nref; ++r) {
            raw[r] = std::exp(logit[r] - mx);
            norm += raw[r];
            dnorm += wf2 * (cr[r] - cn[i]) * raw[r];
        }
        double ni = 1.0 / norm;
        for (int r = 0; r < nref; ++r) {
            double expd = wf2 * (cr[r] - cn[i]) * raw[r];
            double w = raw[r] * ni;
            double dw = expd * ni - raw[r] * dnorm * ni * ni;
            gw[(size_t)i * mr + r] = std::isfinite(w) ? w : 0.0;
            dgw[(size_t)i * mr + r] = std::isfinite(dw) ? dw : 0.0;
        }
    }
    // atomic_c6 (dispersion.rs:1753-1787)
    std::vector<double> c6((size_t)nat * nat), dc6((size_t)nat * nat);
    for (int i = 0; i < nat; ++i)
        for (int j = 0; j <= i; ++j) {
            int zi = g.z[i], zj = g.z[j];
            double cij = 0, dci = 0, dcj = 0;
            for (int ri = 0; ri < ctx.d3_nref[zi - 1]; ++ri)
                for (int rj = 0; rj < ctx.d3_nref[zj - 1]; ++rj) {
                    double ref = d3_c6ref(ctx, ri, rj, zi, zj);
                    cij += gw[(size_t)i * mr + ri] * gw[(size_t)j * mr + rj] * ref;
                    dci += dgw[(size_t)i * mr + ri] * gw[(size_t)j * mr + rj] * ref;
                    dcj += gw[(size_t)i * mr + ri] * dgw[(size_t)j * mr + rj] * ref;
                }
            c6[(size_t)i * nat + j] = c6[(size_t)j * nat + i] = cij;
            if (i == j) dc6[(size_t)i * nat + i] = dci + dcj;
            else { dc6[(size_t)i * nat + j] = dci; dc6[(size_t)j * nat + i] = dcj; }
        }
    double e = 0.0, c2 = 50.0 * 50.0;
    std::vector<double> dedcn(nat, 0.0);
    for (int i = 0; i < nat; ++i)
        for (int j = i + 1; j < nat; ++j) {
            double r2 = g.r2(i, j);
            if (r2 <= PAIR_EPS2 || r2 > c2) continue;  // pair list
            if (r2 <= 1e-14) continue;                  // dispersion.rs:1261
            int zi = g.z[i], zj = g.z[j];
            double r4r2ij = 3.0 * ctx.d3_r4r2[zi - 1] * ctx.d3_r4r2[zj - 1];
            double r0 = a1 * std::sqrt(r4r2ij) + a2;
            double r0_2 = r0 * r0, r0_6 = r0_2 * r0_2 * r0_2, r0_8 = r0_6 * r0_2;
            double r4 = r2 * r2;
            double t6 = 1.0 / (r4 * r2 + r0_6), t8 = 1.0 / (r4 * r4 + r0_8);
            double d6 = -6.0 * r4 * t6 * t6, d8 = -8.0 * r4 * r2 * t8 * t8;
            double ed = s6 * t6 + s8 * r4r2ij * t8;
            double gd = s6 * d6 + s8 * r4r2ij * d8;
            double cij = c6[(size_t)i * nat + j];
            e -= cij * ed;
            if (grad)
                for (int c = 0; c < 3; ++c) {
                    double dg = (g.x[3 * i + c] - g.x[3 * j + c]) * (-cij * gd);
                    grad[3 * i + c] += dg;
                    grad[3 * j + c] -= dg;
                }
            dedcn[i] -= dc6[(size_t)i * nat + j] * ed;
            dedcn[j] -= dc6[(size_t)j * nat + i] * ed;
        }
    if (grad) cn_chain(cnp, dedcn, grad, 1e-14);
    return e;
}

// halogen.rs:64-160 and :716-784 (non-periodic branch)
double halogen(const Ctx& ctx, const Geo& g, double* grad) {
    const double ALP = 6.0, LJ2 = 6.0, EPS = 1e-18, c2 = 20.0 * 20.0;
    double damp = ctx.p.global("xbdamp", 0.44), rads = ctx.p.global("xbrad", 1.3);
    double e = 0.0;
    for (int x = 0; x < g.nat; ++x) {
        int zx = g.z[x];
        if (!(zx == 17 || zx == 35 || zx == 53 || zx == 85)) continue;
        // nearest neighbour of the donor within the cutoff
        int k = -1;
        double best = INFINITY;
        for (int j = 0; j < g.nat; ++j) {
            if (j == x) continue;
            double r2 = g.r2(x, j);
            if (r2 <= PAIR_EPS2 || r2 > c2) continue;
            if (r2 < best) { best = r2; k = j; }
        }
        if (k < 0) continue;
        double cc = ctx.p.el[zx].cxb * 0.1;
        if (cc == 0.0) continue;
        for (int j = 0; j < g.nat; ++j) {
            if (j == x) continue;
            double r2 = g.r2(x, j);
            if (r2 <= PAIR_EPS2 || r2 > c2) continue;
            int zj = g.z[j];
            if (!(zj == 7 || zj == 8 || zj == 15 || zj == 16)) continue;
            double r0 = rads * (ATOMIC_RADII_AA[zx] + ATOMIC_RADII_AA[zj]) * AA2AU;
            double dxj[3], dxk[3], dkj[3];
            for (int c = 0; c < 3; ++c) {
                dxj[c] = g.x[3 * j + c] - g.x[3 * x + c];
                dxk[c] = g.x[3 * k + c] - g.x[3 * x + c];
                dkj[c] = g.x[3 * j + c] - g.x[3 * k + c];
            }
            double d2jx = dxj[0] * dxj[0] + dxj[1] * dxj[1] + dxj[2] * dxj[2];
            double d2kx = dxk[0] * dxk[0] + dxk[1] * dxk[1] + dxk[2] * dxk[2];
            double d2jk = dkj[0] * dkj[0] + dkj[1] * dkj[1] + dkj[2] * dkj[2];
            if (d2jx <= EPS || d2kx <= EPS) continue;
            double rjx = std::sqrt(d2jx) + EPS, rkx = std::sqrt(d2kx) + EPS;
            double xy = std::sqrt(d2kx * d2jx);
            if (xy <= EPS) continue;
            double term = (d2kx + d2jx - d2jk) / xy;
            double ab = 0.5 - 0.25 * term;
            double aterm = std::pow(ab, ALP);
            double t13 = r0 / rjx;
            double t14lj = std::pow(t13, 12.0);
            e += aterm * cc * (t14lj - damp * std::pow(t13, LJ2)) / (1.0 + t14lj);
            if (!grad) continue;
            double t14 = std::pow(r0 / rjx, LJ2);
            double num = t14 * t14 - damp * t14, den = 1.0 + t14 * t14;
            double tlj = num / den;
            double dt = 2.0 * LJ2 * num * t14 * t14 / (rjx * den * den);
            dt += LJ2 * t14 * (damp - 2.0 * t14) / (rjx * den);
            dt *= aterm * cc / rjx;
            double pf = -0.25 * ALP * std::pow(ab, ALP - 1.0) * cc * tlj;
            double djx = (2.0 / rkx - term / rjx) * pf / rjx;
            double dkx = (2.0 / rjx - term / rkx) * pf / rkx;
            double djk = 2.0 * pf / xy;
            for (int c = 0; c < 3; ++c) {
                grad[3 * j + c] += dxj[c] * dt;
                grad[3 * x + c] -= dxj[c] * dt;
                grad[3 * j + c] += dxj[c] * djx;
                grad[3 * x + c] -= dxj[c] * djx;
                grad[3 * k + c] += dxk[c] * dkx;
                grad[3 * x + c] -= dxk[c] * dkx;
                grad[3 * j + c] -= dkj[c] * djk;
                grad[3 * k + c] += dkj[c] * djk;
            }
        }
    }
    return e;
}

// ------------------------------------------------------------------ dense linear algebra
// C = A(m x k) * B(k x n), row-major, optional transposes of A / B.
void gemm(int m, int n, int k, const double* A, bool ta, const double* B, bool tb, double* C) {
    std::vector<double> At, Bt;
    const double* a = A;
    const double* b = B;
    if (ta) {  // make row-major m x k
        At.resize((size_t)m * k);
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < m; ++j) At[(size_t)j * k + i] = A[(size_t)i * m + j];
        a = At.data();
    }
    if (tb) {
        Bt.resize((size_t)k * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < k; ++j) Bt[(size_t)j * n + i] = B[(size_t)i * k + j];
        b = Bt.data();
    }
    const int KB_ = 128;
    for (size_t i = 0; i < (size_t)m * n; ++i) C[i] = 0.0;
    for (int k0 = 0; k0 < k; k0 += KB_) {
        int k1 = std::min(k, k0 + KB_);
        for (int i = 0; i < m; ++i) {
            double* c = C + (size_t)i * n;
            for (int kk = k0; kk < k1; ++kk) {
                double av = a[(size_t)i * k + kk];
                if (av == 0.0) continue;
                const double* br = b + (size_t)kk * n;
                for (int j = 0; j < n; ++j) c[j] += av * br[j];
            }
        }
    }
}

// Symmetric eigendecomposition A = V diag(w) V^T, ascending w; V columns are
// eigenvectors (row-major n x n).  Householder reduction to tridiagonal form
// with accumulated reflectors, then implicit QL/QR sweeps with Wilkinson
// shifts on the tridiagonal matrix (Golub & Van Loan 8.3).
bool eigh(int n, const double* Ain, double* w, double* V) {
    // work on Z = A, stored row-major; build Q explicitly.
    std::vector<double> a(Ain, Ain + (size_t)n * n);
    std::vector<double> d(n), e(n, 0.0);
    // Householder tridiagonalisation: for k = 0..n-3 annihilate a[k+2.., k].
    std::vector<double> Q((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n), pvec(n), wv(n);
    std::vector<std::vector<double>> refl;
    std::vector<double> betas;
    for (int k = 0; k < n - 2; ++k) {
        // x = a[k+1..n-1][k]
        double alpha = 0.0;
        for (int i = k + 1; i < n; ++i) alpha += a[(size_t)i * n + k] * a[(size_t)i * n + k];
        double xnorm = std::sqrt(alpha);
        double x0 = a[(size_t)(k + 1) * n + k];
        if (xnorm == 0.0 || (xnorm == std::fabs(x0) && alpha - x0 * x0 == 0.0)) {
            refl.emplace_back();
            betas.push_back(0.0);
            continue;
        }
        double sgn = x0 >= 0 ? 1.0 : -1.0;
        std::vector<double> u(n, 0.0);
        for (int i = k + 1; i < n; ++i) u[i] = a[(size_t)i * n + k];
        u[k + 1] += sgn * xnorm;
        double unorm2 = 0.0;
        for (int i = k + 1; i < n; ++i) unorm2 += u[i] * u[i];
        double beta = 2.0 / unorm2;
        // A <- H A H with H = I - beta u u^T  (restricted to trailing block)
        // p = beta A u ; K = beta/2 u^T p ; q = p - K u ; A -= u q^T + q u^T
        for (int i = k; i < n; ++i) {
            double s = 0.0;
            const double* row = &a[(size_t)i * n];
            for (int j = k + 1; j < n; ++j) s += row[j] * u[j];
            pvec[i] = beta * s;
        }
        double K = 0.0;
        for (int i = k + 1; i < n; ++i) K += u[i] * pvec[i];
        K *= 0.5 * beta;
        for (int i = k; i < n; ++i) wv[i] = pvec[i] - K * u[i];
        for (int i = k; i < n; ++i) {
            double* row = &a[(size_t)i * n];
            double ui = u[i], wi = wv[i];
            for (int j = k; j < n; ++j) row[j] -= ui * wv[j] + wi * u[j];
        }
        refl.push_back(u);
        betas.push_back(beta);
    }
    for (int i = 0; i < n; ++i) d[i] = a[(size_t)i * n + i];
    for (int i = 0; i + 1 < n; ++i) e[i] = a[(size_t)(i + 1) * n + i];
    // Q = H_0 H_1 ... H_{n-3}: accumulate backwards.  Store Qt (transposed) so
    // that the Givens updates below touch contiguous rows.
    for (int k = (int)refl.size() - 1; k >= 0; --k) {
        if (betas[k] == 0.0) continue;
        const std::vector<double>& u = refl[k];
        // Q <- H_k Q  : Q -= beta u (u^T Q)
        for (int j = k + 1; j < n; ++j) wv[j] = 0.0;
        for (int i = k + 1; i < n; ++i) {
            double ui = u[i];
            const double* row = &Q[(size_t)i * n];
            for (int j = k + 1; j < n; ++j) wv[j] += ui * row[j];
        }
        for (int i = k + 1; i < n; ++i) {
            double f = betas[k] * u[i];
            double* row = &Q[(size_t)i * n];
            for (int j = k + 1; j < n; ++j) row[j] -= f * wv[j];
        }
    }
    // Qt[j][i] = Q[i][j]: row j of Qt is eigenvector-candidate column j.
    std::vector<double> Qt((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Qt[(size_t)j * n + i] = Q[(size_t)i * n + j];
    // implicit QL with Wilkinson-type shift (tridiagonal d, e)
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        while (true) {
            int m;
            for (m = l; m < n - 1; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-300 + 2.220446049250313e-16 * dd * 0.5) break;
            }
            if (m == l) break;
            if (++iter > 200) return false;
            double gq = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(gq, 1.0);
            gq = d[m] - d[l] + e[l] / (gq + (gq >= 0 ? std::fabs(r) : -std::fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            bool early = false;
            for (i = m - 1; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = std::hypot(f, gq);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = gq / r;
                gq = d[i + 1] - p;
                r = (d[i] - gq) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = gq + p;
                gq = c * r - b;
                double* v0 = &Qt[(size_t)i * n];
                double* v1 = &Qt[(size_t)(i + 1) * n];
                for (int kk = 0; kk < n; ++kk) {
                    double t = v1[kk];
                    v1[kk] = s * v0[kk] + c * t;
                    v0[kk] = c * v0[kk] - s * t;
                }
            }
            if (early) continue;
            d[l] -= p;
            e[l] = gq;
            e[m] = 0.0;
        }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    for (int j = 0; j < n; ++j) {
        w[j] = d[idx[j]];
        const double* src = &Qt[(size_t)idx[j] * n];
        for (int i = 0; i < n; ++i) V[(size_t)i * n + j] = src[i];
    }
    return true;
}

// ------------------------------------------------------------------ options / results
}  // namespace

// coulomb.rs:82-107 effective_coulomb_matrix (0-d kernel, coulomb.rs:244): the
// shell hardnesses combine by their harmonic mean (coulomb.rs:239); shells of
// one atom get that mean itself, other pairs 1 / sqrt(R^2 + eta^-2).
void gamma_matrix(const Basis& b, const Geo& g, std::vector<double>& gam) {
    const int nsh = (int)b.sh.size();
    gam.assign((size_t)nsh * nsh, 0.0);
    for (int i = 0; i < nsh; ++i)
        for (int j = 0; j <= i; ++j) {
            double h = 2.0 / (1.0 / b.sh[i].hardness + 1.0 / b.sh[j].hardness);
            double v;
            if (b.sh[i].atom == b.sh[j].atom) v = h;
            else {
                double r2 = g.r2(b.sh[i].atom, b.sh[j].atom);
                v = 1.0 / std::sqrt(r2 + 1.0 / (h * h));
            }
            gam[(size_t)i * nsh + j] = gam[(size_t)j * nsh + i] = v;
        }
}

// ------------------------------------------------------------------ periodic (Gamma point)
// The reference's periodic path (src/pbc/, src/lattice.rs, src/pairlist.rs)
// restated for a fully periodic cell at k = 0.  Every lattice sum below runs
// over explicit image translations T = n0 a + n1 b + n2 c.

// lattice.rs:73-91 Lattice::new (|det| < 1e-14 is SingularCell), :242-261
// image_ranges (ceil(cutoff / plane height) + 1 per axis), :263-296
// image_offsets (sorted by |T|^2, then by the integer triple), :298-376
// reciprocal vectors (2 pi) and their index ranges.
struct Lattice {
    double a[3][3];    // rows: cell vectors (bohr)
    double rec[3][3];  // rows: reciprocal vectors, a_i . rec_j = 2 pi delta_ij
    double vol;
};
struct Image {
    int n[3];
    double t[3];
};

V3 cross3(const double* u, const double* v) {
    return {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
}
double norm3(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

bool make_lattice(const double* cell, Lattice& L) {
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) L.a[i][k] = cell[3 * i + k];
    V3 bc = cross3(L.a[1], L.a[2]), ca = cross3(L.a[2], L.a[0]), ab = cross3(L.a[0], L.a[1]);
    double det = L.a[0][0] * bc.x + L.a[0][1] * bc.y + L.a[0][2] * bc.z;
    if (!(std::fabs(det) >= 1e-14)) return false;
    L.vol = std::fabs(det);
    V3 r[3] = {bc, ca, ab};
    for (int i = 0; i < 3; ++i) {
        L.rec[i][0] = 2.0 * PI * r[i].x / det;
        L.rec[i][1] = 2.0 * PI * r[i].y / det;
        L.rec[i][2] = 2.0 * PI * r[i].z / det;
    }
    return true;
}

// plane heights of a basis (lattice.rs:378-389)
void plane_heights(const double (*v)[3], double h[3]) {
    V3 bc = cross3(v[1], v[2]), ca = cross3(v[2], v[0]), ab = cross3(v[0], v[1]);
    double vol = std::fabs(v[0][0] * bc.x + v[0][1] * bc.y + v[0][2] * bc.z);
    h[0] = vol / std::max(norm3(bc), 1e-12);
    h[1] = vol / std::max(norm3(ca), 1e-12);
    h[2] = vol / std::max(norm3(ab), 1e-12);
}

std::vector<Image> images(const Lattice& L, double cutoff) {
    double h[3];
    plane_heights(L.a, h);
    int rg[3];
    for (int k = 0; k < 3; ++k) rg[k] = cutoff <= 0.0 ? 0 : (int)std::ceil(cutoff / std::max(h[k], 1e-12)) + 1;
    std::vector<Image> out;
    for (int i = -rg[0]; i <= rg[0]; ++i)
        for (int j = -rg[1]; j <= rg[1]; ++j)
            for (int k = -rg[2]; k <= rg[2]; ++k) {
                Image im{{i, j, k}, {0, 0, 0}};
                for (int c = 0; c < 3; ++c) im.t[c] = i * L.a[0][c] + j * L.a[1][c] + k * L.a[2][c];
                out.push_back(im);
            }
    auto n2 = [](const Image& m) { return m.t[0] * m.t[0] + m.t[1] * m.t[1] + m.t[2] * m.t[2]; };
    std::stable_sort(out.begin(), out.end(), [&](const Image& p, const Image& q) {
        double a = n2(p), b = n2(q);
        if (a != b) return a < b;
        return std::lexicographical_compare(p.n, p.n + 3, q.n, q.n + 3);
    });
    return out;
}

// lattice.rs:312-352 reciprocal_vectors_within(cutoff, include_zero = false)
std::vector<Image> recip_within(const Lattice& L, double cutoff) {
    double h[3];
    plane_heights(L.rec, h);
    int rg[3];
    for (int k = 0; k < 3; ++k) rg[k] = cutoff <= 0.0 ? 0 : (int)std::ceil(cutoff / std::max(h[k], 1e-12)) + 1;
    double c2 = cutoff * cutoff, tol = 64.0 * 2.220446049250313e-16 * std::max(c2, 1.0);
    std::vector<Image> out;
    for (int i = -rg[0]; i <= rg[0]; ++i)
        for (int j = -rg[1]; j <= rg[1]; ++j)
            for (int k = -rg[2]; k <= rg[2]; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                Image g{{i, j, k}, {0, 0, 0}};
                for (int c = 0; c < 3; ++c) g.t[c] = i * L.rec[0][c] + j * L.rec[1][c] + k * L.rec[2][c];
                if (g.t[0] * g.t[0] + g.t[1] * g.t[1] + g.t[2] * g.t[2] <= c2 + tol) out.push_back(g);
            }
    return out;
}

bool is_origin(const Image& m) { return m.n[0] == 0 && m.n[1] == 0 && m.n[2] == 0; }
// pairlist.rs:574-584: the first nonzero index of the offset is positive
bool canonical_positive(const Image& m) {
    for (int k = 0; k < 3; ++k) {
        if (m.n[k] > 0) return true;
        if (m.n[k] < 0) return false;
    }
    return false;
}

// pairlist.rs:244-317 for_each_unique_short_range_pair: the origin cell with
// j > i, and every canonical-positive image with every (i, j), i == j included.
// f(i, j, dr = x_j + T - x_i, r2).
template <typename F>
void unique_pairs(const Geo& g, const Lattice& L, double cutoff, F f) {
    double c2 = cutoff * cutoff;
    for (const Image& im : images(L, cutoff)) {
        bool orig = is_origin(im);
        if (!orig && !canonical_positive(im)) continue;
        for (int i = 0; i < g.nat; ++i)
            for (int j = orig ? i + 1 : 0; j < g.nat; ++j) {
                double dr[3];
                for (int c = 0; c < 3; ++c) dr[c] = g.x[3 * j + c] + im.t[c] - g.x[3 * i + c];
                double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                if (r2 <= PAIR_EPS2 || r2 > c2) continue;
                f(i, j, dr, r2);
            }
    }
}

// coordination.rs:51-94 with lattice images: a self-image pair counts twice.
void coordination_pbc(const Geo& g, const Lattice& L, double cutoff, std::vector<double>& cn,
                      std::vector<CnPair>* pairs) {
    cn.assign(g.nat, 0.0);
    if (pairs) pairs->clear();
    const double kcn = 16.0;
    unique_pairs(g, L, cutoff, [&](int i, int j, const double* dr, double r2) {
        double r = std::sqrt(r2);
        double rc = (COV_RADII_AA[g.z[i]] + COV_RADII_AA[g.z[j]]) * (4.0 / 3.0) * AA2AU;
        if (rc <= 1e-12) return;
        double raw = -kcn * (rc / r - 1.0);
        double arg = std::min(80.0, std::max(-80.0, raw));
        double e = std::exp(arg);
        double val = 1.0 / (1.0 + e);
        double dv = (raw < -80.0 || raw > 80.0) ? 0.0 : -kcn * rc * e / (r * r * (1.0 + e) * (1.0 + e));
        double m = (i == j) ? 2.0 : 1.0;
        if (i == j) cn[i] += 2.0 * val;
        else { cn[i] += val; cn[j] += val; }
        if (pairs) {
            CnPair cp;
            cp.i = i; cp.j = j;
            for (int c = 0; c < 3; ++c) cp.rij[c] = -dr[c];
            cp.dcn_dr = m * dv;
            pairs->push_back(cp);
        }
    });
}

// CN chain over periodic pairs (gradient.rs:475-489, pbc/gradient.rs:387-404):
// self-image pairs are skipped.
void cn_chain_pbc(const std::vector<CnPair>& pairs, const std::vector<double>& dedcn, double* grad, double eps) {
    for (auto& cp : pairs) {
        if (cp.i == cp.j) continue;
        double r = std::sqrt(cp.rij[0] * cp.rij[0] + cp.rij[1] * cp.rij[1] + cp.rij[2] * cp.rij[2]);
        if (r <= eps) continue;
        double pref = (dedcn[cp.i] + dedcn[cp.j]) * cp.dcn_dr / r;
        for (int k = 0; k < 3; ++k) {
            grad[3 * cp.i + k] += cp.rij[k] * pref;
            grad[3 * cp.j + k] -= cp.rij[k] * pref;
        }
    }
}

// repulsion.rs:39-59 over image pairs (a self-image pair has energy, no force).
double repulsion_pbc(const Ctx& ctx, const Geo& g, const Lattice& L, double* grad) {
    double e = 0.0;
    double kexp = ctx.p.global("kexp", 1.5);
    double kexpl = ctx.p.global("kexplight", kexp);
    unique_pairs(g, L, 25.0, [&](int i, int j, const double* dr, double r2) {
        double r = std::sqrt(r2);
        const ElementParam &A = ctx.p.el[g.z[i]], &B = ctx.p.el[g.z[j]];
        double alpha = std::sqrt(A.repa * B.repa);
        double zeff = A.repb * B.repb;
        double k = (g.z[i] <= 2 && g.z[j] <= 2) ? kexpl : kexp;
        double rk = std::pow(r, k);
        double en = zeff * std::exp(-alpha * rk) / r;
        double pref = (alpha * rk * k + 1.0) * en / r2;
        e += en;
        if (grad && i != j)
            for (int c = 0; c < 3; ++c) {
                grad[3 * i + c] += dr[c] * pref;
                grad[3 * j + c] -= dr[c] * pref;
            }
    });
    return e;
}

// dispersion.rs:1232-1343 D3(BJ) two-body over image pairs (s9 = 0), with the
// lattice-summed CN (25 bohr) of dispersion.rs:1245-1251.
double dispersion_pbc(const Ctx& ctx, const Geo& g, const Lattice& L, double* grad) {
    const int nat = g.nat, mr = ctx.d3_max_ref;
    double s6 = ctx.p.global("s6", 1.0), s8 = ctx.p.global("s8", 0), a1 = ctx.p.global("a1", 0),
           a2 = ctx.p.global("a2", 0);
    std::vector<double> cn;
    std::vector<CnPair> cnp;
    coordination_pbc(g, L, 25.0, cn, &cnp);
    const double WF = 4.0, wf2 = 2.0 * WF;
    std::vector<double> gw((size_t)nat * mr, 0.0), dgw((size_t)nat * mr, 0.0);
    for (int i = 0; i < nat; ++i) {
        int zi = g.z[i], nref = ctx.d3_nref[zi - 1];
        double mx = -INFINITY;
        std::vector<double> logit(nref), cr(nref), raw(nref);
        for (int r = 0; r < nref; ++r) {
            cr[r] = ctx.d3_refcn[(size_t)(zi - 1) * mr + r];
            logit[r] = -WF * (cn[i] - cr[r]) * (cn[i] - cr[r]);
            mx = std::max(mx, logit[r]);
        }
        double norm = 0, dnorm = 0;
        for (int r = 0; r < nref; ++r) {
            raw[r] = std::exp(logit[r] - mx);
            norm += raw[r];
            dnorm += wf2 * (cr[r] - cn[i]) * raw[r];
        }
        double ni = 1.0 / norm;
        for (int r = 0; r < nref; ++r) {
            double expd = wf2 * (cr[r] - cn[i]) * raw[r];
            double w = raw[r] * ni;
            double dw = expd * ni - raw[r] * dnorm * ni * ni;
            gw[(size_t)i * mr + r] = std::isfinite(w) ? w : 0.0;
            dgw[(size_t)i * mr + r] = std::isfinite(dw) ? dw : 0.0;
        }
    }
    std::vector<double> c6((size_t)nat * nat), dc6((size_t)nat * nat);
    for (int i = 0; i < nat; ++i)
        for (int j = 0; j <= i; ++j) {
            int zi = g.z[i], zj = g.z[j];
            double cij = 0, dci = 0, dcj = 0;
            for (int ri = 0; ri < ctx.d3_nref[zi - 1]; ++ri)
                for (int rj = 0; rj < ctx.d3_nref[zj - 1]; ++rj) {
                    double ref = d3_c6ref(ctx, ri, rj, zi, zj);
                    cij += gw[(size_t)i * mr + ri] * gw[(size_t)j * mr + rj] * ref;
                    dci += dgw[(size_t)i * mr + ri] * gw[(size_t)j * mr + rj] * ref;
                    dcj += gw[(size_t)i * mr + ri] * dgw[(size_t)j * mr + rj] * ref;
                }
            c6[(size_t)i * nat + j] = c6[(size_t)j * nat + i] = cij;
            if (i == j) dc6[(size_t)i * nat + i] = dci + dcj;
            else { dc6[(size_t)i * nat + j] = dci; dc6[(size_t)j * nat + i] = dcj; }
        }
    double e = 0.0;
    std::vector<double> dedcn(nat, 0.0);
    unique_pairs(g, L, 50.0, [&](int i, int j, const double* dr, double r2) {
        if (r2 <= 1e-14) return;
        int zi = g.z[i], zj = g.z[j];
        double r4r2ij = 3.0 * ctx.d3_r4r2[zi - 1] * ctx.d3_r4r2[zj - 1];
        double r0 = a1 * std::sqrt(r4r2ij) + a2;
        double r0_2 = r0 * r0, r0_6 = r0_2 * r0_2 * r0_2, r0_8 = r0_6 * r0_2;
        double r4 = r2 * r2;
        double t6 = 1.0 / (r4 * r2 + r0_6), t8 = 1.0 / (r4 * r4 + r0_8);
        double d6 = -6.0 * r4 * t6 * t6, d8 = -8.0 * r4 * r2 * t8 * t8;
        double ed = s6 * t6 + s8 * r4r2ij * t8;
        double gd = s6 * d6 + s8 * r4r2ij * d8;
        double cij = c6[(size_t)i * nat + j];
        e -= cij * ed;
        if (grad && i != j)
            for (int c = 0; c < 3; ++c) {
                double dgv = -dr[c] * (-cij * gd);
                grad[3 * i + c] += dgv;
                grad[3 * j + c] -= dgv;
            }
        if (i == j) dedcn[i] -= dc6[(size_t)i * nat + i] * ed;
        else {
            dedcn[i] -= dc6[(size_t)i * nat + j] * ed;
            dedcn[j] -= dc6[(size_t)j * nat + i] * ed;
        }
    });
    if (grad) cn_chain_pbc(cnp, dedcn, grad, 1e-14);
    return e;
}

// halogen.rs:64-160 with the periodic triples of :716-740: the donor's
// neighbours are every (atom, image) within 20 bohr except the donor itself
// in the origin cell; the nearest one that is not an image of the donor is
// the bonded neighbour; acceptors may be images.
double halogen_pbc(const Ctx& ctx, const Geo& g, const Lattice& L, double* grad) {
    const double ALP = 6.0, LJ2 = 6.0, EPS = 1e-18, cut = 20.0;
    double damp = ctx.p.global("xbdamp", 0.44), rads = ctx.p.global("xbrad", 1.3);
    double e = 0.0;
    std::vector<Image> ims = images(L, cut);
    struct Nb { int j; double t[3]; double r2; };
    for (int x = 0; x < g.nat; ++x) {
        int zx = g.z[x];
        if (!(zx == 17 || zx == 35 || zx == 53 || zx == 85)) continue;
        std::vector<Nb> nbs;
        for (const Image& im : ims)
            for (int j = 0; j < g.nat; ++j) {
                if (is_origin(im) && j == x) continue;
                double d[3];
                for (int c = 0; c < 3; ++c) d[c] = g.x[3 * j + c] + im.t[c] - g.x[3 * x + c];
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (r2 <= PAIR_EPS2 || r2 > cut * cut) continue;
                nbs.push_back({j, {im.t[0], im.t[1], im.t[2]}, r2});
            }
        int kbest = -1;
        for (int q = 0; q < (int)nbs.size(); ++q)
            if (nbs[q].j != x && (kbest < 0 || nbs[q].r2 < nbs[kbest].r2)) kbest = q;
        if (kbest < 0) continue;
        double cc = ctx.p.el[zx].cxb * 0.1;
        if (cc == 0.0) continue;
        const Nb& K = nbs[kbest];
        int k = K.j;
        for (const Nb& J : nbs) {
            int j = J.j, zj = g.z[j];
            if (!(zj == 7 || zj == 8 || zj == 15 || zj == 16)) continue;
            double r0 = rads * (ATOMIC_RADII_AA[zx] + ATOMIC_RADII_AA[zj]) * AA2AU;
            double dxj[3], dxk[3], dkj[3];
            for (int c = 0; c < 3; ++c) {
                dxj[c] = g.x[3 * j + c] + J.t[c] - g.x[3 * x + c];
                dxk[c] = g.x[3 * k + c] + K.t[c] - g.x[3 * x + c];
                dkj[c] = g.x[3 * j + c] + J.t[c] - g.x[3 * k + c] - K.t[c];
            }
            double d2jx = dxj[0] * dxj[0] + dxj[1] * dxj[1] + dxj[2] * dxj[2];
            double d2kx = dxk[0] * dxk[0] + dxk[1] * dxk[1] + dxk[2] * dxk[2];
            double d2jk = dkj[0] * dkj[0] + dkj[1] * dkj[1] + dkj[2] * dkj[2];
            if (d2jx <= EPS || d2kx <= EPS) continue;
            double rjx = std::sqrt(d2jx) + EPS, rkx = std::sqrt(d2kx) + EPS;
            double xy = std::sqrt(d2kx * d2jx);
            if (xy <= EPS) continue;
            double term = (d2kx + d2jx - d2jk) / xy;
            double ab = 0.5 - 0.25 * term;
            double aterm = std::pow(ab, ALP);
            double t13 = r0 / rjx;
            double t14lj = std::pow(t13, 12.0);
            e += aterm * cc * (t14lj - damp * std::pow(t13, LJ2)) / (1.0 + t14lj);
            if (!grad) continue;
            double t14 = std::pow(r0 / rjx, LJ2);
            double num = t14 * t14 - damp * t14, den = 1.0 + t14 * t14;
            double tlj = num / den;
            double dt = 2.0 * LJ2 * num * t14 * t14 / (rjx * den * den);
            dt += LJ2 * t14 * (damp - 2.0 * t14) / (rjx * den);
            dt *= aterm * cc / rjx;
            double pf = -0.25 * ALP * std::pow(ab, ALP - 1.0) * cc * tlj;
            double djx = (2.0 / rkx - term / rjx) * pf / rjx;
            double dkx = (2.0 / rjx - term / rkx) * pf / rkx;
            double djk = 2.0 * pf / xy;
            for (int c = 0; c < 3; ++c) {
                grad[3 * j + c] += dxj[c] * dt;
                grad[3 * x + c] -= dxj[c] * dt;
                grad[3 * j + c] += dxj[c] * djx;
                grad[3 * x + c] -= dxj[c] * djx;
                grad[3 * k + c] += dxk[c] * dkx;
                grad[3 * x + c] -= dxk[c] * dkx;
                grad[3 * j + c] -= dkj[c] * djk;
                grad[3 * k + c] += dkj[c] * djk;
            }
        }
    }
    return e;
}

// ---- Ewald summation of the Klopman-Ohno interaction (pbc/ewald.rs)
const double EW_TAU = 5.5;           // ewald.rs:39
const double EW_R3 = -0.5;           // ewald.rs:38 binomial coefficient of eta^-2 R^-3
const double SQRTPI = 1.7724538509055160273;

// Exponential integral E1(x), x > 0: the power series
// E1 = -gamma - ln x - sum_k (-x)^k / (k k!) for x <= 1, the continued fraction
// E1 = e^-x / (x + 1 - 1/(x + 3 - 4/(x + 5 - ...))) by the modified Lentz
// method otherwise (Abramowitz & Stegun 5.1.11 and 5.1.22).  Written for this
// file; the reference evaluates the same function (ewald.rs:142-181).
double expint_e1(double x) {
    if (x <= 1.0) {
        double sum = 0.0, term = 1.0;
        for (int k = 1; k <= 60; ++k) {
            term *= -x / k;
            double add = term / k;
            sum += add;
            if (std::fabs(add) < 1e-17 * std::fabs(sum)) break;
        }
        return -0.57721566490153286061 - std::log(x) - sum;
    }
    const double tiny = 1e-300;
    double b = x + 1.0, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i <= 300; ++i) {
        double an = -(double)i * i;
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (std::fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        double del = c * d;
        h *= del;
        if (std::fabs(del - 1.0) < 1e-16) break;
    }
    return h * std::exp(-x);
}

// ewald.rs:48-60: alpha = sqrt(pi) / V^(1/3)
double ewald_alpha(const Lattice& L) { return SQRTPI / std::cbrt(std::max(L.vol, 1e-6)); }

// ewald.rs:200-243 ewald_atom_matrix: Madelung potential of a unit charge and
// its images with the neutralising background, 1/R split by erfc.
void ewald_atom_matrix(const Geo& g, const Lattice& L, double alpha, std::vector<double>& phi) {
    const int nat = g.nat;
    double rcut = EW_TAU / alpha, gcut = 2.0 * alpha * EW_TAU;
    std::vector<Image> ims = images(L, rcut), gs = recip_within(L, gcut);
    double inv4a2 = 1.0 / (4.0 * alpha * alpha), fpv = 4.0 * PI / L.vol;
    double self = -2.0 * alpha / SQRTPI, bg = -PI / (alpha * alpha * L.vol);
    phi.assign((size_t)nat * nat, 0.0);
    for (int a = 0; a < nat; ++a)
        for (int b = 0; b <= a; ++b) {
            double rab[3];
            for (int c = 0; c < 3; ++c) rab[c] = g.x[3 * a + c] - g.x[3 * b + c];
            double real = 0.0;
            for (const Image& im : ims) {
                double d[3] = {rab[0] - im.t[0], rab[1] - im.t[1], rab[2] - im.t[2]};
                double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (r <= 1e-12 || r > rcut) continue;
                real += std::erfc(alpha * r) / r;
            }
            double rec = 0.0;
            for (const Image& G : gs) {
                double g2 = G.t[0] * G.t[0] + G.t[1] * G.t[1] + G.t[2] * G.t[2];
                rec += std::exp(-g2 * inv4a2) / g2 * std::cos(G.t[0] * rab[0] + G.t[1] * rab[1] + G.t[2] * rab[2]);
            }
            double v = real + rec * fpv + bg;
            if (a == b) v += self;
            phi[(size_t)a * nat + b] = phi[(size_t)b * nat + a] = v;
        }
}

// ewald.rs:96-140: short-range KO remainder and the real-space part of the
// eta^-2 R^-3 Ewald term, value and radial derivative.
void ko_short(double r, double eta, double& v, double& dv) {
    double ie2 = 1.0 / (eta * eta), den2 = r * r + ie2, den = std::sqrt(den2);
    double r2 = r * r, r3 = r2 * r, r4 = r2 * r2;
    v = 1.0 / den - 1.0 / r + 0.5 * ie2 / r3;
    dv = -r / std::pow(den2, 1.5) + 1.0 / r2 - 1.5 * ie2 / r4;
}
void r3_real(double r, double eta, double alpha, double& v, double& dv) {
    double ie2 = 1.0 / (eta * eta), ar = alpha * r, ex = std::exp(-ar * ar);
    double q = std::erfc(ar) + (2.0 * ar / SQRTPI) * ex;
    double dq = -(4.0 * alpha * alpha * alpha * r * r / SQRTPI) * ex;
    double r3 = r * r * r, r4 = r3 * r;
    v = ie2 * q / r3;
    dv = ie2 * (dq / r3 - 3.0 * q / r4);
}

// ewald.rs:253-320 periodic_gamma_matrix: Ewald 1/R between atom centres, the
// on-site eta, the generalised Ewald eta^-2 R^-3 term (real, reciprocal with
// E1, the k = 0 logarithm, self) and the directly summed short-range remainder
// (10 bohr).
const double EW_SR_CUT = 10.0;  // pbc/mod.rs:162 sr_cutoff
void ewald_gamma_matrix(const Basis& b, const Geo& g, const Lattice& L, std::vector<double>& gam,
                        double alpha = 0.0) {
    const int nsh = (int)b.sh.size();
    if (alpha <= 0.0) alpha = ewald_alpha(L);
    std::vector<double> phi;
    ewald_atom_matrix(g, L, alpha, phi);
    double r3cut = EW_TAU / alpha;
    std::vector<Image> sri = images(L, EW_SR_CUT), r3i = images(L, r3cut), gs = recip_within(L, 2.0 * alpha * EW_TAU);
    double inv4a2 = 1.0 / (4.0 * alpha * alpha);
    double r3_rec = 2.0 * PI / L.vol;
    const double DIGAMMA_3_2 = 0.03648997397857652;  // psi(3/2) = 2 - gamma - 2 ln 2
    double kk = alpha / SQRTPI;
    double r3_k0 = 4.0 * PI * (std::log(kk) + 0.5 * (std::log(PI) - DIGAMMA_3_2)) / L.vol;
    double r3_self = -(4.0 * PI / 3.0) * kk * kk * kk;
    std::vector<double> e1(gs.size());
    for (size_t q = 0; q < gs.size(); ++q)
        e1[q] = expint_e1((gs[q].t[0] * gs[q].t[0] + gs[q].t[1] * gs[q].t[1] + gs[q].t[2] * gs[q].t[2]) * inv4a2);
    gam.assign((size_t)nsh * nsh, 0.0);
    for (int i = 0; i < nsh; ++i) {
        int ai = b.sh[i].atom;
        for (int j = 0; j <= i; ++j) {
            int aj = b.sh[j].atom;
            double eta = 2.0 / (1.0 / b.sh[i].hardness + 1.0 / b.sh[j].hardness);
            double ie2 = 1.0 / (eta * eta);
            double v = phi[(size_t)ai * g.nat + aj];
            if (ai == aj) v += eta;
            double rab[3];
            for (int c = 0; c < 3; ++c) rab[c] = g.x[3 * ai + c] - g.x[3 * aj + c];
            for (const Image& im : r3i) {
                if (ai == aj && is_origin(im)) continue;
                double d[3] = {rab[0] - im.t[0], rab[1] - im.t[1], rab[2] - im.t[2]};
                double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (r <= 1e-12 || r > r3cut) continue;
                double rv, rdv;
                r3_real(r, eta, alpha, rv, rdv);
                v += EW_R3 * rv;
            }
            for (size_t q = 0; q < gs.size(); ++q)
                v += EW_R3 * r3_rec * ie2 * e1[q] *
                     std::cos(gs[q].t[0] * rab[0] + gs[q].t[1] * rab[1] + gs[q].t[2] * rab[2]);
            v += EW_R3 * r3_k0 * ie2;
            if (ai == aj) v += EW_R3 * r3_self * ie2;
            for (const Image& im : sri) {
                if (ai == aj && is_origin(im)) continue;
                double d[3] = {rab[0] - im.t[0], rab[1] - im.t[1], rab[2] - im.t[2]};
                double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (r <= 1e-12 || r > EW_SR_CUT) continue;
                double sv, sdv;
                ko_short(r, eta, sv, sdv);
                v += sv;
            }
            gam[(size_t)i * nsh + j] = gam[(size_t)j * nsh + i] = v;
        }
    }
}

// pbc/gradient.rs:548-603, :679-779: gradient of 1/2 q^T Gamma q at fixed
// charges (Ewald 1/R with atomic charges, the reciprocal eta^-2 R^-3 part with
// shell charges, the real-space R^-3 and short-range parts).
void ewald_gradient(const Basis& b, const Geo& g, const Lattice& L, const std::vector<double>& q,
                    const std::vector<double>& qat, double* grad) {
    const int nat = g.nat, nsh = (int)b.sh.size();
    double alpha = ewald_alpha(L);
    double rcut = EW_TAU / alpha, gcut = 2.0 * alpha * EW_TAU;
    std::vector<Image> ims = images(L, rcut), gs = recip_within(L, gcut);
    double inv4a2 = 1.0 / (4.0 * alpha * alpha), fpv = 4.0 * PI / L.vol, tas = 2.0 * alpha / SQRTPI;
    for (int a = 0; a < nat; ++a)
        for (int c = 0; c < nat; ++c) {
            double qq = qat[a] * qat[c];
            if (qq == 0.0) continue;
            for (const Image& im : ims) {
                double d[3];
                for (int k = 0; k < 3; ++k) d[k] = g.x[3 * a + k] - g.x[3 * c + k] - im.t[k];
                double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (r <= 1e-12 || r > rcut) continue;
                double dg = -std::erfc(alpha * r) / (r * r) - tas * std::exp(-alpha * alpha * r * r) / r;
                for (int k = 0; k < 3; ++k) grad[3 * a + k] += d[k] * (qq * dg / r);
            }
        }
    for (const Image& G : gs) {
        double g2 = G.t[0] * G.t[0] + G.t[1] * G.t[1] + G.t[2] * G.t[2];
        double wg = std::exp(-g2 * inv4a2) / g2, sc = 0.0, ss = 0.0;
        for (int c = 0; c < nat; ++c) {
            double ph = G.t[0] * g.x[3 * c] + G.t[1] * g.x[3 * c + 1] + G.t[2] * g.x[3 * c + 2];
            sc += qat[c] * std::cos(ph);
            ss += qat[c] * std::sin(ph);
        }
        for (int c = 0; c < nat; ++c) {
            double ph = G.t[0] * g.x[3 * c] + G.t[1] * g.x[3 * c + 1] + G.t[2] * g.x[3 * c + 2];
            double f = std::sin(ph) * sc - std::cos(ph) * ss;
            for (int k = 0; k < 3; ++k) grad[3 * c + k] -= G.t[k] * (fpv * wg * qat[c] * f);
        }
    }
    // reciprocal eta^-2 R^-3 part, per shell pair
    double pref0 = -2.0 * EW_R3 * PI / L.vol;
    std::vector<double> ph(nsh);
    for (const Image& G : gs) {
        double g2 = G.t[0] * G.t[0] + G.t[1] * G.t[1] + G.t[2] * G.t[2];
        double coeff = pref0 * expint_e1(g2 * inv4a2);
        for (int i = 0; i < nsh; ++i) {
            int a = b.sh[i].atom;
            ph[i] = G.t[0] * g.x[3 * a] + G.t[1] * g.x[3 * a + 1] + G.t[2] * g.x[3 * a + 2];
        }
        for (int i = 0; i < nsh; ++i) {
            if (q[i] == 0.0) continue;
            double f = 0.0;
            for (int j = 0; j < nsh; ++j) {
                if (q[j] == 0.0) continue;
                double eta = 2.0 / (1.0 / b.sh[i].hardness + 1.0 / b.sh[j].hardness);
                f += q[j] * std::sin(ph[i] - ph[j]) / (eta * eta);
            }
            int a = b.sh[i].atom;
            for (int k = 0; k < 3; ++k) grad[3 * a + k] += G.t[k] * (coeff * q[i] * f);
        }
    }
    // real-space R^-3 and short-range remainder, per shell pair and image
    double r3cut = rcut, realcut = std::max(r3cut, EW_SR_CUT);
    std::vector<Image> ri = images(L, realcut);
    for (int i = 0; i < nsh; ++i) {
        int ai = b.sh[i].atom;
        for (int j = 0; j < nsh; ++j) {
            int aj = b.sh[j].atom;
            double eta = 2.0 / (1.0 / b.sh[i].hardness + 1.0 / b.sh[j].hardness);
            double qq = q[i] * q[j];
            if (qq == 0.0) continue;
            for (const Image& im : ri) {
                if (ai == aj && is_origin(im)) continue;
                double d[3];
                for (int k = 0; k < 3; ++k) d[k] = g.x[3 * ai + k] - g.x[3 * aj + k] - im.t[k];
                double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (r <= 1e-12 || r > realcut) continue;
                double rad = 0.0, v, dv;
                if (r <= r3cut) { r3_real(r, eta, alpha, v, dv); rad += EW_R3 * dv; }
                if (r <= EW_SR_CUT) { ko_short(r, eta, v, dv); rad += dv; }
                for (int k = 0; k < 3; ++k) grad[3 * ai + k] += d[k] * (qq * rad / r);
            }
        }
    }
}

// pbc/bloch.rs:58-236 BlochBuilder::build and h_s_gamma_real: S(Gamma) and
// H0(Gamma) as sums over every image (30 bohr AO cutoff, pbc/mod.rs:170), the
// same-site block at T = 0 taking the on-site formula, every other
// (atom, image) pair the off-site one at its image distance; atom pairs beyond
// the overlap screen r^2 e_a e_b > 40 (e_a + e_b) (e = the atom's smallest
// primitive exponent) are skipped; finally symmetrised.
const double PBC_AO_CUT = 30.0;
template <typename HS>
void bloch_gamma(const Basis& b, const Geo& g, const Lattice& L, const std::vector<double>& se, HS hscale,
                 std::vector<double>& S, std::vector<double>& H0) {
    const int n = (int)b.ao.size(), nat = g.nat;
    std::vector<double> emin(nat, INFINITY);
    for (auto& a : b.ao)
        for (auto& p : b.sh[a.shell].prims) emin[a.atom] = std::min(emin[a.atom], p.a);
    std::vector<std::vector<int>> aos(nat);
    for (int i = 0; i < n; ++i) aos[b.ao[i].atom].push_back(i);
    S.assign((size_t)n * n, 0.0);
    H0.assign((size_t)n * n, 0.0);
    double c2 = PBC_AO_CUT * PBC_AO_CUT;
    for (const Image& im : images(L, PBC_AO_CUT)) {
        bool orig = is_origin(im);
        for (int a = 0; a < nat; ++a)
            for (int bb = 0; bb < nat; ++bb) {
                V3 ra = g.pos(a);
                V3 rb = {g.x[3 * bb] + im.t[0], g.x[3 * bb + 1] + im.t[1], g.x[3 * bb + 2] + im.t[2]};
                double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
                double r2 = dx * dx + dy * dy + dz * dz;
                bool same = orig && a == bb;
                if (!same) {
                    if (r2 > c2) continue;
                    if (r2 * emin[a] * emin[bb] > 40.0 * (emin[a] + emin[bb])) continue;
                }
                double r = std::sqrt(r2);
                for (int mu : aos[a])
                    for (int nu : aos[bb]) {
                        const Shell &si = b.sh[b.ao[mu].shell], &sj = b.sh[b.ao[nu].shell];
                        double s = overlap_pair(b.ao[mu], b.ao[nu], si.prims, sj.prims, ra, rb);
                        if (s == 0.0) continue;
                        double hij = 0.5 * (se[b.ao[mu].shell] + se[b.ao[nu].shell]);
                        if (!same) {
                            double rr = si.rad + sj.rad > 0.0 ? std::sqrt(r / (si.rad + sj.rad)) : 0.0;
                            hij *= hscale(si, sj) * (1.0 + si.poly * rr) * (1.0 + sj.poly * rr);
                        }
                        S[(size_t)mu * n + nu] += s;
                        H0[(size_t)mu * n + nu] += s * hij;
                    }
            }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            double s = 0.5 * (S[(size_t)i * n + j] + S[(size_t)j * n + i]);
            double h = 0.5 * (H0[(size_t)i * n + j] + H0[(size_t)j * n + i]);
            S[(size_t)i * n + j] = S[(size_t)j * n + i] = s;
            H0[(size_t)i * n + j] = H0[(size_t)j * n + i] = h;
        }
}

// hamiltonian.rs:140-166 hscale: valence pairs k_ll' * pair * (1 + enscale
// 0.01 dEN^2); one valence shell 1/2 (k_ll + kdiff); none kdiff.
double hscale_fn(const Ctx& ctx, const Shell& si, const Shell& sj) {
    double kdiff = ctx.p.global("kdiff", 2.85);
    auto kpair = [&](int li, int lj) {
        if ((li == 0 && lj == 1) || (li == 1 && lj == 0)) return ctx.p.global("ksp", 2.08);
        return 0.5 * (ctx.p.k_shell(li) + ctx.p.k_shell(lj));
    };
    if (si.valence && sj.valence) {
        double den = PAULING_EN[si.z] - PAULING_EN[sj.z];
        den *= den;
        double ens = ctx.p.global("enscale", -0.7) * 0.01;
        return ctx.p.pair_scaling(si.z, sj.z) * kpair(si.l, sj.l) * (1.0 + ens * den);
    }
    if (si.valence) return 0.5 * (kpair(si.l, si.l) + kdiff);
    if (sj.valence) return 0.5 * (kpair(sj.l, sj.l) + kdiff);
    return kdiff;
}

// S and H0 (integrals.rs:133-242, hamiltonian.rs:91-134) for molecules (pairs
// within icut), or S(Gamma), H0(Gamma) for a cell (pbc/bloch.rs:58-236);
// se / dsedcn: the CN-shifted self energies and their CN derivatives.
void s_h0(const Ctx& ctx, const Basis& b, const Geo& g, const Lattice* L, double icut, const std::vector<double>& cn,
          std::vector<double>& S, std::vector<double>& H0, std::vector<double>& se, std::vector<double>& dsedcn) {
    const int n = (int)b.ao.size(), nsh = (int)b.sh.size();
    auto hscale = [&](const Shell& si, const Shell& sj) { return hscale_fn(ctx, si, sj); };
    se.assign(nsh, 0.0);
    dsedcn.assign(nsh, 0.0);
    for (int s = 0; s < nsh; ++s) {
        dsedcn[s] = -b.sh[s].kcn;
        se[s] = b.sh[s].hdiag - b.sh[s].kcn * cn[b.sh[s].atom];
    }
    if (L) {
        bloch_gamma(b, g, *L, se, hscale, S, H0);
        return;
    }
    S.assign((size_t)n * n, 0.0);
    H0.assign((size_t)n * n, 0.0);
    double ic2 = icut * icut;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const AO &A = b.ao[i], &B = b.ao[j];
            if (A.atom != B.atom) {
                double r2 = g.r2(A.atom, B.atom);
                if (r2 <= PAIR_EPS2 || r2 > ic2) continue;
            }
            S[(size_t)i * n + j] =
                overlap_pair(A, B, b.sh[A.shell].prims, b.sh[B.shell].prims, g.pos(A.atom), g.pos(B.atom));
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            const Shell &si = b.sh[b.ao[i].shell], &sj = b.sh[b.ao[j].shell];
            double hij;
            if (si.atom == sj.atom) {
                hij = 0.5 * (se[b.ao[i].shell] + se[b.ao[j].shell]);
            } else {
                double r = std::sqrt(g.r2(si.atom, sj.atom));
                double rr = std::sqrt(r / (si.rad + sj.rad));
                hij = 0.5 * (se[b.ao[i].shell] + se[b.ao[j].shell]) * hscale(si, sj) * (1.0 + si.poly * rr) *
                      (1.0 + sj.poly * rr);
            }
            double v = S[(size_t)i * n + j] * hij;
            H0[(size_t)i * n + j] = v;
            H0[(size_t)j * n + i] = v;
        }
}

// pbc/gradient.rs:250-408 band_and_cn_gradient at k = 0 (P(T) = P, W(T) = W
// for every image), then the electrostatic part (:548-603).  Pairs: the origin
// cell with a < b, every canonical-positive image with every (a, b).
template <typename HS>
void periodic_gradient(const Basis& b, const Geo& g, const Lattice& L, bool cn_on, double cn_cutoff, HS hscale,
                       const std::vector<double>& cn, const std::vector<double>& vsh, const std::vector<double>& q,
                       const std::vector<double>& qat, const std::vector<double>& P, const std::vector<double>& W,
                       double* grad) {
    const int n = (int)b.ao.size(), nat = g.nat, nsh = (int)b.sh.size();
    std::vector<double> sev(nsh), dse(nsh), vao(n), dedcn(nat, 0.0);
    for (int s = 0; s < nsh; ++s) {
        dse[s] = cn_on ? -b.sh[s].kcn : 0.0;
        sev[s] = b.sh[s].hdiag + dse[s] * cn[b.sh[s].atom];
    }
    for (int i = 0; i < n; ++i) vao[i] = vsh[b.ao[i].shell];
    std::vector<double> emin(nat, INFINITY);
    for (auto& a : b.ao)
        for (auto& p : b.sh[a.shell].prims) emin[a.atom] = std::min(emin[a.atom], p.a);
    std::vector<std::vector<int>> aos(nat);
    for (int i = 0; i < n; ++i) aos[b.ao[i].atom].push_back(i);
    double c2 = PBC_AO_CUT * PBC_AO_CUT;
    for (const Image& im : images(L, PBC_AO_CUT)) {
        bool orig = is_origin(im);
        if (!orig && !canonical_positive(im)) continue;
        for (int a = 0; a < nat; ++a)
            for (int bb = 0; bb < nat; ++bb) {
                if (orig && a >= bb) continue;
                V3 ra = g.pos(a);
                V3 rb = {g.x[3 * bb] + im.t[0], g.x[3 * bb + 1] + im.t[1], g.x[3 * bb + 2] + im.t[2]};
                double rv[3] = {ra.x - rb.x, ra.y - rb.y, ra.z - rb.z};
                double r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
                if (r2 <= 1e-12 || r2 > c2) continue;
                if (r2 * emin[a] * emin[bb] > 40.0 * (emin[a] + emin[bb])) continue;
                double ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0}, dca = 0, dcb = 0;
                for (int mu : aos[a])
                    for (int nu : aos[bb]) {
                        int smu = b.ao[mu].shell, snu = b.ao[nu].shell;
                        const Shell &SM = b.sh[smu], &SN = b.sh[snu];
                        double rr = std::sqrt(std::sqrt(r2) / (SM.rad + SN.rad));
                        double fi = 1.0 + SM.poly * rr, fj = 1.0 + SN.poly * rr;
                        double hs = hscale(SM, SN) * fi * fj;
                        double hij = 0.5 * (sev[smu] + sev[snu]) * hs;
                        double dlp[3] = {0, 0, 0};
                        if (std::fabs(fi * fj) > 1e-18) {  // pbc/gradient.rs:781-806
                            double dp = (fi * SN.poly + fj * SM.poly) * 0.5 * rr / r2;
                            for (int k = 0; k < 3; ++k) dlp[k] = rv[k] * (dp / (fi * fj));
                        }
                        double s, da[3], db[3];
                        overlap_pair_deriv(b.ao[mu], b.ao[nu], SM.prims, SN.prims, ra, rb, s, da, db);
                        double p = P[(size_t)mu * n + nu], w = W[(size_t)mu * n + nu];
                        double oc = p * (2.0 * hij - (vao[mu] + vao[nu])) - 2.0 * w;
                        for (int k = 0; k < 3; ++k) {
                            ga[k] += da[k] * oc;
                            gb[k] += db[k] * oc;
                            double pg = dlp[k] * (2.0 * p * hij * s);
                            ga[k] += pg;
                            gb[k] -= pg;
                        }
                        if (cn_on) {
                            dca += dse[smu] * hs * p * s;
                            dcb += dse[snu] * hs * p * s;
                        }
                    }
                for (int k = 0; k < 3; ++k) {
                    grad[3 * a + k] += ga[k];
                    grad[3 * bb + k] += gb[k];
                }
                dedcn[a] += dca;
                dedcn[bb] += dcb;
            }
    }
    if (cn_on) {
        for (int s = 0; s < nsh; ++s)
            for (int i = b.sh[s].first_ao; i < b.sh[s].first_ao + b.sh[s].nao; ++i)
                dedcn[b.sh[s].atom] += dse[s] * P[(size_t)i * n + i];
        std::vector<double> cn2;
        std::vector<CnPair> cp;
        coordination_pbc(g, L, cn_cutoff, cn2, &cp);
        cn_chain_pbc(cp, dedcn, grad, 1e-12);
    }
    ewald_gradient(b, g, L, q, qat, grad);
}

extern "C" {

typedef struct {
    double charge;
    int max_scc;
    double energy_tolerance;
    double charge_tolerance;
    double mixing;
    int scc_broyden;
    int scc_broyden_size;
    double level_shift;
    int nprim;
    double eigen_tolerance;
    double electronic_temperature;
    int enable_dispersion;
    double integral_cutoff;
    double coordination_cutoff;
    int enable_cn_hamiltonian;
    int compute_gradient;
    int include_repulsion;
    int include_dispersion;
    int include_hamiltonian;
    int include_scc;
    int include_halogen;
} oracle_options;

// Output buffers are caller-allocated (sizes from oracle_basis_size); any
// pointer may be NULL to skip that output.
typedef struct {
    double energies[15];  // EnergyTerms::named_values order (electronic.rs:676-694)
    double fermi_level;
    double electronic_temperature;
    double nelec;
    double dipole[3];
    int iterations;
    int converged;
    double* gradient;            // 3*nat
    double* electronic_gradient; // 3*nat
    double* repulsion_gradient;  // 3*nat
    double* dispersion_gradient; // 3*nat
    double* halogen_gradient;    // 3*nat
    double* shell_charges;       // nsh
    double* atomic_charges;      // nat
    double* shell_scc_potential; // nsh
    double* coordination_numbers;// nat
    double* orbital_energies;    // nao
    double* occupations;         // nao
    double* overlap;             // nao*nao
    double* h0;                  // nao*nao
    double* fock;                // nao*nao
    double* density;             // nao*nao
    double* energy_weighted_density; // nao*nao
    char message[256];
} oracle_result;

void oracle_default_options(oracle_options* o) {
    // electronic.rs:395-452, hamiltonian.rs:25-32, gradient.rs:31-41
    o->charge = 0.0;
    o->max_scc = 250;
    o->energy_tolerance = 1e-6;
    o->charge_tolerance = 2e-5;
    o->mixing = 0.4;
    o->scc_broyden = 1;
    o->scc_broyden_size = 250;
    o->level_shift = 0.0;
    o->nprim = 0;
    o->eigen_tolerance = 1e-12;
    o->electronic_temperature = 300.0;
    o->enable_dispersion = 1;
    o->integral_cutoff = 30.0;
    o->coordination_cutoff = 30.0;
    o->enable_cn_hamiltonian = 1;
    o->compute_gradient = 1;
    o->include_repulsion = 1;
    o->include_dispersion = 1;
    o->include_hamiltonian = 1;
    o->include_scc = 1;
    o->include_halogen = 1;
}

// param_file: NULL for the bundled GFN1-xTB file of data_dir
void* oracle_create_params(const char* data_dir, const char* param_file) {
    Ctx* c = new Ctx();
    Err err{0, ""};
    std::string d(data_dir);
    const std::string pf = param_file ? std::string(param_file) : d + "/param_gfn1-xtb.txt";
    if (!parse_params(pf, c->p, err) || !load_sto(d + "/sto_ng.txt", c->sto, err) ||
        !load_d3(d + "/d3_reference.bin", *c, err)) {
        std::fprintf(stderr, "oracle_create: %s\n", err.msg.c_str());
        delete c;
        return nullptr;
    }
    return c;
}

void* oracle_create(const char* data_dir) { return oracle_create_params(data_dir, nullptr); }

void oracle_destroy(void* c) { delete (Ctx*)c; }

int oracle_basis_size(void* vctx, int nat, const int* z, int* nsh, int* nao) {
    Basis b;
    Err err{0, ""};
    if (!build_basis(*(Ctx*)vctx, nat, z, 0, b, err)) return err.code;
    *nsh = (int)b.sh.size();
    *nao = (int)b.ao.size();
    return 0;
}

int oracle_eigh(int n, const double* a, double* w, double* v) { return eigh(n, a, w, v) ? 0 : 1; }

double oracle_dispersion(void* vctx, int nat, const int* z, const double* xyz, double* grad) {
    Geo g{nat, z, xyz};
    if (grad) std::memset(grad, 0, sizeof(double) * 3 * nat);
    return dispersion(*(Ctx*)vctx, g, grad);
}
double oracle_repulsion(void* vctx, int nat, const int* z, const double* xyz, double* grad) {
    Geo g{nat, z, xyz};
    if (grad) std::memset(grad, 0, sizeof(double) * 3 * nat);
    return repulsion(*(Ctx*)vctx, g, grad);
}
double oracle_halogen(void* vctx, int nat, const int* z, const double* xyz, double* grad) {
    Geo g{nat, z, xyz};
    if (grad) std::memset(grad, 0, sizeof(double) * 3 * nat);
    return halogen(*(Ctx*)vctx, g, grad);
}
void oracle_coordination(int nat, const int* z, const double* xyz, double cutoff, double* cn) {
    Geo g{nat, z, xyz};
    std::vector<double> c;
    coordination(g, cutoff, c, nullptr);
    for (int i = 0; i < nat; ++i) cn[i] = c[i];
}

// Overlap matrix and its derivative contracted with a weight matrix:
// dS[a][k] = sum_{mu in a, nu not in a} Wt_{mu,nu} dS_{mu nu}/dR_{a,k} (used by
// FD tests of the integral derivatives).  Any output may be NULL.
int oracle_overlap(void* vctx, int nat, const int* z, const double* xyz, double cutoff, double* S,
                   const double* Wt, double* dS) {
    Ctx& ctx = *(Ctx*)vctx;
    Basis b;
    Err err{0, ""};
    if (!build_basis(ctx, nat, z, 0, b, err)) return err.code;
    Geo g{nat, z, xyz};
    int n = (int)b.ao.size();
    if (S) std::memset(S, 0, sizeof(double) * n * n);
    if (dS) std::memset(dS, 0, sizeof(double) * 3 * nat);
    double c2 = cutoff * cutoff;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const AO &A = b.ao[i], &B = b.ao[j];
            int ai = A.atom, aj = B.atom;
            V3 ca = g.pos(ai), cb = g.pos(aj);
            if (ai != aj) {
                double r2 = g.r2(ai, aj);
                if (r2 <= PAIR_EPS2 || r2 > c2) continue;
            }
            if (S) S[(size_t)i * n + j] = overlap_pair(A, B, b.sh[A.shell].prims, b.sh[B.shell].prims, ca, cb);
            if (dS && Wt && ai != aj) {
                double s, da[3], db[3];
                overlap_pair_deriv(A, B, b.sh[A.shell].prims, b.sh[B.shell].prims, ca, cb, s, da, db);
                for (int k = 0; k < 3; ++k) {
                    dS[3 * ai + k] += Wt[(size_t)i * n + j] * da[k];
                    dS[3 * aj + k] += Wt[(size_t)i * n + j] * db[k];
                }
            }
        }
    return 0;
}

// Parsed parameter values, for the tests that compare the parser with an
// independent reading of the parameter file.  key: "present", "nshell", "gam",
// "gam3", "repa", "repb", "cxb" (element), or "n", "l", "lev", "exp", "poly",
// "lpar" (shell `shell` of the element, file order).  Returns 0 on success.
int oracle_element_value(void* vctx, int z, const char* key, int shell, double* out) {
    const Params& p = ((Ctx*)vctx)->p;
    if (z < 1 || z > 86) return 1;
    const ElementParam& e = p.el[z];
    const std::string k(key);
    if (k == "present") { *out = e.present ? 1.0 : 0.0; return 0; }
    if (k == "nshell") { *out = (double)e.shells.size(); return 0; }
    if (k == "gam") { *out = e.gam; return 0; }
    if (k == "gam3") { *out = e.gam3; return 0; }
    if (k == "repa") { *out = e.repa; return 0; }
    if (k == "repb") { *out = e.repb; return 0; }
    if (k == "cxb") { *out = e.cxb; return 0; }
    if (shell < 0 || shell >= (int)e.shells.size()) return 2;
    const ShellParam& sp = e.shells[shell];
    if (k == "n") { *out = sp.n; return 0; }
    if (k == "l") { *out = sp.l; return 0; }
    if (k == "lev") { *out = sp.level_ev; return 0; }
    if (k == "exp") { *out = sp.slater; return 0; }
    if (k == "poly") { *out = sp.poly; return 0; }
    if (k == "lpar") { *out = sp.lpar; return 0; }
    return 3;
}
// Global parameter (NaN when the file has none) and the pair scaling factor.
double oracle_global(void* vctx, const char* key) {
    const Params& p = ((Ctx*)vctx)->p;
    auto it = p.glob.find(lower(key));
    return it == p.glob.end() ? NAN : it->second;
}
double oracle_pair_scaling(void* vctx, int a, int b) { return ((Ctx*)vctx)->p.pair_scaling(a, b); }

// Shells of a system in basis order: atom, angular momentum, valence flag,
// AO count, hardness and the CN-free self energy (Eh).  Arrays of nsh entries.
int oracle_shells(void* vctx, int nat, const int* z, int* atom, int* l, int* valence, int* nao,
                  double* hardness, double* hdiag) {
    Basis b;
    Err err{0, ""};
    if (!build_basis(*(Ctx*)vctx, nat, z, 0, b, err)) return err.code;
    for (size_t s = 0; s < b.sh.size(); ++s) {
        atom[s] = b.sh[s].atom;
        l[s] = b.sh[s].l;
        valence[s] = b.sh[s].valence ? 1 : 0;
        nao[s] = b.sh[s].nao;
        hardness[s] = b.sh[s].hardness;
        hdiag[s] = b.sh[s].hdiag;
    }
    return 0;
}

// The shell Coulomb matrix gamma (nsh x nsh) that the SCC uses.
int oracle_gamma(void* vctx, int nat, const int* z, const double* xyz, double* out) {
    Basis b;
    Err err{0, ""};
    if (!build_basis(*(Ctx*)vctx, nat, z, 0, b, err)) return err.code;
    Geo g{nat, z, xyz};
    std::vector<double> gam;
    gamma_matrix(b, g, gam);
    std::copy(gam.begin(), gam.end(), out);
    return 0;
}

}  // extern "C"

namespace {

// BroydenMixer (electronic.rs:3203-3309)
struct Broyden {
    int memory, iter = 0;
    double damp;
    std::vector<std::vector<double>> df, u;
    std::vector<double> omega, qlast, dqlast;
    Broyden(int ndim, int mem, double d)
        : memory(mem), damp(d), df(mem, std::vector<double>(ndim, 0.0)), u(mem, std::vector<double>(ndim, 0.0)),
          omega(mem, 0.0), qlast(ndim, 0.0), dqlast(ndim, 0.0) {}
    static double dot(const std::vector<double>& a, const std::vector<double>& b) {
        double s = 0;
        for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
        return s;
    }
    // electronic.rs:3767-3806 Gauss-Jordan with partial pivoting
    static bool solve(std::vector<std::vector<double>> a, std::vector<double>& b) {
        int n = (int)b.size();
        for (int col = 0; col < n; ++col) {
            int piv = col;
            double pa = std::fabs(a[col][col]);
            for (int r = col + 1; r < n; ++r)
                if (std::fabs(a[r][col]) > pa) { piv = r; pa = std::fabs(a[r][col]); }
            if (pa < 1e-18) return false;
            if (piv != col) { std::swap(a[piv], a[col]); std::swap(b[piv], b[col]); }
            double dg = a[col][col];
            for (int j = col; j < n; ++j) a[col][j] /= dg;
            b[col] /= dg;
            for (int r = 0; r < n; ++r) {
                if (r == col) continue;
                double f = a[r][col];
                if (f == 0.0) continue;
                for (int j = col; j < n; ++j) a[r][j] -= f * a[col][j];
                b[r] -= f * b[col];
            }
        }
        return true;
    }
    bool next(const std::vector<double>& q, const std::vector<double>& dq, std::vector<double>& out) {
        iter += 1;
        size_t nd = q.size();
        if (iter == 1) {
            dqlast = dq;
            qlast = q;
            out.resize(nd);
            for (size_t i = 0; i < nd; ++i) out[i] = q[i] + damp * dq[i];
            return true;
        }
        int itn = iter - 1;
        int it1 = (itn - 1) % memory;
        const double omega0 = 0.01, minw = 1.0, maxw = 100000.0, wfac = 0.01;
        double dqn = std::sqrt(dot(dq, dq));
        omega[it1] = dqn > wfac / maxw ? wfac / dqn : maxw;
        if (omega[it1] < minw) omega[it1] = minw;
        std::vector<double> dfc(nd);
        for (size_t i = 0; i < nd; ++i) dfc[i] = dq[i] - dqlast[i];
        double inv = 1.0 / std::max(std::sqrt(dot(dfc, dfc)), 2.220446049250313e-16);
        for (auto& v : dfc) v *= inv;
        df[it1] = dfc;
        for (size_t i = 0; i < nd; ++i) u[it1][i] = damp * df[it1][i] + inv * (q[i] - qlast[i]);
        int start = itn > memory ? itn - memory + 1 : 1;
        std::vector<int> active;
        for (int j = start; j <= itn; ++j) active.push_back((j - 1) % memory);
        int m = (int)active.size();
        std::vector<std::vector<double>> beta(m, std::vector<double>(m, 0.0));
        std::vector<double> c(m);
        for (int r = 0; r < m; ++r) {
            int i = active[r];
            c[r] = omega[i] * dot(df[i], dq);
            for (int col = 0; col < m; ++col) {
                int j = active[col];
                beta[r][col] = omega[i] * omega[j] * dot(df[i], df[j]);
            }
            beta[r][r] += omega0 * omega0;
        }
        if (!solve(beta, c)) return false;
        dqlast = dq;
        qlast = q;
        out.resize(nd);
        for (size_t i = 0; i < nd; ++i) out[i] = q[i] + damp * dq[i];
        for (int r = 0; r < m; ++r) {
            int i = active[r];
            double sc = omega[i] * c[r];
            for (size_t k = 0; k < nd; ++k) out[k] -= sc * u[i][k];
        }
        return true;
    }
};

bool trusted(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x) || std::fabs(x) >= 10.0) return false;
    return true;
}

struct Occ {
    std::vector<double> f;
    double mu, ts;
};

// electronic.rs:3853-3921 fermi_occupations (bisection to the f64 fixed point)
Occ fermi(const std::vector<double>& eps, double nelec, double kt) {
    int n = (int)eps.size();
    Occ o;
    auto focc = [&](double e, double mu) {
        double x = std::min(80.0, std::max(-80.0, (e - mu) / kt));
        return 2.0 / (1.0 + std::exp(x));
    };
    if (kt <= 0.0) {  // aufbau (electronic.rs:3829-3851)
        o.f.assign(n, 0.0);
        double rem = std::max(nelec, 0.0);
        for (int i = 0; i < n; ++i) {
            double fill = std::min(rem, 2.0);
            o.f[i] = fill;
            rem -= fill;
            if (rem <= 0.0) break;
        }
        int homo = 0;
        for (int i = n - 1; i >= 0; --i)
            if (o.f[i] > 1e-12) { homo = i; break; }
        o.mu = n ? eps[homo] : 0.0;
        o.ts = 0.0;
        return o;
    }
    if (nelec <= 0.0) { o.f.assign(n, 0.0); o.mu = n ? eps[0] : 0.0; o.ts = 0.0; return o; }
    if (nelec >= 2.0 * n) { o.f.assign(n, 2.0); o.mu = n ? eps[n - 1] : 0.0; o.ts = 0.0; return o; }
    double mn = INFINITY, mx = -INFINITY;
    for (double e : eps) { mn = std::min(mn, e); mx = std::max(mx, e); }
    double lo = mn - 100.0 * kt - 10.0, hi = mx + 100.0 * kt + 10.0;
    for (int it = 0; it < 200; ++it) {
        double mu = 0.5 * (lo + hi), sum = 0.0;
        for (double e : eps) sum += focc(e, mu);
        double nlo = lo, nhi = hi;
        if (sum < nelec) nlo = mu; else nhi = mu;
        if (nlo == lo && nhi == hi) break;
        lo = nlo;
        hi = nhi;
    }
    o.mu = 0.5 * (lo + hi);
    o.f.resize(n);
    o.ts = 0.0;
    for (int i = 0; i < n; ++i) {
        o.f[i] = focc(eps[i], o.mu);
        double x = std::min(1.0 - 1e-16, std::max(1e-16, 0.5 * o.f[i]));
        o.ts += 2.0 * kt * (x * std::log(x) + (1.0 - x) * std::log(1.0 - x));
    }
    return o;
}

struct Step {
    std::vector<double> fock, P, W, eps, f, q;
    double mu, ts;
};

}  // namespace

extern "C" {

// run_electronic (electronic.rs:800-2031) followed by analytic_gradient_from_result
// (gradient.rs:122-242), stock model, restricted.  With cell != NULL (three
// cell vectors, bohr, row after row) the periodic Gamma-point path that
// run_electronic / analytic_gradient dispatch a lattice to
// (electronic.rs:829-832, gradient.rs:61-66): run_pbc_scc (pbc/scf.rs:140-580),
// pbc_electronic_result (:602-677) and pbc_gradient_from_scc
// (pbc/gradient.rs:75-163).
int oracle_run_cell(void* vctx, int nat, const int* z, const double* xyz, const double* cell,
                    const oracle_options* opt, const double* initial_shell_charges, oracle_result* res) {
    Ctx& ctx = *(Ctx*)vctx;
    Err err{0, ""};
    auto fail = [&](int code, const std::string& m) {
        std::snprintf(res->message, sizeof(res->message), "%s", m.c_str());
        return code;
    };
    Basis b;
    if (!build_basis(ctx, nat, z, opt->nprim, b, err)) return fail(err.code, err.msg);
    Geo g{nat, z, xyz};
    for (int i = 0; i < 3 * nat; ++i)
        if (!std::isfinite(xyz[i])) return fail(4, "non-finite coordinate");
    Lattice lat;
    const Lattice* L = nullptr;
    if (cell) {
        for (int i = 0; i < 9; ++i)
            if (!std::isfinite(cell[i])) return fail(4, "non-finite cell vector");
        if (!make_lattice(cell, lat)) return fail(7, "singular cell");
        L = &lat;
    }
    const int n = (int)b.ao.size(), nsh = (int)b.sh.size();
    if (n == 0) return fail(4, "cannot run GFN1 SCC for an empty basis");
    double nelec = b.nref - opt->charge;
    if (!std::isfinite(nelec) || nelec < -1e-8 || nelec > 2.0 * n + 1e-8)
        return fail(4, "electron count out of range");

    // ---- CN and H0 (hamiltonian.rs:44-134)
    std::vector<double> cn(nat, 0.0);
    std::vector<CnPair> cnpairs;
    if (opt->enable_cn_hamiltonian) {
        if (L) coordination_pbc(g, *L, opt->coordination_cutoff, cn, &cnpairs);
        else coordination(g, opt->coordination_cutoff, cn, &cnpairs);
    }
    std::vector<double> S, H0, se, dsedcn;
    auto hscale = [&](const Shell& si, const Shell& sj) { return hscale_fn(ctx, si, sj); };
    s_h0(ctx, b, g, L, opt->integral_cutoff, cn, S, H0, se, dsedcn);
    double ic2 = opt->integral_cutoff * opt->integral_cutoff;

    // ---- gamma (coulomb.rs:82-107)
    std::vector<double> gam;
    if (L) ewald_gamma_matrix(b, g, *L, gam);  // pbc/ewald.rs:253-320
    else gamma_matrix(b, g, gam);

    // ---- Lowdin orthogonaliser (linalg.rs:193-217)
    std::vector<double> X((size_t)n * n), Xt((size_t)n * n);
    {
        std::vector<double> w(n), V((size_t)n * n);
        if (!eigh(n, S.data(), w.data(), V.data())) return fail(5, "eigensolver failed");
        for (int a = 0; a < n; ++a)
            if (w[a] <= opt->eigen_tolerance) return fail(4, "overlap matrix is not positive definite");
        std::vector<double> sv((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < n; ++a) sv[(size_t)i * n + a] = V[(size_t)i * n + a] / std::sqrt(w[a]);
        gemm(n, n, n, sv.data(), false, V.data(), true, X.data());
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Xt[(size_t)j * n + i] = X[(size_t)i * n + j];
    }

    // ---- geometry-only energies
    double erep = L ? repulsion_pbc(ctx, g, *L, nullptr) : repulsion(ctx, g, nullptr);
    double edisp = !opt->enable_dispersion ? 0.0 : L ? dispersion_pbc(ctx, g, *L, nullptr) : dispersion(ctx, g, nullptr);
    double ehal = L ? halogen_pbc(ctx, g, *L, nullptr) : halogen(ctx, g, nullptr);

    // ---- SCC helpers
    std::vector<int> atom_first_sh(nat, -1);
    for (int s = nsh - 1; s >= 0; --s) atom_first_sh[b.sh[s].atom] = s;
    auto coulomb = [&](const std::vector<double>& q, double& e2, double& e3, std::vector<double>& v,
                       std::vector<double>& qat) {
        v.assign(nsh, 0.0);
        for (int i = 0; i < nsh; ++i) {
            double s = 0;
            for (int j = 0; j < nsh; ++j) s += gam[(size_t)i * nsh + j] * q[j];
            v[i] = s;
        }
        e2 = 0.0;
        for (int i = 0; i < nsh; ++i) e2 += q[i] * v[i];
        e2 *= 0.5;
        qat.assign(nat, 0.0);
        for (int s = 0; s < nsh; ++s) qat[b.sh[s].atom] += q[s];
        e3 = 0.0;
        for (int a = 0; a < nat; ++a) {
            int f0 = atom_first_sh[a];
            if (f0 < 0) continue;
            double g3 = b.sh[f0].gam3;
            e3 += qat[a] * qat[a] * qat[a] * g3 / 3.0;
            double pot = qat[a] * qat[a] * g3;
            for (int s = 0; s < nsh; ++s)
                if (b.sh[s].atom == a) v[s] += pot;
        }
    };
    std::vector<double> refsh(nsh);
    for (int s = 0; s < nsh; ++s) refsh[s] = b.sh[s].ref_occ;

    auto scc_step = [&](const std::vector<double>& qin, double kt, double lshift, const std::vector<double>* prevP,
                        Step& st) -> bool {
        double e2, e3;
        std::vector<double> v, qat;
        coulomb(qin, e2, e3, v, qat);
        std::vector<double> vao(n);
        for (int i = 0; i < n; ++i) vao[i] = v[b.ao[i].shell];
        st.fock = H0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                double val = -S[(size_t)i * n + j] * 0.5 * (vao[i] + vao[j]);
                st.fock[(size_t)i * n + j] += val;
                if (i != j) st.fock[(size_t)j * n + i] += val;
            }
        if (lshift != 0.0 && prevP) {  // electronic.rs:2120-2133
            std::vector<double> sp((size_t)n * n), sps((size_t)n * n);
            gemm(n, n, n, S.data(), false, prevP->data(), false, sp.data());
            gemm(n, n, n, sp.data(), false, S.data(), false, sps.data());
            for (size_t k = 0; k < (size_t)n * n; ++k) st.fock[k] += lshift * (S[k] - 0.5 * sps[k]);
        }
        if (const char* dump = std::getenv("GFN1_ORACLE_DUMP_FOCK")) {
            // debugging aid: append each iteration's Fock matrix (n x n doubles)
            FILE* fp = std::fopen(dump, "ab");
            if (fp) {
                std::fwrite(st.fock.data(), sizeof(double), st.fock.size(), fp);
                std::fclose(fp);
            }
        }
        std::vector<double> fx((size_t)n * n), fo((size_t)n * n), V((size_t)n * n), C((size_t)n * n);
        gemm(n, n, n, st.fock.data(), false, X.data(), false, fx.data());
        gemm(n, n, n, Xt.data(), false, fx.data(), false, fo.data());
        st.eps.resize(n);
        if (!eigh(n, fo.data(), st.eps.data(), V.data())) return false;
        gemm(n, n, n, X.data(), false, V.data(), false, C.data());
        Occ o = fermi(st.eps, nelec, kt);
        st.f = o.f;
        st.mu = o.mu;
        st.ts = o.ts;
        // column_weighted_gram (linalg.rs:243-273)
        auto gram = [&](const std::vector<double>& wts, std::vector<double>& out) {
            std::vector<double> sc((size_t)n * n, 0.0);
            for (int c = 0; c < n; ++c) {
                if (wts[c] == 0.0) continue;
                for (int r = 0; r < n; ++r) sc[(size_t)r * n + c] = C[(size_t)r * n + c] * wts[c];
            }
            out.assign((size_t)n * n, 0.0);
            gemm(n, n, n, sc.data(), false, C.data(), true, out.data());
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < i; ++j) {
                    double avg = 0.5 * (out[(size_t)i * n + j] + out[(size_t)j * n + i]);
                    out[(size_t)i * n + j] = out[(size_t)j * n + i] = avg;
                }
        };
        gram(st.f, st.P);
        std::vector<double> ef(n);
        for (int i = 0; i < n; ++i) ef[i] = st.eps[i] * st.f[i];
        gram(ef, st.W);
        // mulliken_shell_charges (electronic.rs:3160-3180)
        st.q.assign(nsh, 0.0);
        for (int s = 0; s < nsh; ++s) {
            double pop = 0.0;
            for (int i = b.sh[s].first_ao; i < b.sh[s].first_ao + b.sh[s].nao; ++i) {
                double rowsum = 0.0;
                for (int j = 0; j < n; ++j) rowsum += S[(size_t)i * n + j] * st.P[(size_t)i * n + j];
                pop += rowsum;
            }
            st.q[s] = refsh[s] - pop;
        }
        return true;
    };
    auto eel = [&](const std::vector<double>& P) {
        double s = 0;
        for (size_t k = 0; k < (size_t)n * n; ++k) s += H0[k] * P[k];
        return s;
    };

    // ---- SCC loop with the robust-fallback ladder (electronic.rs:1390-1947)
    std::vector<double> q0(nsh, 0.0);
    if (initial_shell_charges)
        for (int s = 0; s < nsh; ++s) q0[s] = initial_shell_charges[s];
    double base_mixing = std::min(1.0, std::max(0.01, opt->mixing));
    bool base_broyden = opt->scc_broyden != 0;
    int mem = std::max(opt->scc_broyden_size, 2);
    int robust = std::max(opt->max_scc, 800);
    struct Rung {
        bool broyden;
        double mix, ls, T;
        int maxit;
    } ladder[4] = {
        {base_broyden, base_mixing, opt->level_shift, opt->electronic_temperature, opt->max_scc},
        {false, 0.20, std::max(opt->level_shift, 0.20), opt->electronic_temperature, robust},
        {false, 0.10, std::max(opt->level_shift, 0.10), opt->electronic_temperature, robust},
        {true, base_mixing, 0.0, std::max(opt->electronic_temperature, 3000.0), std::max(opt->max_scc, 250)},
    };
    bool converged = false;
    int iterations = 0;
    double final_rms = INFINITY, used_T = opt->electronic_temperature;
    Step last;
    bool have_last = false;
    // the periodic SCC runs the caller's scheme once, with no level shift and
    // no fallback (pbc/scf.rs:358-529)
    if (L) ladder[0].ls = 0.0;
    for (int rung = 0; rung < (L ? 1 : 4) && !converged; ++rung) {
        const Rung& R = ladder[rung];
        used_T = R.T;
        std::vector<double> qmix = q0;
        Broyden br(nsh, mem, R.mix);
        std::vector<double> prevP;
        bool have_prev = false;
        double last_e = INFINITY;
        bool have_e = false;
        have_last = false;
        final_rms = INFINITY;
        double kt = std::max(R.T, 0.0) * KB;
        for (int it = 1; it <= R.maxit; ++it) {
            iterations = it;
            Step st;
            if (!scc_step(qmix, kt, R.ls, have_prev ? &prevP : nullptr, st)) return fail(5, "eigensolver failed");
            double ss = 0;
            for (int s = 0; s < nsh; ++s) ss += (st.q[s] - qmix[s]) * (st.q[s] - qmix[s]);
            final_rms = nsh ? std::sqrt(ss / nsh) : 0.0;
            double e2, e3;
            std::vector<double> v, qat;
            coulomb(st.q, e2, e3, v, qat);
            double e = eel(st.P) + e2 + e3 + st.ts;
            double de = have_e ? std::fabs(e - last_e) : INFINITY;
            if (de < opt->energy_tolerance && final_rms < opt->charge_tolerance) {
                converged = true;
                last = st;
                have_last = true;
                break;
            }
            std::vector<double> resid(nsh), nxt;
            for (int s = 0; s < nsh; ++s) resid[s] = st.q[s] - qmix[s];
            bool ok = false;
            if (R.broyden) ok = br.next(qmix, resid, nxt) && trusted(nxt);
            if (!ok) {
                nxt.resize(nsh);
                for (int s = 0; s < nsh; ++s) nxt[s] = qmix[s] + R.mix * resid[s];
            }
            if (R.ls != 0.0) { prevP = st.P; have_prev = true; } else have_prev = false;
            qmix = nxt;
            last_e = e;
            have_e = true;
            last = st;
            have_last = true;
        }
    }
    res->iterations = iterations;
    res->converged = converged ? 1 : 0;
    if (!converged) {
        std::snprintf(res->message, sizeof(res->message), "SCC not converged after %d iterations (rms %.3e)",
                      iterations, final_rms);
        return 6;
    }
    (void)have_last;
    // ---- final assembly (electronic.rs:1957-2013)
    double e_el = eel(last.P), e2, e3;
    std::vector<double> vsh, qat;
    coulomb(last.q, e2, e3, vsh, qat);
    double total_internal = e_el + e2 + e3 + erep + edisp + ehal;
    double total_free = total_internal + last.ts;
    double* E = res->energies;
    for (int k = 0; k < 15; ++k) E[k] = 0.0;
    E[0] = total_free; E[1] = total_internal; E[2] = erep; E[3] = e_el; E[4] = e2; E[5] = e3;
    E[7] = edisp; E[8] = ehal; E[14] = last.ts;
    res->fermi_level = last.mu;
    res->electronic_temperature = used_T;
    res->nelec = nelec;
    for (int k = 0; k < 3; ++k) {
        double d = 0;
        for (int a = 0; a < nat; ++a) d += xyz[3 * a + k] * qat[a];
        res->dipole[k] = d;
    }
    auto put = [](double* dst, const std::vector<double>& src) {
        if (dst) std::memcpy(dst, src.data(), sizeof(double) * src.size());
    };
    put(res->shell_charges, last.q);
    put(res->atomic_charges, qat);
    put(res->shell_scc_potential, vsh);
    put(res->coordination_numbers, cn);
    put(res->orbital_energies, last.eps);
    put(res->occupations, last.f);
    put(res->overlap, S);
    put(res->h0, H0);
    put(res->fock, last.fock);
    put(res->density, last.P);
    put(res->energy_weighted_density, last.W);
    if (!opt->compute_gradient) return 0;

    // ---- gradient (gradient.rs:122-531)
    std::vector<double> gel(3 * nat, 0.0), grep(3 * nat, 0.0), gdisp(3 * nat, 0.0), ghal(3 * nat, 0.0);
    if (L) {
        periodic_gradient(b, g, *L, opt->enable_cn_hamiltonian != 0, opt->coordination_cutoff, hscale, cn, vsh, last.q, qat,
                          last.P, last.W, gel.data());
        repulsion_pbc(ctx, g, *L, grep.data());
        if (opt->enable_dispersion) dispersion_pbc(ctx, g, *L, gdisp.data());
        halogen_pbc(ctx, g, *L, ghal.data());
    } else if (opt->include_hamiltonian || opt->include_scc) {
        bool cn_on = opt->enable_cn_hamiltonian != 0, scc_on = opt->include_scc != 0;
        std::vector<double> dedcn(nat, 0.0), sev(nsh), dse(nsh);
        for (int s = 0; s < nsh; ++s) {
            dse[s] = cn_on ? -b.sh[s].kcn : 0.0;
            sev[s] = b.sh[s].hdiag + dse[s] * cn[b.sh[s].atom];
        }
        std::vector<double> vao(n, 0.0);
        if (scc_on)
            for (int i = 0; i < n; ++i) vao[i] = vsh[b.ao[i].shell];
        std::vector<int> first_sh(nat, -1), nsh_at(nat, 0);
        for (int s = 0; s < nsh; ++s) {
            if (first_sh[b.sh[s].atom] < 0) first_sh[b.sh[s].atom] = s;
            nsh_at[b.sh[s].atom]++;
        }
        for (int an = 0; an < nat; ++an)
            for (int am = an + 1; am < nat; ++am) {
                double r2 = g.r2(an, am);
                if (r2 <= PAIR_EPS2 || r2 > ic2) continue;
                if (r2 <= 1e-12) continue;  // gradient.rs:373
                V3 rmu = g.pos(am), rnu = g.pos(an);
                double rvec[3] = {rmu.x - rnu.x, rmu.y - rnu.y, rmu.z - rnu.z};
                double gmu[3] = {0, 0, 0}, gnu[3] = {0, 0, 0}, dcmu = 0, dcnu = 0;
                for (int smu = first_sh[am]; smu < first_sh[am] + nsh_at[am]; ++smu)
                    for (int snu = first_sh[an]; snu < first_sh[an] + nsh_at[an]; ++snu) {
                        const Shell &SM = b.sh[smu], &SN = b.sh[snu];
                        double radsum = SM.rad + SN.rad;
                        double rr = std::sqrt(std::sqrt(r2) / radsum);
                        double poly = (1.0 + SM.poly * rr) * (1.0 + SN.poly * rr);
                        double hs = hscale(SM, SN) * poly;
                        double hij = 0.5 * (sev[smu] + sev[snu]) * hs;
                        // gradient.rs:1157-1175 log-derivative of the polynomial
                        double dlp[3] = {0, 0, 0};
                        {
                            double fi = 1.0 + SM.poly * rr, fj = 1.0 + SN.poly * rr;
                            double pl = fi * fj;
                            if (std::fabs(pl) > 1e-18) {
                                double dp = (fi * SN.poly + fj * SM.poly) * 0.5 * rr / r2;
                                for (int k = 0; k < 3; ++k) dlp[k] = rvec[k] * (dp / pl);
                            }
                        }
                        for (int mu = SM.first_ao; mu < SM.first_ao + SM.nao; ++mu)
                            for (int nu = SN.first_ao; nu < SN.first_ao + SN.nao; ++nu) {
                                double s, da[3], db[3];
                                overlap_pair_deriv(b.ao[mu], b.ao[nu], SM.prims, SN.prims, rmu, rnu, s, da, db);
                                double p = last.P[(size_t)mu * n + nu], w = last.W[(size_t)mu * n + nu];
                                double shift = scc_on ? vao[mu] + vao[nu] : 0.0;
                                double oc = p * (2.0 * hij - shift) - 2.0 * w;
                                double hp = p * hij;
                                for (int k = 0; k < 3; ++k) {
                                    gmu[k] += da[k] * oc;
                                    gnu[k] += db[k] * oc;
                                    double pg = dlp[k] * (2.0 * hp * s);
                                    gmu[k] += pg;
                                    gnu[k] -= pg;
                                }
                                if (cn_on) {
                                    dcmu += dse[smu] * hs * p * s;
                                    dcnu += dse[snu] * hs * p * s;
                                }
                            }
                    }
                for (int k = 0; k < 3; ++k) {
                    gel[3 * am + k] += gmu[k];
                    gel[3 * an + k] += gnu[k];
                }
                dedcn[am] += dcmu;
                dedcn[an] += dcnu;
            }
        if (cn_on) {
            for (int s = 0; s < nsh; ++s)
                for (int i = b.sh[s].first_ao; i < b.sh[s].first_ao + b.sh[s].nao; ++i)
                    dedcn[b.sh[s].atom] += dse[s] * last.P[(size_t)i * n + i];
            std::vector<double> cn2;
            std::vector<CnPair> cp2;
            coordination(g, opt->coordination_cutoff, cn2, &cp2);
            cn_chain(cp2, dedcn, gel.data(), 1e-12);
        }
    }
    if (opt->include_scc && !L) {  // gradient.rs:493-516
        for (int i = 0; i < nsh; ++i) {
            int ai = b.sh[i].atom;
            for (int j = 0; j < i; ++j) {
                int aj = b.sh[j].atom;
                if (ai == aj) continue;
                double rv[3] = {xyz[3 * ai] - xyz[3 * aj], xyz[3 * ai + 1] - xyz[3 * aj + 1],
                                xyz[3 * ai + 2] - xyz[3 * aj + 2]};
                double r = std::sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
                if (r <= 1e-12) continue;
                double h = 2.0 / (1.0 / b.sh[i].hardness + 1.0 / b.sh[j].hardness);
                double den = r * r + 1.0 / (h * h);
                double pref = -1.0 / (den * std::sqrt(den)) * last.q[i] * last.q[j];
                for (int k = 0; k < 3; ++k) {
                    gel[3 * ai + k] += rv[k] * pref;
                    gel[3 * aj + k] -= rv[k] * pref;
                }
            }
        }
    }
    if (opt->include_repulsion && !L) repulsion(ctx, g, grep.data());
    if (opt->include_dispersion && opt->enable_dispersion && !L) dispersion(ctx, g, gdisp.data());
    if (opt->include_halogen && !L) halogen(ctx, g, ghal.data());
    std::vector<double> gt(3 * nat);
    for (int k = 0; k < 3 * nat; ++k) gt[k] = gel[k] + grep[k] + gdisp[k] + ghal[k];
    put(res->gradient, gt);
    put(res->electronic_gradient, gel);
    put(res->repulsion_gradient, grep);
    put(res->dispersion_gradient, gdisp);
    put(res->halogen_gradient, ghal);
    return 0;
}

int oracle_run(void* vctx, int nat, const int* z, const double* xyz, const oracle_options* opt,
               const double* initial_shell_charges, oracle_result* res) {
    return oracle_run_cell(vctx, nat, z, xyz, nullptr, opt, initial_shell_charges, res);
}

// ---- pieces of the periodic path, for the tests that pin them
// Ewald matrix of unit point charges (nat x nat) at splitting alpha (<= 0: the default).
int oracle_ewald_atoms(int nat, const double* xyz, const double* cell, double alpha, double* out) {
    Lattice L;
    if (!make_lattice(cell, L)) return 7;
    std::vector<int> z(nat, 1);
    Geo g{nat, z.data(), xyz};
    std::vector<double> phi;
    ewald_atom_matrix(g, L, alpha > 0.0 ? alpha : ewald_alpha(L), phi);
    std::copy(phi.begin(), phi.end(), out);
    return 0;
}

// Periodic shell Coulomb matrix Gamma (nsh x nsh) at splitting alpha (<= 0: the default).
int oracle_gamma_cell(void* vctx, int nat, const int* z, const double* xyz, const double* cell, double alpha,
                      double* out) {
    Basis b;
    Err err{0, ""};
    if (!build_basis(*(Ctx*)vctx, nat, z, 0, b, err)) return err.code;
    Lattice L;
    if (!make_lattice(cell, L)) return 7;
    Geo g{nat, z, xyz};
    std::vector<double> gam;
    ewald_gamma_matrix(b, g, L, gam, alpha);
    std::copy(gam.begin(), gam.end(), out);
    return 0;
}

// Gradient of 1/2 q^T Gamma q at fixed shell charges q (3 nat).
int oracle_gamma_cell_gradient(void* vctx, int nat, const int* z, const double* xyz, const double* cell,
                               const double* q, double* grad) {
    Basis b;
    Err err{0, ""};
    if (!build_basis(*(Ctx*)vctx, nat, z, 0, b, err)) return err.code;
    Lattice L;
    if (!make_lattice(cell, L)) return 7;
    Geo g{nat, z, xyz};
    std::vector<double> qs(q, q + b.sh.size()), qat(nat, 0.0);
    for (size_t s = 0; s < b.sh.size(); ++s) qat[b.sh[s].atom] += qs[s];
    std::fill(grad, grad + 3 * nat, 0.0);
    ewald_gradient(b, g, L, qs, qat, grad);
    return 0;
}

// Periodic classical terms: which = 0 repulsion, 1 dispersion, 2 halogen; grad nullable.
double oracle_classical_cell(void* vctx, int nat, const int* z, const double* xyz, const double* cell, int which,
                             double* grad) {
    Ctx& ctx = *(Ctx*)vctx;
    Lattice L;
    if (!make_lattice(cell, L)) return NAN;
    Geo g{nat, z, xyz};
    if (grad) std::fill(grad, grad + 3 * nat, 0.0);
    if (which == 0) return repulsion_pbc(ctx, g, L, grad);
    if (which == 1) return dispersion_pbc(ctx, g, L, grad);
    return halogen_pbc(ctx, g, L, grad);
}

// S and H0 of a molecule (cell NULL, pairs within 30 bohr) or S(Gamma),
// H0(Gamma) of a cell, with or without the CN shift of the self energies.
int oracle_hs(void* vctx, int nat, const int* z, const double* xyz, const double* cell, int cn_on, double* S,
              double* H0) {
    Ctx& ctx = *(Ctx*)vctx;
    Basis b;
    Err err{0, ""};
    if (!build_basis(ctx, nat, z, 0, b, err)) return err.code;
    Geo g{nat, z, xyz};
    Lattice lat;
    const Lattice* L = nullptr;
    if (cell) {
        if (!make_lattice(cell, lat)) return 7;
        L = &lat;
    }
    std::vector<double> cn(nat, 0.0);
    if (cn_on) {
        if (L) coordination_pbc(g, *L, 30.0, cn, nullptr);
        else coordination(g, 30.0, cn, nullptr);
    }
    std::vector<double> s, h, se, ds;
    s_h0(ctx, b, g, L, 30.0, cn, s, h, se, ds);
    std::copy(s.begin(), s.end(), S);
    std::copy(h.begin(), h.end(), H0);
    return 0;
}

// Lattice-summed coordination numbers.
int oracle_coordination_cell(int nat, const int* z, const double* xyz, const double* cell, double cutoff,
                             double* cn) {
    Lattice L;
    if (!make_lattice(cell, L)) return 7;
    Geo g{nat, z, xyz};
    std::vector<double> c;
    coordination_pbc(g, L, cutoff, c, nullptr);
    std::copy(c.begin(), c.end(), cn);
    return 0;
}

// Run many independent systems (the CPU baseline for batched throughput):
// instances are distributed over OpenMP threads; returns total_free per
// instance in energies_out and the flattened gradients in grad_out (offsets by
// atom_offsets).  status_out gets each instance's return code.
int oracle_run_batch(void* vctx, int ninst, const int* atom_offsets, const int* z, const double* xyz,
                     const oracle_options* opt, double* energies_out, double* grad_out, int* iters_out,
                     int* status_out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int i = 0; i < ninst; ++i) {
        int a0 = atom_offsets[i], na = atom_offsets[i + 1] - a0;
        oracle_result r;
        std::memset(&r, 0, sizeof(r));
        r.gradient = grad_out ? grad_out + 3 * a0 : nullptr;
        status_out[i] = oracle_run(vctx, na, z + a0, xyz + 3 * a0, opt, nullptr, &r);
        energies_out[i] = r.energies[0];
        iters_out[i] = r.iterations;
    }
    return 0;
}

}  // extern "C"
