// Host-side GFN1-xTB model data.  Parameter-file semantics follow
// src/params.rs:272-341 and :826-872; basis construction follows
// src/basis.rs:121-344 and src/sto.rs:790-881 of the reference.
#include "model.hpp"

#include <dlfcn.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace gfn1 {
namespace {

constexpr double kPi = 3.14159265358979323846;

// xTB atomic radii in Angstrom (data_tables.rs:14-21), index Z.
const double kAtomicRadiusAA[kMaxZ + 1] = {
    0.0,  0.32, 0.37, 1.3,  0.99, 0.84, 0.75, 0.71, 0.64, 0.6,  0.62, 1.6,  1.4,  1.24, 1.14, 1.09, 1.04, 1.0,
    1.01, 2.0,  1.74, 1.59, 1.48, 1.44, 1.3,  1.29, 1.24, 1.18, 1.17, 1.22, 1.2,  1.23, 1.2,  1.2,  1.18, 1.17,
    1.16, 2.15, 1.9,  1.76, 1.64, 1.56, 1.46, 1.38, 1.36, 1.34, 1.3,  1.36, 1.4,  1.42, 1.4,  1.4,  1.37, 1.36,
    1.36, 2.38, 2.06, 1.94, 1.84, 1.9,  1.88, 1.86, 1.85, 1.83, 1.82, 1.81, 1.8,  1.79, 1.77, 1.77, 1.78, 1.74,
    1.64, 1.58, 1.5,  1.41, 1.36, 1.32, 1.3,  1.3,  1.32, 1.44, 1.45, 1.5,  1.42, 1.48, 1.46};
// Pyykko-Atsumi (2009) covalent radii in Angstrom as used by mctc-lib (data_tables.rs:29-37).
const double kCovalentRadiusAA[kMaxZ + 1] = {
    0.0,  0.32, 0.46, 1.20, 0.94, 0.77, 0.75, 0.71, 0.63, 0.64, 0.67, 1.40, 1.25, 1.13, 1.04, 1.10, 1.02, 0.99,
    0.96, 1.76, 1.54, 1.33, 1.22, 1.21, 1.10, 1.07, 1.04, 1.00, 0.99, 1.01, 1.09, 1.12, 1.09, 1.15, 1.10, 1.14,
    1.17, 1.89, 1.67, 1.47, 1.39, 1.32, 1.24, 1.15, 1.13, 1.13, 1.08, 1.15, 1.23, 1.28, 1.26, 1.26, 1.23, 1.32,
    1.31, 2.09, 1.76, 1.62, 1.47, 1.58, 1.57, 1.56, 1.55, 1.51, 1.52, 1.51, 1.50, 1.49, 1.49, 1.48, 1.53, 1.46,
    1.37, 1.31, 1.23, 1.18, 1.16, 1.11, 1.12, 1.13, 1.32, 1.30, 1.30, 1.36, 1.31, 1.38, 1.42};
// Pauling electronegativities (data_tables.rs:60-69).
const double kPaulingEN[kMaxZ + 1] = {
    0.0,  2.20, 3.00, 0.98, 1.57, 2.04, 2.55, 3.04, 3.44, 3.98, 4.50, 0.93, 1.31, 1.61, 1.90, 2.19, 2.58, 3.16,
    3.50, 0.82, 1.00, 1.36, 1.54, 1.63, 1.66, 1.55, 1.83, 1.88, 1.91, 1.90, 1.65, 1.81, 2.01, 2.18, 2.55, 2.96,
    3.00, 0.82, 0.95, 1.22, 1.33, 1.60, 2.16, 1.90, 2.20, 2.28, 2.20, 1.93, 1.69, 1.78, 1.96, 2.05, 2.10, 2.66,
    2.60, 0.79, 0.89, 1.10, 1.12, 1.13, 1.14, 1.15, 1.17, 1.18, 1.20, 1.21, 1.22, 1.23, 1.24, 1.25, 1.26, 1.27,
    1.30, 1.50, 2.36, 1.90, 2.20, 2.20, 2.28, 2.54, 2.00, 1.62, 2.33, 2.02, 2.00, 2.20, 2.20};

// GFN1 reference shell occupations {s, p, d} (basis.rs:443-531), listed per period.
struct Occ3 {
    double s, p, d;
};
Occ3 reference_occupation(int z) {
    static Occ3 table[kMaxZ + 1];
    static bool filled = false;
    if (!filled) {
        auto put = [](int zz, double s, double p, double d) { table[zz] = {s, p, d}; };
        put(1, 1, 0, 0);
        put(2, 2, 0, 0);
        const int sp_periods[2] = {3, 11};  // Li..Ne, Na..Ar
        for (int first : sp_periods) {
            put(first, 1, 0, 0);
            put(first + 1, 2, 0, 0);
            for (int np = 1; np <= 6; ++np) put(first + 1 + np, 2, np, 0);
        }
        const int d_periods[2] = {19, 37};  // K..Kr, Rb..Xe
        for (int first : d_periods) {
            put(first, 1, 0, 0);
            put(first + 1, 2, 0, 0);
            for (int nd = 1; nd <= 9; ++nd) put(first + 1 + nd, 2, 0, nd);
            put(first + 11, 2, 0, 0);
            for (int np = 1; np <= 6; ++np) put(first + 11 + np, 2, np, 0);
        }
        put(55, 1, 0, 0);
        put(56, 2, 0, 0);
        for (int zz = 57; zz <= 71; ++zz) put(zz, 2, 0, 1);  // f-in-core lanthanoids
        for (int nd = 2; nd <= 9; ++nd) put(70 + nd, 2, 0, nd);
        put(80, 2, 0, 0);
        for (int np = 1; np <= 6; ++np) put(80 + np, 2, np, 0);
        filled = true;
    }
    return table[z];
}

// Element class for the CN shift (basis.rs:296-302).
int cn_shift_class(int z) {
    if (z == 1 || z == 2 || (z >= 6 && z <= 10) || (z >= 14 && z <= 18) || (z >= 32 && z <= 36) ||
        (z >= 50 && z <= 54) || (z >= 82 && z <= 86))
        return 1;
    if (z >= 21 && z <= 24) return 2;
    return 0;
}

// GFN1 primitive-count rule (basis.rs:251-279).
int primitive_count(int z, int n, int l) {
    if (z <= 2) return l == 0 ? (n == 1 ? 4 : 3) : (l == 1 ? 3 : 4);
    return l <= 1 ? 6 : 4;
}

struct StoEntry {
    std::vector<double> alpha, coeff;
};

struct StoData {
    std::map<std::pair<int, int>, StoEntry> by_type;  // (Slater type 1..15, ng)
    StoEntry six_s, six_p;
};

Status read_sto(const std::string& path, StoData& d) {
    std::ifstream f(path);
    if (!f) return Status::error(kErrIo, "cannot read " + path);
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream in(line);
        std::string tag;
        int ng = 0;
        in >> tag >> ng;
        StoEntry e;
        e.alpha.resize(ng);
        e.coeff.resize(ng);
        for (int i = 0; i < ng; ++i) in >> e.alpha[i];
        for (int i = 0; i < ng; ++i) in >> e.coeff[i];
        if (tag == "6s") d.six_s = e;
        else if (tag == "6p") d.six_p = e;
        else d.by_type[{std::atoi(tag.c_str()), ng}] = e;
    }
    return Status{};
}

// Normalised STO-nG primitives for a Slater shell (sto.rs:790-881).
Status contract_slater(const StoData& d, int ng, int n, int l, double zeta, std::vector<double>& a,
                       std::vector<double>& c) {
    if ((n > 5 || n <= l) && !(n == 6 && ng == 6 && l <= 1))
        return Status::error(kErrInvalidInput, "unsupported STO-nG shell");
    const StoEntry* e = nullptr;
    if (n == 6 && ng == 6) {
        e = l == 0 ? &d.six_s : &d.six_p;
    } else {
        static const int type_base[5] = {0, 4, 7, 9, 10};
        auto it = d.by_type.find({type_base[l] + n, ng});
        if (it == d.by_type.end()) return Status::error(kErrInvalidInput, "no STO-nG table entry");
        e = &it->second;
    }
    static const double double_factorial[3] = {1.0, 1.0, 3.0};  // (2l-1)!!
    a.resize(ng);
    c.resize(ng);
    for (int i = 0; i < ng; ++i) {
        a[i] = e->alpha[i] * zeta * zeta;
        c[i] = e->coeff[i] * std::pow(2.0 / kPi * a[i], 0.75) * std::pow(std::sqrt(4.0 * a[i]), l) /
               std::sqrt(double_factorial[l]);
    }
    return Status{};
}

// <g_i|g_j> of two normalised s-type contractions on one centre.
double same_centre_s_overlap(const std::vector<double>& a1, const std::vector<double>& c1,
                             const std::vector<double>& a2, const std::vector<double>& c2) {
    double s = 0.0;
    for (size_t i = 0; i < a1.size(); ++i)
        for (size_t j = 0; j < a2.size(); ++j) s += c1[i] * c2[j] * std::pow(std::sqrt(kPi / (a1[i] + a2[j])), 3);
    return s;
}

std::string lowercase(std::string s) {
    for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

std::string strip(const std::string& s) {
    size_t b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return "";
    return s.substr(b, s.find_last_not_of(" \t\r\n") - b + 1);
}

struct RawElement {
    std::string ao;
    std::map<std::string, std::vector<double>> kv;
};

}  // namespace

std::string default_data_dir() {
    if (const char* env = std::getenv("GFN1_AMD_DATA")) return env;
    Dl_info info;
    if (dladdr((void*)&default_data_dir, &info) && info.dli_fname) {
        std::string so(info.dli_fname);
        size_t slash = so.rfind('/');
        std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
        return dir + "/../data";
    }
    return "data";
}

double Model::pair_scaling(int za, int zb) const {
    auto key = za <= zb ? std::make_pair(za, zb) : std::make_pair(zb, za);
    auto it = pairpar_.find(key);
    if (it != pairpar_.end()) return it->second;
    auto drow = [](int z) {  // params.rs:694-711
        if (z >= 21 && z < 30) return 1.1;
        if (z >= 39 && z < 48) return 1.2;
        if (z >= 57 && z < 80) return 1.2;
        return 0.0;
    };
    double a = drow(za), b = drow(zb);
    return (a > 0.0 && b > 0.0) ? 0.5 * (a + b) : 1.0;
}

double Model::hscale(int za, int sa, int zb, int sb) const {
    const ShellType& A = shell(za, sa);
    const ShellType& B = shell(zb, sb);
    auto kl = [&](int l) { return l == 0 ? glob_.ks : l == 1 ? glob_.kp : glob_.kd; };
    auto kpair = [&](int la, int lb) {
        if ((la == 0 && lb == 1) || (la == 1 && lb == 0)) return glob_.ksp;
        return 0.5 * (kl(la) + kl(lb));
    };
    if (A.valence && B.valence) {
        double den = elements_[za].en - elements_[zb].en;
        return pair_scaling(za, zb) * kpair(A.l, B.l) * (1.0 + glob_.enscale * 0.01 * den * den);
    }
    if (A.valence) return 0.5 * (kpair(A.l, A.l) + glob_.kdiff);
    if (B.valence) return 0.5 * (kpair(B.l, B.l) + glob_.kdiff);
    return glob_.kdiff;
}

Status Model::load(const char* spec, Model& m) { return load_impl(spec, nullptr, m); }

// Gfn1Parameters::from_str (params.rs:272): parameters from text, no file identity.
Status Model::load_text(const char* text, Model& m) { return load_impl(nullptr, text ? text : "", m); }

Status Model::load_impl(const char* spec, const char* text, Model& m) {
    m = Model();
    m.data_dir_ = default_data_dir();
    std::string path;
    const std::string inline_text = text ? text : "";
    std::string s = spec ? strip(spec) : "";
    if (text) {
        m.source_ = "inline text";
        m.source_kind_ = 3;
        path = "(inline text)";
    } else if (s.empty()) {
        const char* env = std::getenv("GFN1_XTB_PARAM");
        if (env && *env) {
            path = strip(env);
            m.source_ = path + " (via GFN1_XTB_PARAM)";
        }
    } else if (lowercase(s) == "builtin:si" || lowercase(s) == "builtin:gfn1-si" || lowercase(s) == "builtin-si") {
        // the bundled GFN1(Si)-xTB reparametrization (params.rs:192-196, 210)
        path = m.data_dir_ + "/param_gfn1-si-xtb.txt";
        m.source_ = "builtin";
        m.source_kind_ = 2;
    } else if (lowercase(s) != "builtin" && lowercase(s) != "builtin:gfn1") {
        path = s;
        m.source_ = path;
    }
    if (path.empty()) {
        path = m.data_dir_ + "/param_gfn1-xtb.txt";
        m.source_ = "builtin";
        m.source_kind_ = 1;
    }

    std::ifstream file;
    std::istringstream text_in;
    if (m.source_kind_ == 3) {
        text_in.str(inline_text);
    } else {
        file.open(path);
        if (!file) return Status::error(kErrIo, "cannot read parameter file " + path);
    }
    std::istream& f = m.source_kind_ == 3 ? static_cast<std::istream&>(text_in) : static_cast<std::istream&>(file);
    std::map<std::string, double> glob;
    std::map<int, RawElement> raw;
    int section = 0;  // 1 info, 2 globpar, 3 pairpar, 4 element
    int z = 0;
    std::string line;
    int line_no = 0;
    while (std::getline(f, line)) {
        ++line_no;
        std::string t = strip(line);
        if (t.empty() || t[0] == '#') continue;
        if (t.rfind("$info", 0) == 0) { section = 1; continue; }
        if (t.rfind("$globpar", 0) == 0) { section = 2; continue; }
        if (t.rfind("$pairpar", 0) == 0) { section = 3; continue; }
        if (t.rfind("$Z=", 0) == 0) {
            z = std::atoi(t.c_str() + 3);
            if (z < 1 || z > 255)
                return Status::error(kErrParse, "line " + std::to_string(line_no) + ": invalid element number");
            section = 4;
            raw[z];
            continue;
        }
        if (t.rfind("$end", 0) == 0) { section = 0; continue; }
        std::istringstream in(t);
        if (section == 2) {
            std::string key;
            double v;
            if (!(in >> key >> v))
                return Status::error(kErrParse, "line " + std::to_string(line_no) + ": global parameter line must be `key value`");
            glob[lowercase(key)] = v;
        } else if (section == 3) {
            int a, b;
            double v;
            if (!(in >> a >> b >> v))
                return Status::error(kErrParse, "line " + std::to_string(line_no) + ": pair parameter line must be `ZA ZB value`");
            m.pairpar_[a <= b ? std::make_pair(a, b) : std::make_pair(b, a)] = v;
        } else if (section == 4) {
            if (t.rfind("ao=", 0) == 0) {
                raw[z].ao = strip(t.substr(3));
                continue;
            }
            size_t eq = t.find('=');
            if (eq == std::string::npos)
                return Status::error(kErrParse, "line " + std::to_string(line_no) + ": element parameter line must contain `=`");
            std::string key = strip(t.substr(0, eq));
            for (char& ch : key) ch = (char)std::toupper((unsigned char)ch);
            std::istringstream vs(t.substr(eq + 1));
            std::vector<double> vals;
            double v;
            while (vs >> v) vals.push_back(v);
            raw[z].kv[key] = vals;
        } else if (section == 1) {
            // "key value" (params.rs:743-751); the name labels the parametrization
            std::istringstream is(t);
            std::string key, value;
            is >> key;
            std::getline(is, value);
            if (lowercase(key) == "name") m.info_name_ = strip(value);
            continue;
        } else {
            return Status::error(kErrParse, "line " + std::to_string(line_no) + ": data outside a section");
        }
    }
    // params.rs:555-584 required globals
    const char* required[] = {"ks", "kp", "kd", "ksp", "kdiff", "enscale", "ipeashift", "cns", "cnp", "cnd1",
                              "cnd2", "alphaj", "a1", "a2", "s8", "s9", "kexp", "kexplight", "xbdamp", "xbrad"};
    for (const char* k : required)
        if (!glob.count(k)) return Status::error(kErrMissingGlobal, std::string("missing global parameter ") + k);
    auto g = [&](const char* k, double d) {
        auto it = glob.find(k);
        return it == glob.end() ? d : it->second;
    };
    m.glob_ = Globals{g("ks", 1.85),   g("kp", 2.25),    g("kd", 2.0),        g("ksp", 2.08), g("kdiff", 2.85),
                      g("enscale", -0.7), g("a1", 0.0),  g("a2", 0.0),        g("s6", 1.0),   g("s8", 0.0),
                      g("s9", 0.0),    g("kexp", 1.5),   g("kexplight", 1.5), g("xbdamp", 0.44), g("xbrad", 1.3)};

    StoData sto;
    Status st = read_sto(m.data_dir_ + "/sto_ng.txt", sto);
    if (!st.ok()) return st;

    // D3 reference (format documented in data/make_data.py)
    {
        std::ifstream d3(m.data_dir_ + "/d3_reference.bin", std::ios::binary);
        if (!d3) return Status::error(kErrIo, "cannot read " + m.data_dir_ + "/d3_reference.bin");
        int32_t hdr[3];
        d3.read((char*)hdr, sizeof(hdr));
        if (hdr[0] != 0x44334331 || hdr[1] != kMaxZ || hdr[2] != kMaxRef)
            return Status::error(kErrParse, "unexpected D3 reference layout");
        std::vector<int32_t> nref(kMaxZ);
        std::vector<double> refcn(kMaxZ * kMaxRef), r4(kMaxZ);
        d3.read((char*)nref.data(), kMaxZ * 4);
        d3.read((char*)refcn.data(), refcn.size() * 8);
        d3.read((char*)r4.data(), r4.size() * 8);
        m.d3_c6_.resize((size_t)kMaxZ * (kMaxZ + 1) / 2 * kMaxRef * kMaxRef);
        d3.read((char*)m.d3_c6_.data(), m.d3_c6_.size() * 8);
        if (!d3) return Status::error(kErrParse, "truncated D3 reference");
        for (int zz = 1; zz <= kMaxZ; ++zz) {
            ElementType& e = m.elements_[zz];
            e.d3_nref = nref[zz - 1];
            for (int r = 0; r < kMaxRef; ++r) e.d3_refcn[r] = refcn[(zz - 1) * kMaxRef + r];
            e.d3_r4r2 = std::sqrt(0.5 * r4[zz - 1] * std::sqrt((double)zz));
        }
    }

    for (auto& kvz : raw) {
        int zz = kvz.first;
        if (zz > kMaxZ) continue;
        RawElement& re = kvz.second;
        auto scalar = [&](const std::string& k, double d) {
            auto it = re.kv.find(k);
            return (it == re.kv.end() || it->second.empty()) ? d : it->second[0];
        };
        if (re.ao.empty()) return Status::error(kErrParse, "Z=" + std::to_string(zz) + " is missing ao= entry");
        if (!re.kv.count("LEV") || !re.kv.count("EXP") || !re.kv.count("GAM") || !re.kv.count("REPA") ||
            !re.kv.count("REPB"))
            return Status::error(kErrParse, "Z=" + std::to_string(zz) + " is missing a required entry");
        // parse "1s2s" / "3d4s4p"
        std::vector<std::pair<int, int>> labels;
        for (size_t i = 0; i < re.ao.size();) {
            if (std::isspace((unsigned char)re.ao[i])) { ++i; continue; }
            int n = 0;
            while (i < re.ao.size() && std::isdigit((unsigned char)re.ao[i])) n = n * 10 + (re.ao[i++] - '0');
            if (i >= re.ao.size()) return Status::error(kErrParse, "AO label missing angular momentum");
            char lc = (char)std::tolower((unsigned char)re.ao[i++]);
            int l = lc == 's' ? 0 : lc == 'p' ? 1 : lc == 'd' ? 2 : lc == 'f' ? 3 : lc == 'g' ? 4 : -1;
            if (l < 0) return Status::error(kErrParse, "unsupported angular momentum");
            labels.push_back({n, l});
        }
        const auto& lev = re.kv["LEV"];
        const auto& ex = re.kv["EXP"];
        if (lev.size() != labels.size() || ex.size() != labels.size())
            return Status::error(kErrParse, "Z=" + std::to_string(zz) + " AO label / LEV / EXP count mismatch");
        if ((int)labels.size() > kMaxShellsPerAtom) return Status::error(kErrParse, "too many shells");
        ElementType& e = m.elements_[zz];
        e.present = 1;
        e.nshell = (int)labels.size();
        e.gam3 = 0.1 * scalar("GAM3", 0.0);
        e.repa = scalar("REPA", 0.0);
        e.repb = scalar("REPB", 0.0);
        e.xb_strength = 0.1 * scalar("CXB", 0.0);
        e.atomic_radius = kAtomicRadiusAA[zz] * kAngstromToBohr;
        e.covalent_radius = (4.0 / 3.0) * kCovalentRadiusAA[zz] * kAngstromToBohr;
        e.en = kPaulingEN[zz];
        double gam = scalar("GAM", 0.0);
        const char suffix[] = "SPDFG";
        bool seen[5] = {false, false, false, false, false};
        std::vector<double> ref_a[5], ref_c[5];
        Occ3 occ = reference_occupation(zz);
        int cls = cn_shift_class(zz);
        e.nao = 0;
        for (int s = 0; s < e.nshell; ++s) {
            int n = labels[s].first, l = labels[s].second;
            if (l > 2)
                return Status::error(kErrInvalidInput, "integral engine implements s, p, d shells only");
            ShellType& t = m.shells_[zz * kMaxShellsPerAtom + s];
            t.l = l;
            t.nao = 2 * l + 1;
            t.valence = seen[l] ? 0 : 1;
            std::vector<double> a, c;
            st = contract_slater(sto, primitive_count(zz, n, l), n, l, ex[s], a, c);
            if (!st.ok()) return st;
            if (seen[l]) {
                // orthogonalise against the first shell of this l (basis.rs:318-344)
                double ov = same_centre_s_overlap(ref_a[l], ref_c[l], a, c);
                for (size_t i = 0; i < ref_a[l].size(); ++i) {
                    a.push_back(ref_a[l][i]);
                    c.push_back(-ov * ref_c[l][i]);
                }
                double nrm = same_centre_s_overlap(a, c, a, c);
                for (double& ci : c) ci /= std::sqrt(nrm);
            } else {
                ref_a[l] = a;
                ref_c[l] = c;
            }
            seen[l] = true;
            if ((int)a.size() > kMaxPrim) return Status::error(kErrInvalidInput, "too many primitives");
            t.nprim = (int)a.size();
            for (int i = 0; i < t.nprim; ++i) {
                t.alpha[i] = a[i];
                t.coeff[i] = c[i];
            }
            double level = lev[s];
            t.hdiag = level * kEvToHartreeModel;
            double fac = 0.0;
            if (cls == 1 || cls == 2) {
                if (l == 0) fac = g("cns", 0.6);
                else if (l == 1) fac = g("cnp", -0.3);
                else fac = cls == 1 ? g("cnd1", -0.5) : g("cnd2", 0.5);
            }
            t.kcn = -level * kEvToHartreeModel * fac * 0.01;
            t.poly = 0.01 * scalar(std::string("POLY") + suffix[l], 0.0);
            // hardness uses the LPAR of the first shell with this l (params.rs:153-160)
            double lpar = 0.0;
            for (int s2 = 0; s2 < e.nshell; ++s2)
                if (labels[s2].second == l) {
                    lpar = scalar(std::string("LPAR") + suffix[l], 0.0);
                    break;
                }
            t.hardness = gam * (1.0 + 0.1 * lpar);
            t.ref_occ = t.valence ? (l == 0 ? occ.s : l == 1 ? occ.p : occ.d) : 0.0;
            e.nao += t.nao;
        }
        e.neutral_electrons = occ.s + occ.p + occ.d;
        e.min_exponent = INFINITY;
        for (int s = 0; s < e.nshell; ++s) {
            const ShellType& t = m.shells_[zz * kMaxShellsPerAtom + s];
            for (int i = 0; i < t.nprim; ++i) e.min_exponent = std::min(e.min_exponent, t.alpha[i]);
        }
    }
    return Status{};
}

}  // namespace gfn1

namespace gfn1 {

std::string Model::source_description() const {
    const std::string name = info_name_.empty() ? "unnamed parametrization" : info_name_;
    // the bundled files are xtb's, commit 2b5cd48 (LGPL-3.0-or-later)
    const std::string where =
        (source_kind_ == 1 || source_kind_ == 2) ? "builtin, grimme-lab/xtb@2b5cd48, LGPL-3.0-or-later" : source_;
    std::string refs;
    if (source_kind_ == 1)
        refs = "GFN1-xTB DOI 10.1021/acs.jctc.7b00118, Ln f-in-core DOI 10.1021/acs.inorgchem.7b01950";
    else if (source_kind_ == 2)
        refs = "GFN1-xTB DOI 10.1021/acs.jctc.7b00118, GFN1(Si)-xTB DOI 10.1021/acs.jcim.1c01170";
    return refs.empty() ? name + " (" + where + ")" : name + " (" + where + "; refs: " + refs + ")";
}

}  // namespace gfn1

