// Host-side GFN1-xTB model: parameter file, element tables, STO-nG contraction
// and the per-element shell tables uploaded to the device.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace gfn1 {

constexpr int kMaxZ = 86;
constexpr int kMaxShellsPerAtom = 3;
constexpr int kMaxPrim = 7;   // H 2s after same-centre orthogonalisation (3 + 4)
constexpr int kMaxRef = 7;    // D3 reference CNs per element

// Legacy eV -> Eh used for reading the parameter file (basis.rs:18).
constexpr double kEvToHartreeModel = 1.0 / 27.21138505;
constexpr double kBoltzmannHartree = 3.1668115634556e-6;   // constants.rs:26
constexpr double kAngstromToBohr = 1.8897261246257702;     // system.rs:8
constexpr double kHartreeToEv = 27.211386245988;           // constants.rs:17 (reporting)

enum ErrorCode : int {
    kOk = 0,
    kErrIo = 1,
    kErrParse = 2,
    kErrInvalidInput = 3,
    kErrMissingElement = 4,
    kErrMissingGlobal = 5,
    kErrLinearAlgebra = 6,
    kErrSingularCell = 7,
    kErrSccNotConverged = 8,
    kErrDevice = 100,
};

struct Status {
    int code = kOk;
    std::string message;
    bool ok() const { return code == kOk; }
    static Status error(int c, std::string m) { return Status{c, std::move(m)}; }
};

// One shell of an element as the device sees it: everything a kernel needs to
// evaluate integrals and model terms without consulting the parameter file.
struct ShellType {
    int32_t l = 0;
    int32_t nprim = 0;
    int32_t valence = 0;
    int32_t nao = 0;
    double alpha[kMaxPrim] = {};
    double coeff[kMaxPrim] = {};
    double hdiag = 0.0;     // Eh
    double kcn = 0.0;       // CN shift of the self energy (basis.rs:281-293)
    double poly = 0.0;      // shell polynomial coefficient (x 0.01)
    double hardness = 0.0;  // GAM (1 + 0.1 LPAR_l)
    double ref_occ = 0.0;   // reference occupation of the shell
};

// Per-element data (index Z, 0 unused).
struct ElementType {
    int32_t present = 0;
    int32_t nshell = 0;
    int32_t nao = 0;
    int32_t d3_nref = 0;
    double gam3 = 0.0;      // 0.1 GAM3 (params.rs:149)
    double repa = 0.0, repb = 0.0;
    double xb_strength = 0.0;  // 0.1 CXB (halogen.rs:775)
    double atomic_radius = 0.0;    // bohr (xTB radii, data_tables.rs:195)
    double covalent_radius = 0.0;  // bohr, 4/3 x Pyykko-Atsumi (data_tables.rs:208)
    double en = 0.0;               // Pauling electronegativity
    double d3_r4r2 = 0.0;          // sqrt(0.5 <r4>/<r2> sqrt(Z))
    double d3_refcn[kMaxRef] = {};
    double neutral_electrons = 0.0;
    double min_exponent = 0.0;     // smallest primitive exponent of the element's shells (basis.rs:361-371)
};

struct Globals {
    double ks, kp, kd, ksp, kdiff, enscale, a1, a2, s6, s8, s9, kexp, kexplight, xbdamp, xbrad;
};

class Model {
   public:
    // spec: nullptr/"" -> $GFN1_XTB_PARAM or the bundled file; "builtin" ->
    // bundled; otherwise a path (params.rs:204-230).
    static Status load(const char* spec, Model& out);
    static Status load_text(const char* text, Model& out);

    const ShellType& shell(int z, int s) const { return shells_[z * kMaxShellsPerAtom + s]; }
    const ElementType& element(int z) const { return elements_[z]; }
    const std::vector<ShellType>& shell_table() const { return shells_; }
    const std::vector<ElementType>& element_table() const { return elements_; }
    const std::vector<double>& d3_c6() const { return d3_c6_; }
    const Globals& globals() const { return glob_; }
    double pair_scaling(int za, int zb) const;
    // H0 shell-pair scaling K_ll' (hamiltonian.rs:140-165) without the EN factor
    // split: returns the full factor for the two shells.
    double hscale(int za, int sa, int zb, int sb) const;
    const std::string& source() const { return source_; }
    // Gfn1Parameters::source_description (params.rs:236-270): name, source and
    // literature references of the loaded parametrization.
    std::string source_description() const;
    const std::string& data_dir() const { return data_dir_; }

   private:
    std::vector<ShellType> shells_ = std::vector<ShellType>((kMaxZ + 1) * kMaxShellsPerAtom);
    std::vector<ElementType> elements_ = std::vector<ElementType>(kMaxZ + 1);
    std::vector<double> d3_c6_;
    std::map<std::pair<int, int>, double> pairpar_;
    Globals glob_{};
    std::string source_, data_dir_;
    std::string info_name_;
    int source_kind_ = 0;  // 0 file, 1 bundled GFN1-xTB, 2 bundled GFN1(Si)-xTB, 3 inline text
    static Status load_impl(const char* spec, const char* text, Model& out);
};

// Location of the bundled data directory (next to the shared library).
std::string default_data_dir();

}  // namespace gfn1
