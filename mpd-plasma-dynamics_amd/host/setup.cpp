// setup.cpp -- geometry plug-ins, grid, field arrays and the expression IC.
#include "setup.h"

#include <cmath>
#include <mutex>
#include <stdexcept>

namespace mpd {

// ------------------------------------------------------------- registry

GeometryRegistry& GeometryRegistry::Instance() {
    static GeometryRegistry inst;
    return inst;
}

void GeometryRegistry::RegisterGeometry(std::string name, Factory factory) {
    std::lock_guard<std::mutex> lock(mu_);
    factories_[std::move(name)] = std::move(factory);
}

std::unique_ptr<IGeometry> GeometryRegistry::Create(const std::string& type, const YNode& params) const {
    Factory f;
    {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = factories_.find(type);
        if (it == factories_.end()) {
            throw std::runtime_error("GeometryRegistry: unknown geometry type '" + type +
                                     "'.\n  Did you forget to call register_all_geometries(), or mistype the name in "
                                     "config.yaml?");
        }
        f = it->second;
    }
    return f(params);
}

std::vector<std::string> GeometryRegistry::Names() const {
    std::lock_guard<std::mutex> lock(mu_);
    std::vector<std::string> n;
    for (const auto& kv : factories_) n.push_back(kv.first);
    return n;
}

namespace {

double Param(const YNode& p, const char* key, double def) {
    if (p.IsNull()) return def;
    if (!p.IsMap()) throw std::runtime_error("geometry: 'params' must be a YAML map");
    const YNode* v = p.Get(key);
    return v ? v->AsDouble() : def;
}

// Straight coaxial channel: constant wall radii.
class StraightCoaxial : public IGeometry {
public:
    explicit StraightCoaxial(const YNode& p)
        : rin_(Param(p, "r_inner", 0.2)), rout_(Param(p, "r_outer", 0.8)) {
        if (!(rout_ > rin_)) throw std::runtime_error("straight_coaxial: r_outer must exceed r_inner");
    }
    double RInner(double) const override { return rin_; }
    double ROuter(double) const override { return rout_; }
    double DrInnerDz(double) const override { return 0.0; }
    double DrOuterDz(double) const override { return 0.0; }
    std::string Name() const override { return "straight_coaxial"; }

private:
    double rin_, rout_;
};

// Converging-diverging channel: r_in = r_inner + a_inner sin(pi z),
// r_out = r_outer - a_outer z^2.
class ConvergingDiverging : public IGeometry {
public:
    explicit ConvergingDiverging(const YNode& p)
        : ri_(Param(p, "r_inner", 0.2)), ai_(Param(p, "a_inner", 0.05)),
          ro_(Param(p, "r_outer", 0.8)), ao_(Param(p, "a_outer", 0.1)) {}
    double RInner(double z) const override { return ri_ + ai_ * std::sin(M_PI * z); }
    double ROuter(double z) const override { return ro_ - ao_ * z * z; }
    double DrInnerDz(double z) const override { return ai_ * M_PI * std::cos(M_PI * z); }
    double DrOuterDz(double z) const override { return -2.0 * ao_ * z; }
    std::string Name() const override { return "converging_diverging"; }

private:
    double ri_, ai_, ro_, ao_;
};

// short_cathode_cosine: the reference's plug-in,
// src/geometries/short_cathode_cosine.cpp:9-71 (same parameters, defaults,
// validation and formulas).
class ShortCathodeCosine : public IGeometry {
public:
    explicit ShortCathodeCosine(const YNode& p) {
        if (!p.IsNull() && !p.IsMap()) throw std::runtime_error("ShortCathodeCosineGeometry: 'params' must be a YAML map");
        zc_ = Param(p, "z_center", 0.31);
        hw_ = Param(p, "transition_half_width", 0.015);
        rib_ = Param(p, "r_inner_before", 0.200);
        ria_ = Param(p, "r_inner_after", 0.005);
        ro_ = Param(p, "r_outer", 0.800);
        if (hw_ <= 0.0) throw std::runtime_error("ShortCathodeCosineGeometry: transition_half_width must be positive");
        if (ro_ <= rib_ || ro_ <= ria_)
            throw std::runtime_error(
                "ShortCathodeCosineGeometry: r_outer must be greater than both r_inner_before and r_inner_after");
        z_start_ = zc_ - hw_;
        z_end_ = zc_ + hw_;
        delta_r_ = ria_ - rib_;
        inv_width_ = 1.0 / (z_end_ - z_start_);
    }
    double RInner(double z) const override {
        if (z < z_start_) return rib_;
        if (z < z_end_) {
            const double xi = (z - z_start_) * inv_width_;
            return rib_ + delta_r_ * 0.5 * (1.0 - std::cos(M_PI * xi));
        }
        return ria_;
    }
    double ROuter(double) const override { return ro_; }
    double DrInnerDz(double z) const override {
        if (z < z_start_ || z >= z_end_) return 0.0;
        const double xi = (z - z_start_) * inv_width_;
        return delta_r_ * 0.5 * M_PI * std::sin(M_PI * xi) * inv_width_;
    }
    double DrOuterDz(double) const override { return 0.0; }
    std::string Name() const override { return "short_cathode_cosine"; }

private:
    double zc_, hw_, rib_, ria_, ro_, z_start_, z_end_, delta_r_, inv_width_;
};

}  // namespace

void RegisterAllGeometries() {
    // Once per process, whichever thread gets here first (a batch sets up
    // its instances on several threads).
    static std::once_flag once;
    std::call_once(once, [] {
        GeometryRegistry& reg = GeometryRegistry::Instance();
        reg.RegisterGeometry("straight_coaxial", [](const YNode& p) { return std::make_unique<StraightCoaxial>(p); });
        reg.RegisterGeometry("converging_diverging",
                             [](const YNode& p) { return std::make_unique<ConvergingDiverging>(p); });
        reg.RegisterGeometry("short_cathode_cosine",
                             [](const YNode& p) { return std::make_unique<ShortCathodeCosine>(p); });
    });
}

// ----------------------------------------------------------------- grid

void HostGrid::Build(const SimConfig& cfg, const IGeometry& g) {
    L = cfg.L_max;
    M_max = cfg.M_max;
    Mn = M_max + 1;
    rows = L + 2;
    cols = Mn + 2;
    dz = cfg.dz;
    dy = cfg.dy;
    r_0 = (g.RInner(0.0) + g.ROuter(0.0)) / 2.0;
    rin.assign(rows, 0.0);
    rout.assign(rows, 0.0);
    drin.assign(rows, 0.0);
    drout.assign(rows, 0.0);
    R.assign(rows, 0.0);
    dr.assign(rows, 0.0);
    for (int l = 0; l < rows; ++l) {
        const int l_global = l - 1;
        const double z = l_global * dz;
        R[l] = g.ROuter(z) - g.RInner(z);
        dr[l] = R[l] / cfg.M_max;
        rin[l] = g.RInner(z);
        rout[l] = g.ROuter(z);
        drin[l] = g.DrInnerDz(z);
        drout[l] = g.DrOuterDz(z);
    }
}

void HostGrid::BuildFromRows(const SimConfig& cfg, const double* r_in, const double* r_out, const double* dr_in_dz,
                             const double* dr_out_dz) {
    L = cfg.L_max;
    M_max = cfg.M_max;
    Mn = M_max + 1;
    rows = L + 2;
    cols = Mn + 2;
    dz = cfg.dz;
    dy = cfg.dy;
    // Grid::Grid's r_0 = (RInner(0) + ROuter(0)) / 2: z = 0 is row l = 1.
    r_0 = (r_in[1] + r_out[1]) / 2.0;
    rin.assign(r_in, r_in + rows);
    rout.assign(r_out, r_out + rows);
    drin.assign(dr_in_dz, dr_in_dz + rows);
    drout.assign(dr_out_dz, dr_out_dz + rows);
    R.assign(rows, 0.0);
    dr.assign(rows, 0.0);
    for (int l = 0; l < rows; ++l) {
        R[l] = rout[l] - rin[l];
        dr[l] = R[l] / cfg.M_max;
    }
}

// --------------------------------------------------------------- fields

void HostFields::Allocate(int r, int c) {
    rows = r;
    cols = c;
    for (auto& a : u0) a.assign((size_t)r * c, 0.0);
    for (auto& a : phys) a.assign((size_t)r * c, 0.0);
}

void HostFields::InitConservative(const HostGrid& g) {
    for (int l = 1; l < rows - 1; ++l) {
        for (int m = 1; m < cols - 1; ++m) {
            const double r = g.r(l, m);
            const size_t i = (size_t)l * cols + m;
            u0[0][i] = phys[F_RHO][i] * r;
            u0[1][i] = phys[F_RHO][i] * phys[F_VZ][i] * r;
            u0[2][i] = phys[F_RHO][i] * phys[F_VR][i] * r;
            u0[3][i] = phys[F_RHO][i] * phys[F_VPHI][i] * r;
            u0[4][i] = phys[F_RHO][i] * phys[F_E][i] * r;
            u0[5][i] = phys[F_HPHI][i];
            u0[6][i] = phys[F_HZ][i] * r;
            u0[7][i] = phys[F_HR][i] * r;
        }
    }
}

// ------------------------------------------------------------------- IC

double NoiseU01(unsigned long long seed, unsigned long long idx) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ULL * (idx + 1ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

namespace {
enum { IS_GAMMA, IS_BETA, IS_HZ0, IS_R0, IS_Z, IS_R, IS_RZ, IS_RHO, IS_VZ, IS_VR, IS_VPHI, IS_HZ, IS_HR, IS_HPHI, IS_E,
       IS_COUNT };
const char* const kIcNames[IS_COUNT] = {"gamma", "beta", "H_z0", "r_0", "z", "r", "r_z", "rho",
                                        "v_z", "v_r", "v_phi", "H_z", "H_r", "H_phi", "e"};
const char* const kIcFields[8] = {"rho", "v_z", "v_r", "v_phi", "H_z", "H_r", "H_phi", "e"};
const char* const kIcDefaults[8] = {"1.0", "0.0", "0.0", "0.0", "H_z0", "H_z * r_z", "(1 - 0.9 * z) * r_0 / r",
                                    "beta / (2 * (gamma - 1))"};
}  // namespace

ExpressionIC::ExpressionIC(const YNode* params, double noise_amplitude, unsigned long long noise_seed)
    : noise_amp_(noise_amplitude), noise_seed_(noise_seed) {
    for (const char* n : kIcNames) names_.emplace_back(n);
    // User constants (src/ics/expression_ic.cpp:231-257), checked against the
    // reserved names (src/ics/expression_ic.cpp:26-39).
    if (params && params->IsMap()) {
        if (const YNode* vars = params->Get("vars")) {
            if (!vars->IsNull()) {
                if (!vars->IsMap()) throw std::runtime_error("ExpressionIC: 'vars' must be a YAML map of name: value pairs.");
                for (const auto& kv : vars->map) {
                    for (const char* res : kIcNames) {
                        if (kv.first == res)
                            throw std::runtime_error("ExpressionIC: user variable '" + kv.first +
                                                     "' shadows the built-in symbol '" + res +
                                                     "'. Choose a different name.");
                    }
                    if (kv.first == "pi")
                        throw std::runtime_error("ExpressionIC: user variable 'pi' shadows the built-in symbol 'pi'.");
                    if (!kv.second.IsScalar())
                        throw std::runtime_error("ExpressionIC: value for var '" + kv.first + "' must be a scalar number.");
                    names_.push_back(kv.first);
                    user_values_.push_back(kv.second.AsDouble());
                }
            }
        }
    }
    for (int k = 0; k < 8; ++k) {
        std::string src = kIcDefaults[k];
        if (params && params->IsMap()) {
            if (const YNode* n = params->Get(kIcFields[k])) {
                if (!n->IsNull()) {
                    if (!n->IsScalar())
                        throw std::runtime_error(std::string("ExpressionIC: field '") + kIcFields[k] +
                                                 "' must be a scalar expression string or a plain number.");
                    src = n->scalar;
                }
            }
        }
        try {
            ex_[k] = CompileExpr(src, names_);
        } catch (const std::exception& e) {
            throw std::runtime_error(std::string("ExpressionIC: failed to compile expression for '") + kIcFields[k] +
                                     "': " + e.what());
        }
    }
}

void ExpressionIC::Apply(HostFields& f, const HostGrid& g, const SimConfig& cfg, int l_start) const {
    std::vector<double> sym(names_.size(), 0.0);
    for (size_t i = 0; i < user_values_.size(); ++i) sym[IS_COUNT + i] = user_values_[i];
    sym[IS_GAMMA] = cfg.gamma;
    sym[IS_BETA] = cfg.beta;
    sym[IS_HZ0] = cfg.H_z0;
    sym[IS_R0] = g.r_0;
    const double gamma = cfg.gamma;
    for (int l = 1; l < f.rows - 1; ++l) {
        const int l_global = l_start + l - 1;
        const double z = l_global * cfg.dz;
        for (int m = 1; m < f.cols - 1; ++m) {
            sym[IS_Z] = z;
            sym[IS_R] = g.r(l, m);
            sym[IS_RZ] = g.r_z(l, m);
            sym[IS_RHO] = EvalExpr(ex_[0], sym.data());
            if (noise_amp_ != 0.0) {
                const unsigned long long idx = (unsigned long long)l_global * (unsigned long long)g.Mn +
                                               (unsigned long long)(m - 1);
                sym[IS_RHO] = sym[IS_RHO] * (1.0 + noise_amp_ * (2.0 * NoiseU01(noise_seed_, idx) - 1.0));
            }
            sym[IS_VZ] = EvalExpr(ex_[1], sym.data());
            sym[IS_VR] = EvalExpr(ex_[2], sym.data());
            sym[IS_VPHI] = EvalExpr(ex_[3], sym.data());
            sym[IS_HZ] = EvalExpr(ex_[4], sym.data());
            sym[IS_HR] = EvalExpr(ex_[5], sym.data());
            sym[IS_HPHI] = EvalExpr(ex_[6], sym.data());
            sym[IS_E] = EvalExpr(ex_[7], sym.data());
            const size_t i = (size_t)l * f.cols + m;
            f.phys[F_RHO][i] = sym[IS_RHO];
            f.phys[F_VZ][i] = sym[IS_VZ];
            f.phys[F_VR][i] = sym[IS_VR];
            f.phys[F_VPHI][i] = sym[IS_VPHI];
            f.phys[F_HZ][i] = sym[IS_HZ];
            f.phys[F_HR][i] = sym[IS_HR];
            f.phys[F_HPHI][i] = sym[IS_HPHI];
            f.phys[F_E][i] = sym[IS_E];
            const double p = (gamma - 1.0) * sym[IS_RHO] * sym[IS_E];
            f.phys[F_P][i] = p;
            f.phys[F_PT][i] = p + 0.5 * (sym[IS_HZ] * sym[IS_HZ] + sym[IS_HR] * sym[IS_HR] + sym[IS_HPHI] * sym[IS_HPHI]);
        }
    }
}

}  // namespace mpd
