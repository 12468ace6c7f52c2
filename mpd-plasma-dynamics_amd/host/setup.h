// setup.h -- the host-side plug-in points the reference keeps around the
// step: IGeometry + GeometryRegistry (include/igeometry.hpp:16-34,
// include/geometry_registry.hpp:16-34), the mesh (include/grid.hpp:20-47),
// the field arrays (include/fields.hpp:24-79) and IInitialCondition +
// ExpressionIC (include/iinitial_condition.hpp:21-39,
// include/ics/expression_ic.hpp:124-143).
#pragma once

#include <mutex>
#include <functional>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "config.h"
#include "expr.h"
#include "yaml_lite.h"

namespace mpd {

class IGeometry {
public:
    virtual ~IGeometry() = default;
    virtual double RInner(double z) const = 0;
    virtual double ROuter(double z) const = 0;
    virtual double DrInnerDz(double z) const = 0;
    virtual double DrOuterDz(double z) const = 0;
    virtual std::string Name() const = 0;
};

class GeometryRegistry {
public:
    using Factory = std::function<std::unique_ptr<IGeometry>(const YNode&)>;
    static GeometryRegistry& Instance();
    void RegisterGeometry(std::string name, Factory factory);
    std::unique_ptr<IGeometry> Create(const std::string& type, const YNode& params) const;
    std::vector<std::string> Names() const;

private:
    GeometryRegistry() = default;
    // Batches build their instances' host setups on several threads at once.
    mutable std::mutex mu_;
    std::unordered_map<std::string, Factory> factories_;
};

// Registers straight_coaxial, converging_diverging and short_cathode_cosine.
void RegisterAllGeometries();

// Grid::Build (src/grid.cpp:20-42) for one rank owning the whole domain,
// stored per row: the wall radii and slopes at z_l = (l - 1) dz for
// l = 0..L+1 and dr[l] = (r_out - r_in) / M_max.  r(l, m) and r_z(l, m)
// follow from them with the reference's formula.
struct HostGrid {
    int L = 0, M_max = 0, Mn = 0, rows = 0, cols = 0;
    double dz = 0.0, dy = 0.0, r_0 = 0.0;
    std::vector<double> rin, rout, drin, drout, R, dr;

    void Build(const SimConfig& cfg, const IGeometry& g);
    // The same grid from per-row wall data computed by the caller (r_in,
    // r_out and their z-derivatives at z_l = (l - 1) dz, l = 0..L+1: what
    // Grid::Build evaluates an IGeometry at).
    void BuildFromRows(const SimConfig& cfg, const double* r_in, const double* r_out, const double* dr_in_dz,
                       const double* dr_out_dz);
    double r(int l, int m) const {
        const double frac = (m - 1) * dy;
        return (1.0 - frac) * rin[l] + frac * rout[l];
    }
    double r_z(int l, int m) const {
        const double frac = (m - 1) * dy;
        return (1.0 - frac) * drin[l] + frac * drout[l];
    }
};

// Physical field ids (include/fields.hpp member order).
enum { F_RHO = 0, F_VZ, F_VR, F_VPHI, F_E, F_P, F_PT, F_HZ, F_HR, F_HPHI };

// Host copy of the field arrays, each rows x cols (ghost ring included).
struct HostFields {
    int rows = 0, cols = 0;
    std::vector<double> u0[8];
    std::vector<double> phys[10];
    void Allocate(int rows, int cols);
    double& at(std::vector<double>& a, int l, int m) { return a[(size_t)l * cols + m]; }
    // Fields::InitConservative, src/fields.cpp:40-55.
    void InitConservative(const HostGrid& g);
};

class IInitialCondition {
public:
    virtual ~IInitialCondition() = default;
    virtual void Apply(HostFields& f, const HostGrid& g, const SimConfig& cfg, int l_start) const = 0;
    virtual std::string Name() const = 0;
};

// ExpressionIC (src/ics/expression_ic.cpp): per-cell formulas for rho, v_z,
// v_r, v_phi, H_z, H_r, H_phi, e in that order, then p and P.  Evaluated
// serially with symbol values that persist from cell to cell, as exprtk's do.
class ExpressionIC : public IInitialCondition {
public:
    ExpressionIC(const YNode* params, double noise_amplitude, unsigned long long noise_seed);
    void Apply(HostFields& f, const HostGrid& g, const SimConfig& cfg, int l_start) const override;
    std::string Name() const override { return "expression"; }

private:
    std::vector<std::string> names_;
    std::vector<double> user_values_;
    CompiledExpr ex_[8];
    double noise_amp_;
    unsigned long long noise_seed_;
};

// Seeded uniform U in [0,1): SplitMix64 of (seed, cell index).
double NoiseU01(unsigned long long seed, unsigned long long idx);

}  // namespace mpd
