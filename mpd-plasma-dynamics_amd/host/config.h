// config.h -- the reference's configuration schema (include/config.hpp:1-177,
// include/field_condition.hpp:94-107): same structs, same keys, same defaults,
// same config.yaml grammar, loaded with the YAML reader of yaml_lite.h.
//
// Schema extensions (absent from the reference, documented in DESIGN.md):
//   time.max_steps                 step cap of the time loop (0 = none)
//   initial_conditions.rho_noise   {amplitude, seed}: rho *= 1 + amp (2U - 1)
//   gpu                            kernel selection / tile shape options
#pragma once

#include <string>
#include <vector>

#include "yaml_lite.h"

namespace mpd {

enum class FieldCondType { Neumann, Dirichlet, Expression, AxisLF };

struct FieldCond {
    FieldCondType type = FieldCondType::Neumann;
    double value = 0.0;
    std::string expr_str;
};

struct BCSegmentConfig {
    int global_lo = -1;
    int global_hi = -1;
    FieldCond rho, v_z, v_r, v_phi, e, H_z, H_r, H_phi;
};

struct BCFaceConfig {
    std::vector<BCSegmentConfig> segments;
};

struct GeometryConfig {
    std::string type = "short_cathode_cosine";
    YNode params;
};

struct ICConfig {
    std::string vtk_file;
    bool vtk_exact_z = false;  // VtkIC axial mapping (see host/vtk.h)
    YNode params;  // the initial_conditions node (expression IC)
    bool has_params = false;
};

struct GpuConfig {
    std::string kernel = "tiled";  // "tiled" | "simple"
    bool fast_div = true;          // correctly rounded constant division
    int tile_rows = 0;             // rows per wave of the tiled kernel (0 = auto)
    int steps_per_graph = 0;       // 0 = auto
};

struct SimConfig {
    double gamma = 1.67;
    double beta = 0.05;
    double H_z0 = 0.25;

    double T = 0.5;
    double dt = 0.000025;

    bool adaptive_dt = false;
    double cfl_number = 0.5;
    double dt_growth_factor = 1.1;
    double dt_min = 1.0e-9;
    double dt_max = 1.0e-3;

    int L_max = 800;
    int M_max = 400;

    double convergence_threshold = 0.0;
    int check_frequency = 100;

    std::string output_dir = "output";
    std::string run_name = "run";
    int vtk_step = 100;

    int openmp_threads = 0;
    int mpi_dims_l = 0;
    int mpi_dims_m = 0;

    GeometryConfig geometry;
    ICConfig initial_conditions;

    BCFaceConfig bc_l_lo, bc_l_hi, bc_m_lo, bc_m_hi;

    int max_steps = 0;
    double rho_noise_amplitude = 0.0;
    unsigned long long rho_noise_seed = 0;
    GpuConfig gpu;

    double dz{};
    double dy{};

    SimConfig() { Init(); }

    // SimConfig::Load (src/config.cpp:261-384).
    void Load(const std::string& path);
    void LoadFromString(const std::string& yaml_text);
    void LoadNode(const YNode& root);

    // SimConfig::Init (src/config.cpp:15-18).
    void Init() {
        dz = 1.0 / L_max;
        dy = 1.0 / M_max;
    }
};

// The field-condition grammar of src/config.cpp:52-202.
FieldCond ParseFieldCond(const YNode* n);
BCFaceConfig ParseFace(const YNode* n);

}  // namespace mpd
