// config.cpp -- SimConfig loading with the reference's keys, grammar and error
// behaviour (src/config.cpp:52-384), on the YAML reader of yaml_lite.h.
#include "config.h"

#include <stdexcept>

namespace mpd {

FieldCond ParseFieldCond(const YNode* n) {
    FieldCond fc;
    if (!n || n->IsNull()) return fc;  // default: Neumann, offset 0

    if (n->IsScalar()) {
        const std::string& s = n->scalar;
        if (s == "neumann") return fc;
        if (s == "dirichlet") {
            fc.type = FieldCondType::Dirichlet;
            return fc;
        }
        if (s == "axis_lf") {
            fc.type = FieldCondType::AxisLF;
            return fc;
        }
        if (s == "wall_tangent") {
            throw std::runtime_error(
                "config: 'wall_tangent' is no longer a built-in preset.\n"
                "  Replace it with an expression that encodes the wall-tangent\n"
                "  condition explicitly:\n"
                "    On M_LO (inner wall) — uses interior-neighbour slope:\n"
                "      v_r: \"v_z_nb * r_z_nb\"\n"
                "      H_r: \"H_z_nb * r_z_nb\"\n"
                "    On M_HI (outer wall) — uses wall-cell slope:\n"
                "      v_r: \"v_z * r_z\"\n"
                "      H_r: \"H_z * r_z\"");
        }
        if (s == "hphi_r0_over_r") {
            throw std::runtime_error(
                "config: 'hphi_r0_over_r' is no longer a built-in preset.\n"
                "  Replace it with the equivalent expression:\n"
                "    H_phi: \"r_0 / r\"");
        }
        fc.type = FieldCondType::Expression;
        fc.expr_str = s;
        return fc;
    }

    if (n->IsMap()) {
        if (const YNode* v = n->Get("neumann")) {
            if (!v->IsScalar())
                throw std::runtime_error("config: 'neumann' offset must be a scalar number (e.g. { neumann: 0.5 })");
            fc.type = FieldCondType::Neumann;
            fc.value = v->AsDouble();
            return fc;
        }
        if (const YNode* v = n->Get("dirichlet")) {
            if (!v->IsScalar())
                throw std::runtime_error("config: 'dirichlet' value must be a scalar number (e.g. { dirichlet: 1.0 })");
            fc.type = FieldCondType::Dirichlet;
            fc.value = v->AsDouble();
            return fc;
        }
        if (const YNode* v = n->Get("expr")) {
            if (!v->IsScalar())
                throw std::runtime_error(
                    "config: 'expr' value must be a scalar expression string (e.g. { expr: \"v_z * r_z\" })");
            fc.type = FieldCondType::Expression;
            fc.expr_str = v->scalar;
            return fc;
        }
        if (const YNode* tn = n->Get("type")) {
            const std::string t = tn->AsString();
            if (t == "neumann") {
                fc.type = FieldCondType::Neumann;
                if (const YNode* g = n->Get("gradient")) fc.value = g->AsDouble();
                else if (const YNode* g2 = n->Get("value")) fc.value = g2->AsDouble();
                return fc;
            }
            if (t == "dirichlet") {
                fc.type = FieldCondType::Dirichlet;
                if (const YNode* v = n->Get("value")) fc.value = v->AsDouble();
                return fc;
            }
            if (t == "axis_lf") {
                fc.type = FieldCondType::AxisLF;
                return fc;
            }
            if (t == "expression" || t == "expr") {
                const YNode* e = n->Get("expr");
                if (!e)
                    throw std::runtime_error(
                        "config: { type: expression } requires an 'expr' key containing the formula string\n"
                        "  Example: { type: expression, expr: \"v_z * r_z\" }");
                fc.type = FieldCondType::Expression;
                fc.expr_str = e->AsString();
                return fc;
            }
            if (t == "wall_tangent")
                throw std::runtime_error(
                    "config: { type: wall_tangent } is no longer supported.\n"
                    "  See the 'wall_tangent' migration note above.");
            if (t == "hphi_r0_over_r")
                throw std::runtime_error(
                    "config: { type: hphi_r0_over_r } is no longer supported.\n  Use: H_phi: \"r_0 / r\"");
            throw std::runtime_error("config: unknown field condition type '" + t +
                                     "'.\n  Valid types: neumann, dirichlet, expression, axis_lf.\n"
                                     "  For an expression use a quoted string directly:  v_r: \"v_z * r_z\"");
        }
        throw std::runtime_error(
            "config: field condition map must contain one of: 'neumann', 'dirichlet', 'expr', or 'type'");
    }

    throw std::runtime_error(
        "config: field condition must be:\n"
        "  - a keyword scalar:           neumann / dirichlet / axis_lf\n"
        "  - a quoted expression string: \"v_z * r_z\"\n"
        "  - a short map:                { dirichlet: 1.0 } / { neumann: 0.5 } / { expr: \"...\" }\n"
        "  - a verbose map:              { type: dirichlet, value: 1.0 }");
}

BCFaceConfig ParseFace(const YNode* node) {
    BCFaceConfig face;
    if (!node || node->IsNull()) return face;
    if (!node->IsSequence())
        throw std::runtime_error("config: each boundary-condition face must be a YAML sequence of segment maps");
    for (const YNode& item : node->seq) {
        if (!item.IsMap()) throw std::runtime_error("config: each BC segment must be a YAML map");
        BCSegmentConfig seg;
        if (const YNode* rng = item.Get("range")) {
            if (!rng->IsSequence() || rng->size() != 2)
                throw std::runtime_error(
                    "config: BC segment 'range' must be a two-element sequence [global_lo, global_hi]  (use -1 as "
                    "sentinel for face start/end)");
            seg.global_lo = rng->seq[0].AsInt();
            seg.global_hi = rng->seq[1].AsInt();
        }
        seg.rho = ParseFieldCond(item.Get("rho"));
        seg.v_z = ParseFieldCond(item.Get("v_z"));
        seg.v_r = ParseFieldCond(item.Get("v_r"));
        seg.v_phi = ParseFieldCond(item.Get("v_phi"));
        seg.e = ParseFieldCond(item.Get("e"));
        seg.H_z = ParseFieldCond(item.Get("H_z"));
        seg.H_r = ParseFieldCond(item.Get("H_r"));
        seg.H_phi = ParseFieldCond(item.Get("H_phi"));
        face.segments.push_back(std::move(seg));
    }
    return face;
}

void SimConfig::Load(const std::string& path) {
    YNode root;
    try {
        root = LoadYamlFile(path);
    } catch (const std::exception& e) {
        throw std::runtime_error(std::string("Failed to load config file '") + path + "': " + e.what());
    }
    LoadNode(root);
}

void SimConfig::LoadFromString(const std::string& text) { LoadNode(ParseYaml(text)); }

void SimConfig::LoadNode(const YNode& cfg) {
    if (const YNode* n = cfg.Get("physics")) {
        if (const YNode* v = n->Get("gamma")) gamma = v->AsDouble();
        if (const YNode* v = n->Get("beta")) beta = v->AsDouble();
        if (const YNode* v = n->Get("H_z0")) H_z0 = v->AsDouble();
    }
    if (const YNode* n = cfg.Get("time")) {
        if (const YNode* v = n->Get("T")) T = v->AsDouble();
        if (const YNode* v = n->Get("dt")) dt = v->AsDouble();
        if (const YNode* v = n->Get("max_steps")) max_steps = v->AsInt();
    }
    if (const YNode* n = cfg.Get("adaptive_dt")) {
        if (const YNode* v = n->Get("enabled")) adaptive_dt = v->AsBool();
        if (const YNode* v = n->Get("cfl_number")) cfl_number = v->AsDouble();
        if (const YNode* v = n->Get("growth_factor")) dt_growth_factor = v->AsDouble();
        if (const YNode* v = n->Get("dt_min")) dt_min = v->AsDouble();
        if (const YNode* v = n->Get("dt_max")) dt_max = v->AsDouble();
    }
    if (const YNode* n = cfg.Get("grid")) {
        if (const YNode* v = n->Get("L_max")) L_max = v->AsInt();
        if (const YNode* v = n->Get("M_max")) M_max = v->AsInt();
    }
    if (const YNode* n = cfg.Get("convergence")) {
        if (const YNode* v = n->Get("threshold")) convergence_threshold = v->AsDouble();
        if (const YNode* v = n->Get("check_frequency")) check_frequency = v->AsInt();
    }
    if (const YNode* n = cfg.Get("output")) {
        if (const YNode* v = n->Get("directory")) output_dir = v->AsString();
        if (const YNode* v = n->Get("run_name")) run_name = v->AsString();
        if (const YNode* v = n->Get("vtk_step")) vtk_step = v->AsInt();
    }
    if (const YNode* n = cfg.Get("parallel")) {
        if (const YNode* v = n->Get("openmp_threads")) openmp_threads = v->AsInt();
        if (const YNode* v = n->Get("mpi_dims_l")) mpi_dims_l = v->AsInt();
        if (const YNode* v = n->Get("mpi_dims_m")) mpi_dims_m = v->AsInt();
    }
    if (const YNode* n = cfg.Get("geometry")) {
        if (const YNode* v = n->Get("type")) geometry.type = v->AsString();
        if (const YNode* v = n->Get("params")) geometry.params = *v;
    }
    if (const YNode* n = cfg.Get("initial_conditions")) {
        if (n->Get("type"))
            throw std::runtime_error(
                "config: 'initial_conditions.type' is no longer supported — ExpressionIC is always used.\n"
                "  Remove the 'type' key.  Field expressions now live directly\n"
                "  under 'initial_conditions:', not under 'initial_conditions.params:'.");
        if (n->Get("params"))
            throw std::runtime_error(
                "config: 'initial_conditions.params' is no longer supported.\n"
                "  Move field expressions one level up, directly under 'initial_conditions:'.");
        if (const YNode* vf = n->Get("vtk_file")) {
            if (!vf->IsScalar())
                throw std::runtime_error(
                    "config: 'initial_conditions.vtk_file' must be a scalar file-path string.");
            initial_conditions.vtk_file = vf->scalar;
            // Extension: map axial cells exactly onto a same-size frame (VtkIC).
            if (const YNode* x = n->Get("vtk_exact_z")) initial_conditions.vtk_exact_z = x->AsBool();
        } else {
            initial_conditions.params = *n;
            initial_conditions.has_params = true;
        }
        if (const YNode* nz = n->Get("rho_noise")) {
            if (!nz->IsMap()) throw std::runtime_error("config: 'rho_noise' must be a map {amplitude, seed}");
            if (const YNode* a = nz->Get("amplitude")) rho_noise_amplitude = a->AsDouble();
            if (const YNode* s = nz->Get("seed")) rho_noise_seed = (unsigned long long)std::stoull(s->AsString());
        }
    }
    if (const YNode* n = cfg.Get("boundary_conditions")) {
        if (const YNode* v = n->Get("l_lo")) bc_l_lo = ParseFace(v);
        if (const YNode* v = n->Get("l_hi")) bc_l_hi = ParseFace(v);
        if (const YNode* v = n->Get("m_lo")) bc_m_lo = ParseFace(v);
        if (const YNode* v = n->Get("m_hi")) bc_m_hi = ParseFace(v);
    }
    if (const YNode* n = cfg.Get("gpu")) {
        if (const YNode* v = n->Get("kernel")) gpu.kernel = v->AsString();
        if (const YNode* v = n->Get("fast_div")) gpu.fast_div = v->AsBool();
        if (const YNode* v = n->Get("tile_rows")) gpu.tile_rows = v->AsInt();
        if (const YNode* v = n->Get("steps_per_graph")) gpu.steps_per_graph = v->AsInt();
    }
    Init();
}

}  // namespace mpd
