// vtk.cpp -- legacy VTK structured grids: the frame of IOManager::WriteVtk
// (src/io_manager.cpp:218-315), its writer and reader, and VtkIC
// (src/ics/vtk_ic.cpp:42-239).
#include "vtk.h"

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>

namespace mpd {

const VtkArray* VtkFrame::Find(const std::string& name) const {
    for (const VtkArray& a : arrays)
        if (a.name == name) return &a;
    return nullptr;
}

VtkFrame MakeFrame(const SimConfig& cfg, const HostGrid& g, const std::vector<double>* phys) {
    const int L_g = cfg.L_max, M_max = cfg.M_max;
    const int ni = L_g + 1, nj = M_max + 1;
    const int cols = g.cols;
    const size_t n = (size_t)ni * nj;
    VtkFrame f;
    f.ni = ni;
    f.nj = nj;
    f.points.resize(3 * n);
    // Global array [l_g][m_g] of one rank = local [l_g + 1][m_g + 1].
    auto fld = [&](int k, int li, int j) { return phys[k][(size_t)(li + 1) * cols + (j + 1)]; };
    auto rr = [&](int li, int j) { return g.r(li + 1, j + 1); };
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const int li = std::min(i, L_g - 1);
            const size_t k = (size_t)i + (size_t)j * ni;
            f.points[3 * k] = i * cfg.dz;
            f.points[3 * k + 1] = rr(li, j);
            f.points[3 * k + 2] = 0.0;
        }
    auto scalar = [&](const char* name, auto value) {
        VtkArray a;
        a.name = name;
        a.ncomp = 1;
        a.data.resize(n);
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) a.data[(size_t)i + (size_t)j * ni] = value(std::min(i, L_g - 1), j);
        f.arrays.push_back(std::move(a));
    };
    auto vector3 = [&](const char* name, int kz, int kr, int kphi) {
        VtkArray a;
        a.name = name;
        a.ncomp = 3;
        a.data.resize(3 * n);
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const int li = std::min(i, L_g - 1);
                const size_t k = (size_t)i + (size_t)j * ni;
                a.data[3 * k] = fld(kz, li, j);
                a.data[3 * k + 1] = fld(kr, li, j);
                a.data[3 * k + 2] = fld(kphi, li, j);
            }
        f.arrays.push_back(std::move(a));
    };
    scalar("Rho", [&](int li, int j) { return fld(F_RHO, li, j); });
    scalar("Energy", [&](int li, int j) { return fld(F_E, li, j); });
    scalar("MPI_Rank", [&](int, int) { return 0.0; });
    scalar("Hphi_times_r", [&](int li, int j) { return fld(F_HPHI, li, j) * rr(li, j); });
    vector3("Velocity", F_VZ, F_VR, F_VPHI);
    vector3("MagneticField", F_HZ, F_HR, F_HPHI);
    return f;
}

namespace {

void PutBE(std::ostream& os, const double* v, size_t n) {
    std::vector<unsigned char> buf(8 * n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t b;
        std::memcpy(&b, &v[i], 8);
        for (int k = 0; k < 8; ++k) buf[8 * i + k] = (unsigned char)(b >> (56 - 8 * k));
    }
    os.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)buf.size());
}

void PutAscii(std::ostream& os, const double* v, size_t n, int per_line) {
    char tmp[40];
    for (size_t i = 0; i < n; ++i) {
        std::snprintf(tmp, sizeof(tmp), "%.17g", v[i]);
        os << tmp << (((i + 1) % per_line == 0 || i + 1 == n) ? '\n' : ' ');
    }
}

}  // namespace

void WriteVtk(const std::string& path, const VtkFrame& f, bool binary) {
    std::ofstream os(path, std::ios::binary);
    if (!os) throw std::runtime_error("WriteVtk: cannot open '" + path + "' for writing");
    const size_t n = (size_t)f.ni * f.nj;
    os << "# vtk DataFile Version 5.1\n"
       << "vtk output\n"
       << (binary ? "BINARY\n" : "ASCII\n")
       << "DATASET STRUCTURED_GRID\n"
       << "DIMENSIONS " << f.ni << " " << f.nj << " 1\n"
       << "POINTS " << n << " double\n";
    if (binary) {
        PutBE(os, f.points.data(), f.points.size());
        os << "\n";
    } else {
        PutAscii(os, f.points.data(), f.points.size(), 9);
    }
    os << "POINT_DATA " << n << "\n"
       << "FIELD FieldData " << f.arrays.size() << "\n";
    for (const VtkArray& a : f.arrays) {
        os << a.name << " " << a.ncomp << " " << n << " double\n";
        if (binary) {
            PutBE(os, a.data.data(), a.data.size());
            os << "\n";
        } else {
            PutAscii(os, a.data.data(), a.data.size(), 9);
        }
    }
    if (!os) throw std::runtime_error("WriteVtk: write to '" + path + "' failed");
}

namespace {

// Token / line reader over an in-memory file with binary payloads.
struct Cursor {
    std::string s;
    size_t pos = 0;
    std::string path;

    [[noreturn]] void Fail(const std::string& what) const {
        throw std::runtime_error("ReadVtk: " + what + " in '" + path + "'");
    }
    std::string Line() {
        if (pos >= s.size()) Fail("unexpected end of file");
        const size_t e = s.find('\n', pos);
        std::string l = s.substr(pos, (e == std::string::npos ? s.size() : e) - pos);
        pos = (e == std::string::npos) ? s.size() : e + 1;
        if (!l.empty() && l.back() == '\r') l.pop_back();
        return l;
    }
    // Next non-empty line.
    std::string Content() {
        for (;;) {
            std::string l = Line();
            if (l.find_first_not_of(" \t") != std::string::npos) return l;
        }
    }
    bool AtEnd() {
        size_t p = pos;
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
        return p >= s.size();
    }
    std::string Word() {
        while (pos < s.size() && std::isspace((unsigned char)s[pos])) ++pos;
        if (pos >= s.size()) Fail("unexpected end of file");
        const size_t b = pos;
        while (pos < s.size() && !std::isspace((unsigned char)s[pos])) ++pos;
        return s.substr(b, pos - b);
    }
    // n values of the given type, binary big-endian or ASCII.
    std::vector<double> Values(size_t n, const std::string& type, bool binary) {
        std::vector<double> v(n);
        const bool dbl = (type == "double");
        if (!dbl && type != "float") Fail("unsupported data type '" + type + "'");
        if (binary) {
            const size_t w = dbl ? 8 : 4;
            if (pos + n * w > s.size()) Fail("truncated binary data");
            const unsigned char* p = reinterpret_cast<const unsigned char*>(s.data() + pos);
            for (size_t i = 0; i < n; ++i) {
                uint64_t b = 0;
                for (size_t k = 0; k < w; ++k) b = (b << 8) | p[w * i + k];
                if (dbl) {
                    std::memcpy(&v[i], &b, 8);
                } else {
                    const uint32_t b32 = (uint32_t)b;
                    float x;
                    std::memcpy(&x, &b32, 4);
                    v[i] = x;
                }
            }
            pos += n * w;
            if (pos < s.size() && s[pos] == '\n') ++pos;
        } else {
            for (size_t i = 0; i < n; ++i) {
                const std::string t = Word();
                char* end = nullptr;
                v[i] = std::strtod(t.c_str(), &end);
                if (end == t.c_str() || *end) Fail("bad number '" + t + "'");
            }
        }
        return v;
    }
};

}  // namespace

VtkFrame ReadVtk(const std::string& path) {
    std::ifstream is(path, std::ios::binary);
    if (!is) throw std::runtime_error("ReadVtk: cannot open '" + path + "'");
    Cursor c;
    c.path = path;
    c.s.assign(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>());
    if (c.Line().rfind("# vtk DataFile Version", 0) != 0) c.Fail("not a legacy VTK file");
    c.Line();  // title
    const std::string fmt = c.Content();
    const bool binary = (fmt == "BINARY");
    if (!binary && fmt != "ASCII") c.Fail("format must be ASCII or BINARY");
    VtkFrame f;
    size_t n = 0;
    bool have_dims = false, have_points = false;
    {
        std::istringstream ds(c.Content());
        std::string kw, kind;
        ds >> kw >> kind;
        if (kw != "DATASET" || kind != "STRUCTURED_GRID") c.Fail("dataset must be STRUCTURED_GRID");
    }
    while (!c.AtEnd()) {
        const std::string kw = c.Word();
        if (kw == "DIMENSIONS") {
            f.ni = std::atoi(c.Word().c_str());
            f.nj = std::atoi(c.Word().c_str());
            const int nk = std::atoi(c.Word().c_str());
            if (f.ni < 1 || f.nj < 1 || nk != 1) c.Fail("DIMENSIONS must be ni nj 1");
            n = (size_t)f.ni * f.nj;
            have_dims = true;
            c.Line();
        } else if (kw == "POINTS") {
            const size_t np = (size_t)std::atoll(c.Word().c_str());
            const std::string type = c.Word();
            c.Line();
            if (!have_dims || np != n) c.Fail("POINTS count does not match DIMENSIONS");
            f.points = c.Values(3 * np, type, binary);
            have_points = true;
        } else if (kw == "POINT_DATA") {
            if ((size_t)std::atoll(c.Word().c_str()) != n) c.Fail("POINT_DATA count does not match");
            c.Line();
        } else if (kw == "FIELD") {
            c.Word();  // field data name
            const int narr = std::atoi(c.Word().c_str());
            c.Line();
            for (int a = 0; a < narr; ++a) {
                VtkArray arr;
                arr.name = c.Word();
                arr.ncomp = std::atoi(c.Word().c_str());
                const size_t nt = (size_t)std::atoll(c.Word().c_str());
                const std::string type = c.Word();
                c.Line();
                if (nt != n || arr.ncomp < 1) c.Fail("array '" + arr.name + "' has the wrong size");
                arr.data = c.Values(nt * arr.ncomp, type, binary);
                f.arrays.push_back(std::move(arr));
            }
        } else if (kw == "SCALARS") {
            VtkArray arr;
            std::istringstream ls(c.Line());
            std::string type;
            ls >> arr.name >> type;
            if (!(ls >> arr.ncomp)) arr.ncomp = 1;
            const std::string lut = c.Content();
            if (lut.rfind("LOOKUP_TABLE", 0) != 0) c.Fail("SCALARS without LOOKUP_TABLE");
            arr.data = c.Values(n * arr.ncomp, type, binary);
            f.arrays.push_back(std::move(arr));
        } else if (kw == "VECTORS") {
            VtkArray arr;
            std::istringstream ls(c.Line());
            std::string type;
            ls >> arr.name >> type;
            arr.ncomp = 3;
            arr.data = c.Values(3 * n, type, binary);
            f.arrays.push_back(std::move(arr));
        } else if (kw == "METADATA") {
            // Skip to the blank line that ends the block.
            c.Line();
            for (;;) {
                const std::string l = c.Line();
                if (l.find_first_not_of(" \t") == std::string::npos) break;
            }
        } else if (kw == "CELL_DATA") {
            break;  // cell data is not written by the solver and not read
        } else {
            c.Fail("unexpected keyword '" + kw + "'");
        }
    }
    if (!have_points) c.Fail("no POINTS section");
    return f;
}

VtkIC::VtkIC(const std::string& path, bool exact_z) : exact_z_(exact_z) {
    frame_ = ReadVtk(path);
    if (frame_.ni < 2 || frame_.nj < 2)
        throw std::runtime_error("VtkIC: VTK grid too small (" + std::to_string(frame_.ni) + " x " +
                                 std::to_string(frame_.nj) + ").  Expected at least 2 x 2.");
    auto comp = [&](const char* name, int k, int need) {
        const VtkArray* a = frame_.Find(name);
        if (!a)
            throw std::runtime_error(std::string("VtkIC: point-data array '") + name + "' not found in '" + path +
                                     "'.\n  The file must have been written by IOManager::WriteFrame.");
        if (a->ncomp < need)
            throw std::runtime_error(std::string("VtkIC: array '") + name + "' has only " +
                                     std::to_string(a->ncomp) + " component(s); requested component " +
                                     std::to_string(k) + ".");
        std::vector<double> out((size_t)frame_.ni * frame_.nj);
        for (size_t p = 0; p < out.size(); ++p) out[p] = a->data[p * a->ncomp + k];
        return out;
    };
    rho_ = comp("Rho", 0, 1);
    e_ = comp("Energy", 0, 1);
    vz_ = comp("Velocity", 0, 1);
    vr_ = comp("Velocity", 1, 2);
    vphi_ = comp("Velocity", 2, 3);
    hz_ = comp("MagneticField", 0, 1);
    hr_ = comp("MagneticField", 1, 2);
    hphi_ = comp("MagneticField", 2, 3);
}

// VtkIC::Impl::Interp, src/ics/vtk_ic.cpp:137-160.
double VtkIC::Interp(const std::vector<double>& d, double i_frac, double j_frac) const {
    const int ni = frame_.ni, nj = frame_.nj;
    i_frac = std::clamp(i_frac, 0.0, (double)(ni - 1));
    j_frac = std::clamp(j_frac, 0.0, (double)(nj - 1));
    const int i0 = std::min((int)i_frac, ni - 2);
    const int j0 = std::min((int)j_frac, nj - 2);
    const double ti = i_frac - (double)i0;
    const double tj = j_frac - (double)j0;
    const double v00 = d[(size_t)i0 + (size_t)j0 * ni];
    const double v10 = d[(size_t)(i0 + 1) + (size_t)j0 * ni];
    const double v01 = d[(size_t)i0 + (size_t)(j0 + 1) * ni];
    const double v11 = d[(size_t)(i0 + 1) + (size_t)(j0 + 1) * ni];
    return (1.0 - ti) * (1.0 - tj) * v00 + ti * (1.0 - tj) * v10 + (1.0 - ti) * tj * v01 + ti * tj * v11;
}

// VtkIC::Apply, src/ics/vtk_ic.cpp:171-239 (one rank: m_start = 0).
void VtkIC::Apply(HostFields& f, const HostGrid&, const SimConfig& cfg, int l_start) const {
    const double l_max_new = (double)cfg.L_max;
    const double m_max_new = (double)cfg.M_max;
    const int iden = exact_z_ ? frame_.ni - 2 : frame_.ni - 1;
    const double i_scale = (cfg.L_max > 1) ? (double)iden / (l_max_new - 1.0) : 0.0;
    const double j_scale = (cfg.M_max > 0) ? (double)(frame_.nj - 1) / m_max_new : 0.0;
    const double gamma = cfg.gamma;
    for (int l = 1; l < f.rows - 1; ++l) {
        const int l_global = l_start + l - 1;
        const double i_frac = (double)l_global * i_scale;
        for (int m = 1; m < f.cols - 1; ++m) {
            const int m_global = m - 1;
            const double j_frac = (double)m_global * j_scale;
            const size_t k = (size_t)l * f.cols + m;
            f.phys[F_RHO][k] = Interp(rho_, i_frac, j_frac);
            f.phys[F_VZ][k] = Interp(vz_, i_frac, j_frac);
            f.phys[F_VR][k] = Interp(vr_, i_frac, j_frac);
            f.phys[F_VPHI][k] = Interp(vphi_, i_frac, j_frac);
            f.phys[F_HZ][k] = Interp(hz_, i_frac, j_frac);
            f.phys[F_HR][k] = Interp(hr_, i_frac, j_frac);
            f.phys[F_HPHI][k] = Interp(hphi_, i_frac, j_frac);
            f.phys[F_E][k] = Interp(e_, i_frac, j_frac);
            f.phys[F_P][k] = (gamma - 1.0) * f.phys[F_RHO][k] * f.phys[F_E][k];
            f.phys[F_PT][k] = f.phys[F_P][k] + 0.5 * (f.phys[F_HZ][k] * f.phys[F_HZ][k] +
                                                      f.phys[F_HR][k] * f.phys[F_HR][k] +
                                                      f.phys[F_HPHI][k] * f.phys[F_HPHI][k]);
        }
    }
}

}  // namespace mpd
