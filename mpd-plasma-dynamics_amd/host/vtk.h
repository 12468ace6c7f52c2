// vtk.h -- legacy VTK structured-grid files on either side of the step: the
// frames IOManager::WriteVtk writes (src/io_manager.cpp:218-315) and the
// VTK restart initial condition VtkIC (include/ics/vtk_ic.hpp,
// src/ics/vtk_ic.cpp:42-239).  No VTK library: the legacy format is written
// and parsed here directly.
#pragma once

#include <string>
#include <vector>

#include "config.h"
#include "setup.h"

namespace mpd {

// One point-data array of a frame: ncomp components per point, point k = i +
// j * ni (i axial, j radial).
struct VtkArray {
    std::string name;
    int ncomp = 1;
    std::vector<double> data;
};

// A structured grid of ni x nj x 1 points with point coordinates (x, y, z)
// and point-data arrays.
struct VtkFrame {
    int ni = 0, nj = 0;
    std::vector<double> points;  // 3 per point
    std::vector<VtkArray> arrays;
    const VtkArray* Find(const std::string& name) const;
};

// The frame IOManager::WriteVtk builds from the gathered fields: ni = L_max +
// 1 points in z (x = i dz; the last point repeats row L_max - 1), nj = M_max +
// 1 in r (y = r), arrays Rho, Energy, MPI_Rank (0: one rank), Hphi_times_r,
// Velocity (v_z, v_r, v_phi) and MagneticField (H_z, H_r, H_phi).  phys holds
// the 10 physical arrays of one rank owning the domain (rows x cols, ghost
// ring included, Fields member order).
VtkFrame MakeFrame(const SimConfig& cfg, const HostGrid& g, const std::vector<double>* phys);

// Legacy file, "BINARY" (big-endian doubles) as vtkStructuredGridWriter's
// binary mode writes it, or ASCII.  Point-data arrays go in one FIELD block.
void WriteVtk(const std::string& path, const VtkFrame& f, bool binary = true);

// Reads a legacy structured grid (BINARY or ASCII; point data as FIELD,
// SCALARS or VECTORS sections).  Throws std::runtime_error on anything else.
VtkFrame ReadVtk(const std::string& path);

// VtkIC: bilinear interpolation of the eight physical fields of a frame onto
// this grid, then p and P as ExpressionIC derives them.  The reference maps
// axial cell l_g to fractional point index l_g (ni - 1) / (L_max - 1)
// (src/ics/vtk_ic.cpp:185-188), which for a same-size restart is not the
// exact lookup its comment describes; exact_z maps l_g (ni - 2) / (L_max - 1)
// instead (the last point repeats the last row), which is exact there.
class VtkIC : public IInitialCondition {
public:
    VtkIC(const std::string& path, bool exact_z);
    void Apply(HostFields& f, const HostGrid& g, const SimConfig& cfg, int l_start) const override;
    std::string Name() const override { return "vtk"; }

private:
    VtkFrame frame_;
    bool exact_z_;
    std::vector<double> rho_, vz_, vr_, vphi_, hz_, hr_, hphi_, e_;
    double Interp(const std::vector<double>& d, double i_frac, double j_frac) const;
};

}  // namespace mpd
