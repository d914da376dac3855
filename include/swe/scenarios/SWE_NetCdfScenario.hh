// SWE_NetCdfScenario.hh -- the file-driven scenario of the reference
// (SWE_AsagiScenario.hh:42-151): bathymetry and an initial displacement of the
// sea floor read from gridded files, OUTFLOW boundaries, and the domain given
// by the bathymetry grid's extent.
//
//   water height  h(x, y) = bathy(x, y) > 0 ? 0 : -bathy(x, y)
//   bathymetry    b(x, y) = bathy(x, y) + displ(x, y), the displacement taken
//                           only strictly inside its grid's extent (:104-131)
//   boundaries    OUTFLOW; positions = the bathymetry extent (:133-146)
//
// The reference reads the grids through ASAGI (TUM-I5/ASAGI, a library absent
// from the reference tree; version not pinned there).  This class reads
// netCDF classic files itself (NetCdfReader) and restates the grid lookup as
// ASAGI documents it for cell-centred values without interpolation: the
// variable `z` (ASAGI's default name) over dimensions (y, x) with coordinate
// variables named after the dimensions; value (i, j) sits at the coordinates
// (x_i, y_j) and covers the cell of width dx around it; a point takes the
// value of the cell containing it, found in double precision as
// floor((x - (x_0 - dx/2)) / dx), clamped to the grid; the extent runs from
// x_0 - dx/2 to x_{n-1} + dx/2.  The same lookup runs on the device in
// swe_hip_init_grid_scenario.
#ifndef SWE_HIP_NETCDF_SCENARIO_HH
#define SWE_HIP_NETCDF_SCENARIO_HH

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "scenarios/SWE_Scenario.hh"
#include "tools/NetCdfReader.hh"

class GriddedField {
public:
    int nx = 0, ny = 0;
    double x0 = 0, y0 = 0, dx = 1, dy = 1;  // first cell centre, spacing
    std::vector<float> values;               // ny rows of nx, x fastest

    // reads variable `var` (dimensions (y, x)) and its coordinate variables
    static GriddedField load(const std::string& file, const std::string& var = "z") {
        NetCdfReader r(file);
        const NetCdfReader::Var& z = r.var(var);
        if (z.dims.size() != 2 || z.record)
            throw std::runtime_error(file + ": '" + var + "' must have two fixed dimensions (y, x)");
        const std::string ny = r.dimensions()[z.dims[0]].name, nx = r.dimensions()[z.dims[1]].name;
        const std::vector<float> xs = r.read(r.var(nx)), ys = r.read(r.var(ny));
        GriddedField g;
        g.nx = (int)xs.size();
        g.ny = (int)ys.size();
        if (g.nx < 2 || g.ny < 2) throw std::runtime_error(file + ": a grid needs at least 2 x 2 values");
        g.x0 = xs[0];
        g.y0 = ys[0];
        g.dx = ((double)xs[g.nx - 1] - (double)xs[0]) / (g.nx - 1);
        g.dy = ((double)ys[g.ny - 1] - (double)ys[0]) / (g.ny - 1);
        if (!(g.dx > 0) || !(g.dy > 0)) throw std::runtime_error(file + ": coordinates must increase");
        g.values = r.read(z);
        return g;
    }

    // the extent as the reference keeps it: float (SWE_AsagiScenario.hh:148-149)
    float xMin() const { return (float)(x0 - 0.5 * dx); }
    float xMax() const { return (float)(x0 + (nx - 0.5) * dx); }
    float yMin() const { return (float)(y0 - 0.5 * dy); }
    float yMax() const { return (float)(y0 + (ny - 0.5) * dy); }

    bool inside(float x, float y) const { return x > xMin() && x < xMax() && y > yMin() && y < yMax(); }

    float at(float x, float y) const {
        long i = (long)std::floor(((double)x - (x0 - 0.5 * dx)) / dx);
        long j = (long)std::floor(((double)y - (y0 - 0.5 * dy)) / dy);
        i = i < 0 ? 0 : (i >= nx ? nx - 1 : i);
        j = j < 0 ? 0 : (j >= ny ? ny - 1 : j);
        return values[(size_t)j * nx + i];
    }
};

class SWE_NetCdfScenario : public SWE_Scenario {
public:
    // displacementFile may be empty: no displacement
    SWE_NetCdfScenario(const std::string& bathymetryFile, const std::string& displacementFile = "")
        : bathy(GriddedField::load(bathymetryFile)) {
        if (!displacementFile.empty()) displ.reset(new GriddedField(GriddedField::load(displacementFile)));
    }

    float getWaterHeight(float x, float y) override {
        const float bv = bathy.at(x, y);
        return (bv > 0.f) ? 0.f : -bv;
    }
    float getBathymetry(float x, float y) override {
        const float bv = bathy.at(x, y);
        float dv = 0.f;
        if (displ && displ->inside(x, y)) dv = displ->at(x, y);
        return bv + dv;
    }
    BoundaryType getBoundaryType(Boundary) override { return OUTFLOW; }
    float getBoundaryPos(Boundary boundary) override {
        if (boundary == BND_LEFT) return bathy.xMin();
        if (boundary == BND_RIGHT) return bathy.xMax();
        if (boundary == BND_BOTTOM) return bathy.yMin();
        return bathy.yMax();
    }

    const GriddedField& bathymetryGrid() const { return bathy; }
    const GriddedField* displacementGrid() const { return displ.get(); }

private:
    GriddedField bathy;
    std::unique_ptr<GriddedField> displ;
};

#endif
