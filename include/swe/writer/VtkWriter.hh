// VtkWriter.hh -- VTK XML structured-grid output (<base>.<step>.vts, ASCII),
// the format of the reference's VtkWriter (src/writer/VtkWriter.cpp:45-120):
// grid points (offset + i) * d, cell data h, hu, hv, b over the interior.
#ifndef SWE_HIP_VTKWRITER_HH
#define SWE_HIP_VTKWRITER_HH

#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "writer/Writer.hh"

class VtkWriter : public Writer {
public:
    VtkWriter(const std::string& baseName, const Float2D& b, const BoundarySize& boundarySize, int nX, int nY,
              float dX, float dY, int offsetX = 0, int offsetY = 0)
        : Writer(baseName, b, boundarySize, nX, nY), dX(dX), dY(dY), offsetX(offsetX), offsetY(offsetY) {}

    void writeTimeStep(const Float2D& h, const Float2D& hu, const Float2D& hv, float) override {
        std::ofstream f(generateFileName().c_str());
        if (!f.good()) throw std::runtime_error("VtkWriter: cannot open " + generateFileName());
        f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"StructuredGrid\">\n";
        f << "<StructuredGrid WholeExtent=\"" << offsetX << " " << offsetX + nX << " " << offsetY << " "
          << offsetY + nY << " 0 0\">\n";
        f << "<Piece Extent=\"" << offsetX << " " << offsetX + nX << " " << offsetY << " " << offsetY + nY
          << " 0 0\">\n";
        f << "<Points>\n<DataArray NumberOfComponents=\"3\" type=\"Float32\" format=\"ascii\">\n";
        for (int j = 0; j < nY + 1; j++)
            for (int i = 0; i < nX + 1; i++) f << (offsetX + i) * dX << " " << (offsetY + j) * dY << " 0\n";
        f << "</DataArray>\n</Points>\n<CellData>\n";
        cells(f, "h", h);
        cells(f, "hu", hu);
        cells(f, "hv", hv);
        cells(f, "b", b);
        f << "</CellData>\n</Piece>\n</StructuredGrid>\n</VTKFile>\n";
        timeStep++;
    }

private:
    void cells(std::ofstream& f, const char* name, const Float2D& a) {
        f << "<DataArray Name=\"" << name << "\" type=\"Float32\" format=\"ascii\">\n";
        for (int j = 1; j < nY + 1; j++)
            for (int i = 1; i < nX + 1; i++) f << a[i][j] << "\n";
        f << "</DataArray>\n";
    }
    std::string generateFileName() {
        std::ostringstream name;
        name << fileName << '.' << timeStep << ".vts";
        return name.str();
    }
    float dX, dY;
    int offsetX, offsetY;
};

#endif
