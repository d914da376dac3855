// Writer.hh -- output interface with the reference's signature
// (src/writer/Writer.hh:36-100): constructed with a file name, a reference to
// the bathymetry, the ghost-layer widths and the interior size; writes one
// time step from host Float2D arrays.  Blocks whose state lives on the GPU
// hand out host arrays through their accessors, which synchronise lazily.
#ifndef SWE_HIP_WRITER_HH
#define SWE_HIP_WRITER_HH

#include <string>

#include "tools/Float2D.hh"

struct BoundarySize {
    int boundarySize[4];  // left, right, bottom, top
    int& operator[](unsigned int i) { return boundarySize[i]; }
    int operator[](unsigned int i) const { return boundarySize[i]; }
};

class Writer {
public:
    Writer(const std::string& fileName, const Float2D& b, const BoundarySize& boundarySize, int nX, int nY)
        : fileName(fileName), b(b), boundarySize(boundarySize), nX(nX), nY(nY), timeStep(0) {}
    virtual ~Writer() {}
    virtual void writeTimeStep(const Float2D& h, const Float2D& hu, const Float2D& hv, float time) = 0;

protected:
    const std::string fileName;
    const Float2D& b;
    const BoundarySize boundarySize;
    const int nX, nY;
    size_t timeStep;
};

#endif
