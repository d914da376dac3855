// Float2DBuffer.hh -- ghost-layer receive buffer: its own storage when local
// time stepping is on, otherwise an alias of the field it serves
// (src/tools/Float2DBuffer.hh:13-41).
#ifndef SWE_HIP_FLOAT2DBUFFER_HH
#define SWE_HIP_FLOAT2DBUFFER_HH

#include <memory>

#include "tools/Float2D.hh"
#include "tools/Float2DNative.hh"

class Float2DBuffer : public Float2D {
public:
    Float2DBuffer() : Float2D(0, 0) {}
    Float2DBuffer(int c, int r, bool localTimestepping, Float2DNative& field) : Float2D(c, r) {
        if (localTimestepping)
            data = std::shared_ptr<float>(new float[(size_t)c * r](), std::default_delete<float[]>());
        else
            data = field.getPointer();
        rawData = data.get();
    }
    ~Float2DBuffer() {}

private:
    std::shared_ptr<float> data;
};

#endif
