// Float2DNative.hh -- owning Float2D (shared ownership of a heap array), as
// src/tools/Float2DNative.hh:8-28.
#ifndef SWE_HIP_FLOAT2DNATIVE_HH
#define SWE_HIP_FLOAT2DNATIVE_HH

#include <memory>

#include "tools/Float2D.hh"

class Float2DNative : public Float2D {
public:
    Float2DNative() : Float2D(0, 0) {}
    Float2DNative(int c, int r) : Float2D(c, r), data(new float[(size_t)c * r](), std::default_delete<float[]>()) {
        rawData = data.get();
    }
    ~Float2DNative() {}
    std::shared_ptr<float> getPointer() { return data; }

private:
    std::shared_ptr<float> data;
};

#endif
