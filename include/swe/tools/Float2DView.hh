// Float2DView.hh -- non-owning Float2D over an existing array (for example
// the pinned host arrays of an asynchronous snapshot), so that the writers,
// which take Float2D, can write from it.  No counterpart in the reference.
#ifndef SWE_HIP_FLOAT2DVIEW_HH
#define SWE_HIP_FLOAT2DVIEW_HH

#include "tools/Float2D.hh"

class Float2DView : public Float2D {
public:
    Float2DView(int c, int r, const float* p) : Float2D(c, r) { rawData = const_cast<float*>(p); }
    ~Float2DView() {}
};

#endif
