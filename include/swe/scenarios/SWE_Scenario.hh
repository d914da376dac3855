// SWE_Scenario.hh -- scenario interface (src/scenarios/SWE_Scenario.hh:42-66):
// initial water height, bathymetry and velocities at a point, boundary types
// and positions.  The default is a constant scenario.
#ifndef SWE_HIP_SCENARIO_HH
#define SWE_HIP_SCENARIO_HH

#include "types/Boundary.hh"

class SWE_Scenario {
public:
    virtual ~SWE_Scenario() {}
    virtual float getWaterHeight(float x, float y) { return 0; }
    virtual float getBathymetry(float x, float y) { return 0; }
    virtual float getVeloc_u(float x, float y) { return 0; }
    virtual float getVeloc_v(float x, float y) { return 0; }
    virtual BoundaryType getBoundaryType(Boundary boundary) { return OUTFLOW; }
    virtual float getBoundaryPos(Boundary boundary) {
        return (boundary == BND_LEFT || boundary == BND_BOTTOM) ? 0.0f : 1.0f;
    }
    virtual float waterHeightAtRest() { return 0; }
    virtual float endSimulation() { return 0; }
};

#endif
