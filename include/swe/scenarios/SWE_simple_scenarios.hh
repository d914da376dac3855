// SWE_simple_scenarios.hh -- the reference's analytic scenarios
// (src/scenarios/SWE_simple_scenarios.hh:37-285: same class names, domains,
// depths and boundary types, evaluated in single precision) plus the
// centred flat-bed dam break of the 256^2 f-wave workload.
#ifndef SWE_HIP_SIMPLE_SCENARIOS_HH
#define SWE_HIP_SIMPLE_SCENARIOS_HH

#include <cmath>

#include "scenarios/SWE_Scenario.hh"

namespace swe_scenario_detail {
inline float radius(float x, float y, float cx, float cy) {
    return std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy));
}
inline float box_pos(Boundary b, float hi) { return (b == BND_LEFT || b == BND_BOTTOM) ? 0.f : hi; }
}  // namespace swe_scenario_detail

// half of the domain wet around a raised disc, walls (SWE_simple_scenarios.hh:37-71)
class SWE_HalfDomainDry : public SWE_Scenario {
public:
    float getBathymetry(float, float) override { return -250.f; }
    float getWaterHeight(float x, float y) override {
        return swe_scenario_detail::radius(x, y, 250.f, 250.f) < 100.f ? 15.f : 10.f;
    }
    float endSimulation() override { return 15.f; }
    BoundaryType getBoundaryType(Boundary) override { return WALL; }
    float getBoundaryPos(Boundary b) override { return swe_scenario_detail::box_pos(b, 1000.f); }
};

// radial dam break: 15 m within 100 m of (250, 250), 10 m elsewhere, b = 10
// (SWE_simple_scenarios.hh:77-108)
class SWE_RadialDamBreakScenario : public SWE_Scenario {
public:
    float getBathymetry(float, float) override { return 10.f; }
    float getWaterHeight(float x, float y) override {
        return swe_scenario_detail::radius(x, y, 250.f, 250.f) < 100.f ? 15.f : 10.f;
    }
    float endSimulation() override { return 15.f; }
    BoundaryType getBoundaryType(Boundary) override { return OUTFLOW; }
    float getBoundaryPos(Boundary b) override { return swe_scenario_detail::box_pos(b, 1000.f); }
};

// the scenario of the reference's parallel drivers: raised disc of
// bathymetry under a dam break on [0, 2000]^2 (SWE_simple_scenarios.hh:110-152)
class SWE_RadialBathymetryDamBreakScenario : public SWE_Scenario {
public:
    float getBathymetry(float x, float y) override {
        return swe_scenario_detail::radius(x, y, 1000.f, 1000.f) < 300.f ? 20.f : 10.f;
    }
    float getWaterHeight(float x, float y) override {
        return swe_scenario_detail::radius(x, y, 500.f, 500.f) < 300.f ? 40.f : 15.f;
    }
    float endSimulation() override { return 15.f; }
    BoundaryType getBoundaryType(Boundary) override { return OUTFLOW; }
    float getBoundaryPos(Boundary b) override { return swe_scenario_detail::box_pos(b, 2000.f); }
};

// uniform depth over a raised disc (SWE_simple_scenarios.hh:157-196)
class SWE_BathymetryDamBreakScenario : public SWE_Scenario {
public:
    float getBathymetry(float x, float y) override {
        return swe_scenario_detail::radius(x, y, 500.f, 500.f) < 50.f ? -255.f : -260.f;
    }
    float getWaterHeight(float, float) override { return 260.f; }
    float endSimulation() override { return 15.f; }
    BoundaryType getBoundaryType(Boundary) override { return OUTFLOW; }
    float getBoundaryPos(Boundary b) override { return swe_scenario_detail::box_pos(b, 1000.f); }
};

// flat surface over a bump (SWE_simple_scenarios.hh:204-216)
class SWE_SeaAtRestScenario : public SWE_Scenario {
public:
    float getWaterHeight(float x, float y) override {
        return swe_scenario_detail::radius(x, y, 0.5f, 0.5f) < 0.1f ? 9.9f : 10.0f;
    }
    float getBathymetry(float x, float y) override {
        return swe_scenario_detail::radius(x, y, 0.5f, 0.5f) < 0.1f ? 0.1f : 0.0f;
    }
};

// inclined surface (SWE_simple_scenarios.hh:222-253)
class SWE_SplashingPoolScenario : public SWE_Scenario {
public:
    float getBathymetry(float, float) override { return -250.f; }
    float getWaterHeight(float x, float y) override { return 250.0f + (5.0f - (x + y) / 200); }
    float endSimulation() override { return 15.f; }
    float getBoundaryPos(Boundary b) override { return swe_scenario_detail::box_pos(b, 1000.f); }
};

// cone with a raised centre (SWE_simple_scenarios.hh:261-284)
class SWE_SplashingConeScenario : public SWE_Scenario {
public:
    float getWaterHeight(float x, float y) override {
        const float r = swe_scenario_detail::radius(x, y, 0.5f, 0.5f);
        float h = 4.0f - 4.5f * (r / 0.5f);
        if (r < 0.1f) h = h + 1.0f;
        return (h > 0.0f) ? h : 0.0f;
    }
    float getBathymetry(float x, float y) override {
        const float r = swe_scenario_detail::radius(x, y, 0.5f, 0.5f);
        return 1.0f + 9.0f * ((r < 0.5f) ? r : 0.5f);
    }
    float waterHeightAtRest() override { return 4.0f; }
    float endSimulation() override { return 0.5f; }
    BoundaryType getBoundaryType(Boundary) override { return OUTFLOW; }
};

// 256^2 f-wave workload: 15 m within 100 m of the centre of [0, 1000]^2,
// 10 m elsewhere, flat b = 0, outflow
class SWE_CentredDamBreakScenario : public SWE_Scenario {
public:
    float getBathymetry(float, float) override { return 0.f; }
    float getWaterHeight(float x, float y) override {
        return swe_scenario_detail::radius(x, y, 500.f, 500.f) < 100.f ? 15.f : 10.f;
    }
    float endSimulation() override { return 25.f; }
    BoundaryType getBoundaryType(Boundary) override { return OUTFLOW; }
    float getBoundaryPos(Boundary b) override { return swe_scenario_detail::box_pos(b, 1000.f); }
};

#endif
