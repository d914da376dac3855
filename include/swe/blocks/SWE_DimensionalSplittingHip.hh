// SWE_DimensionalSplittingHip.hh -- drop-in SWE_Block subclass whose step
// runs on an MI355X through the C ABI of libswe_hip.so (include/swe_hip.h).
//
// Same driving interface as the reference's blocks:
//   setGhostLayer()           SWE_DimensionalSplittingMPIOverdecomp.cpp:269-304
//   receiveGhostLayer()       SWE_DimensionalSplittingMPIOverdecomp.cpp:400-479
//   computeNumericalFluxes()  SWE_DimensionalSplittingMpi.cpp:383-476 (flavour U)
//                             SWE_DimensionalSplitting.cpp:190-264   (flavour S)
//   getMaxTimestep()          SWE_Block.hh:594-597 (the maxTimestep member)
//   updateUnknowns(float dt)  SWE_DimensionalSplittingMpi.cpp:484-518 / SWE_DimensionalSplitting.cpp:272-295
// and the constructor arguments of the reference blocks
// (nx, ny, dx, dy, originX, originY[, localTimestepping]).
//
// The device holds the authoritative state.  Host arrays (h, hu, hv, b) are
// refreshed lazily: every accessor of the base class (getWaterHeight() ...)
// triggers the download of a stale field; direct writes to the host arrays
// must be followed by synchAfterWrite() (initScenario does this itself).
// In-rank neighbours (CONNECT_WITHIN_RANK) exchange ghost layers device to
// device; with local time stepping the ghost lines go through the host,
// where the base class's state machine interpolates them.
//
// Errors of the C ABI are raised as std::runtime_error (the reference's
// methods return void).
#ifndef SWE_HIP_DIMENSIONAL_SPLITTING_HIP_HH
#define SWE_HIP_DIMENSIONAL_SPLITTING_HIP_HH

#include <array>
#include <stdexcept>
#include <string>

#include "../../swe_hip.h"
#include "blocks/SWE_Block.hh"
#include "scenarios/SWE_NetCdfScenario.hh"
#include "tools/Float2DNative.hh"

// The synch* hooks are virtual in this library's SWE_Block; compiled against
// the reference's SWE_Block (which only documents them) they are plain
// methods that a driver calls before host reads.
#ifdef SWE_HIP_BLOCK_MIRROR
#define SWE_HIP_HOOK override
#else
#define SWE_HIP_HOOK
#endif

class SWE_DimensionalSplittingHip : public SWE_Block<Float2DNative> {
public:
    SWE_DimensionalSplittingHip(int cellCountHorizontal, int cellCountVertical, float cellSizeHorizontal,
                                float cellSizeVertical, float originX, float originY, bool localTimestepping = false,
                                int solver = SWE_SOLVER_AUGRIE, int flavour = SWE_FLAVOUR_U);
    ~SWE_DimensionalSplittingHip() override;
    SWE_DimensionalSplittingHip(const SWE_DimensionalSplittingHip&) = delete;
    SWE_DimensionalSplittingHip& operator=(const SWE_DimensionalSplittingHip&) = delete;

    // step interface
    void setGhostLayer();
    void receiveGhostLayer();
    void computeNumericalFluxes() override;
    void updateUnknowns(float dt) override;

    // scenario / boundaries (host side of the base class, then upload)
    void initScenario(SWE_Scenario& scenario, BoundaryType boundaries[]) override;
    void setBoundaryType(Boundary boundary, BoundaryType type) override;
    // initScenario of the file-driven scenario with the sampling done on the
    // device (swe_hip_init_grid_scenario); the same arrays as initScenario
    void initScenarioOnDevice(const SWE_NetCdfScenario& scenario, BoundaryType boundaries[]);
    // initScenario of the analytic scenario `scenario` (SWE_SCENARIO_* of
    // swe_hip.h, the classes of SWE_simple_scenarios.hh) sampled on the device
    // (swe_hip_init_scenario); the same arrays as initScenario with that
    // scenario object
    void initScenarioOnDevice(int scenario, BoundaryType boundaries[]);

    // in-rank neighbours (SWE_DimensionalSplittingMPIOverdecomp.cpp:119-139)
    void connectLocalNeighbours(std::array<SWE_DimensionalSplittingHip*, 4> neighbourBlocks);
    // in-rank bathymetry exchange (SWE_DimensionalSplittingMPIOverdecomp.cpp:208-235)
    void sendBathymetry();

    // lazy synchronisation
    void synchAfterWrite() SWE_HIP_HOOK;
    void synchWaterHeightAfterWrite() SWE_HIP_HOOK { synchAfterWrite(); }
    void synchDischargeAfterWrite() SWE_HIP_HOOK { synchAfterWrite(); }
    void synchBathymetryAfterWrite() SWE_HIP_HOOK { synchAfterWrite(); }
    void synchGhostLayerAfterWrite() SWE_HIP_HOOK;
    void synchBeforeRead() SWE_HIP_HOOK;
    void synchWaterHeightBeforeRead() SWE_HIP_HOOK { synchBeforeRead(); }
    void synchDischargeBeforeRead() SWE_HIP_HOOK { synchBeforeRead(); }
    void synchBathymetryBeforeRead() SWE_HIP_HOOK { synchBeforeRead(); }
    void synchCopyLayerBeforeRead() SWE_HIP_HOOK { synchBeforeRead(); }

    swe_hip_block* handle() { return dev_; }
    int solver() const { return solver_; }
    int flavour() const { return flavour_; }

    SWE_DimensionalSplittingHip* left = nullptr;
    SWE_DimensionalSplittingHip* right = nullptr;
    SWE_DimensionalSplittingHip* bottom = nullptr;
    SWE_DimensionalSplittingHip* top = nullptr;

private:
    void check(int rc, const char* what) const;
    SWE_DimensionalSplittingHip* neighbour(int side) const;
    void receiveLayerToBuffer(int side, SWE_DimensionalSplittingHip* from);

    swe_hip_block* dev_ = nullptr;
    int solver_;
    int flavour_;
    bool hostStale_ = false;  // the device holds newer h, hu, hv
};

#endif
