// SWE_DimensionalSplittingHip.cpp -- the drop-in block on top of the C ABI.
#include "blocks/SWE_DimensionalSplittingHip.hh"

#include <cmath>
#include <vector>

SWE_DimensionalSplittingHip::SWE_DimensionalSplittingHip(int nx_, int ny_, float dx_, float dy_, float ox, float oy,
                                                         bool lts, int solver, int flavour)
    : SWE_Block(nx_, ny_, dx_, dy_, ox, oy, lts), solver_(solver), flavour_(flavour) {
    check(swe_hip_create(nx_, ny_, dx_, dy_, ox, oy, solver, flavour, &dev_), "swe_hip_create");
    for (int s = 0; s < 4; s++) check(swe_hip_set_boundary(dev_, s, boundaryType[s]), "swe_hip_set_boundary");
}

SWE_DimensionalSplittingHip::~SWE_DimensionalSplittingHip() {
    if (dev_) swe_hip_destroy(dev_);
}

void SWE_DimensionalSplittingHip::check(int rc, const char* what) const {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + swe_hip_last_error());
}

SWE_DimensionalSplittingHip* SWE_DimensionalSplittingHip::neighbour(int side) const {
    switch (side) {
        case BND_LEFT:
            return left;
        case BND_RIGHT:
            return right;
        case BND_BOTTOM:
            return bottom;
        default:
            return top;
    }
}

void SWE_DimensionalSplittingHip::initScenario(SWE_Scenario& scenario, BoundaryType boundaries[]) {
    for (int s = 0; s < 4; s++) setBoundaryType(static_cast<Boundary>(s), boundaries[s]);
    hostStale_ = false;
    SWE_Block::initScenario(scenario, boundaries);  // samples and applies the boundaries on the host
    synchAfterWrite();
}

void SWE_DimensionalSplittingHip::initScenarioOnDevice(const SWE_NetCdfScenario& scenario, BoundaryType boundaries[]) {
    for (int s = 0; s < 4; s++) setBoundaryType(static_cast<Boundary>(s), boundaries[s]);
    auto grid = [](const GriddedField& g) {
        swe_hip_grid d;
        d.nx = g.nx;
        d.ny = g.ny;
        d.x0 = g.x0;
        d.y0 = g.y0;
        d.dx = g.dx;
        d.dy = g.dy;
        d.values = g.values.data();
        return d;
    };
    const swe_hip_grid gb = grid(scenario.bathymetryGrid());
    swe_hip_grid gd{};
    const swe_hip_grid* pd = nullptr;
    if (scenario.displacementGrid()) {
        gd = grid(*scenario.displacementGrid());
        pd = &gd;
    }
    check(swe_hip_init_grid_scenario(dev_, &gb, pd), "swe_hip_init_grid_scenario");
    hostStale_ = true;  // the host arrays are filled on the next read
}

void SWE_DimensionalSplittingHip::initScenarioOnDevice(int scenario, BoundaryType boundaries[]) {
    for (int s = 0; s < 4; s++) setBoundaryType(static_cast<Boundary>(s), boundaries[s]);
    check(swe_hip_init_scenario(dev_, scenario), "swe_hip_init_scenario");
    hostStale_ = true;  // the host arrays are filled on the next read
}

void SWE_DimensionalSplittingHip::setBoundaryType(Boundary boundary, BoundaryType type) {
    boundaryType[boundary] = type;
    // CONNECT and CONNECT_WITHIN_RANK sides are filled by neighbour exchange;
    // for the device boundary kernel they behave as PASSIVE
    check(swe_hip_set_boundary(dev_, boundary, type), "swe_hip_set_boundary");
}

void SWE_DimensionalSplittingHip::connectLocalNeighbours(std::array<SWE_DimensionalSplittingHip*, 4> nb) {
    for (int s = 0; s < 4; s++) {
        if (boundaryType[s] != CONNECT_WITHIN_RANK) continue;
        switch (s) {
            case BND_LEFT:
                left = nb[s];
                break;
            case BND_RIGHT:
                right = nb[s];
                break;
            case BND_BOTTOM:
                bottom = nb[s];
                break;
            default:
                top = nb[s];
                break;
        }
    }
}

void SWE_DimensionalSplittingHip::sendBathymetry() {
    for (int s = 0; s < 4; s++) {
        if (boundaryType[s] != CONNECT_WITHIN_RANK) continue;
        check(swe_hip_copy_layer(dev_, s, neighbour(s)->dev_, SWE_FIELD_B), "swe_hip_copy_layer(b)");
        hostStale_ = true;
    }
}

// The neighbour's copy layer goes into this block's receive buffer on the
// host (LTS) -- the buffer is where the reference's MPI receives land.
void SWE_DimensionalSplittingHip::receiveLayerToBuffer(int side, SWE_DimensionalSplittingHip* from) {
    const int n = (side == BND_LEFT || side == BND_RIGHT) ? ny : nx;
    std::vector<float> line(3 * (size_t)n);
    check(swe_hip_read_layer(from->dev_, side ^ 1, 0, SWE_FIELD_H | SWE_FIELD_HU | SWE_FIELD_HV, line.data()),
          "swe_hip_read_layer");
    for (int i = 1; i <= n; i++) {
        const int x = side == BND_LEFT ? 0 : side == BND_RIGHT ? nx + 1 : i;
        const int y = side == BND_BOTTOM ? 0 : side == BND_TOP ? ny + 1 : i;
        bufferH[x][y] = line[i - 1];
        bufferHu[x][y] = line[n + i - 1];
        bufferHv[x][y] = line[2 * (size_t)n + i - 1];
    }
}

void SWE_DimensionalSplittingHip::setGhostLayer() {
    check(swe_hip_apply_bc(dev_), "swe_hip_apply_bc");
    hostStale_ = true;
    const int order[4] = {BND_RIGHT, BND_LEFT, BND_TOP, BND_BOTTOM};  // as the reference
    for (int s : order) {
        if (boundaryType[s] != CONNECT_WITHIN_RANK || !isReceivable(static_cast<Boundary>(s))) continue;
        SWE_DimensionalSplittingHip* nb = neighbour(s);
        borderTimestep[s] = nb->getTotalLocalTimestep();
        if (localTimestepping)
            receiveLayerToBuffer(s, nb);
        else
            check(swe_hip_copy_layer(dev_, s, nb->dev_, SWE_FIELD_H | SWE_FIELD_HU | SWE_FIELD_HV),
                  "swe_hip_copy_layer");
    }
}

void SWE_DimensionalSplittingHip::receiveGhostLayer() {
    // no remote ranks on one GPU: only the LTS state machine remains
    checkAllGhostlayers();
#ifndef SWE_HIP_BLOCK_MIRROR
    // the reference's SWE_Block has no hook after checkAllGhostlayers
    if (localTimestepping) synchGhostLayerAfterWrite();
#endif
}

// Connected ghost lines changed on the host (copied or interpolated by the
// LTS state machine): send them to the device.
void SWE_DimensionalSplittingHip::synchGhostLayerAfterWrite() {
    if (hostStale_) {
        // bring the rest of the host mirror up to date first, keeping the
        // ghost lines just written on the host
        std::vector<std::vector<float>> keep;
        for (int s = 0; s < 4; s++) {
            if (boundaryType[s] != CONNECT && boundaryType[s] != CONNECT_WITHIN_RANK) continue;
            const int n = (s == BND_LEFT || s == BND_RIGHT) ? ny : nx;
            std::vector<float> line(3 * (size_t)n);
            for (int i = 1; i <= n; i++) {
                const int x = s == BND_LEFT ? 0 : s == BND_RIGHT ? nx + 1 : i;
                const int y = s == BND_BOTTOM ? 0 : s == BND_TOP ? ny + 1 : i;
                line[i - 1] = h[x][y];
                line[n + i - 1] = hu[x][y];
                line[2 * (size_t)n + i - 1] = hv[x][y];
            }
            check(swe_hip_write_layer(dev_, s, 1, SWE_FIELD_H | SWE_FIELD_HU | SWE_FIELD_HV, line.data()),
                  "swe_hip_write_layer");
        }
        return;
    }
    synchAfterWrite();
}

void SWE_DimensionalSplittingHip::computeNumericalFluxes() {
    if (!allGhostlayersInSync()) return;
    float dt = 0.f;
    check(swe_hip_fluxes(dev_, &dt), "swe_hip_fluxes");
    maxTimestep = dt;
}

void SWE_DimensionalSplittingHip::updateUnknowns(float dt) {
    if (!allGhostlayersInSync()) return;
    // flavour U uses maxTimestep (which a driver may have lowered to the
    // global minimum), as SWE_DimensionalSplittingMpi.cpp:487 does; flavour
    // S requires dt == maxTimestep (SWE_DimensionalSplitting.cpp:278)
    if (flavour_ == SWE_FLAVOUR_S && std::abs(dt - maxTimestep) >= 0.00001f)
        throw std::runtime_error("updateUnknowns: flavour S needs dt == maxTimestep");
    check(swe_hip_update(dev_, maxTimestep), "swe_hip_update");
    hostStale_ = true;
}

void SWE_DimensionalSplittingHip::synchAfterWrite() {
    check(swe_hip_upload(dev_, h.getRawPointer(), hu.getRawPointer(), hv.getRawPointer(), b.getRawPointer()),
          "swe_hip_upload");
    hostStale_ = false;
}

void SWE_DimensionalSplittingHip::synchBeforeRead() {
    if (!hostStale_) return;
    check(swe_hip_download(dev_, h.getRawPointer(), hu.getRawPointer(), hv.getRawPointer(), b.getRawPointer()),
          "swe_hip_download");
    hostStale_ = false;
}
