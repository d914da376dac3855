// swe_overdecomp_hip.cpp -- many blocks in one process, neighbours connected
// in memory (CONNECT_WITHIN_RANK), global or local time stepping: the loop
// structure of the reference's overdecomposed driver
// (src/examples/swe_mpi_overdecomp.cpp:147-343) with every block's step on
// the GPU and no MPI (one rank).
//
//   swe_overdecomp_hip -t 80 -n 20 -x 2048 -y 2048 -b 28 [-l 1] [--rows R]
//                      [--solver augrie] [--input blocks.bin] [--dump out.bin]
//                      [-w 1 -o base [--writer netcdf|vtk]] [--batched]
//
// --batched hands the whole loop to the batched engine (swe_hip_batch.h):
// the blocks, scenario, checkpoints and output stay as below, but ghost
// exchange, the LTS state machine, dt and the updates of every block run on
// the device, one call per checkpoint interval (LTS) or per step (global dt).
// The results are bit-identical to the block-by-block loop.
//
// With -w 1 every block writes its state at t = 0 and at every checkpoint to
// its own file <base>_<column><row>, netCDF by default, as the reference's
// blocks do (SWE_DimensionalSplittingMPIOverdecomp.cpp:98-116,
// swe_mpi_overdecomp.cpp:254-258, 348-352; file names of
// tools/help.hh:100-105).
//
// Block k sits at column k / rows, row k % rows; neighbours follow
// getNeighbours (swe_mpi_overdecomp.cpp:41-49).  Without --input the blocks
// are an even split of the scenario (remainder cells to the last row/column,
// swe_mpi_overdecomp.cpp:164-178); with --input the block sizes and initial
// states come from a snapshot file.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>

#include "../../include/swe_hip_batch.h"
#include "blocks/SWE_DimensionalSplittingHip.hh"
#include "scenarios/SWE_simple_scenarios.hh"
#include "swe_drivers_common.hh"
#include "writer/NetCdfWriter.hh"
#include "writer/VtkWriter.hh"

namespace {

std::array<int, 4> get_neighbours(int px, int py, int cols, int rows, int k) {
    std::array<int, 4> n;
    n[BND_LEFT] = (px > 0) ? k - rows : -1;
    n[BND_RIGHT] = (px < cols - 1) ? k + rows : -1;
    n[BND_BOTTOM] = (py > 0) ? k - 1 : -1;
    n[BND_TOP] = (py < rows - 1) ? k + 1 : -1;
    return n;
}

// initial state copied from a snapshot instead of a scenario
void batch_check(int rc, const char* what) {
    if (rc != SWE_HIP_OK) throw std::runtime_error(std::string(what) + ": " + swe_hip_last_error());
}

// The driver loop below, run by the batched engine.  Host arrays of the
// blocks are the upload source and the download target (writers and dumps
// read them); the blocks' own device state is not used.
void run_batched(std::vector<std::unique_ptr<SWE_DimensionalSplittingHip>>& blocks, const std::vector<int>& bnd,
                 const std::vector<int>& nbr, const DriverArgs& a, bool lts, const std::vector<float>& cp,
                 std::vector<std::unique_ptr<Writer>>& writers, std::vector<long>& updates, float& t,
                 long& iterations, long& macro, double& wall) {
    const int n = (int)blocks.size();
    std::vector<int> nx(n), ny(n);
    std::vector<float> dx(n), dy(n);
    std::vector<size_t> off(n + 1, 0);
    for (int k = 0; k < n; k++) {
        blocks[k]->synchBeforeRead();  // host arrays current; nothing marks them stale again
        nx[k] = blocks[k]->nx;
        ny[k] = blocks[k]->ny;
        dx[k] = blocks[k]->dx;
        dy[k] = blocks[k]->dy;
        off[k + 1] = off[k] + (size_t)(nx[k] + 2) * (ny[k] + 2);
    }
    std::vector<float> f[4];
    for (auto& v : f) v.resize(off[n]);
    auto gather = [&]() {
        for (int k = 0; k < n; k++) {
            const float* p[4] = {blocks[k]->h.getRawPointer(), blocks[k]->hu.getRawPointer(),
                                 blocks[k]->hv.getRawPointer(), blocks[k]->b.getRawPointer()};
            for (int q = 0; q < 4; q++) std::copy(p[q], p[q] + (off[k + 1] - off[k]), f[q].begin() + off[k]);
        }
    };
    auto scatter = [&]() {
        for (int k = 0; k < n; k++) {
            float* p[3] = {blocks[k]->h.getRawPointer(), blocks[k]->hu.getRawPointer(), blocks[k]->hv.getRawPointer()};
            for (int q = 0; q < 3; q++) std::copy(f[q].begin() + off[k], f[q].begin() + off[k + 1], p[q]);
        }
    };
    swe_hip_batch* bt = nullptr;
    batch_check(swe_hip_batch_create(n, nx.data(), ny.data(), dx.data(), dy.data(), bnd.data(), nbr.data(), a.solver,
                                      lts ? SWE_BATCH_LTS : SWE_BATCH_GLOBAL, lts ? a.lts : 1.f, a.duration, &bt),
                "swe_hip_batch_create");
    std::unique_ptr<swe_hip_batch, int (*)(swe_hip_batch*)> guard(bt, swe_hip_batch_destroy);
    gather();
    batch_check(swe_hip_batch_upload(bt, f[0].data(), f[1].data(), f[2].data(), f[3].data()), "swe_hip_batch_upload");
    bool stop = false;
    for (size_t i = 0; i < cp.size() && !stop; i++) {
        // the units this checkpoint interval needs, with the loop's float time
        auto t0 = std::chrono::steady_clock::now();
        if (lts) {
            int units = 0;
            float tt = t;
            while (tt < cp[i] && !(a.max_steps > 0 && macro + units >= a.max_steps)) {
                tt += a.lts;
                units++;
            }
            if (a.max_steps > 0 && macro + units >= a.max_steps) stop = true;
            if (units > 0) {
                std::vector<int> its(units);
                batch_check(swe_hip_batch_run(bt, units, nullptr, its.data(), nullptr), "swe_hip_batch_run");
                for (int v : its) iterations += v;
            }
            t = tt;
            macro += units;
        } else {
            while (t < cp[i]) {
                float dt = 0.f;
                batch_check(swe_hip_batch_run(bt, 1, &dt, nullptr, nullptr), "swe_hip_batch_run");
                t += dt;
                iterations++;
                macro++;
                if (a.max_steps > 0 && macro >= a.max_steps) {
                    stop = true;
                    break;
                }
            }
        }
        wall += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        batch_check(swe_hip_batch_download(bt, f[0].data(), f[1].data(), f[2].data(), nullptr),
                    "swe_hip_batch_download");
        scatter();
        std::printf("Write timestep (%fs)\n", t);
        for (int k = 0; k < (int)writers.size(); k++)
            writers[k]->writeTimeStep(blocks[k]->h, blocks[k]->hu, blocks[k]->hv, t);
    }
    std::vector<int64_t> upd(n);
    batch_check(swe_hip_batch_stats(bt, upd.data(), nullptr), "swe_hip_batch_stats");
    for (int k = 0; k < n; k++) updates[k] = (long)upd[k];
}

}  // namespace

int main(int argc, char** argv) {
    try {
        // --rows and --batched are read here (not reference flags)
        int rowsArg = -1;
        bool batched = false;
        std::vector<char*> rest;
        for (int i = 0; i < argc; i++) {
            if (std::string(argv[i]) == "--rows" && i + 1 < argc) {
                rowsArg = std::atoi(argv[++i]);
                continue;
            }
            if (std::string(argv[i]) == "--batched") {
                batched = true;
                continue;
            }
            rest.push_back(argv[i]);
        }
        DriverArgs a = parse_args((int)rest.size(), rest.data(), "radial-bathymetry");
        const bool lts = a.lts > 0.f;
        std::unique_ptr<SWE_Scenario> scenario(new SWE_RadialBathymetryDamBreakScenario());
        std::vector<float> cp = checkpoint_instants(a.duration, a.checkpoints);

        std::vector<BlockState> init;
        if (!a.input.empty()) init = read_snapshot(a.input);
        const int total = init.empty() ? a.blocks : (int)init.size();
        int rows = rowsArg;
        if (rows <= 0) {
            rows = (int)std::sqrt((double)total);
            while (total % rows != 0) rows--;
        }
        if (total % rows != 0) throw std::invalid_argument("block count must be a multiple of --rows");
        const int cols = total / rows;

        const int width = (int)(scenario->getBoundaryPos(BND_RIGHT) - scenario->getBoundaryPos(BND_LEFT));
        const int height = (int)(scenario->getBoundaryPos(BND_TOP) - scenario->getBoundaryPos(BND_BOTTOM));
        float dx = (float)width / a.nx;
        float dy = (float)height / a.ny;
        if (!init.empty()) {
            int nxTot = 0, nyTot = 0;
            for (int c = 0; c < cols; c++) nxTot += init[c * rows].nx;
            for (int r = 0; r < rows; r++) nyTot += init[r].ny;
            dx = (float)width / nxTot;
            dy = (float)height / nyTot;
        }

        std::vector<std::unique_ptr<SWE_DimensionalSplittingHip>> blocks(total);
        const int nxBlock = a.nx / cols, nyBlock = a.ny / rows;
        for (int k = 0; k < total; k++) {
            const int px = k / rows, py = k % rows;
            int nxl, nyl;
            float ox, oy;
            if (init.empty()) {
                nxl = (px < cols - 1) ? nxBlock : a.nx - (cols - 1) * nxBlock;
                nyl = (py < rows - 1) ? nyBlock : a.ny - (rows - 1) * nyBlock;
                ox = scenario->getBoundaryPos(BND_LEFT) + px * dx * nxBlock;
                oy = scenario->getBoundaryPos(BND_BOTTOM) + py * dy * nyBlock;
            } else {
                nxl = init[k].nx;
                nyl = init[k].ny;
                int xo = 0, yo = 0;
                for (int c = 0; c < px; c++) xo += init[c * rows].nx;
                for (int r = 0; r < py; r++) yo += init[r].ny;
                ox = xo * dx;
                oy = yo * dy;
            }
            blocks[k].reset(new SWE_DimensionalSplittingHip(nxl, nyl, dx, dy, ox, oy, lts, a.solver, a.flavour));
        }
        std::vector<int> allBnd(4 * total), allNbr(4 * total);
        for (int k = 0; k < total; k++) {
            const int px = k / rows, py = k % rows;
            std::array<int, 4> nb = get_neighbours(px, py, cols, rows, k);
            BoundaryType bnd[4];
            std::array<SWE_DimensionalSplittingHip*, 4> nbb{};
            for (int s = 0; s < 4; s++) {
                if (nb[s] >= 0) {
                    bnd[s] = CONNECT_WITHIN_RANK;
                    nbb[s] = blocks[nb[s]].get();
                } else {
                    bnd[s] = scenario->getBoundaryType(static_cast<Boundary>(s));
                }
            }
            for (int s = 0; s < 4; s++) {
                allBnd[4 * k + s] = bnd[s];
                allNbr[4 * k + s] = nb[s];
            }
            if (init.empty() && !a.hostInit)
                blocks[k]->initScenarioOnDevice(SWE_SCENARIO_RADIAL_BATHYMETRY_DAM_BREAK, bnd);
            else if (init.empty())
                blocks[k]->initScenario(*scenario, bnd);
            else
                load_snapshot(*blocks[k], init[k], bnd);
            blocks[k]->connectLocalNeighbours(nbb);
            blocks[k]->setRank(k);
            blocks[k]->setDuration(a.duration);
        }
        for (auto& b : blocks) b->sendBathymetry();
        if (!a.dump0.empty()) {
            std::vector<BlockState> s0;
            for (auto& b : blocks) s0.push_back(snapshot_of(*b));
            write_snapshot(a.dump0, s0);
        }

        std::vector<std::unique_ptr<Writer>> writers;
        auto write_all = [&](float time) {
            for (int k = 0; k < (int)writers.size(); k++)
                writers[k]->writeTimeStep(blocks[k]->getWaterHeight(), blocks[k]->getMomentumHorizontal(),
                                          blocks[k]->getMomentumVertical(), time);
        };
        if (a.write) {
            for (int k = 0; k < total; k++) {
                const int px = k / rows, py = k % rows;
                const std::string name = a.out + "_" + std::to_string(px) + std::to_string(py);
                SWE_DimensionalSplittingHip& b = *blocks[k];
                if (a.writer == "vtk")
                    writers.emplace_back(new VtkWriter(name, b.getBathymetry(), {{1, 1, 1, 1}}, b.nx, b.ny, b.dx,
                                                       b.dy, (int)(b.originX / b.dx), (int)(b.originY / b.dy)));
                else
                    writers.emplace_back(new NetCdfWriter(name, b.getBathymetry(), {{1, 1, 1, 1}}, b.nx, b.ny, b.dx,
                                                          b.dy, b.originX, b.originY));
            }
            write_all(0.f);
        }

        float maxLocalTimestep = 0.f;
        if (lts) {
            // swe_mpi_overdecomp.cpp:235-245: the CFL estimate is computed and
            // then replaced by the command-line macro step
            for (auto& b : blocks) b->computeMaxTimestep(0.01f, 0.4f);
            maxLocalTimestep = a.lts;
            for (auto& b : blocks) b->setMaxLocalTimestep(maxLocalTimestep);
        }

        std::vector<long> updates(total, 0);
        float t = 0.f, timestep = 0.f;
        long iterations = 0, macro = 0;
        double wall = 0.0;
        bool stop = false;
        if (batched) {
            if (a.flavour != SWE_FLAVOUR_U) throw std::invalid_argument("--batched runs flavour U");
            run_batched(blocks, allBnd, allNbr, a, lts, cp, writers, updates, t, iterations, macro, wall);
            stop = true;
        }
        for (int i = 0; i < a.checkpoints && !stop; i++) {
            auto t0 = std::chrono::steady_clock::now();
            while (t < cp[i]) {
                bool synchronized = true;
                do {
#pragma omp parallel for schedule(dynamic, 1)
                    for (int k = 0; k < total; k++) blocks[k]->setGhostLayer();
#pragma omp parallel for schedule(dynamic, 1)
                    for (int k = 0; k < total; k++) blocks[k]->receiveGhostLayer();
#pragma omp parallel for schedule(dynamic, 1)
                    for (int k = 0; k < total; k++) blocks[k]->computeNumericalFluxes();
                    if (!lts) {
                        timestep = blocks[0]->maxTimestep;
                        for (auto& b : blocks) timestep = std::min(timestep, b->maxTimestep);
                        for (auto& b : blocks) b->maxTimestep = timestep;
                    } else {
                        for (auto& b : blocks)
                            if (b->allGhostlayersInSync()) b->maxTimestep = b->getRoundTimestep(b->maxTimestep);
                    }
#pragma omp parallel for schedule(dynamic, 1)
                    for (int k = 0; k < total; k++) {
                        if (blocks[k]->allGhostlayersInSync()) updates[k]++;
                        blocks[k]->updateUnknowns(timestep);
                    }
                    iterations++;
                    if (lts) {
                        synchronized = true;
                        for (auto& b : blocks)
                            if (!b->hasMaxLocalTimestep()) synchronized = false;
                    }
                    if (iterations > 100000000L) throw std::runtime_error("LTS made no progress");
                } while (lts && !synchronized);
                t += lts ? maxLocalTimestep : timestep;
                macro++;
                if (lts)
                    for (auto& b : blocks) b->resetStepSizeCounter();
                if (a.max_steps > 0 && macro >= a.max_steps) {
                    stop = true;
                    break;
                }
            }
            // the last updates are still pending or queued: the clock stops
            // once every block has finished them
            for (auto& b : blocks) swe_hip_synchronize(b->handle());
            wall += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::printf("Write timestep (%fs)\n", t);
            write_all(t);
        }
        long cellUpdates = 0;
        for (int k = 0; k < total; k++) cellUpdates += updates[k] * (long)blocks[k]->nx * blocks[k]->ny;
        if (!a.dump.empty()) {
            std::vector<BlockState> s;
            for (auto& b : blocks) s.push_back(snapshot_of(*b));
            write_snapshot(a.dump, s);
        }
        std::printf("blocks %d (%d x %d) macro-steps %ld iterations %ld simulated %.6f s wall %.3f s "
                    "%.2f MCell-updates/s\n",
                    total, cols, rows, macro, iterations, t, wall, cellUpdates / wall / 1e6);
        std::printf("updates");
        for (long u : updates) std::printf(" %ld", u);
        std::printf("\n");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "swe_overdecomp_hip: %s\n", e.what());
        return 1;
    }
}
