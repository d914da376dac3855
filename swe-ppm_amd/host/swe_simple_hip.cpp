// swe_simple_hip.cpp -- single-block driver with swe_simple's structure
// (reference src/examples/swe_simple.cpp:56-274): scenario, checkpoints,
// and the loop setGhostLayer -> computeNumericalFluxes -> getMaxTimestep ->
// updateUnknowns -> t += dt, with the block's step on the GPU.
//
//   swe_simple_hip -t 25 -n 5 -x 256 -y 256 -o out [-w 1] [--writer vtk|netcdf]
//                  [--solver fwave] [--flavour u|s] [--scenario radial|centred|...]
//                  [-b bathymetry.nc [-d displacement.nc] [--host-init]] [--input f] [--dump f]
//
// -b / -d (--bathymetry-file / --displacement-file) select the file-driven
// scenario (the reference's ASAGI build, swe_simple.cpp:68-71, 109-113),
// sampled on the device unless --host-init asks for the host loop.
// --input starts from a state snapshot of one block instead (sizes from the
// snapshot, domain and boundary types from the scenario).  --warmup K runs K
// untimed steps first and starts again from the initial state; --lookahead 0
// turns off the look-ahead launches of the per-call API; --async-write has
// the checkpoints written by a writer thread from device snapshots.
//
// With -w 1 the state is written at t = 0 and at every checkpoint: VTK files
// <out>.<k>.vts (the default) or one netCDF file <out>.nc, the two writers
// the reference chooses between at build time (swe_simple.cpp:44-48, 172-185).
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>

#include "blocks/SWE_DimensionalSplittingHip.hh"
#include "scenarios/SWE_simple_scenarios.hh"
#include "swe_drivers_common.hh"
#include "tools/Float2DView.hh"
#include "writer/NetCdfWriter.hh"
#include "writer/VtkWriter.hh"

// --async-write: checkpoints are snapshotted on the device and written by a
// thread of their own from the snapshot's host arrays while the loop runs on
// (two snapshots in flight at most, swe_hip_snapshot).
class AsyncWriter {
public:
    AsyncWriter(Writer& w, swe_hip_block* blk, int nx, int ny) : w_(w), blk_(blk), nx_(nx), ny_(ny) {
        thread_ = std::thread([this] { loop(); });
    }
    ~AsyncWriter() {  // on an exception path: stop without throwing
        {
            std::lock_guard<std::mutex> g(mu_);
            done_ = true;
            cv_.notify_one();
        }
        if (thread_.joinable()) thread_.join();
    }
    void write(float t) {
        int slot = 0;
        if (swe_hip_snapshot(blk_, &slot) != SWE_HIP_OK) throw std::runtime_error(swe_hip_last_error());
        std::lock_guard<std::mutex> g(mu_);
        jobs_.push_back({slot, t});
        cv_.notify_one();
    }
    // waits for every queued frame; rethrows the writer thread's failure
    void finish() {
        {
            std::lock_guard<std::mutex> g(mu_);
            done_ = true;
            cv_.notify_one();
        }
        if (thread_.joinable()) thread_.join();
        if (!error_.empty()) {
            const std::string e = error_;
            error_.clear();
            throw std::runtime_error(e);
        }
    }

private:
    struct Job {
        int slot;
        float t;
    };
    void loop() {
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return done_ || !jobs_.empty(); });
                if (jobs_.empty()) return;
                j = jobs_.front();
                jobs_.pop_front();
            }
            const float *h, *hu, *hv;
            if (swe_hip_snapshot_wait(blk_, j.slot, &h, &hu, &hv) != SWE_HIP_OK) {
                error_ = swe_hip_last_error();
                return;
            }
            if (error_.empty()) {
                try {
                    w_.writeTimeStep(Float2DView(nx_ + 2, ny_ + 2, h), Float2DView(nx_ + 2, ny_ + 2, hu),
                                     Float2DView(nx_ + 2, ny_ + 2, hv), j.t);
                } catch (const std::exception& e) {
                    error_ = e.what();
                }
            }
            swe_hip_snapshot_release(blk_, j.slot);
        }
    }
    Writer& w_;
    swe_hip_block* blk_;
    int nx_, ny_;
    std::thread thread_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Job> jobs_;
    bool done_ = false;
    std::string error_;
};

// the scenario object and its device-sampling id (SWE_SCENARIO_*)
static std::unique_ptr<SWE_Scenario> make_scenario(const std::string& name, int* id) {
    if (name == "radial") return *id = SWE_SCENARIO_RADIAL_DAM_BREAK, std::unique_ptr<SWE_Scenario>(new SWE_RadialDamBreakScenario());
    if (name == "centred")
        return *id = SWE_SCENARIO_CENTRED_DAM_BREAK, std::unique_ptr<SWE_Scenario>(new SWE_CentredDamBreakScenario());
    if (name == "radial-bathymetry")
        return *id = SWE_SCENARIO_RADIAL_BATHYMETRY_DAM_BREAK,
               std::unique_ptr<SWE_Scenario>(new SWE_RadialBathymetryDamBreakScenario());
    if (name == "bathymetry")
        return *id = SWE_SCENARIO_BATHYMETRY_DAM_BREAK,
               std::unique_ptr<SWE_Scenario>(new SWE_BathymetryDamBreakScenario());
    if (name == "half-dry") return *id = SWE_SCENARIO_HALF_DOMAIN_DRY, std::unique_ptr<SWE_Scenario>(new SWE_HalfDomainDry());
    if (name == "splashing-pool")
        return *id = SWE_SCENARIO_SPLASHING_POOL, std::unique_ptr<SWE_Scenario>(new SWE_SplashingPoolScenario());
    if (name == "splashing-cone")
        return *id = SWE_SCENARIO_SPLASHING_CONE, std::unique_ptr<SWE_Scenario>(new SWE_SplashingConeScenario());
    if (name == "sea-at-rest")
        return *id = SWE_SCENARIO_SEA_AT_REST, std::unique_ptr<SWE_Scenario>(new SWE_SeaAtRestScenario());
    throw std::invalid_argument("unknown scenario " + name);
}

int main(int argc, char** argv) {
    try {
        // -b and -d are the reference's short flags for the input files here
        std::vector<std::string> words(argv, argv + argc);
        std::vector<char*> args;
        for (auto& w : words) {
            if (w == "-b") w = "--bathymetry-file";
            if (w == "-d") w = "--displacement-file";
            args.push_back(&w[0]);
        }
        DriverArgs a = parse_args((int)args.size(), args.data(), "radial");
        // --input: the initial state from a snapshot (one block) in the
        // scenario's domain and with its boundary types
        std::vector<BlockState> init;
        if (!a.input.empty()) {
            init = read_snapshot(a.input);
            if (init.size() != 1) throw std::invalid_argument("--input needs a snapshot of one block");
            a.nx = init[0].nx;
            a.ny = init[0].ny;
        }
        std::unique_ptr<SWE_Scenario> scenario;
        SWE_NetCdfScenario* fileScenario = nullptr;
        int scenarioId = -1;
        if (!a.bathymetry.empty()) {
            fileScenario = new SWE_NetCdfScenario(a.bathymetry, a.displacement);
            scenario.reset(fileScenario);
        } else {
            scenario = make_scenario(a.scenario, &scenarioId);
        }
        std::vector<float> cp = checkpoint_instants(a.duration, a.checkpoints);

        const int width = (int)(scenario->getBoundaryPos(BND_RIGHT) - scenario->getBoundaryPos(BND_LEFT));
        const int height = (int)(scenario->getBoundaryPos(BND_TOP) - scenario->getBoundaryPos(BND_BOTTOM));
        const float dx = (float)width / a.nx;
        const float dy = (float)height / a.ny;
        const float ox = scenario->getBoundaryPos(BND_LEFT);
        const float oy = scenario->getBoundaryPos(BND_BOTTOM);
        BoundaryType boundaries[4];
        for (int s = 0; s < 4; s++) boundaries[s] = scenario->getBoundaryType(static_cast<Boundary>(s));

        SWE_DimensionalSplittingHip sim(a.nx, a.ny, dx, dy, ox, oy, false, a.solver, a.flavour);
        if (!a.lookahead && swe_hip_set_lookahead(sim.handle(), 0) != SWE_HIP_OK)
            throw std::runtime_error(swe_hip_last_error());
        // sampled on the device unless --host-init asks for the host loop
        auto initialise = [&]() {
            if (!init.empty())
                load_snapshot(sim, init[0], boundaries);
            else if (fileScenario && !a.hostInit)
                sim.initScenarioOnDevice(*fileScenario, boundaries);
            else if (scenarioId >= 0 && !a.hostInit)
                sim.initScenarioOnDevice(scenarioId, boundaries);
            else
                sim.initScenario(*scenario, boundaries);
        };
        initialise();
        if (a.warmup > 0) {  // untimed steps (first launches load the kernels), then the initial state again
            for (long s = 0; s < a.warmup; s++) {
                sim.setGhostLayer();
                sim.computeNumericalFluxes();
                sim.updateUnknowns(sim.getMaxTimestep());
            }
            initialise();
        }
        if (!a.dump0.empty()) write_snapshot(a.dump0, {snapshot_of(sim)});

        std::unique_ptr<Writer> writer;
        if (a.write) {
            if (a.writer == "netcdf")
                writer.reset(new NetCdfWriter(a.out, sim.getBathymetry(), {{1, 1, 1, 1}}, a.nx, a.ny, dx, dy, ox, oy));
            else
                writer.reset(new VtkWriter(a.out, sim.getBathymetry(), {{1, 1, 1, 1}}, a.nx, a.ny, dx, dy, ox, oy));
        }
        std::unique_ptr<AsyncWriter> async;
        if (writer && a.asyncWrite) async.reset(new AsyncWriter(*writer, sim.handle(), a.nx, a.ny));
        auto write_frame = [&](float time) {
            if (async)
                async->write(time);
            else if (writer)
                writer->writeTimeStep(sim.getWaterHeight(), sim.getMomentumHorizontal(), sim.getMomentumVertical(),
                                      time);
        };
        write_frame(0.f);

        float t = 0.f;
        long iterations = 0;
        double wall = 0.0;
        bool stop = false;
        for (int i = 0; i < a.checkpoints && !stop; i++) {
            auto t0 = std::chrono::steady_clock::now();
            while (t < cp[i]) {
                sim.setGhostLayer();
                sim.computeNumericalFluxes();
                const float timestep = sim.getMaxTimestep();
                sim.updateUnknowns(timestep);
                t += timestep;
                iterations++;
                if (a.max_steps > 0 && iterations >= a.max_steps) {
                    stop = true;
                    break;
                }
            }
            // the last update is still queued: the clock stops once it is done
            swe_hip_synchronize(sim.handle());
            wall += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::printf("Write timestep (%fs)\n", t);
            write_frame(t);
        }
        if (async) async->finish();
        if (!a.dump.empty()) write_snapshot(a.dump, {snapshot_of(sim)});
        std::printf("iterations %ld simulated %.6f s wall %.4f s  %.2f MCell-updates/s\n", iterations, t, wall,
                    (double)a.nx * a.ny * iterations / wall / 1e6);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "swe_simple_hip: %s\n", e.what());
        return 1;
    }
}
