// swe_drivers_common.hh -- small shared pieces of the example drivers:
// command-line parsing with the reference's flag names
// (src/tools/args.hh; swe_simple.cpp:72-76, swe_mpi_overdecomp.cpp:63-72),
// solver names, checkpoint instants and binary state snapshots.
#ifndef SWE_HIP_DRIVERS_COMMON_HH
#define SWE_HIP_DRIVERS_COMMON_HH

#include <getopt.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/swe_hip.h"
#include "types/Boundary.hh"

struct DriverArgs {
    float duration = 1.f;    // -t
    int checkpoints = 1;     // -n
    int nx = 256, ny = 256;  // -x -y
    std::string out = "swe";  // -o
    int blocks = 1;          // -b
    float lts = 0.f;         // -l (macro step; 0 = global time stepping)
    int write = 0;           // -w
    int solver = SWE_SOLVER_AUGRIE;
    int flavour = SWE_FLAVOUR_U;
    std::string scenario;
    std::string input;   // initial block states (binary snapshot) instead of a scenario
    std::string dump;    // final states
    std::string dump0;   // initial states
    long max_steps = -1;  // stop after this many steps (global stepping)
    std::string writer;   // --writer vtk|netcdf ("" = the driver's default)
    std::string bathymetry;    // --bathymetry-file: file-driven scenario
    std::string displacement;  // --displacement-file
    bool hostInit = false;     // --host-init: sample the scenario on the host (initScenario), not the device
    long warmup = 0;           // --warmup: untimed steps, after which the initial state is loaded again
    int lookahead = 1;         // --lookahead 0: no look-ahead launches (swe_hip_set_lookahead)
    bool asyncWrite = false;   // --async-write: checkpoints written from snapshots by a writer thread
};

inline int solver_from_name(const std::string& s) {
    if (s == "hlle") return SWE_SOLVER_HLLE;
    if (s == "fwave") return SWE_SOLVER_FWAVE;
    if (s == "augrie") return SWE_SOLVER_AUGRIE;
    if (s == "hybrid") return SWE_SOLVER_HYBRID;
    throw std::invalid_argument("unknown solver " + s);
}

inline DriverArgs parse_args(int argc, char** argv, const char* default_scenario) {
    DriverArgs a;
    a.scenario = default_scenario;
    static option opts[] = {{"simulation-duration", required_argument, nullptr, 't'},
                            {"checkpoint-count", required_argument, nullptr, 'n'},
                            {"resolution-horizontal", required_argument, nullptr, 'x'},
                            {"resolution-vertical", required_argument, nullptr, 'y'},
                            {"output-basepath", required_argument, nullptr, 'o'},
                            {"blocks", required_argument, nullptr, 'b'},
                            {"local-timestepping", required_argument, nullptr, 'l'},
                            {"write", required_argument, nullptr, 'w'},
                            {"solver", required_argument, nullptr, 1},
                            {"flavour", required_argument, nullptr, 2},
                            {"scenario", required_argument, nullptr, 3},
                            {"input", required_argument, nullptr, 4},
                            {"dump", required_argument, nullptr, 5},
                            {"dump-initial", required_argument, nullptr, 6},
                            {"max-steps", required_argument, nullptr, 7},
                            {"writer", required_argument, nullptr, 8},
                            {"bathymetry-file", required_argument, nullptr, 9},
                            {"displacement-file", required_argument, nullptr, 10},
                            {"host-init", no_argument, nullptr, 11},
                            {"warmup", required_argument, nullptr, 12},
                            {"lookahead", required_argument, nullptr, 13},
                            {"async-write", no_argument, nullptr, 14},
                            {nullptr, 0, nullptr, 0}};
    int c;
    while ((c = getopt_long(argc, argv, "t:n:x:y:o:b:l:w:", opts, nullptr)) != -1) {
        switch (c) {
            case 't':
                a.duration = std::strtof(optarg, nullptr);
                break;
            case 'n':
                a.checkpoints = std::atoi(optarg);
                break;
            case 'x':
                a.nx = std::atoi(optarg);
                break;
            case 'y':
                a.ny = std::atoi(optarg);
                break;
            case 'o':
                a.out = optarg;
                break;
            case 'b':
                a.blocks = std::atoi(optarg);
                break;
            case 'l':
                a.lts = std::strtof(optarg, nullptr);
                break;
            case 'w':
                a.write = std::atoi(optarg);
                break;
            case 1:
                a.solver = solver_from_name(optarg);
                break;
            case 2:
                a.flavour = (std::string(optarg) == "s") ? SWE_FLAVOUR_S : SWE_FLAVOUR_U;
                break;
            case 3:
                a.scenario = optarg;
                break;
            case 4:
                a.input = optarg;
                break;
            case 5:
                a.dump = optarg;
                break;
            case 6:
                a.dump0 = optarg;
                break;
            case 7:
                a.max_steps = std::atol(optarg);
                break;
            case 8:
                a.writer = optarg;
                if (a.writer != "vtk" && a.writer != "netcdf") throw std::invalid_argument("--writer: vtk or netcdf");
                break;
            case 9:
                a.bathymetry = optarg;
                break;
            case 10:
                a.displacement = optarg;
                break;
            case 11:
                a.hostInit = true;
                break;
            case 12:
                a.warmup = std::atol(optarg);
                break;
            case 13:
                a.lookahead = std::atoi(optarg) != 0;
                break;
            case 14:
                a.asyncWrite = true;
                break;
            default:
                throw std::invalid_argument("bad command line");
        }
    }
    if (a.nx < 1 || a.ny < 1 || a.checkpoints < 1 || a.blocks < 1)
        throw std::invalid_argument("sizes, checkpoint count and block count must be positive");
    return a;
}

// checkpoint instants as swe_simple.cpp:116-123
inline std::vector<float> checkpoint_instants(float duration, int n) {
    std::vector<float> t(n);
    const float delta = duration / n;
    t[0] = delta;
    for (int i = 1; i < n; i++) t[i] = t[i - 1] + delta;
    return t;
}

// Snapshot file: int32 block count, then per block int32 nx, ny and the
// float32 arrays h, hu, hv, b of (nx+2)*(ny+2) values in Float2D order.
struct BlockState {
    int nx, ny;
    std::vector<float> h, hu, hv, b;
};

inline void write_snapshot(const std::string& path, const std::vector<BlockState>& blocks) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const int32_t n = (int32_t)blocks.size();
    std::fwrite(&n, 4, 1, f);
    for (const BlockState& s : blocks) {
        const int32_t d[2] = {s.nx, s.ny};
        std::fwrite(d, 4, 2, f);
        for (const std::vector<float>* a : {&s.h, &s.hu, &s.hv, &s.b}) std::fwrite(a->data(), 4, a->size(), f);
    }
    std::fclose(f);
}

inline std::vector<BlockState> read_snapshot(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot read " + path);
    int32_t n = 0;
    if (std::fread(&n, 4, 1, f) != 1 || n < 1) throw std::runtime_error("bad snapshot " + path);
    std::vector<BlockState> out(n);
    for (BlockState& s : out) {
        int32_t d[2];
        if (std::fread(d, 4, 2, f) != 2) throw std::runtime_error("bad snapshot " + path);
        s.nx = d[0];
        s.ny = d[1];
        const size_t m = (size_t)(s.nx + 2) * (s.ny + 2);
        for (std::vector<float>* a : {&s.h, &s.hu, &s.hv, &s.b}) {
            a->resize(m);
            if (std::fread(a->data(), 4, m, f) != m) throw std::runtime_error("bad snapshot " + path);
        }
    }
    std::fclose(f);
    return out;
}

// A snapshot's state into a block's host arrays, the boundary types set,
// then uploaded (the block's initScenario without a scenario object).
template <typename Block>
void load_snapshot(Block& blk, const BlockState& s, BoundaryType bnd[4]) {
    if (s.nx != blk.nx || s.ny != blk.ny) throw std::runtime_error("snapshot size differs from the block's");
    std::copy(s.h.begin(), s.h.end(), blk.h.getRawPointer());
    std::copy(s.hu.begin(), s.hu.end(), blk.hu.getRawPointer());
    std::copy(s.hv.begin(), s.hv.end(), blk.hv.getRawPointer());
    std::copy(s.b.begin(), s.b.end(), blk.b.getRawPointer());
    for (int i = 0; i < 4; i++) blk.setBoundaryType(static_cast<Boundary>(i), bnd[i]);
    blk.synchAfterWrite();
}

template <typename Block>
BlockState snapshot_of(Block& blk) {
    BlockState s;
    s.nx = blk.nx;
    s.ny = blk.ny;
    const size_t m = (size_t)(s.nx + 2) * (s.ny + 2);
    const float* p[4] = {blk.getWaterHeight().getRawPointer(), blk.getMomentumHorizontal().getRawPointer(),
                         blk.getMomentumVertical().getRawPointer(), blk.getBathymetry().getRawPointer()};
    std::vector<float>* q[4] = {&s.h, &s.hu, &s.hv, &s.b};
    for (int k = 0; k < 4; k++) q[k]->assign(p[k], p[k] + m);
    return s;
}

#endif
