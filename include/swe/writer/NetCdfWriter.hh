// NetCdfWriter.hh -- netCDF output with the data model of the reference's
// NetCdfWriter (src/writer/NetCdfWriter.cpp:49-240): file <base>.nc,
// dimensions time (unlimited), x, y; float variables time(time) with
// long_name / units, x(x) and y(y) at the cell centres, ah, hu, hv
// (time, y, x) and b (y, x); the CF-1.5 global attributes; b written with the
// first time step; the file brought up to date after every step.
//
// The reference creates netCDF-4 (HDF5) files through libnetcdf.  This writer
// needs no library: it writes the classic format with 64-bit offsets
// (Unidata's "netCDF classic format specification", version byte 2), which
// every netCDF reader opens with the same variables, so files over 2 GiB
// (long 4096^2 runs) stay valid.  Layout: header; the non-record variables
// x, y, b; then one record per time step holding time, ah, hu, hv.  All
// values big-endian.
#ifndef SWE_HIP_NETCDFWRITER_HH
#define SWE_HIP_NETCDFWRITER_HH

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "writer/Writer.hh"

class NetCdfWriter : public Writer {
public:
    NetCdfWriter(const std::string& baseName, const Float2D& b, const BoundarySize& boundarySize, int nX, int nY,
                 float dX, float dY, float originX = 0.f, float originY = 0.f, unsigned int flush = 0)
        : Writer(baseName + ".nc", b, boundarySize, nX, nY), flush(flush) {
        if (nX <= 0 || nY <= 0) throw std::invalid_argument("NetCdfWriter: empty grid");
        file = std::fopen(fileName.c_str(), "w+b");  // an existing file is replaced
        if (!file) throw std::runtime_error("NetCdfWriter: cannot create " + fileName);
        writeHeader();
        // cell-centre coordinates, accumulated in float as the reference does (:137-149)
        std::vector<float> xs(nX), ys(nY);
        float pos = originX + 0.5f * dX;
        for (int i = 0; i < nX; i++) {
            xs[i] = pos;
            pos += dX;
        }
        pos = originY + 0.5f * dY;
        for (int j = 0; j < nY; j++) {
            ys[j] = pos;
            pos += dY;
        }
        putFloats(xBegin, xs.data(), xs.size());
        putFloats(yBegin, ys.data(), ys.size());
        sync();
    }

    ~NetCdfWriter() override {
        if (file) std::fclose(file);
    }

    NetCdfWriter(const NetCdfWriter&) = delete;
    NetCdfWriter& operator=(const NetCdfWriter&) = delete;

    // one time step: the interior of h, hu, hv (ghost widths from the
    // boundary sizes) as record timeStep; b with the first step
    void writeTimeStep(const Float2D& h, const Float2D& hu, const Float2D& hv, float time) override {
        if (timeStep == 0) putGrid(bBegin, b);
        const uint64_t rec = recBegin + (uint64_t)timeStep * recSize;
        putFloats(rec, &time, 1);
        putGrid(rec + 4, h);
        putGrid(rec + 4 + gridBytes(), hu);
        putGrid(rec + 4 + 2 * gridBytes(), hv);
        timeStep++;
        // numrecs: the header's record count, written last so a reader never
        // sees a record that is not complete
        uint8_t n[4];
        be32(n, (uint32_t)timeStep);
        seek(4);
        put(n, 4);
        sync();
    }

private:
    // ---- classic-format header -------------------------------------------
    enum : uint32_t { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12, NC_CHAR = 2, NC_FLOAT = 5 };

    struct Var {
        const char* name;
        std::vector<uint32_t> dims;
        std::vector<std::pair<const char*, const char*>> atts;
        bool record;
    };

    uint64_t gridBytes() const { return (uint64_t)nX * nY * 4; }

    void writeHeader() {
        static const std::pair<const char*, const char*> globals[] = {
            {"Conventions", "CF-1.5"},
            {"title", "Computed tsunami solution"},
            {"history", "SWE"},
            {"institution",
             "Technische Universitaet Muenchen, Department of Informatics, Chair of Scientific Computing"},
            {"source", "Bathymetry and displacement data."},
            {"references", "http://www5.in.tum.de/SWE"},
            {"comment",
             "SWE is free software and licensed under the GNU General Public License. Remark: In general this does "
             "not hold for the used input data."},
        };
        enum { DIM_TIME = 0, DIM_X = 1, DIM_Y = 2 };
        const std::vector<Var> vars = {
            {"time", {DIM_TIME}, {{"long_name", "Time"}, {"units", "seconds since simulation start"}}, true},
            {"x", {DIM_X}, {}, false},
            {"y", {DIM_Y}, {}, false},
            {"ah", {DIM_TIME, DIM_Y, DIM_X}, {}, true},
            {"hu", {DIM_TIME, DIM_Y, DIM_X}, {}, true},
            {"hv", {DIM_TIME, DIM_Y, DIM_X}, {}, true},
            {"b", {DIM_Y, DIM_X}, {}, false},
        };
        const uint64_t vsize[] = {4, 4ull * nX, 4ull * nY, gridBytes(), gridBytes(), gridBytes(), gridBytes()};
        for (uint64_t s : vsize)
            if (s > 0xFFFFFFFCull) throw std::invalid_argument("NetCdfWriter: grid too large for one variable");

        std::vector<uint8_t> hd;
        auto u32 = [&](uint32_t v) {
            uint8_t b4[4];
            be32(b4, v);
            hd.insert(hd.end(), b4, b4 + 4);
        };
        auto u64 = [&](uint64_t v) {
            u32((uint32_t)(v >> 32));
            u32((uint32_t)v);
        };
        auto padded = [&](const void* p, size_t n) {
            const uint8_t* c = static_cast<const uint8_t*>(p);
            hd.insert(hd.end(), c, c + n);
            while (hd.size() % 4) hd.push_back(0);
        };
        auto name = [&](const char* s) {
            u32((uint32_t)std::strlen(s));
            padded(s, std::strlen(s));
        };
        auto text_atts = [&](const std::pair<const char*, const char*>* a, size_t n) {
            if (n == 0) {
                u32(0);
                u32(0);
                return;
            }
            u32(NC_ATTRIBUTE);
            u32((uint32_t)n);
            for (size_t k = 0; k < n; k++) {
                name(a[k].first);
                u32(NC_CHAR);
                u32((uint32_t)std::strlen(a[k].second));
                padded(a[k].second, std::strlen(a[k].second));
            }
        };

        hd.insert(hd.end(), {'C', 'D', 'F', 2});
        u32(0);  // numrecs
        u32(NC_DIMENSION);
        u32(3);
        name("time");
        u32(0);  // the record dimension
        name("x");
        u32((uint32_t)nX);
        name("y");
        u32((uint32_t)nY);
        text_atts(globals, sizeof(globals) / sizeof(globals[0]));
        u32(NC_VARIABLE);
        u32((uint32_t)vars.size());
        std::vector<size_t> beginAt(vars.size());
        for (size_t v = 0; v < vars.size(); v++) {
            name(vars[v].name);
            u32((uint32_t)vars[v].dims.size());
            for (uint32_t d : vars[v].dims) u32(d);
            text_atts(vars[v].atts.data(), vars[v].atts.size());
            u32(NC_FLOAT);
            u32((uint32_t)vsize[v]);
            beginAt[v] = hd.size();
            u64(0);  // begin, patched below
        }
        // data offsets: the non-record variables in order, then the records
        uint64_t off = hd.size();
        uint64_t begins[7];
        for (size_t v = 0; v < vars.size(); v++)
            if (!vars[v].record) {
                begins[v] = off;
                off += vsize[v];
            }
        recBegin = off;
        recSize = 0;
        for (size_t v = 0; v < vars.size(); v++)
            if (vars[v].record) {
                begins[v] = recBegin + recSize;
                recSize += vsize[v];
            }
        for (size_t v = 0; v < vars.size(); v++) {
            uint8_t b8[8];
            be32(b8, (uint32_t)(begins[v] >> 32));
            be32(b8 + 4, (uint32_t)begins[v]);
            std::memcpy(&hd[beginAt[v]], b8, 8);
        }
        xBegin = begins[1];
        yBegin = begins[2];
        bBegin = begins[6];
        seek(0);
        put(hd.data(), hd.size());
    }

    // ---- data ----------------------------------------------------------------
    // the interior of a Float2D as a (y, x) slab, x fastest
    void putGrid(uint64_t at, const Float2D& a) {
        std::vector<float> slab((size_t)nX * nY);
        const int x0 = boundarySize[0], y0 = boundarySize[2];
        for (int j = 0; j < nY; j++)
            for (int i = 0; i < nX; i++) slab[(size_t)j * nX + i] = a[i + x0][j + y0];
        putFloats(at, slab.data(), slab.size());
    }

    void putFloats(uint64_t at, const float* v, size_t n) {
        std::vector<uint8_t> buf(4 * n);
        for (size_t k = 0; k < n; k++) {
            uint32_t u;
            std::memcpy(&u, &v[k], 4);
            be32(&buf[4 * k], u);
        }
        seek(at);
        put(buf.data(), buf.size());
    }

    static void be32(uint8_t* p, uint32_t v) {
        p[0] = (uint8_t)(v >> 24);
        p[1] = (uint8_t)(v >> 16);
        p[2] = (uint8_t)(v >> 8);
        p[3] = (uint8_t)v;
    }

    void seek(uint64_t at) {
        if (fseeko(file, (off_t)at, SEEK_SET) != 0) throw std::runtime_error("NetCdfWriter: seek failed in " + fileName);
    }
    void put(const void* p, size_t n) {
        if (std::fwrite(p, 1, n, file) != n) throw std::runtime_error("NetCdfWriter: write failed in " + fileName);
    }
    void sync() {
        if (std::fflush(file) != 0) throw std::runtime_error("NetCdfWriter: flush failed in " + fileName);
    }

    std::FILE* file = nullptr;
    unsigned int flush;  // kept for the reference's signature; every step is flushed
    uint64_t xBegin = 0, yBegin = 0, bBegin = 0, recBegin = 0, recSize = 0;
};

#endif
