// NetCdfReader.hh -- reader for netCDF classic files (format versions 1 and
// 2: 32- and 64-bit offsets; Unidata's "netCDF classic format
// specification"), enough to load the gridded inputs of a file-driven
// scenario without libnetcdf: dimensions, variables with their dimensions
// and type, and the values of a non-record variable (or of one record of a
// record variable) converted to float.  Attributes are parsed and skipped.
#ifndef SWE_HIP_NETCDFREADER_HH
#define SWE_HIP_NETCDFREADER_HH

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

class NetCdfReader {
public:
    struct Dim {
        std::string name;
        uint64_t length;  // 0 for the record dimension
    };
    struct Var {
        std::string name;
        std::vector<int> dims;
        int type;        // 1 byte, 2 char, 3 short, 4 int, 5 float, 6 double
        uint64_t vsize;  // bytes (one record's worth for a record variable)
        uint64_t begin;  // file offset of the data
        bool record;
    };

    explicit NetCdfReader(const std::string& path) : path(path) {
        f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("NetCdfReader: cannot open " + path);
        try {
            parseHeader();
        } catch (...) {
            std::fclose(f);
            throw;
        }
    }
    ~NetCdfReader() {
        if (f) std::fclose(f);
    }
    NetCdfReader(const NetCdfReader&) = delete;
    NetCdfReader& operator=(const NetCdfReader&) = delete;

    const std::vector<Dim>& dimensions() const { return dims; }
    const std::vector<Var>& variables() const { return vars; }
    uint64_t records() const { return numrecs; }

    const Var* find(const std::string& name) const {
        for (const Var& v : vars)
            if (v.name == name) return &v;
        return nullptr;
    }
    const Var& var(const std::string& name) const {
        const Var* v = find(name);
        if (!v) throw std::runtime_error("NetCdfReader: no variable '" + name + "' in " + path);
        return *v;
    }
    // number of values of a variable (one record of a record variable)
    uint64_t count(const Var& v) const {
        uint64_t n = 1;
        for (int d : v.dims)
            if (dims[d].length) n *= dims[d].length;
        return n;
    }
    // the values in file order (last dimension fastest), as float
    std::vector<float> read(const Var& v, uint64_t record = 0) const {
        if (v.record && record >= numrecs) throw std::runtime_error("NetCdfReader: record out of range");
        const uint64_t n = count(v);
        const int w = width(v.type);
        std::vector<uint8_t> raw(n * w);
        const uint64_t at = v.begin + (v.record ? record * recsize : 0);
        if (fseeko(f, (off_t)at, SEEK_SET) != 0 || std::fread(raw.data(), 1, raw.size(), f) != raw.size())
            throw std::runtime_error("NetCdfReader: short read in " + path);
        std::vector<float> out(n);
        for (uint64_t k = 0; k < n; k++) {
            const uint8_t* p = &raw[k * w];
            switch (v.type) {
                case 1:
                    out[k] = (float)(int8_t)p[0];
                    break;
                case 3:
                    out[k] = (float)(int16_t)(uint16_t)(p[0] << 8 | p[1]);
                    break;
                case 4:
                    out[k] = (float)(int32_t)be32(p);
                    break;
                case 5: {
                    const uint32_t u = be32(p);
                    std::memcpy(&out[k], &u, 4);
                    break;
                }
                case 6: {
                    const uint64_t u = (uint64_t)be32(p) << 32 | be32(p + 4);
                    double d;
                    std::memcpy(&d, &u, 8);
                    out[k] = (float)d;
                    break;
                }
                default:
                    throw std::runtime_error("NetCdfReader: variable '" + v.name + "' is not numeric");
            }
        }
        return out;
    }

private:
    static uint32_t be32(const uint8_t* p) {
        return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
    }
    static int width(int type) {
        switch (type) {
            case 1:
            case 2:
                return 1;
            case 3:
                return 2;
            case 4:
            case 5:
                return 4;
            case 6:
                return 8;
            default:
                throw std::runtime_error("NetCdfReader: unknown type");
        }
    }

    void bytes(void* p, size_t n) {
        if (std::fread(p, 1, n, f) != n) throw std::runtime_error("NetCdfReader: truncated header in " + path);
    }
    uint32_t u32() {
        uint8_t b[4];
        bytes(b, 4);
        return be32(b);
    }
    uint64_t offset() {
        if (version == 1) return u32();
        const uint64_t hi = u32();
        return hi << 32 | u32();
    }
    std::string name() {
        const uint32_t n = u32();
        if (n > (1u << 20)) throw std::runtime_error("NetCdfReader: bad name length in " + path);
        std::string s(n, '\0');
        if (n) bytes(&s[0], n);
        skipPad(n);
        return s;
    }
    void skipPad(uint64_t n) {
        const uint64_t pad = (4 - n % 4) % 4;
        uint8_t b[4];
        if (pad) bytes(b, (size_t)pad);
    }
    void skipAttributes() {
        const uint32_t tag = u32();
        const uint32_t n = u32();
        if (tag == 0 && n == 0) return;
        if (tag != 12) throw std::runtime_error("NetCdfReader: bad attribute list in " + path);
        for (uint32_t k = 0; k < n; k++) {
            name();
            const int type = (int)u32();
            const uint64_t len = (uint64_t)u32() * width(type);
            std::vector<uint8_t> skip((size_t)len);
            if (len) bytes(skip.data(), (size_t)len);
            skipPad(len);
        }
    }
    void parseHeader() {
        uint8_t magic[4];
        bytes(magic, 4);
        if (magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F' || (magic[3] != 1 && magic[3] != 2))
            throw std::runtime_error("NetCdfReader: " + path + " is not a netCDF classic file (netCDF-4 is not read)");
        version = magic[3];
        numrecs = u32();
        if (numrecs == 0xFFFFFFFFu) numrecs = 0;  // streaming: unknown count
        uint32_t tag = u32(), n = u32();
        if (!(tag == 0 && n == 0)) {
            if (tag != 10) throw std::runtime_error("NetCdfReader: bad dimension list in " + path);
            for (uint32_t k = 0; k < n; k++) {
                Dim d;
                d.name = name();
                d.length = u32();
                dims.push_back(d);
            }
        }
        skipAttributes();  // global attributes
        tag = u32();
        n = u32();
        if (!(tag == 0 && n == 0)) {
            if (tag != 11) throw std::runtime_error("NetCdfReader: bad variable list in " + path);
            for (uint32_t k = 0; k < n; k++) {
                Var v;
                v.name = name();
                const uint32_t nd = u32();
                for (uint32_t j = 0; j < nd; j++) {
                    const uint32_t id = u32();
                    if (id >= dims.size()) throw std::runtime_error("NetCdfReader: bad dimension id in " + path);
                    v.dims.push_back((int)id);
                }
                skipAttributes();
                v.type = (int)u32();
                v.vsize = u32();
                v.begin = offset();
                v.record = !v.dims.empty() && dims[v.dims[0]].length == 0;
                vars.push_back(v);
            }
        }
        // one record: the record variables' slabs in order, each padded to 4
        // bytes unless it is the only record variable
        int nrec = 0;
        uint64_t single = 0;
        for (const Var& v : vars)
            if (v.record) {
                nrec++;
                single = count(v) * width(v.type);
                recsize += v.vsize;
            }
        if (nrec == 1) recsize = single;
    }

    std::string path;
    std::FILE* f = nullptr;
    int version = 1;
    uint64_t numrecs = 0, recsize = 0;
    std::vector<Dim> dims;
    std::vector<Var> vars;
};

#endif
