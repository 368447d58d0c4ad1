// On-disk particle state.
//
// Two formats, both read and written from host arrays in the caller's
// particle order (no device involved), and from/to a sim's layer:
//
// * DMPM_FILE_NATIVE -- this library's checkpoint format (layout documented
//   in include/dmpm.h): a 64-byte header, a table of named fields, then each
//   field as n x ncomp values of the file's dtype, particle-major, 3x3
//   matrices row-major.  float32 or float64.
//
// * DMPM_FILE_CEREAL -- the reference's PointCloud binary file, written by
//   cereal's BinaryOutputArchive (PointCloud::WriteEntirePointCloudToBinaryFile,
//   PointCloud.cpp:382-394): a uint64 point count, then per point the fields
//   of MaterialPoint::serialize (MaterialPoint.h:55-71) in order x, v, F, C, P,
//   m, vol, lam, mu, dFc, dLdx, dLdv, dLdF, momentum, vector, vector_max.  An
//   Eigen 3-vector is written as int32 rows (3) then its 3 floats, a 3x3
//   matrix as int32 rows, int32 cols (3, 3) then its 9 floats row by row
//   (cereal_eigen.h); scalars are raw float32.  432 bytes per point, native
//   (little-endian) byte order.
#include "../../include/dmpm.h"
#include "dmpm_kernels.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct IoError : std::runtime_error {
    int code;
    IoError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <typename F>
int io_guard(F&& f) {
    try {
        f();
        return DMPM_OK;
    } catch (const IoError& e) {
        dmpm::set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        dmpm::set_last_error(e.what());
        return DMPM_ERR_ARG;
    }
}

struct File {
    FILE* f = nullptr;
    std::string path;
    File(const char* p, const char* mode) : path(p ? p : "") {
        if (!p) throw IoError(DMPM_ERR_ARG, "null path");
        f = std::fopen(p, mode);
        if (!f) throw IoError(DMPM_ERR_ARG, "cannot open " + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void write(const void* p, size_t bytes) {
        if (bytes && std::fwrite(p, 1, bytes, f) != bytes) throw IoError(DMPM_ERR_ARG, "write failed: " + path);
    }
    void read(void* p, size_t bytes) {
        if (bytes && std::fread(p, 1, bytes, f) != bytes)
            throw IoError(DMPM_ERR_ARG, "truncated or unreadable file: " + path);
    }
    long long size() {
        const long pos = std::ftell(f);
        std::fseek(f, 0, SEEK_END);
        const long long s = std::ftell(f);
        std::fseek(f, pos, SEEK_SET);
        return s;
    }
};

// The fields of a particle file, in the order of the cereal record.
struct FieldDef {
    const char* name;
    int ncomp;
    void* dmpm_particle_arrays::*member;
    bool required;  // must be given on write and present on read
};
const FieldDef kFields[] = {
    {"x", 3, &dmpm_particle_arrays::x, true},
    {"v", 3, &dmpm_particle_arrays::v, true},
    {"F", 9, &dmpm_particle_arrays::F, true},
    {"C", 9, &dmpm_particle_arrays::C, true},
    {"P", 9, &dmpm_particle_arrays::P, false},
    {"m", 1, &dmpm_particle_arrays::m, true},
    {"vol", 1, &dmpm_particle_arrays::vol, true},
    {"lam", 1, &dmpm_particle_arrays::lam, true},
    {"mu", 1, &dmpm_particle_arrays::mu, true},
    {"dFc", 9, &dmpm_particle_arrays::dFc, false},
    {"dLdx", 3, &dmpm_particle_arrays::dLdx, false},
    {"dLdv", 3, &dmpm_particle_arrays::dLdv, false},
    {"dLdF", 9, &dmpm_particle_arrays::dLdF, false},
    {"momentum", 9, &dmpm_particle_arrays::adam_m, false},
    {"vector", 9, &dmpm_particle_arrays::adam_v, false},
    {"vector_max", 9, &dmpm_particle_arrays::adam_vmax, false},
};
constexpr int kNumFields = sizeof(kFields) / sizeof(kFields[0]);
constexpr size_t kCerealPointBytes = 432;
const char kMagic[8] = {'D', 'M', 'P', 'M', 'S', 'T', 'A', 'T'};

size_t esize(int dtype) { return dtype == DMPM_DTYPE_F64 ? 8 : 4; }

double get(const void* a, int dtype, size_t i) {
    return dtype == DMPM_DTYPE_F64 ? ((const double*)a)[i] : (double)((const float*)a)[i];
}
void put(void* a, int dtype, size_t i, double v) {
    if (dtype == DMPM_DTYPE_F64) ((double*)a)[i] = v;
    else ((float*)a)[i] = (float)v;
}

void check_dtype(int dtype) {
    if (dtype != DMPM_DTYPE_F32 && dtype != DMPM_DTYPE_F64) throw IoError(DMPM_ERR_ARG, "unknown dtype");
}

// ---- native ----------------------------------------------------------------
struct NativeHeader {
    char magic[8];
    uint32_t version;
    uint32_t dtype;
    uint64_t n;
    uint32_t nfields;
    int32_t layer;
    uint8_t reserved[32];
};
static_assert(sizeof(NativeHeader) == 64, "native header is 64 bytes");
struct NativeField {
    char name[16];
    uint32_t ncomp;
    uint32_t reserved;
};
static_assert(sizeof(NativeField) == 24, "native field entry is 24 bytes");

void write_native(const char* path, int n, const dmpm_particle_arrays& a, int dtype, int layer) {
    File f(path, "wb");
    NativeHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = 1;
    h.dtype = (uint32_t)dtype;
    h.n = (uint64_t)n;
    h.nfields = kNumFields;
    h.layer = layer;
    f.write(&h, sizeof(h));
    for (const auto& d : kFields) {
        NativeField e{};
        std::strncpy(e.name, d.name, sizeof(e.name) - 1);
        e.ncomp = (uint32_t)d.ncomp;
        f.write(&e, sizeof(e));
    }
    const size_t es = esize(dtype);
    std::vector<char> zero;
    for (const auto& d : kFields) {
        const void* src = a.*d.member;
        const size_t bytes = (size_t)n * d.ncomp * es;
        if (!src) {
            zero.assign(bytes, 0);
            src = zero.data();
        }
        f.write(src, bytes);
    }
}

NativeHeader read_native_header(File& f) {
    NativeHeader h{};
    f.read(&h, sizeof(h));
    if (std::memcmp(h.magic, kMagic, 8) != 0) throw IoError(DMPM_ERR_ARG, "not a dmpm state file: " + f.path);
    if (h.version != 1) throw IoError(DMPM_ERR_ARG, "unsupported dmpm state file version");
    if (h.dtype != DMPM_DTYPE_F32 && h.dtype != DMPM_DTYPE_F64) throw IoError(DMPM_ERR_ARG, "bad dtype in file");
    if (h.n > (uint64_t)INT32_MAX) throw IoError(DMPM_ERR_ARG, "particle count too large");
    if (h.nfields == 0 || h.nfields > 256) throw IoError(DMPM_ERR_ARG, "bad field count in file");
    return h;
}

void read_native(const char* path, int n, dmpm_particle_arrays& a, int dtype) {
    File f(path, "rb");
    const NativeHeader h = read_native_header(f);
    if ((int)h.n != n) throw IoError(DMPM_ERR_ARG, "particle count differs from the file's");
    std::vector<NativeField> table(h.nfields);
    f.read(table.data(), table.size() * sizeof(NativeField));
    const size_t fes = esize((int)h.dtype);
    long long expect = (long long)sizeof(NativeHeader) + (long long)table.size() * sizeof(NativeField);
    for (const auto& e : table) expect += (long long)n * e.ncomp * fes;
    if (f.size() != expect) throw IoError(DMPM_ERR_ARG, "file size does not match its header: " + f.path);
    std::vector<char> buf;
    std::vector<bool> seen(kNumFields, false);
    for (const auto& e : table) {
        char name[17] = {0};
        std::memcpy(name, e.name, 16);
        const size_t bytes = (size_t)n * e.ncomp * fes;
        buf.resize(bytes);
        f.read(buf.data(), bytes);
        for (int k = 0; k < kNumFields; ++k) {
            const auto& d = kFields[k];
            if (std::strcmp(name, d.name) != 0) continue;
            if ((int)e.ncomp != d.ncomp) throw IoError(DMPM_ERR_ARG, std::string("bad width of field ") + name);
            seen[k] = true;
            void* dst = a.*d.member;
            if (!dst) break;
            const size_t cnt = (size_t)n * d.ncomp;
            if ((int)h.dtype == dtype) std::memcpy(dst, buf.data(), cnt * fes);
            else
                for (size_t i = 0; i < cnt; ++i) put(dst, dtype, i, get(buf.data(), (int)h.dtype, i));
        }
    }
    for (int k = 0; k < kNumFields; ++k) {
        const auto& d = kFields[k];
        if (seen[k]) continue;
        if (d.required) throw IoError(DMPM_ERR_ARG, std::string("file lacks field ") + d.name);
        if (void* dst = a.*d.member) std::memset(dst, 0, (size_t)n * d.ncomp * esize(dtype));
    }
}

// ---- cereal ----------------------------------------------------------------
void write_cereal(const char* path, int n, const dmpm_particle_arrays& a, int dtype) {
    File f(path, "wb");
    const uint64_t count = (uint64_t)n;
    f.write(&count, sizeof(count));
    std::vector<char> rec(kCerealPointBytes);
    for (int p = 0; p < n; ++p) {
        char* w = rec.data();
        for (const auto& d : kFields) {
            const void* src = a.*d.member;
            if (d.ncomp == 3) {
                const int32_t rows = 3;
                std::memcpy(w, &rows, 4);
                w += 4;
            } else if (d.ncomp == 9) {
                const int32_t rc[2] = {3, 3};
                std::memcpy(w, rc, 8);
                w += 8;
            }
            for (int c = 0; c < d.ncomp; ++c) {
                const float v = src ? (float)get(src, dtype, (size_t)p * d.ncomp + c) : 0.0f;
                std::memcpy(w, &v, 4);
                w += 4;
            }
        }
        f.write(rec.data(), rec.size());
    }
}

int cereal_count(File& f) {
    uint64_t count = 0;
    f.read(&count, sizeof(count));
    const long long body = f.size() - 8;
    if (count > (uint64_t)INT32_MAX || body != (long long)(count * kCerealPointBytes))
        throw IoError(DMPM_ERR_ARG, "not a PointCloud binary file of 432-byte points: " + f.path);
    return (int)count;
}

void read_cereal(const char* path, int n, dmpm_particle_arrays& a, int dtype) {
    File f(path, "rb");
    if (cereal_count(f) != n) throw IoError(DMPM_ERR_ARG, "particle count differs from the file's");
    std::vector<char> rec(kCerealPointBytes);
    for (int p = 0; p < n; ++p) {
        f.read(rec.data(), rec.size());
        const char* r = rec.data();
        for (const auto& d : kFields) {
            if (d.ncomp == 3) {
                int32_t rows;
                std::memcpy(&rows, r, 4);
                r += 4;
                if (rows != 3) throw IoError(DMPM_ERR_ARG, std::string("bad vector tag in field ") + d.name);
            } else if (d.ncomp == 9) {
                int32_t rc[2];
                std::memcpy(rc, r, 8);
                r += 8;
                if (rc[0] != 3 || rc[1] != 3)
                    throw IoError(DMPM_ERR_ARG, std::string("bad matrix tag in field ") + d.name);
            }
            void* dst = a.*d.member;
            for (int c = 0; c < d.ncomp; ++c) {
                float v;
                std::memcpy(&v, r, 4);
                r += 4;
                if (dst) put(dst, dtype, (size_t)p * d.ncomp + c, (double)v);
            }
        }
    }
}

void check_arrays_for_write(const dmpm_particle_arrays* a) {
    if (!a) throw IoError(DMPM_ERR_ARG, "null arrays");
    for (const auto& d : kFields)
        if (d.required && !(a->*d.member)) throw IoError(DMPM_ERR_ARG, std::string("missing field ") + d.name);
}

}  // namespace

extern "C" {

int dmpm_file_info(const char* path, int format, int* n, int* dtype) {
    return io_guard([&] {
        if (!n) throw IoError(DMPM_ERR_ARG, "null output");
        File f(path, "rb");
        if (format == DMPM_FILE_NATIVE) {
            const NativeHeader h = read_native_header(f);
            *n = (int)h.n;
            if (dtype) *dtype = (int)h.dtype;
        } else if (format == DMPM_FILE_CEREAL) {
            *n = cereal_count(f);
            if (dtype) *dtype = DMPM_DTYPE_F32;
        } else {
            throw IoError(DMPM_ERR_ARG, "unknown file format");
        }
    });
}

int dmpm_write_file(const char* path, int format, int n, const dmpm_particle_arrays* a, int dtype) {
    return io_guard([&] {
        check_dtype(dtype);
        if (n < 0) throw IoError(DMPM_ERR_ARG, "negative particle count");
        check_arrays_for_write(a);
        if (format == DMPM_FILE_NATIVE) write_native(path, n, *a, dtype, -1);
        else if (format == DMPM_FILE_CEREAL) write_cereal(path, n, *a, dtype);
        else throw IoError(DMPM_ERR_ARG, "unknown file format");
    });
}

int dmpm_read_file(const char* path, int format, int n, dmpm_particle_arrays* a, int dtype) {
    return io_guard([&] {
        check_dtype(dtype);
        if (!a) throw IoError(DMPM_ERR_ARG, "null arrays");
        if (format == DMPM_FILE_NATIVE) read_native(path, n, *a, dtype);
        else if (format == DMPM_FILE_CEREAL) read_cereal(path, n, *a, dtype);
        else throw IoError(DMPM_ERR_ARG, "unknown file format");
    });
}

int dmpm_write_layer(dmpm_sim* sim, int layer, const char* path, int format) {
    return io_guard([&] {
        if (format != DMPM_FILE_NATIVE && format != DMPM_FILE_CEREAL)
            throw IoError(DMPM_ERR_ARG, "unknown file format");
        int n = 0;
        if (dmpm_num_particles(sim, &n) != DMPM_OK) throw IoError(DMPM_ERR_ARG, dmpm_last_error());
        dmpm_config cfg;
        if (dmpm_get_config(sim, &cfg) != DMPM_OK) throw IoError(DMPM_ERR_ARG, dmpm_last_error());
        const int dtype = format == DMPM_FILE_NATIVE && cfg.precision == DMPM_PREC_F64 ? DMPM_DTYPE_F64
                                                                                        : DMPM_DTYPE_F32;
        const size_t es = esize(dtype);
        std::vector<std::vector<char>> store(kNumFields);
        dmpm_particle_arrays a{};
        struct Src {
            int field;
            bool grad;
        };
        const Src src[kNumFields] = {
            {DMPM_FIELD_X, false},      {DMPM_FIELD_V, false},      {DMPM_FIELD_F, false},
            {DMPM_FIELD_C, false},      {DMPM_FIELD_P, false},      {DMPM_FIELD_M, false},
            {DMPM_FIELD_VOL, false},    {DMPM_FIELD_LAM, false},    {DMPM_FIELD_MU, false},
            {DMPM_FIELD_DFC, false},    {DMPM_FIELD_X, true},       {DMPM_FIELD_V, true},
            {DMPM_FIELD_F, true},       {DMPM_FIELD_ADAM_M, false}, {DMPM_FIELD_ADAM_V, false},
            {DMPM_FIELD_ADAM_VMAX, false},
        };
        for (int k = 0; k < kNumFields; ++k) {
            store[k].assign((size_t)n * kFields[k].ncomp * es, 0);
            void* out = store[k].data();
            int rc;
            if (src[k].grad) {
                // gradients exist for the last layer and the control layer of the
                // last backward pass; other layers are written as zeros
                rc = dmpm_get_layer_grad(sim, layer, src[k].field, out, dtype, 0);
                if (rc == DMPM_ERR_STATE) rc = DMPM_OK;
            } else {
                rc = dmpm_get_layer(sim, layer, src[k].field, out, dtype, 0);
            }
            if (rc != DMPM_OK) throw IoError(rc, dmpm_last_error());
            a.*kFields[k].member = out;
        }
        if (format == DMPM_FILE_NATIVE) write_native(path, n, a, dtype, layer);
        else write_cereal(path, n, a, dtype);
    });
}

int dmpm_read_particles(dmpm_sim* sim, const char* path, int format) {
    return io_guard([&] {
        int n = 0, fdtype = DMPM_DTYPE_F32;
        if (dmpm_file_info(path, format, &n, &fdtype) != DMPM_OK) throw IoError(DMPM_ERR_ARG, dmpm_last_error());
        if (n <= 0) throw IoError(DMPM_ERR_ARG, "file holds no particles");
        const int dtype = fdtype;
        const size_t es = esize(dtype);
        std::vector<std::vector<char>> store(kNumFields);
        dmpm_particle_arrays a{};
        for (int k = 0; k < kNumFields; ++k) {
            store[k].assign((size_t)n * kFields[k].ncomp * es, 0);
            a.*kFields[k].member = store[k].data();
        }
        if (dmpm_read_file(path, format, n, &a, dtype) != DMPM_OK) throw IoError(DMPM_ERR_ARG, dmpm_last_error());
        int rc = dmpm_set_particles(sim, n, a.x, a.v, a.F, a.C, a.dFc, a.m, a.vol, a.lam, a.mu, dtype, 0);
        if (rc != DMPM_OK) throw IoError(rc, dmpm_last_error());
        bool adam = false;
        for (void* p : {a.adam_m, a.adam_v, a.adam_vmax}) {
            const size_t cnt = (size_t)n * 9;
            for (size_t i = 0; i < cnt && !adam; ++i) adam = get(p, dtype, i) != 0.0;
        }
        if (adam) {
            rc = dmpm_set_adam_state(sim, 0, a.adam_m, a.adam_v, a.adam_vmax, dtype, 0);
            if (rc != DMPM_OK) throw IoError(rc, dmpm_last_error());
        }
    });
}

}  // extern "C"
