// ncwriter.cpp -- the NetCDF output of the asynchronous writer
// (src/async_output.cpp:18-54 defines the file, :108-141 appends a save),
// written in the classic 64-bit-offset NetCDF format.
//
// The reference creates a NetCDF-4 file through libnetcdf (HDF5 storage).
// Neither library is linked here, so the same dimensions, variables,
// attributes and records go into the classic format: big-endian, a header
// followed by one record per save, readable by every NetCDF library.
//
//   dimensions: time (unlimited), x = ncols, y = nrows
//   variables:  double time(time)        units = "seconds"
//               float  H(time, y, x)     units = "m", long_name = "water_depth",
//                                        _FillValue = (float)nodata
//               float  Vx(time, y, x)    units = "m/s", long_name = "x_direction_velo"
//               float  Vy(time, y, x)    units = "m/s", long_name = "y_direction_velo"
//               (Vx, Vy only with velocity output; their long_name is the
//               first 16 characters of the text, the length the reference
//               passes, async_output.cpp:40,44)
//   global:     double xllcorner, yllcorner, cellsize
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "adaptflood.h"

namespace af {
void set_error(const char* fmt, ...);
}

struct af_nc {
    FILE* fp = nullptr;
    std::string path;
    int32_t nx = 0, ny = 0;
    bool velocity = false;
    uint64_t rec_begin = 0;  // file offset of record 0
    uint64_t recsize = 0;    // bytes per record (all record variables)
    uint32_t numrecs = 0;
};

namespace {

// classic-format tags and types
constexpr uint32_t kDimension = 0x0A, kVariable = 0x0B, kAttribute = 0x0C;
constexpr uint32_t kChar = 2, kFloat = 5, kDouble = 6;

struct Header {
    std::string b;
    void u32(uint32_t v) {
        const char c[4] = {(char)(v >> 24), (char)(v >> 16), (char)(v >> 8), (char)v};
        b.append(c, 4);
    }
    void u64(uint64_t v) {
        u32((uint32_t)(v >> 32));
        u32((uint32_t)v);
    }
    void f64(double d) {
        uint64_t v;
        memcpy(&v, &d, 8);
        u64(v);
    }
    void f32(float f) {
        uint32_t v;
        memcpy(&v, &f, 4);
        u32(v);
    }
    void pad() {
        while (b.size() % 4) b.push_back('\0');
    }
    void name(const char* s) {
        u32((uint32_t)strlen(s));
        b.append(s);
        pad();
    }
    void text_att(const char* n, const char* s, size_t len) {
        name(n);
        u32(kChar);
        u32((uint32_t)len);
        b.append(s, len);
        pad();
    }
    void double_att(const char* n, double v) {
        name(n);
        u32(kDouble);
        u32(1);
        f64(v);
    }
};

// One record variable of the header; `begin` is patched once the header
// size is known.
struct VarDef {
    const char* name;
    std::vector<uint32_t> dims;
    uint32_t type;
    uint64_t vsize;
    size_t begin_at = 0;  // byte position of the begin field in the header
};

void put_be_floats(std::vector<char>& out, const double* v, size_t n) {
    out.resize(n * 4);
    for (size_t i = 0; i < n; i++) {
        const float f = (float)v[i];
        uint32_t u;
        memcpy(&u, &f, 4);
        out[4 * i] = (char)(u >> 24);
        out[4 * i + 1] = (char)(u >> 16);
        out[4 * i + 2] = (char)(u >> 8);
        out[4 * i + 3] = (char)u;
    }
}

}  // namespace

extern "C" {

int af_nc_create(const char* path, int32_t ncols, int32_t nrows, double xll, double yll,
                 double cellsize, double nodata, int32_t velocity, af_nc** out) {
    if (!path || !out || ncols <= 0 || nrows <= 0) {
        af::set_error("nc_create: bad arguments");
        return AF_E_INVALID;
    }
    const uint64_t grid = (uint64_t)ncols * (uint64_t)nrows * 4;
    if (grid >= (1ull << 31)) {  // readers take the 32-bit slab size as signed
        af::set_error("nc_create: a %d x %d float grid exceeds the 2 GiB record slab the "
                      "64-bit-offset format holds", ncols, nrows);
        return AF_E_INVALID;
    }
    *out = nullptr;
    Header h;
    h.b.append("CDF\x02", 4);
    h.u32(0);  // numrecs
    h.u32(kDimension);
    h.u32(3);
    h.name("time");
    h.u32(0);  // the unlimited (record) dimension
    h.name("x");
    h.u32((uint32_t)ncols);
    h.name("y");
    h.u32((uint32_t)nrows);
    h.u32(kAttribute);
    h.u32(3);
    h.double_att("xllcorner", xll);
    h.double_att("yllcorner", yll);
    h.double_att("cellsize", cellsize);

    std::vector<VarDef> vars = {{"time", {0}, kDouble, 8}, {"H", {0, 2, 1}, kFloat, grid}};
    if (velocity) {
        vars.push_back({"Vx", {0, 2, 1}, kFloat, grid});
        vars.push_back({"Vy", {0, 2, 1}, kFloat, grid});
    }
    h.u32(kVariable);
    h.u32((uint32_t)vars.size());
    for (VarDef& v : vars) {
        h.name(v.name);
        h.u32((uint32_t)v.dims.size());
        for (uint32_t d : v.dims) h.u32(d);
        const std::string n = v.name;
        if (n == "time") {
            h.u32(kAttribute);
            h.u32(1);
            h.text_att("units", "seconds", 7);
        } else if (n == "H") {
            h.u32(kAttribute);
            h.u32(3);
            h.text_att("units", "m", 1);
            h.text_att("long_name", "water_depth", 11);
            h.name("_FillValue");
            h.u32(kFloat);
            h.u32(1);
            h.f32((float)nodata);
        } else {
            h.u32(kAttribute);
            h.u32(2);
            h.text_att("units", "m/s", 3);
            h.text_att("long_name", n == "Vx" ? "x_direction_velocity" : "y_direction_velocity",
                       16);
        }
        h.u32(v.type);
        h.u32((uint32_t)v.vsize);
        v.begin_at = h.b.size();
        h.u64(0);  // begin, patched below
    }
    // records follow the header directly: time, H[, Vx, Vy] of record 0, ...
    uint64_t off = h.b.size(), recsize = 0;
    for (VarDef& v : vars) {
        Header be;
        be.u64(off + recsize);
        memcpy(&h.b[v.begin_at], be.b.data(), 8);
        recsize += v.vsize;
    }
    FILE* fp = fopen(path, "wb");
    if (!fp) {
        af::set_error("Cannot open %s for writing", path);
        return AF_E_IO;
    }
    if (fwrite(h.b.data(), 1, h.b.size(), fp) != h.b.size() || fflush(fp) != 0) {
        fclose(fp);
        af::set_error("write error on %s", path);
        return AF_E_IO;
    }
    af_nc* nc = new af_nc;
    nc->fp = fp;
    nc->path = path;
    nc->nx = ncols;
    nc->ny = nrows;
    nc->velocity = velocity != 0;
    nc->rec_begin = off;
    nc->recsize = recsize;
    *out = nc;
    return AF_OK;
}

int af_nc_append(af_nc* nc, double t, const double* H, const double* Vx, const double* Vy) {
    if (!nc || !nc->fp || !H || (nc->velocity && (!Vx || !Vy))) {
        af::set_error("nc_append: bad arguments");
        return AF_E_INVALID;
    }
    const uint64_t at = nc->rec_begin + (uint64_t)nc->numrecs * nc->recsize;
    bool ok = fseeko(nc->fp, (off_t)at, SEEK_SET) == 0;
    Header tb;
    tb.f64(t);
    ok = ok && fwrite(tb.b.data(), 1, 8, nc->fp) == 8;
    std::vector<char> buf;
    // the grid goes out in slabs of rows, converted to big-endian floats
    const size_t rows_per_slab = std::max<size_t>(1, (size_t)(1 << 22) / (size_t)nc->nx);
    const double* fields[3] = {H, Vx, Vy};
    const int nfields = nc->velocity ? 3 : 1;
    for (int f = 0; f < nfields && ok; f++) {
        for (size_t r0 = 0; r0 < (size_t)nc->ny && ok; r0 += rows_per_slab) {
            const size_t rows = std::min(rows_per_slab, (size_t)nc->ny - r0);
            put_be_floats(buf, fields[f] + r0 * nc->nx, rows * nc->nx);
            ok = fwrite(buf.data(), 1, buf.size(), nc->fp) == buf.size();
        }
    }
    // numrecs (header offset 4) follows the records written
    Header nb;
    nb.u32(nc->numrecs + 1);
    ok = ok && fseeko(nc->fp, 4, SEEK_SET) == 0 && fwrite(nb.b.data(), 1, 4, nc->fp) == 4 &&
         fflush(nc->fp) == 0;
    if (!ok) {
        af::set_error("write error on %s", nc->path.c_str());
        return AF_E_IO;
    }
    nc->numrecs++;
    return AF_OK;
}

int af_nc_close(af_nc* nc) {
    if (!nc) return AF_OK;
    int rc = AF_OK;
    if (nc->fp && fclose(nc->fp) != 0) {
        af::set_error("write error on %s", nc->path.c_str());
        rc = AF_E_IO;
    }
    delete nc;
    return rc;
}

}  // extern "C"
