// io.cpp -- readers for the reference's input formats and its ESRI-ASCII
// writer, with the reference's parsing behaviour:
//   .par        read_par     src/io.cpp:9-47
//   .dem.ascii  read_dem     src/io.cpp:132-161
//   .bci        read_bci     src/io.cpp:164-254
//   .bdy        read_bdy     src/io.cpp:257-296
//   rain        read_rain    src/io.cpp:103-129
//   interpolation            src/io.cpp:299-311
//   write_asc                src/io.cpp:314-334
// Differences are limited to inputs on which the reference has undefined
// behaviour (over-long tokens, a missing header number): those are rejected
// with AF_E_IO instead.
#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <algorithm>
#include <cmath>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/adaptflood.h"

namespace af {
void set_error(const char* fmt, ...);
}

namespace {

struct Point {
    int x = 0, y = 0;
    int type = AF_BC_CLOSED;
    double value = 0.0;
    std::string name;
    std::vector<double> times, values;
};

}  // namespace

struct af_case {
    af_par par;
    bool have_dem = false;
    int ncols = 0, nrows = 0;
    double xll = 0, yll = 0, cellsize = 0, nodata = -9999.0;
    std::vector<double> dem;
    std::vector<Point> points;
    std::vector<char> line_sides;  // FREE sides in file order
    bool rain_enabled = false;
    std::vector<double> rain_t, rain_r;
    // stage gauges: grid coordinates and names, in file order
    std::vector<int32_t> gx, gy;
    std::vector<std::string> gname;
    // flattened views for af_case_scenario
    std::vector<int32_t> px, py, ptype, poff, plen;
    std::vector<double> pvalue, st, sv;
};

namespace {

void copy_str(char* dst, size_t cap, const std::string& s) {
    size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
    memcpy(dst, s.data(), n);
    dst[n] = 0;
}

// std::stod / std::stoi semantics: leading whitespace, longest numeric
// prefix, failure when nothing converts.
bool to_double(const std::string& v, double* out) {
    const char* b = v.c_str();
    char* e = nullptr;
    errno = 0;
    double d = strtod(b, &e);
    if (e == b || errno == ERANGE) return false;
    *out = d;
    return true;
}
bool to_int(const std::string& v, int* out) {
    const char* b = v.c_str();
    char* e = nullptr;
    errno = 0;
    long l = strtol(b, &e, 10);
    if (e == b || errno == ERANGE || l < -2147483648L || l > 2147483647L) return false;
    *out = (int)l;
    return true;
}

}  // namespace

extern "C" {

// types.hpp:18-38
void af_par_defaults(af_par* p) {
    memset(p, 0, sizeof(*p));
    copy_str(p->dem_file, sizeof(p->dem_file), "dem.ascii");
    copy_str(p->bci_file, sizeof(p->bci_file), "bci.bci");
    copy_str(p->bdy_file, sizeof(p->bdy_file), "bdy.bdy");
    copy_str(p->output_dir, sizeof(p->output_dir), "results");
    p->sim_time = 3600.0;
    p->save_int = 600.0;
    p->mass_int = 60.0;
    p->stage_int = 60.0;
    p->init_tstep = 10.0;
    p->cfl = 0.7;
    p->fpn = 0.035;
    p->depth_thresh = 0.001;
    p->nodata = -9999.0;
    p->output_format = 0;
    p->voutput = 0;
    p->async_output = 1;
}

int af_read_par(const char* path, af_par* p) {
    if (!path || !p) {
        af::set_error("read_par: null argument");
        return AF_E_INVALID;
    }
    std::ifstream f(path);
    if (!f.is_open()) {
        af::set_error("Cannot open parameter file %s", path);
        return AF_E_IO;
    }
    // keys the file leaves out keep the defaults of types.hpp:18-38, as the
    // reference's default-constructed Parameters does
    af_par_defaults(p);
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        lineno++;
        if (line.empty() || line[0] == '#') continue;  // io.cpp:19
        std::istringstream iss(line);
        std::string key, value;
        if (!(iss >> key)) continue;
        iss >> value;  // may stay empty
        double d = 0.0;
        int iv = 0;
        auto num = [&](double* dst) -> bool {
            if (!to_double(value, &d)) {
                af::set_error("%s:%d: key %s needs a number, got '%s'", path, lineno, key.c_str(),
                              value.c_str());
                return false;
            }
            *dst = d;
            return true;
        };
        if (key == "DEMfile") copy_str(p->dem_file, sizeof(p->dem_file), value);
        else if (key == "bcifile") copy_str(p->bci_file, sizeof(p->bci_file), value);
        else if (key == "bdyfile") copy_str(p->bdy_file, sizeof(p->bdy_file), value);
        else if (key == "dirroot") copy_str(p->output_dir, sizeof(p->output_dir), value);
        else if (key == "sim_time") { if (!num(&p->sim_time)) return AF_E_IO; }
        else if (key == "saveint") { if (!num(&p->save_int)) return AF_E_IO; }
        else if (key == "massint") { if (!num(&p->mass_int)) return AF_E_IO; }
        else if (key == "initial_tstep") { if (!num(&p->init_tstep)) return AF_E_IO; }
        else if (key == "cfl") { if (!num(&p->cfl)) return AF_E_IO; }
        else if (key == "fpfric") { if (!num(&p->fpn)) return AF_E_IO; }
        else if (key == "depth_thresh") { if (!num(&p->depth_thresh)) return AF_E_IO; }
        else if (key == "rainfile") copy_str(p->rain_file, sizeof(p->rain_file), value);
        else if (key == "stagefile") copy_str(p->stage_file, sizeof(p->stage_file), value);
        else if (key == "stageint") { if (!num(&p->stage_int)) return AF_E_IO; }
        else if (key == "output_format") {
            if (!to_int(value, &iv)) {
                af::set_error("%s:%d: output_format needs an integer", path, lineno);
                return AF_E_IO;
            }
            p->output_format = iv;
        } else if (key == "voutput") p->voutput = 1;
        else if (key == "async_output") {
            if (!to_int(value, &iv)) {
                af::set_error("%s:%d: async_output needs an integer", path, lineno);
                return AF_E_IO;
            }
            p->async_output = (iv != 0);
        }
        // any other key is ignored, as in the reference
    }
    return AF_OK;
}

af_case* af_case_new(void) {
    af_case* c = new af_case();
    af_par_defaults(&c->par);
    return c;
}

void af_case_free(af_case* c) { delete c; }

// read_dem: six positional header pairs (names ignored), then nrows rows of
// ncols values, north row first.  Values that fail to parse stay 0, as
// fscanf leaves the zero-initialised vector untouched.
namespace {

inline bool is_ws(char ch) {
    return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f';
}

// Exact powers of ten: 10^k is a double exactly for k <= 22.
constexpr double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                               1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                               1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

// strtod for the plain decimal tokens a DEM holds ([+-]digits[.digits]
// followed by whitespace or the end of the piece), when the digits form an
// integer w <= 2^53 with at most 22 of them after the point: then w and
// 10^f are doubles exactly and w / 10^f is one correctly rounded division,
// which is the correctly rounded value of the token, as strtod returns it
// (Clinger's fast path).  Returns false, leaving *x unset, for every other
// token (exponents, more digits, nan/inf, hex, trailing text), which the
// caller gives to strtod itself.
inline bool parse_plain_decimal(const char* q, const char* e, double* x, const char** after) {
    const char* p = q;
    bool neg = false;
    if (p < e && (*p == '+' || *p == '-')) neg = (*p++ == '-');
    unsigned long long w = 0;
    int sig = 0, frac = 0, digits = 0;
    bool point = false;
    for (; p < e; p++) {
        const char ch = *p;
        if (ch >= '0' && ch <= '9') {
            digits++;
            if (point) frac++;
            if (w == 0 && ch == '0') continue;  // leading zeros add no significant digit
            if (++sig > 16) return false;       // 10^16 > 2^53: beyond the exact range
            w = w * 10 + (unsigned long long)(ch - '0');
        } else if (ch == '.' && !point) {
            point = true;
        } else {
            break;
        }
    }
    if (digits == 0 || (p < e && !is_ws(*p))) return false;
    if (w > (1ull << 53) || frac > 22) return false;
    const double v = frac ? (double)w / kPow10[frac] : (double)w;
    *x = neg ? -v : v;
    *after = p;
    return true;
}

// The value loop of read_dem (io.cpp:150-155): fscanf("%lf") per cell, in
// order, until a token does not convert (every later fscanf fails on that
// same token, so the remaining cells keep 0).  The text is cut at
// whitespace into one piece per thread; each piece is parsed with strtod
// (the conversion fscanf uses) and the pieces are concatenated in order up
// to the first piece that stopped on a bad token.
void parse_values(const char* beg, const char* end, std::vector<double>& out) {
    const size_t n = out.size();
    const size_t bytes = (size_t)(end - beg);
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (bytes < (1u << 20)) nt = 1;
    std::vector<const char*> cut(nt + 1);
    cut[0] = beg;
    cut[nt] = end;
    for (unsigned t = 1; t < nt; t++) {
        const char* p = beg + bytes * t / nt;
        if (p < cut[t - 1]) p = cut[t - 1];
        while (p < end && !is_ws(*p)) p++;
        cut[t] = p;
    }
    std::vector<std::vector<double>> part(nt);
    std::vector<char> stopped(nt, 0);
    auto work = [&](unsigned t) {
        const char* q = cut[t];
        const char* e = cut[t + 1];
        std::vector<double>& v = part[t];
        v.reserve((size_t)(e - q) / 6 + 1);
        while (true) {
            while (q < e && is_ws(*q)) q++;
            if (q >= e) break;
            double x;
            const char* fast = nullptr;
            if (parse_plain_decimal(q, e, &x, &fast)) {
                v.push_back(x);
                q = fast;
                continue;
            }
            char* after = nullptr;
            x = strtod(q, &after);
            if (after == q) {
                stopped[t] = 1;
                break;
            }
            v.push_back(x);
            q = after;
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++) th.emplace_back(work, t);
        for (auto& x : th) x.join();
    }
    size_t k = 0;
    for (unsigned t = 0; t < nt && k < n; t++) {
        const size_t m = std::min(part[t].size(), n - k);
        std::copy(part[t].begin(), part[t].begin() + m, out.begin() + k);
        k += m;
        if (stopped[t]) break;
    }
}

// "%.4f " of x, as printf writes it.  printf rounds the exact binary value
// to 4 decimals.  For |x| < 1e5, y = |x| * 1e4 < 2^30 is the exact product
// to within half an ulp of y, at most 6e-8; so whenever y is more than 1e-6
// away from a half-integer, the nearest integer of y is printf's last digit
// and the digits are formed here directly.  Near-halves, larger values,
// infinities and NaN go to snprintf.
int fmt4(char* out, size_t cap, double x) {
    const double ax = fabs(x);
    if (!(ax < 1e5)) return snprintf(out, cap, "%.4f ", x);
    const double y = ax * 10000.0;
    const double fl = floor(y);
    const double fr = y - fl;
    if (fabs(fr - 0.5) < 1e-6) return snprintf(out, cap, "%.4f ", x);
    const unsigned long long q = (unsigned long long)fl + (fr > 0.5 ? 1 : 0);
    char tmp[32];
    int n = 0;
    unsigned long long ip = q / 10000, fp = q % 10000;
    // fraction digits (reversed), then the point, then the integer digits
    for (int d = 0; d < 4; d++) {
        tmp[n++] = (char)('0' + fp % 10);
        fp /= 10;
    }
    tmp[n++] = '.';
    do {
        tmp[n++] = (char)('0' + ip % 10);
        ip /= 10;
    } while (ip);
    int k = 0;
    if (std::signbit(x)) out[k++] = '-';  // printf keeps the sign of -0.0 and of tiny negatives
    while (n) out[k++] = tmp[--n];
    out[k++] = ' ';
    out[k] = 0;
    return k;
}

}  // namespace

int af_case_read_dem(af_case* c, const char* path) {
    if (!c || !path) {
        af::set_error("read_dem: null argument");
        return AF_E_INVALID;
    }
    FILE* fp = fopen(path, "rb");
    if (!fp) {
        af::set_error("Cannot open DEM file %s", path);
        return AF_E_IO;
    }
    std::vector<char> buf;
    if (fseek(fp, 0, SEEK_END) == 0) {
        const long size = ftell(fp);
        rewind(fp);
        if (size > 0) {
            buf.resize((size_t)size);
            buf.resize(fread(buf.data(), 1, (size_t)size, fp));
        }
    }
    fclose(fp);
    buf.push_back(0);
    const char* s = buf.data();
    auto skip_ws = [&]() { while (*s && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r' || *s == '\v' || *s == '\f')) s++; };
    auto token = [&]() -> bool {  // %s
        skip_ws();
        if (!*s) return false;
        while (*s && !(*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r' || *s == '\v' || *s == '\f')) s++;
        return true;
    };
    auto read_int = [&](int* out) -> bool {  // %d
        skip_ws();
        char* e = nullptr;
        long v = strtol(s, &e, 10);
        if (e == s) return false;
        s = e;
        *out = (int)v;
        return true;
    };
    auto read_dbl = [&](double* out) -> bool {  // %lf
        skip_ws();
        char* e = nullptr;
        double v = strtod(s, &e);
        if (e == s) return false;
        s = e;
        *out = v;
        return true;
    };
    int ncols = 0, nrows = 0;
    double xll = 0, yll = 0, cs = 0, nod = 0;
    bool ok = token() && read_int(&ncols) && token() && read_int(&nrows) && token() &&
              read_dbl(&xll) && token() && read_dbl(&yll) && token() && read_dbl(&cs) &&
              token() && read_dbl(&nod);
    if (!ok || ncols <= 0 || nrows <= 0) {
        af::set_error("%s: malformed DEM header", path);
        return AF_E_IO;
    }
    c->ncols = ncols;
    c->nrows = nrows;
    c->xll = xll;
    c->yll = yll;
    c->cellsize = cs;
    c->nodata = nod;
    c->dem.assign((size_t)ncols * nrows, 0.0);
    parse_values(s, buf.data() + buf.size() - 1, c->dem);
    c->have_dem = true;
    return AF_OK;
}

// read_bci: 255-character line chunks (fgets into 256 bytes), point lines
// `P x y TYPE [extra]` in world coordinates, side lines `S a b FREE`.
int af_case_read_bci(af_case* c, const char* path) {
    if (!c || !path) {
        af::set_error("read_bci: null argument");
        return AF_E_INVALID;
    }
    if (!c->have_dem) {
        af::set_error("read_bci: read the DEM first");
        return AF_E_STATE;
    }
    FILE* fp = fopen(path, "r");
    if (!fp) {
        af::set_error("Cannot open BCI file %s", path);
        return AF_E_IO;
    }
    const double tly = c->yll + c->nrows * c->cellsize;  // io.cpp:170
    char line[256];
    while (fgets(line, sizeof(line), fp)) {
        if (line[0] == '#' || line[0] == '\n' || line[0] == '\r') continue;
        char type = line[0];
        if (type == 'P') {
            double x = 0, y = 0;
            char bctype[16] = {0}, extra[64] = {0};
            int n = sscanf(line, "%c %lf %lf %15s %63s", &type, &x, &y, bctype, extra);
            if (n < 4) continue;
            Point pb;
            pb.x = (int)((x - c->xll) / c->cellsize);  // io.cpp:185-186, truncation
            pb.y = (int)((tly - y) / c->cellsize);
            pb.value = 0.0;
            if (strcmp(bctype, "QVAR") == 0 && n >= 5) {
                pb.type = AF_BC_QVAR;
                pb.name = extra;
            } else if (strcmp(bctype, "QFIX") == 0 && n >= 5) {
                pb.type = AF_BC_QFIX;
                pb.value = atof(extra);
            } else if (strcmp(bctype, "FREE") == 0) {
                pb.type = AF_BC_FREE;
            } else if (strcmp(bctype, "HFIX") == 0 && n >= 5) {
                pb.type = AF_BC_HFIX;
                pb.value = atof(extra);
            } else if (strcmp(bctype, "HVAR") == 0 && n >= 5) {
                pb.type = AF_BC_HVAR;
                pb.name = extra;
            } else {
                continue;  // unknown type: skipped (io.cpp:214-216)
            }
            c->points.push_back(pb);
        } else if (type == 'N' || type == 'S' || type == 'E' || type == 'W') {
            int a = 0, b = 0;
            char bctype[16] = {0};
            int n = sscanf(line, "%c %d %d %15s", &type, &a, &b, bctype);
            bool is_free = false;
            if (n >= 4 && strcmp(bctype, "FREE") == 0) {
                is_free = true;
            } else if (n >= 2) {
                char bctype2[16] = {0};
                if (sscanf(line, "%c %15s", &type, bctype2) == 2 && strcmp(bctype2, "FREE") == 0)
                    is_free = true;
            }
            if (is_free) c->line_sides.push_back(type);
        }
    }
    fclose(fp);
    return AF_OK;
}

// read_bdy: first line skipped; a line that starts with a QVAR/HVAR point's
// name followed by end-of-line, space, CR or LF starts a section whose next
// line holds the count and the count lines after it `value time`.  Only the
// first point (in file order) with that name receives the section.
int af_case_read_bdy(af_case* c, const char* path) {
    if (!c || !path) {
        af::set_error("read_bdy: null argument");
        return AF_E_INVALID;
    }
    std::ifstream f(path);
    if (!f.is_open()) {
        af::set_error("Cannot open BDY file %s", path);
        return AF_E_IO;
    }
    // A section header is a line that starts with a point's name followed by
    // a space, CR or the end of the line; the first QVAR/HVAR point (in file
    // order) with that name takes the series (io.cpp:266-289).  Names hold
    // no whitespace, so that is the line's text up to its first space or CR
    // looked up among the names: one hash lookup per line instead of a scan
    // of every point.
    std::unordered_map<std::string, size_t> first;
    for (size_t k = 0; k < c->points.size(); k++) {
        const Point& pb = c->points[k];
        if ((pb.type != AF_BC_QVAR && pb.type != AF_BC_HVAR) || pb.name.empty()) continue;
        first.emplace(pb.name, k);
    }
    std::string line;
    std::getline(f, line);
    double val = 0.0, time = 0.0;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        const size_t cut = line.find_first_of(" \n\r");
        const auto hit = first.find(cut == std::string::npos ? line : line.substr(0, cut));
        if (hit != first.end()) {
            Point& pb = c->points[hit->second];
            {
                if (!std::getline(f, line)) line.clear();
                int n = 0;
                if (sscanf(line.c_str(), "%d", &n) != 1) n = 0;
                for (int i = 0; i < n; i++) {
                    if (!std::getline(f, line)) line.clear();
                    double v2, t2;
                    int got = sscanf(line.c_str(), "%lf %lf", &v2, &t2);
                    if (got >= 1) val = v2;
                    if (got >= 2) time = t2;
                    pb.times.push_back(time);
                    pb.values.push_back(val);
                }
            }
        }
    }
    return AF_OK;
}

// read_rain: first line skipped; rows `rate time`, rate in m/s.
int af_case_read_rain(af_case* c, const char* path) {
    if (!c || !path) {
        af::set_error("read_rain: null argument");
        return AF_E_INVALID;
    }
    std::ifstream f(path);
    if (!f.is_open()) {
        c->rain_enabled = false;
        af::set_error("Cannot open rain file %s, rainfall disabled", path);
        return AF_E_IO;
    }
    std::string line;
    std::getline(f, line);
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        double rate, time;
        if (sscanf(line.c_str(), "%lf %lf", &rate, &time) == 2) {
            c->rain_t.push_back(time);
            c->rain_r.push_back(rate);
        }
    }
    if (!c->rain_t.empty()) c->rain_enabled = true;
    return AF_OK;
}

// The input sequence of main.cpp:24-92.
// read_stage_file (src/io.cpp:50-85): "x y [name]" per line, empty and
// '#' lines skipped, world coordinates truncated to grid cells from the top
// left corner, unnamed gauges called Stage<k>.  A file with no gauge is an
// error (the reference disables stage output then).
int af_case_read_stage(af_case* c, const char* path) {
    if (!c || !path) {
        af::set_error("read_stage: null argument");
        return AF_E_INVALID;
    }
    if (!c->have_dem) {
        af::set_error("read_stage: the DEM header is needed first");
        return AF_E_STATE;
    }
    std::ifstream f(path);
    if (!f.is_open()) {
        af::set_error("Cannot open stage file %s", path);
        return AF_E_IO;
    }
    c->gx.clear();
    c->gy.clear();
    c->gname.clear();
    const double tly = c->yll + c->nrows * c->cellsize;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        double x, y;
        if (!(iss >> x >> y)) continue;
        std::string name;
        iss >> name;
        if (name.empty()) name = "Stage" + std::to_string(c->gname.size() + 1);
        c->gx.push_back((int32_t)((x - c->xll) / c->cellsize));
        c->gy.push_back((int32_t)((tly - y) / c->cellsize));
        c->gname.push_back(name);
    }
    if (c->gname.empty()) {
        af::set_error("no stage points in %s", path);
        return AF_E_IO;
    }
    return AF_OK;
}

int af_case_stage(af_case* c, int32_t i, int32_t* grid_x, int32_t* grid_y, char* name,
                  int32_t cap, int32_t* n) {
    if (!c) return AF_E_INVALID;
    if (n) *n = (int32_t)c->gname.size();
    if (i < 0) return AF_OK;
    if (i >= (int32_t)c->gname.size()) {
        af::set_error("stage %d of %zu", i, c->gname.size());
        return AF_E_INVALID;
    }
    if (grid_x) *grid_x = c->gx[i];
    if (grid_y) *grid_y = c->gy[i];
    if (name && cap > 0) copy_str(name, (size_t)cap, c->gname[i]);
    return AF_OK;
}

int af_case_load(af_case* c, const char* par_path) {
    if (!c || !par_path) {
        af::set_error("case_load: null argument");
        return AF_E_INVALID;
    }
    af_par par;
    const int rc = af_read_par(par_path, &par);
    if (rc) return rc;
    return af_case_load_par(c, &par);
}

// The inputs a parameter set names, in the order of main.cpp:48-113.
int af_case_load_par(af_case* c, const af_par* par) {
    if (!c || !par) {
        af::set_error("case_load: null argument");
        return AF_E_INVALID;
    }
    c->par = *par;
    int rc;
    if ((rc = af_case_read_dem(c, c->par.dem_file))) return rc;
    if ((rc = af_case_read_bci(c, c->par.bci_file))) return rc;
    if ((rc = af_case_read_bdy(c, c->par.bdy_file))) return rc;
    if (c->par.rain_file[0]) (void)af_case_read_rain(c, c->par.rain_file);  // optional
    // optional, as main.cpp:95-113: a missing or empty file only disables the output
    if (c->par.stage_file[0] && af_case_read_stage(c, c->par.stage_file) != AF_OK) {
        c->gx.clear();
        c->gy.clear();
        c->gname.clear();
    }
    return AF_OK;
}

int af_case_par(af_case* c, af_par* p) {
    if (!c || !p) return AF_E_INVALID;
    *p = c->par;
    return AF_OK;
}

int af_case_dem_header(af_case* c, int32_t* ncols, int32_t* nrows, double* xll, double* yll,
                       double* cellsize, double* nodata) {
    if (!c || !c->have_dem) {
        af::set_error("no DEM loaded");
        return AF_E_STATE;
    }
    if (ncols) *ncols = c->ncols;
    if (nrows) *nrows = c->nrows;
    if (xll) *xll = c->xll;
    if (yll) *yll = c->yll;
    if (cellsize) *cellsize = c->cellsize;
    if (nodata) *nodata = c->nodata;
    return AF_OK;
}

int af_case_line_bcs(af_case* c, char* sides, int32_t cap, int32_t* n) {
    if (!c || !n) return AF_E_INVALID;
    *n = (int32_t)c->line_sides.size();
    for (int32_t i = 0; i < *n && i < cap && sides; i++) sides[i] = c->line_sides[i];
    return AF_OK;
}

int af_case_point(af_case* c, int32_t i, int32_t* x, int32_t* y, int32_t* type, double* value,
                  int32_t* series_len, char* name, int32_t cap) {
    if (!c || i < 0 || i >= (int32_t)c->points.size()) {
        af::set_error("point %d out of range", i);
        return AF_E_INVALID;
    }
    const Point& p = c->points[i];
    if (x) *x = p.x;
    if (y) *y = p.y;
    if (type) *type = p.type;
    if (value) *value = p.value;
    if (series_len) *series_len = (int32_t)p.times.size();
    if (name && cap > 0) copy_str(name, (size_t)cap, p.name);
    return AF_OK;
}

int af_case_point_series(af_case* c, int32_t i, double* times, double* values, int32_t cap) {
    if (!c || i < 0 || i >= (int32_t)c->points.size()) return AF_E_INVALID;
    const Point& p = c->points[i];
    for (size_t k = 0; k < p.times.size() && (int32_t)k < cap; k++) {
        if (times) times[k] = p.times[k];
        if (values) values[k] = p.values[k];
    }
    return AF_OK;
}

int af_case_rain(af_case* c, int32_t* enabled, double* times, double* rates, int32_t cap,
                 int32_t* n) {
    if (!c) return AF_E_INVALID;
    if (enabled) *enabled = c->rain_enabled ? 1 : 0;
    if (n) *n = (int32_t)c->rain_t.size();
    for (size_t k = 0; k < c->rain_t.size() && (int32_t)k < cap; k++) {
        if (times) times[k] = c->rain_t[k];
        if (rates) rates[k] = c->rain_r[k];
    }
    return AF_OK;
}

// Scenario for the step: constants of main.cpp:115-127, line flags of
// main.cpp:62-70, points in file order (main.cpp:77-85 keeps file order
// within the flow and the other group, which the per-cell ordering of the
// fused step preserves).
int af_case_scenario(af_case* c, af_scenario_desc* d) {
    if (!c || !d) return AF_E_INVALID;
    if (!c->have_dem) {
        af::set_error("no DEM loaded");
        return AF_E_STATE;
    }
    memset(d, 0, sizeof(*d));
    d->nx = c->ncols;
    d->ny = c->nrows;
    d->dx = c->cellsize;
    d->nodata = c->nodata;
    d->dem = c->dem.data();
    d->fpn = c->par.fpn;
    d->depth_thresh = c->par.depth_thresh;
    d->max_hflow = 10.0;
    d->init_tstep = c->par.init_tstep;
    d->cfl = c->par.cfl;
    d->g = 9.81;
    d->theta = 1.0;
    d->sim_time = c->par.sim_time;
    d->mass_int = c->par.mass_int;
    for (char s : c->line_sides) {
        if (s == 'W') d->west_free = 1;
        else if (s == 'E') d->east_free = 1;
        else if (s == 'N') d->north_free = 1;
        else if (s == 'S') d->south_free = 1;
    }
    c->px.clear(); c->py.clear(); c->ptype.clear(); c->poff.clear(); c->plen.clear();
    c->pvalue.clear(); c->st.clear(); c->sv.clear();
    for (const Point& p : c->points) {
        c->px.push_back(p.x);
        c->py.push_back(p.y);
        c->ptype.push_back(p.type);
        c->pvalue.push_back(p.value);
        c->poff.push_back((int32_t)c->st.size());
        c->plen.push_back((int32_t)p.times.size());
        c->st.insert(c->st.end(), p.times.begin(), p.times.end());
        c->sv.insert(c->sv.end(), p.values.begin(), p.values.end());
    }
    d->n_points = (int32_t)c->points.size();
    d->pt_x = c->px.data();
    d->pt_y = c->py.data();
    d->pt_type = c->ptype.data();
    d->pt_value = c->pvalue.data();
    d->pt_series_off = c->poff.data();
    d->pt_series_len = c->plen.data();
    d->n_series = (int32_t)c->st.size();
    d->series_t = c->st.data();
    d->series_v = c->sv.data();
    d->rain_enabled = c->rain_enabled ? 1 : 0;
    d->n_rain = (int32_t)c->rain_t.size();
    d->rain_t = c->rain_t.data();
    d->rain_r = c->rain_r.data();
    d->h0 = nullptr;
    return AF_OK;
}

double af_interp_series(const double* times, const double* values, int32_t n, double empty_value,
                        double t) {
    if (n <= 0) return empty_value;
    if (t <= times[0]) return values[0];
    if (t >= times[n - 1]) return values[n - 1];
    for (int32_t i = 0; i < n - 1; i++) {
        if (t >= times[i] && t < times[i + 1]) {
            double ratio = (t - times[i]) / (times[i + 1] - times[i]);
            return values[i] + ratio * (values[i + 1] - values[i]);
        }
    }
    return values[n - 1];
}

namespace {

// The grid body shared by both ESRI-ASCII writers: "%.4f " per value and a
// newline per row, north row first.
int write_asc_body(FILE* fp, const char* path, const double* data, int32_t ncols, int32_t nrows) {
    // rows are formatted ("%.4f " per value, a newline per row) by several
    // threads into per-block buffers and written in order
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if ((long long)ncols * nrows < (1 << 20)) nt = 1;
    const int32_t rows_per_batch = (int32_t)std::max<long long>(1, (1LL << 22) / ncols) * nt;
    std::vector<std::string> blk(nt);
    for (int32_t j0 = 0; j0 < nrows; j0 += rows_per_batch) {
        const int32_t j1 = std::min(nrows, j0 + rows_per_batch);
        auto work = [&](unsigned t) {
            const int32_t a = j0 + (int32_t)((long long)(j1 - j0) * t / nt);
            const int32_t b = j0 + (int32_t)((long long)(j1 - j0) * (t + 1) / nt);
            std::string& o = blk[t];
            o.clear();
            char tmp[64];
            for (int32_t j = a; j < b; j++) {
                for (int32_t i = 0; i < ncols; i++) {
                    const int len = fmt4(tmp, sizeof tmp, data[i + (long long)j * ncols]);
                    o.append(tmp, (size_t)len);
                }
                o.push_back('\n');
            }
        };
        if (nt == 1) {
            work(0);
        } else {
            std::vector<std::thread> th;
            for (unsigned t = 0; t < nt; t++) th.emplace_back(work, t);
            for (auto& x : th) x.join();
        }
        for (unsigned t = 0; t < nt; t++) fwrite(blk[t].data(), 1, blk[t].size(), fp);
    }
    const bool ok = (fclose(fp) == 0);
    if (!ok) {
        af::set_error("write error on %s", path);
        return AF_E_IO;
    }
    return AF_OK;
}

FILE* open_asc(const char* path) {
    FILE* fp = fopen(path, "w");
    if (!fp) af::set_error("Cannot open %s for writing", path);
    return fp;
}

}  // namespace

// write_asc (io.cpp:314-334): corners "%.1f", NODATA_value always -9999.
int af_write_asc(const char* path, const double* data, int32_t ncols, int32_t nrows, double xll,
                 double yll, double cellsize) {
    if (!path || !data || ncols <= 0 || nrows <= 0) return AF_E_INVALID;
    FILE* fp = open_asc(path);
    if (!fp) return AF_E_IO;
    fprintf(fp, "ncols         %d\n", ncols);
    fprintf(fp, "nrows         %d\n", nrows);
    fprintf(fp, "xllcorner     %.1f\n", xll);
    fprintf(fp, "yllcorner     %.1f\n", yll);
    fprintf(fp, "cellsize      %.1f\n", cellsize);
    fprintf(fp, "NODATA_value  -9999\n");
    return write_asc_body(fp, path, data, ncols, nrows);
}

// AsyncOutputManager::write_asc_file (async_output.cpp:158-176): corners
// "%.6f", NODATA_value the grid's own value as "%.0f".
int af_write_asc_async(const char* path, const double* data, int32_t ncols, int32_t nrows,
                       double xll, double yll, double cellsize, double nodata) {
    if (!path || !data || ncols <= 0 || nrows <= 0) return AF_E_INVALID;
    FILE* fp = open_asc(path);
    if (!fp) return AF_E_IO;
    fprintf(fp, "ncols         %d\n", ncols);
    fprintf(fp, "nrows         %d\n", nrows);
    fprintf(fp, "xllcorner     %.6f\n", xll);
    fprintf(fp, "yllcorner     %.6f\n", yll);
    fprintf(fp, "cellsize      %.1f\n", cellsize);
    fprintf(fp, "NODATA_value  %.0f\n", nodata);
    return write_asc_body(fp, path, data, ncols, nrows);
}

}  // extern "C"
