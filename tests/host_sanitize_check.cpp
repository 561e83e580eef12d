// Drives the host-side readers and writers of the library (io.cpp,
// ncwriter.cpp) on the inputs named on the command line, for a build with
// the address and undefined-behaviour sanitizers.  Usage:
//   host_sanitize_check CASE.par TOKENS.dem.ascii OUTDIR
// Prints one line per call with its return code; the sanitizers abort on
// the first error they find.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "adaptflood.h"

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s CASE.par TOKENS.dem.ascii OUTDIR\n", argv[0]);
        return 2;
    }
    const std::string out = argv[3];
    int bad = 0;

    af_case* c = af_case_new();
    int rc = af_case_load(c, argv[1]);
    printf("case_load %d\n", rc);
    bad |= rc != 0;
    af_scenario_desc d;
    rc = af_case_scenario(c, &d);
    printf("case_scenario %d (%d x %d)\n", rc, d.nx, d.ny);
    bad |= rc != 0;
    af_case_free(c);

    // a DEM whose last token ends the file without a newline, with tokens
    // for both conversion routes
    c = af_case_new();
    rc = af_case_read_dem(c, argv[2]);
    printf("read_dem %d\n", rc);
    bad |= rc != 0;
    af_case_free(c);

    // a missing file is an error, not a crash
    c = af_case_new();
    rc = af_case_read_dem(c, (out + "/missing.dem.ascii").c_str());
    printf("read_dem missing %d\n", rc);
    bad |= rc == 0;
    af_case_free(c);

    const int nx = 37, ny = 23;
    std::vector<double> H(nx * ny), V(nx * ny);
    for (int i = 0; i < nx * ny; i++) {
        H[i] = 0.001 * i - 0.3;
        V[i] = std::sin(i * 0.1);
    }
    H[5] = -9999.0;
    H[6] = 1e6;
    H[7] = -0.0;
    H[8] = 0.00005;
    H[9] = 123.45675;
    H[10] = NAN;
    H[11] = INFINITY;
    rc = af_write_asc((out + "/h.asc").c_str(), H.data(), nx, ny, 0.0, 0.0, 5.0);
    printf("write_asc %d\n", rc);
    bad |= rc != 0;
    rc = af_write_asc_async((out + "/h_async.asc").c_str(), H.data(), nx, ny, 1.5, 2.5, 5.0,
                            -9999.0);
    printf("write_asc_async %d\n", rc);
    bad |= rc != 0;
    af_nc* nc = nullptr;
    rc = af_nc_create((out + "/out.nc").c_str(), nx, ny, 0.0, 0.0, 5.0, -9999.0, 1, &nc);
    printf("nc_create %d\n", rc);
    bad |= rc != 0;
    for (int k = 0; k < 3 && nc; k++) {
        rc = af_nc_append(nc, 60.0 * k, H.data(), V.data(), V.data());
        bad |= rc != 0;
    }
    printf("nc_append %d\n", rc);
    if (nc) {
        rc = af_nc_close(nc);
        printf("nc_close %d\n", rc);
        bad |= rc != 0;
    }
    printf(bad ? "FAILED\n" : "all calls returned as expected\n");
    return bad;
}
