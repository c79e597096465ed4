// load_gauge.cpp -- gauge input (src/load_gauge.cpp:29-181), one process.
#include "load_gauge.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "gauge_io.h"
#include "operator.h"

namespace {
double be_double(const unsigned char *p)
{
    uint64_t u = 0;
    for (int k = 0; k < 8; ++k)
        u = (u << 8) | p[k];
    double d;
    std::memcpy(&d, &u, sizeof d);
    return d;
}
} // namespace

int load_gauge(std::complex<double> *gauge[4], std::string &filename, int *subgs, int *site_vec)
{
    const long V = (long)subgs[0] * subgs[1] * subgs[2] * subgs[3];
    if (filename == "UNIT") {
        for (int mu = 0; mu < 4; ++mu)
            for (long c = 0; c < V * 9; c += 9) {
                gauge[mu][c + 0] = 1.0;
                gauge[mu][c + 4] = 1.0;
                gauge[mu][c + 8] = 1.0;
            }
        return 0;
    }
    if (filename == "RANDOM")
        return 0; // the reference's stub leaves the links untouched
    FILE *fh = std::fopen(filename.c_str(), "rb");
    if (!fh) {
        std::cerr << "error: can not open file: " << filename << std::endl;
        std::exit(EXIT_FAILURE);
    }
    // NERSC (text header) and ILDG (LIME container) files are recognised by
    // their first bytes and read by gauge_io; anything else is the
    // reference's raw format
    unsigned char head[12] = {0};
    const size_t got = std::fread(head, 1, sizeof head, fh);
    const bool nersc = got == 12 && std::memcmp(head, "BEGIN_HEADER", 12) == 0;
    const bool lime = got >= 4 && head[0] == 0x45 && head[1] == 0x67 && head[2] == 0x89 &&
                      head[3] == 0xab;
    if (nersc || lime) {
        std::fclose(fh);
        gauge_file_info info;
        const int rc = nersc ? read_nersc_gauge(gauge, filename, site_vec, &info)
                             : read_ildg_gauge(gauge, filename, site_vec, &info);
        if (rc != GAUGE_IO_OK) {
            std::cerr << "error: " << (nersc ? "NERSC" : "ILDG") << " gauge file " << filename
                      << " not read (status " << rc << ")" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        return 0;
    }
    std::rewind(fh);
    // 72 big-endian doubles per site, lexicographic over the full lattice
    std::vector<unsigned char> row((size_t)72 * 8 * site_vec[0]);
    int x[4];
    for (x[3] = 0; x[3] < site_vec[3]; ++x[3])
        for (x[2] = 0; x[2] < site_vec[2]; ++x[2])
            for (x[1] = 0; x[1] < site_vec[1]; ++x[1]) {
                if (std::fread(row.data(), row.size(), 1, fh) != 1) {
                    std::cerr << "error: gauge file too short: " << filename << std::endl;
                    std::fclose(fh);
                    std::exit(EXIT_FAILURE);
                }
                for (x[0] = 0; x[0] < site_vec[0]; ++x[0]) {
                    const long eo = site2eoindex(x, site_vec);
                    const unsigned char *s = row.data() + (size_t)72 * 8 * x[0];
                    for (int mu = 0; mu < 4; ++mu)
                        for (int c = 0; c < 9; ++c)
                            gauge[mu][eo * 9 + c] = std::complex<double>(
                                be_double(s + 8 * (mu * 18 + 2 * c)),
                                be_double(s + 8 * (mu * 18 + 2 * c + 1)));
                }
            }
    std::fclose(fh);
    return 0;
}
