/**
 * gauge_io.h -- gauge configuration files beyond the reference's raw reader
 * (load_gauge.h): the NERSC archive format and ILDG (LIME container), the
 * formats of the reference's utilities refs/utils_nersc_lat.cpp:19-80
 * (read_nersc_lattice), :82-168 (save_nersc_lattice), :185-207 (payload
 * location), :230-297 (read_ildg_lattice), plus the gauge observables those
 * files carry.
 *
 * Links cross this interface in the reference's layout: gauge[mu] is
 * complex<double>[V*9], site eo(x) (include/operator.h:62-67), row-major 3x3.
 * Files store sites lexicographically (x fastest), per site the four
 * directions x, y, z, t, per link the rows of the 3x3 matrix (all three, or
 * the first two with row 2 = conj(row0 x row1)).
 *
 * The reference's readers locate the payload from the file size and check
 * nothing; these parse the NERSC header and the LIME records, and verify the
 * header's LINK_TRACE and PLAQUETTE against the links read.
 */
#ifndef GAUGE_IO_H
#define GAUGE_IO_H

#include <complex>
#include <cstdint>
#include <string>

struct gauge_file_info {
    int dims[4] = {0, 0, 0, 0};
    int rows = 3;              // stored rows per link: 3, or 2 (NERSC 4D_SU3_GAUGE)
    int bytes_per_real = 8;    // 4 or 8
    bool big_endian = true;
    bool has_plaquette = false, has_link_trace = false, has_checksum = false;
    double header_plaquette = 0.0, header_link_trace = 0.0;
    uint32_t header_checksum = 0;
    double plaquette = 0.0;    // computed from the links read
    double link_trace = 0.0;
    uint32_t checksum = 0;     // 32-bit sum of the payload words (file byte order)
};

// status codes (0 = ok)
enum {
    GAUGE_IO_OK = 0,
    GAUGE_IO_EOPEN = -1,      // file cannot be opened / written
    GAUGE_IO_EFORMAT = -2,    // header or record structure not understood
    GAUGE_IO_EDIMS = -3,      // lattice size in the file differs from site_vec
    GAUGE_IO_ESHORT = -4,     // payload shorter than the lattice
    GAUGE_IO_ECHECK = -5,     // header link trace / plaquette disagree with the data
    GAUGE_IO_ECHECKSUM = -6,  // header checksum disagrees with the payload
};

// Average plaquette (1/(6V)) sum_x sum_{mu<nu} Re tr P_{mu nu}(x) / 3, periodic.
double gauge_plaquette(std::complex<double> *const gauge[4], const int *site_vec);
// Average link trace (1/(4V)) sum_{x,mu} Re tr U_mu(x) / 3.
double gauge_link_trace(std::complex<double> *const gauge[4], const int *site_vec);

// NERSC: text header BEGIN_HEADER ... END_HEADER (DATATYPE 4D_SU3_GAUGE or
// 4D_SU3_GAUGE_3x3, FLOATING_POINT IEEE{32,64}{BIG,LITTLE}, DIMENSION_1..4),
// then the payload.  A header checksum is compared with the 32-bit word sum
// of the payload when check_checksum is set (conventions differ between
// writers, so it is off by default); LINK_TRACE and PLAQUETTE are always
// checked (tolerance 1e-9, 1e-5 for 32-bit payloads).
int read_nersc_gauge(std::complex<double> *gauge[4], const std::string &filename,
                     const int *site_vec, gauge_file_info *info = nullptr,
                     bool check_checksum = false);
// rows 2 or 3, bytes_per_real 4 or 8; header with LINK_TRACE, PLAQUETTE and
// CHECKSUM (the 32-bit word sum of the payload as written).
int write_nersc_gauge(std::complex<double> *const gauge[4], const std::string &filename,
                      const int *site_vec, int rows = 2, int bytes_per_real = 8,
                      bool big_endian = true);

// ILDG: LIME records (144-byte headers, payloads padded to 8 bytes);
// "ildg-format" gives precision and lattice size, "ildg-binary-data" holds
// the big-endian 3x3 links.
int read_ildg_gauge(std::complex<double> *gauge[4], const std::string &filename,
                    const int *site_vec, gauge_file_info *info = nullptr);
int write_ildg_gauge(std::complex<double> *const gauge[4], const std::string &filename,
                     const int *site_vec, int bytes_per_real = 8);

// The same with plain pointers, for FFI callers.  format 0 = NERSC, 1 = ILDG;
// L = {Lx, Ly, Lz, Lt}; gauge[mu] = V*9 complex doubles (re, im interleaved).
// ipcc_gauge_read fills, when not null, values = {plaquette, link trace,
// header plaquette, header link trace} and meta = {rows, bytes per real, big
// endian, header has plaquette, has link trace, has checksum, payload
// checksum, header checksum}; it returns a GAUGE_IO_* status.
extern "C" {
int ipcc_gauge_read(int format, const char *path, const int *L, double **gauge, double *values,
                    int *meta, int check_checksum);
int ipcc_gauge_write(int format, const char *path, const int *L, double **gauge, int rows,
                     int bytes_per_real, int big_endian);
double ipcc_gauge_plaquette(const int *L, double **gauge);
double ipcc_gauge_link_trace(const int *L, double **gauge);
}

#endif
