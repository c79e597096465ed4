/**
 * load_gauge.h -- gauge input with the reference's interface
 * (include/load_gauge.h:13, src/load_gauge.cpp:29-181):
 *   "UNIT"    identity links
 *   "RANDOM"  left untouched (the reference's stub; zero links for a fresh
 *             lattice_gauge)
 *   path      raw big-endian FP64 file, 72 doubles per site in lexicographic
 *             order, [mu][row][col][re,im] per site, reordered to even-odd;
 *             a file that starts with a NERSC header or a LIME record is read
 *             as NERSC or ILDG (gauge_io.h), with its checks.
 * Returns 0, or exits with EXIT_FAILURE when the file cannot be opened or is
 * too short (the reference exits on open failure, src/load_gauge.cpp:104-108).
 */
#ifndef LOAD_GAUGE_H
#define LOAD_GAUGE_H

#include <complex>
#include <string>

#include "lattice_gauge.h"

int load_gauge(std::complex<double> *gauge[4], std::string &filename, int *subgs, int *site_vec);

#endif
