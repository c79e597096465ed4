/**
 * lattice_gauge.h -- gauge field container with the reference's interface
 * (include/lattice_gauge.h:11-26): four complex<double>[V*9] arrays, one per
 * direction x, y, z, t, at eo(x)*9 + row*3 + col.
 */
#ifndef LATTICE_LATTICE_GAUGE_H
#define LATTICE_LATTICE_GAUGE_H

#include <complex>

class lattice_gauge {
  public:
    int *site_vec;
    std::complex<double> *A[4];
    int *subgs;
    int size; // complex entries per direction: prod(subgs) * 9

    lattice_gauge(std::complex<double> *chroma_gauge[4], int *subgs1, int *site_vec1);
    // owning and zero-initialised (src/lattice_gauge.cpp:34-43)
    lattice_gauge(int *subgs1, int *site_vec1);
    ~lattice_gauge();
    // U_dd(site)[ll][mm]; one process holds the whole lattice, so no communication
    // (the reference gathers it over MPI, src/lattice_gauge.cpp:53-122)
    std::complex<double> peeksite(const int *site, int ll = 0, int mm = 0, int dd = 0);

  private:
    bool mem_flag;
};

#endif
