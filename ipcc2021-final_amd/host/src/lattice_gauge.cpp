// lattice_gauge.cpp -- host container (src/lattice_gauge.cpp:22-50).
#include "lattice_gauge.h"

#include "operator.h"

using std::complex;

lattice_gauge::lattice_gauge(complex<double> *chroma_gauge[4], int *subgs1, int *site_vec1)
    : site_vec(site_vec1), subgs(subgs1), mem_flag(false)
{
    for (int i = 0; i < 4; i++)
        A[i] = chroma_gauge[i];
    size = subgs[0] * subgs[1] * subgs[2] * subgs[3] * 3 * 3;
}

lattice_gauge::lattice_gauge(int *subgs1, int *site_vec1)
    : site_vec(site_vec1), subgs(subgs1), mem_flag(true)
{
    size = subgs[0] * subgs[1] * subgs[2] * subgs[3] * 3 * 3;
    for (int i = 0; i < 4; i++)
        A[i] = new complex<double>[size]();
}

lattice_gauge::~lattice_gauge()
{
    if (mem_flag)
        for (int i = 0; i < 4; i++)
            delete[] A[i];
}

complex<double> lattice_gauge::peeksite(const int *site, int ll, int mm, int dd)
{
    int x[4] = {site[0], site[1], site[2], site[3]};
    return A[dd][site2eoindex(x, subgs) * 9 + ll * 3 + mm];
}
