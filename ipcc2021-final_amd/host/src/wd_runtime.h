// wd_runtime.h -- glue between the reference-style C++ API and libwilson.so:
// one cached wd_ctx per lattice shape, RAII device fields and gauge fields.
#pragma once

#include "lattice_fermion.h"
#include "lattice_gauge.h"
#include "wilson.h"

namespace wdrt {

// Context for the lattice (periodic boundaries, the reference's only choice).
// subgs must equal site_vec: one GPU holds the whole lattice.  The device is
// HIP ordinal $WD_DEVICE (default 0).  Throws std::runtime_error on failure.
wd_ctx *context(const int *subgs, const int *site_vec);

// Throws std::runtime_error naming `what` when rc < 0.
void check(wd_ctx *ctx, int rc, const char *what);

struct Field {
    wd_ctx *ctx;
    wd_field *f = nullptr;
    explicit Field(wd_ctx *c, int prec = WD_F64);
    Field(wd_ctx *c, const lattice_fermion &src);
    ~Field();
    void upload(const lattice_fermion &src);
    void download(lattice_fermion &dst) const;
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;
};

struct Gauge {
    wd_ctx *ctx;
    wd_gauge *g = nullptr;
    Gauge(wd_ctx *c, const lattice_gauge &U);
    ~Gauge();
    Gauge(const Gauge &) = delete;
    Gauge &operator=(const Gauge &) = delete;
};

} // namespace wdrt
