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

// The device links for one operator call: the copy a live gauge_resident
// keeps for U, else a fresh upload released at the end of the call.
struct GaugeRef {
    wd_gauge *g = nullptr;
    GaugeRef(wd_ctx *c, const lattice_gauge &U);
    GaugeRef(const GaugeRef &) = delete;
    GaugeRef &operator=(const GaugeRef &) = delete;

private:
    Gauge *own = nullptr;

public:
    ~GaugeRef();
};

// Two F64 single-system device fields per context, kept between operator
// calls (a fresh allocation per call cost the drop-in API more than the
// kernel); slot 0 and slot 1 are distinct.  Valid until the context ends.
wd_field *scratch_field(wd_ctx *ctx, int slot);
void upload(wd_ctx *ctx, wd_field *f, const lattice_fermion &src);
void download(wd_ctx *ctx, const wd_field *f, lattice_fermion &dst);

// gauge_resident's registry (keyed by the lattice_gauge object)
void keep_gauge(const lattice_gauge &U);
void release_gauge(const lattice_gauge &U);

} // namespace wdrt
