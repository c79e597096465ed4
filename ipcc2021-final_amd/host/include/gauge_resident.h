// gauge_resident.h -- keeps a lattice_gauge's links on the GPU between
// operator calls.  An extension of this build: the reference has no
// counterpart (its operators read the host links on every call).
//
// The reference's operator API takes the links as host arrays on every call
// (DslashNew, Dslashoffd, ...), so each call of this build's drop-in API
// uploads the whole gauge field (1.2 GB at 32^3x64).  A caller's own loop of
// operator applications can keep the device copy for a stretch of calls:
//
//     {
//         gauge_resident keep(U);     // uploads U once
//         for (...) DslashNew(src, dest, U, mass, false);   // no gauge upload
//     }                               // device copy released
//
// While a gauge_resident of U lives, the caller must not change U.A (the
// calls keep using the copy taken at construction); a second guard on the
// same U takes a fresh copy.  Guards on different gauge fields may coexist.
// Not thread-safe, like the rest of the API.
#pragma once

#include "lattice_gauge.h"

class gauge_resident {
public:
    explicit gauge_resident(lattice_gauge &U);
    ~gauge_resident();
    gauge_resident(const gauge_resident &) = delete;
    gauge_resident &operator=(const gauge_resident &) = delete;

private:
    lattice_gauge &U_;
};
