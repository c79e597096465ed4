// wd_runtime.cpp -- context cache and RAII wrappers over libwilson.so.
#include "wd_runtime.h"

#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>

namespace wdrt {

namespace {
std::mutex g_mu;
std::map<std::tuple<int, int, int, int>, wd_ctx *> g_ctx;
std::map<std::pair<wd_ctx *, int>, wd_field *> g_scratch;
// resident gauges: lattice_gauge object -> its device copy
std::map<const lattice_gauge *, Gauge *> g_resident;
} // namespace

void check(wd_ctx *ctx, int rc, const char *what)
{
    if (rc < 0)
        throw std::runtime_error(std::string(what) + ": " + wd_strerror(rc) + ": " +
                                 (ctx ? wd_last_error(ctx) : ""));
}

wd_ctx *context(const int *subgs, const int *site_vec)
{
    for (int mu = 0; mu < 4; ++mu)
        if (subgs[mu] != site_vec[mu])
            throw std::runtime_error(
                "the lattice is partitioned (subgs != site_vec): one GPU holds the whole "
                "lattice in this build");
    std::lock_guard<std::mutex> lock(g_mu);
    const auto key = std::make_tuple(site_vec[0], site_vec[1], site_vec[2], site_vec[3]);
    auto it = g_ctx.find(key);
    if (it != g_ctx.end())
        return it->second;
    const char *dev = std::getenv("WD_DEVICE");
    wd_ctx *c = nullptr;
    // Creating a device context draws from the C library's random() state on
    // this runtime (observed: a CGinvert whose start vector was drawn after the
    // first context creation differed from the reference's).  The caller's
    // rand() sequence is part of the reference's semantics (src/main.cpp:77-85,
    // src/invert.cpp:42-44), so glibc's state is switched to a scratch array
    // for the creation and switched back untouched afterwards.
    static char scratch[256];
    char *prev = initstate(1u, scratch, sizeof scratch);
    const int rc = wd_create(&c, site_vec, nullptr, dev ? std::atoi(dev) : 0);
    setstate(prev);
    if (rc != WD_OK)
        throw std::runtime_error(std::string("wd_create: ") + wd_strerror(rc));
    g_ctx[key] = c;
    return c;
}

Field::Field(wd_ctx *c, int prec) : ctx(c)
{
    check(ctx, wd_field_create(ctx, &f, prec, 1), "wd_field_create");
}

Field::Field(wd_ctx *c, const lattice_fermion &src) : Field(c)
{
    upload(src);
}

Field::~Field()
{
    wd_field_destroy(f);
}

void Field::upload(const lattice_fermion &src)
{
    check(ctx, wd_field_upload_ref(ctx, f, reinterpret_cast<const double *>(src.A), 0),
          "wd_field_upload_ref");
}

void Field::download(lattice_fermion &dst) const
{
    check(ctx, wd_field_download_ref(ctx, f, reinterpret_cast<double *>(dst.A), 0),
          "wd_field_download_ref");
}

Gauge::Gauge(wd_ctx *c, const lattice_gauge &U) : ctx(c)
{
    check(ctx, wd_gauge_create(ctx, &g, WD_F64, WD_RECON_18), "wd_gauge_create");
    const double *A[4];
    for (int mu = 0; mu < 4; ++mu)
        A[mu] = reinterpret_cast<const double *>(U.A[mu]);
    check(ctx, wd_gauge_upload_ref(ctx, g, A), "wd_gauge_upload_ref");
}

Gauge::~Gauge()
{
    wd_gauge_destroy(g);
}

GaugeRef::GaugeRef(wd_ctx *c, const lattice_gauge &U)
{
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_resident.find(&U);
        if (it != g_resident.end() && it->second->ctx == c) {
            g = it->second->g;
            return;
        }
    }
    own = new Gauge(c, U);
    g = own->g;
}

GaugeRef::~GaugeRef()
{
    delete own;
}

void keep_gauge(const lattice_gauge &U)
{
    Gauge *copy = new Gauge(context(U.subgs, U.site_vec), U);
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_resident.find(&U);
    if (it != g_resident.end()) {
        delete it->second;
        it->second = copy;
    } else {
        g_resident[&U] = copy;
    }
}

void release_gauge(const lattice_gauge &U)
{
    Gauge *copy = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_resident.find(&U);
        if (it == g_resident.end())
            return;
        copy = it->second;
        g_resident.erase(it);
    }
    delete copy;
}

wd_field *scratch_field(wd_ctx *ctx, int slot)
{
    std::lock_guard<std::mutex> lock(g_mu);
    wd_field *&f = g_scratch[std::make_pair(ctx, slot)];
    if (!f)
        check(ctx, wd_field_create(ctx, &f, WD_F64, 1), "wd_field_create");
    return f;
}

void upload(wd_ctx *ctx, wd_field *f, const lattice_fermion &src)
{
    check(ctx, wd_field_upload_ref(ctx, f, reinterpret_cast<const double *>(src.A), 0),
          "wd_field_upload_ref");
}

void download(wd_ctx *ctx, const wd_field *f, lattice_fermion &dst)
{
    check(ctx, wd_field_download_ref(ctx, f, reinterpret_cast<double *>(dst.A), 0),
          "wd_field_download_ref");
}

} // namespace wdrt

// ---- gauge_resident (host/include/gauge_resident.h) ---------------------------
#include "gauge_resident.h"

gauge_resident::gauge_resident(lattice_gauge &U) : U_(U)
{
    wdrt::keep_gauge(U_);
}

gauge_resident::~gauge_resident()
{
    wdrt::release_gauge(U_);
}
