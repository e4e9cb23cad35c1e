// "SDFO" 0.1.0 binary files (SdfIO.cpp:60-552): byte-identical writer and a
// reader with the reference's validation rules and messages.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "sdfo_internal.hpp"

namespace sdfo {

namespace {

constexpr uint32_t kVersion = 100; // 0 * 10000 + 1 * 100 + 0
constexpr uint32_t kFlagHessian = 1u, kFlagSparse = 2u, kFlagContact = 4u;
constexpr uint32_t kLayoutDense = 1, kLayoutSparse = 2, kLayoutDenseContact = 3,
                   kLayoutSparseContact = 4;

// 128-byte header, little endian, packed (SdfIO.cpp:62-80).
struct Header
{
    unsigned char b[128];
    Header() { std::memset(b, 0, sizeof(b)); }
    template <typename T>
    void put(int off, T v) { std::memcpy(b + off, &v, sizeof(T)); }
    template <typename T>
    T get(int off) const
    {
        T v;
        std::memcpy(&v, b + off, sizeof(T));
        return v;
    }
    // field offsets
    static constexpr int kVer = 4, kNx = 8, kBmin = 20, kBmax = 44, kVoxel = 68,
                         kFlags = 92, kReserved = 96;
};

Header makeHeader(const GridSpec& s, uint32_t flags)
{
    Header h;
    std::memcpy(h.b, "SDFO", 4);
    h.put<uint32_t>(Header::kVer, kVersion);
    for (int a = 0; a < 3; ++a)
    {
        h.put<uint32_t>(Header::kNx + 4 * a, uint32_t(s.n[a]));
        h.put<double>(Header::kBmin + 8 * a, s.bmin[a]);
        h.put<double>(Header::kBmax + 8 * a, s.bmax[a]);
        h.put<double>(Header::kVoxel + 8 * a, s.voxel[a]);
    }
    h.put<uint32_t>(Header::kFlags, flags);
    return h;
}

GridSpec specOf(const Header& h)
{
    GridSpec s;
    for (int a = 0; a < 3; ++a)
    {
        s.n[a] = int(h.get<uint32_t>(Header::kNx + 4 * a));
        s.bmin[a] = h.get<double>(Header::kBmin + 8 * a);
        s.bmax[a] = h.get<double>(Header::kBmax + 8 * a);
        s.voxel[a] = h.get<double>(Header::kVoxel + 8 * a);
    }
    return s;
}

template <typename T>
std::vector<T> download(const T* d, size_t count)
{
    std::vector<T> h(count);
    if (count) SDFO_HIP(hipMemcpy(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}

template <typename T>
T* upload(sdfo_grid& g, const std::vector<T>& h)
{
    if (h.empty()) return nullptr;
    T* d = g.alloc<T>(h.size());
    SDFO_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

bool readHeader(std::ifstream& in, Header& h)
{
    in.read(reinterpret_cast<char*>(h.b), 128);
    return in.gcount() == 128;
}

void validateCommon(const Header& h)
{
    if (std::memcmp(h.b, "SDFO", 4) != 0)
        throw Error(SDFO_ERR_FORMAT, "Invalid SdfOracle data magic number");
    const uint32_t v = h.get<uint32_t>(Header::kVer);
    if (v != kVersion)
        throw Error(SDFO_ERR_FORMAT, "Unsupported SdfOracle data version: " + std::to_string(v) +
                                         " (only 0.1.0 is supported)");
}

// Host views of a grid's arrays (the file layout's element types).
struct DenseHost
{
    GridSpec spec;
    const double *phi, *n3, *h6, *wit3;
    int stride;
    const uint8_t* cc;
    const double *cphi, *cn3;
    const int32_t* cbr;
};

struct SparseHost
{
    GridSpec spec;
    int bd, halo, nb;
    double band;
    bool hasHessian;
    int stride;
    const int32_t* bijk;
    const float *phi, *n3, *h6;
    const uint8_t* cc;
    const float *cphi, *cn3;
    const int32_t* cbr;
};

// saveSdf (SdfIO.cpp:93-183) from host arrays.
void writeDense(const DenseHost& d, std::ofstream& out)
{
    const size_t N = size_t(d.spec.total());
    const bool contact = d.stride > 0;
    Header h = makeHeader(d.spec, kFlagHessian | (contact ? kFlagContact : 0u));
    h.put<uint32_t>(Header::kReserved + 0, contact ? kLayoutDenseContact : kLayoutDense);
    if (contact) h.put<uint32_t>(Header::kReserved + 4, uint32_t(d.stride));
    out.write(reinterpret_cast<const char*>(h.b), 128);
    auto wr = [&](const void* p, size_t bytes) {
        out.write(reinterpret_cast<const char*>(p), std::streamsize(bytes));
    };
    const std::vector<double> zeros6(d.h6 ? 0 : 6 * N, 0.0);
    wr(d.phi, N * 8);
    wr(d.n3, 3 * N * 8);
    wr(d.h6 ? d.h6 : zeros6.data(), 6 * N * 8);
    wr(d.wit3, 3 * N * 8);
    if (contact)
    {
        const size_t NS = N * size_t(d.stride);
        std::vector<uint32_t> c32(d.cc, d.cc + N);
        wr(c32.data(), N * 4);
        wr(d.cphi, NS * 8);
        wr(d.cn3, 3 * NS * 8);
        wr(d.cbr, NS * 4);
    }
}

void saveDense(const sdfo_grid& g, std::ofstream& out)
{
    const size_t N = size_t(g.spec.total());
    DenseHost d{};
    d.spec = g.spec;
    d.stride = g.stride;
    const std::vector<double> phi = download(g.phi, N), n3 = download(g.n3, 3 * N),
                              h6 = download(g.h6, 6 * N), w3 = download(g.witness3, 3 * N);
    d.phi = phi.data();
    d.n3 = n3.data();
    d.h6 = h6.data();
    d.wit3 = w3.data();
    std::vector<uint8_t> cc;
    std::vector<double> cphi, cn3;
    std::vector<int32_t> cbr;
    if (g.stride > 0)
    {
        const size_t NS = N * size_t(g.stride);
        cc = download(g.ccount, N);
        cphi = download(g.cphi, NS);
        cn3 = download(g.cn3, 3 * NS);
        cbr = download(g.cbranch, NS);
        d.cc = cc.data();
        d.cphi = cphi.data();
        d.cn3 = cn3.data();
        d.cbr = cbr.data();
    }
    writeDense(d, out);
}

// saveSparseSdf (SdfIO.cpp:185-275) from host arrays.
void writeSparse(const SparseHost& d, std::ofstream& out)
{
    const int bd = d.bd;
    const size_t bvc = size_t(bd) * bd * bd;
    const bool contact = d.stride > 0 && d.nb > 0;
    Header h = makeHeader(d.spec, kFlagSparse | (d.hasHessian ? kFlagHessian : 0u) |
                                      (contact ? kFlagContact : 0u));
    h.put<uint32_t>(Header::kReserved + 0, contact ? kLayoutSparseContact : kLayoutSparse);
    h.put<uint32_t>(Header::kReserved + 4, uint32_t(bd));
    h.put<uint32_t>(Header::kReserved + 8, uint32_t(d.halo));
    h.put<uint32_t>(Header::kReserved + 12, uint32_t(d.nb));
    h.put<double>(Header::kReserved + 16, d.band);
    if (contact) h.put<uint32_t>(Header::kReserved + 24, uint32_t(d.stride));
    out.write(reinterpret_cast<const char*>(h.b), 128);
    const size_t S = size_t(d.stride);
    const std::vector<float> zeros6(d.hasHessian && !d.h6 ? 6 * bvc : 0, 0.0f);
    for (int b = 0; b < d.nb; ++b)
    {
        int32_t bh[4] = {d.bijk[3 * b], d.bijk[3 * b + 1], d.bijk[3 * b + 2], 0};
        out.write(reinterpret_cast<const char*>(bh), 16);
        out.write(reinterpret_cast<const char*>(&d.phi[b * bvc]), std::streamsize(bvc * 4));
        out.write(reinterpret_cast<const char*>(&d.n3[3 * b * bvc]), std::streamsize(3 * bvc * 4));
        if (d.hasHessian)
            out.write(reinterpret_cast<const char*>(d.h6 ? &d.h6[6 * b * bvc] : zeros6.data()),
                      std::streamsize(6 * bvc * 4));
        if (contact)
        {
            out.write(reinterpret_cast<const char*>(&d.cc[b * bvc]), std::streamsize(bvc));
            out.write(reinterpret_cast<const char*>(&d.cphi[b * bvc * S]), std::streamsize(bvc * S * 4));
            out.write(reinterpret_cast<const char*>(&d.cn3[3 * b * bvc * S]),
                      std::streamsize(3 * bvc * S * 4));
            out.write(reinterpret_cast<const char*>(&d.cbr[b * bvc * S]), std::streamsize(bvc * S * 4));
        }
    }
}

void saveSparse(const sdfo_grid& g, std::ofstream& out)
{
    const size_t bvc = size_t(g.blockDim) * g.blockDim * g.blockDim;
    const size_t cnt = size_t(g.nb) * bvc;
    SparseHost d{};
    d.spec = g.spec;
    d.bd = g.blockDim;
    d.halo = g.halo;
    d.nb = g.nb;
    d.band = g.band;
    d.hasHessian = g.hasHessian;
    d.stride = g.stride;
    const std::vector<int32_t> bijk = download(g.bijk, size_t(g.nb) * 3);
    const std::vector<float> phi = download(g.sphi, cnt), n3 = download(g.sn3, 3 * cnt);
    std::vector<float> h6;
    if (g.hasHessian && g.sh6) h6 = download(g.sh6, 6 * cnt);
    d.bijk = bijk.data();
    d.phi = phi.data();
    d.n3 = n3.data();
    d.h6 = h6.empty() ? nullptr : h6.data();
    const size_t S = size_t(g.stride);
    std::vector<uint8_t> cc;
    std::vector<float> cphi, cn3;
    std::vector<int32_t> cbr;
    if (g.stride > 0 && g.nb > 0)
    {
        cc = download(g.sccount, cnt);
        cphi = download(g.scphi, cnt * S);
        cn3 = download(g.scn3, 3 * cnt * S);
        cbr = download(g.scbranch, cnt * S);
        d.cc = cc.data();
        d.cphi = cphi.data();
        d.cn3 = cn3.data();
        d.cbr = cbr.data();
    }
    writeSparse(d, out);
}

// loadSdf (SdfIO.cpp:279-394).
std::unique_ptr<sdfo_grid> loadDense(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw Error(SDFO_ERR_IO, "Cannot open SdfOracle data file: " + path);
    Header h;
    if (!readHeader(in, h)) throw Error(SDFO_ERR_FORMAT, "Failed to read SdfOracle data header");
    validateCommon(h);
    const uint32_t flags = h.get<uint32_t>(Header::kFlags);
    if (flags & kFlagSparse)
        throw Error(SDFO_ERR_FORMAT, "SdfOracle data file uses sparse layout; call loadSparseSdf");
    const uint32_t layout = h.get<uint32_t>(Header::kReserved);
    if (layout != kLayoutDense && layout != kLayoutDenseContact)
        throw Error(SDFO_ERR_FORMAT, "SdfOracle data file is not dense 0.1.0 layout");
    const bool contactFlag = (flags & kFlagContact) != 0;
    if ((layout == kLayoutDenseContact) != contactFlag)
        throw Error(SDFO_ERR_FORMAT, "Dense SDF 0.1.0 contact-aware flag/layout mismatch");
    auto g = std::make_unique<sdfo_grid>();
    g->spec = specOf(h);
    for (int a = 0; a < 3; ++a)
        if (g->spec.n[a] <= 0) throw Error(SDFO_ERR_FORMAT, "SdfOracle data file has an empty grid");
    const size_t N = size_t(g->spec.total());
    // SdfData::resize defaults (SdfData.h:52-61).
    std::vector<double> phi(N, 0.0), n3(3 * N, 0.0), h6(6 * N, 0.0), wit(3 * N, 0.0);
    for (size_t i = 0; i < N; ++i) n3[3 * i] = 1.0;
    auto rd = [&](void* dst, size_t bytes) { in.read(reinterpret_cast<char*>(dst), std::streamsize(bytes)); };
    rd(phi.data(), N * 8);
    rd(n3.data(), 3 * N * 8);
    if (flags & kFlagHessian) rd(h6.data(), 6 * N * 8);
    rd(wit.data(), 3 * N * 8);
    std::vector<uint8_t> cc;
    std::vector<double> cphi, cn3;
    std::vector<int32_t> cbr;
    int stride = 0;
    if (contactFlag)
    {
        stride = int(h.get<uint32_t>(Header::kReserved + 4));
        if (stride <= 0 || stride > 8) throw Error(SDFO_ERR_FORMAT, "Invalid contact-aware candidate stride");
        cc.assign(N, 0);
        for (size_t i = 0; i < N; ++i)
        {
            uint32_t c = 0;
            rd(&c, 4);
            if (c > uint32_t(stride)) throw Error(SDFO_ERR_FORMAT, "Invalid contact-aware candidate count");
            cc[i] = uint8_t(c);
        }
        const size_t NS = N * size_t(stride);
        cphi.assign(NS, 0.0);
        cn3.assign(3 * NS, 0.0);
        for (size_t i = 0; i < NS; ++i) cn3[3 * i] = 1.0;
        cbr.assign(NS, -1);
        rd(cphi.data(), NS * 8);
        rd(cn3.data(), 3 * NS * 8);
        rd(cbr.data(), NS * 4);
    }
    if (!in) throw Error(SDFO_ERR_FORMAT, "Failed to read dense SDF payload");
    SDFO_HIP(hipGetDevice(&g->device));
    g->phi = upload(*g, phi);
    g->n3 = upload(*g, n3);
    g->h6 = upload(*g, h6);
    g->witness3 = upload(*g, wit);
    g->stride = stride;
    if (stride > 0)
    {
        g->ccount = upload(*g, cc);
        g->cphi = upload(*g, cphi);
        g->cn3 = upload(*g, cn3);
        g->cbranch = upload(*g, cbr);
    }
    finishDense(*g, nullptr);
    SDFO_HIP(hipDeviceSynchronize());
    return g;
}

// Uploads host sparse arrays (blocks in file order) into a grid whose spec,
// blockDim, halo, band, hasHessian, stride and nb are set.
void uploadSparse(sdfo_grid& g, const std::vector<int32_t>& bijk, const std::vector<float>& phi,
                  const std::vector<float>& n3, const std::vector<float>& h6,
                  const std::vector<uint8_t>& cc, const std::vector<float>& cphi,
                  const std::vector<float>& cn3, const std::vector<int32_t>& cbr)
{
    const int bd = g.blockDim;
    SDFO_HIP(hipGetDevice(&g.device));
    g.nbx = (g.spec.n[0] + bd - 1) / bd;
    g.nby = (g.spec.n[1] + bd - 1) / bd;
    g.nbz = (g.spec.n[2] + bd - 1) / bd;
    g.bijk = upload(g, bijk);
    g.sphi = upload(g, phi);
    g.sn3 = upload(g, n3);
    if (g.hasHessian) g.sh6 = upload(g, h6);
    if (g.stride > 0)
    {
        g.sccount = upload(g, cc);
        g.scphi = upload(g, cphi);
        g.scn3 = upload(g, cn3);
        g.scbranch = upload(g, cbr);
    }
    finishSparse(g, bijk, nullptr);
    SDFO_HIP(hipDeviceSynchronize());
}

// loadSparseSdf (SdfIO.cpp:396-523).
std::unique_ptr<sdfo_grid> loadSparse(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw Error(SDFO_ERR_IO, "Cannot open SdfOracle data file: " + path);
    Header h;
    if (!readHeader(in, h)) throw Error(SDFO_ERR_FORMAT, "Failed to read SdfOracle data header");
    validateCommon(h);
    const uint32_t flags = h.get<uint32_t>(Header::kFlags);
    const uint32_t layout = h.get<uint32_t>(Header::kReserved);
    const bool contactFlag = (flags & kFlagContact) != 0;
    if ((layout == kLayoutSparseContact) != contactFlag)
        throw Error(SDFO_ERR_FORMAT, "Sparse SDF 0.1.0 contact-aware flag/layout mismatch");
    const bool contact = contactFlag && layout == kLayoutSparseContact;
    if ((flags & kFlagSparse) == 0 || !(layout == kLayoutSparse || contact))
        throw Error(SDFO_ERR_FORMAT, "SdfOracle data file is not sparse block 0.1.0");
    auto g = std::make_unique<sdfo_grid>();
    g->sparse = true;
    g->spec = specOf(h);
    g->blockDim = int(h.get<uint32_t>(Header::kReserved + 4));
    g->halo = int(h.get<uint32_t>(Header::kReserved + 8));
    const uint32_t nblocks = h.get<uint32_t>(Header::kReserved + 12);
    g->band = h.get<double>(Header::kReserved + 16);
    g->hasHessian = (flags & kFlagHessian) != 0;
    g->stride = contact ? int(h.get<uint32_t>(Header::kReserved + 24)) : 0;
    if (g->blockDim <= 0) throw Error(SDFO_ERR_FORMAT, "Sparse SDF blockDim must be positive");
    if (contact && (g->stride <= 0 || g->stride > 8))
        throw Error(SDFO_ERR_FORMAT, "Invalid sparse contact-aware candidate stride");
    for (int a = 0; a < 3; ++a)
        if (g->spec.n[a] <= 0) throw Error(SDFO_ERR_FORMAT, "SdfOracle data file has an empty grid");
    const int bd = g->blockDim;
    if (bd > 64) throw Error(SDFO_ERR_UNSUPPORTED, "Sparse SDF blockDim above 64 is not supported");
    const size_t bvc = size_t(bd) * bd * bd;
    const size_t S = size_t(g->stride);
    std::vector<int32_t> bijk;
    std::vector<float> phi, n3, h6, cphi, cn3;
    std::vector<uint8_t> cc;
    std::vector<int32_t> cbr;
    auto rd = [&](void* dst, size_t bytes) { in.read(reinterpret_cast<char*>(dst), std::streamsize(bytes)); };
    for (uint32_t b = 0; b < nblocks; ++b)
    {
        int32_t bh[4];
        rd(bh, 16);
        if (!in) throw Error(SDFO_ERR_FORMAT, "Failed to read sparse SDF block header");
        bijk.insert(bijk.end(), bh, bh + 3);
        const size_t o = phi.size();
        phi.resize(o + bvc);
        n3.resize(3 * (o + bvc));
        rd(&phi[o], bvc * 4);
        rd(&n3[3 * o], 3 * bvc * 4);
        if (g->hasHessian)
        {
            h6.resize(6 * (o + bvc));
            rd(&h6[6 * o], 6 * bvc * 4);
        }
        if (contact)
        {
            cc.resize(o + bvc);
            cphi.resize((o + bvc) * S);
            cn3.resize(3 * (o + bvc) * S);
            cbr.resize((o + bvc) * S);
            rd(&cc[o], bvc);
            rd(&cphi[o * S], bvc * S * 4);
            rd(&cn3[3 * o * S], 3 * bvc * S * 4);
            rd(&cbr[o * S], bvc * S * 4);
            for (size_t q = o; q < o + bvc; ++q)
                if (cc[q] > S) throw Error(SDFO_ERR_FORMAT, "Invalid sparse contact-aware candidate count");
        }
        if (!in) throw Error(SDFO_ERR_FORMAT, "Failed to read sparse SDF block payload");
    }
    g->nb = int(nblocks);
    uploadSparse(*g, bijk, phi, n3, h6, cc, cphi, cn3, cbr);
    return g;
}

bool headerOf(const char* path, Header& h)
{
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) return false;
    return readHeader(in, h);
}

// isSparseSdf (SdfIO.cpp:539-552).
bool isSparse(const char* path)
{
    Header h;
    if (!headerOf(path, h)) return false;
    if (std::memcmp(h.b, "SDFO", 4) != 0 || h.get<uint32_t>(Header::kVer) != kVersion) return false;
    const uint32_t layout = h.get<uint32_t>(Header::kReserved);
    return (h.get<uint32_t>(Header::kFlags) & kFlagSparse) != 0 &&
           (layout == kLayoutSparse || layout == kLayoutSparseContact);
}

template <typename F>
int run(F&& f)
{
    try
    {
        setLastError("");
        f();
        return SDFO_OK;
    }
    catch (const Error& e)
    {
        setLastError(e.what());
        return e.code;
    }
    catch (const std::exception& e)
    {
        setLastError(e.what());
        return SDFO_ERR_IO;
    }
}

} // namespace
} // namespace sdfo

using namespace sdfo;

extern "C" {

int sdfo_grid_save(const sdfo_grid* grid, const char* path)
{
    return run([&] {
        if (!grid || !path) throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL argument");
        if (grid->sparse && grid->blockDim <= 0)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "Sparse SDF blockDim must be positive");
        int cur = 0;
        SDFO_HIP(hipGetDevice(&cur));
        if (cur != grid->device) SDFO_HIP(hipSetDevice(grid->device));
        std::ofstream out(path, std::ios::binary);
        if (!out.is_open()) throw Error(SDFO_ERR_IO, std::string("Cannot open file for writing: ") + path);
        if (grid->sparse)
            saveSparse(*grid, out);
        else
            saveDense(*grid, out);
        out.flush();
        const bool ok = out.good();
        if (cur != grid->device) (void)hipSetDevice(cur);
        if (!ok) throw Error(SDFO_ERR_IO, std::string("Failed to write ") + path);
    });
}

int sdfo_grid_load(const char* path, sdfo_grid** out)
{
    return run([&] {
        if (!path || !out) throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL argument");
        *out = nullptr;
        // SdfOracle::loadFromFile (SdfOracle.cpp:165-172).
        *out = (isSparse(path) ? loadSparse(path) : loadDense(path)).release();
    });
}

int sdfo_is_valid_file(const char* path)
{
    if (!path) return 0;
    Header h;
    if (!headerOf(path, h)) return 0;
    return (std::memcmp(h.b, "SDFO", 4) == 0 && h.get<uint32_t>(Header::kVer) == kVersion) ? 1 : 0;
}

int sdfo_is_sparse_file(const char* path) { return (path && isSparse(path)) ? 1 : 0; }

int sdfo_save_dense_host(const char* path, const double* bmin3, const double* bmax3,
                         const double* voxel3, const int32_t* shape3, const double* phi,
                         const double* n3, const double* h6, const double* witness3, int32_t stride,
                         const uint8_t* ccount, const double* cphi, const double* cn3,
                         const int32_t* cbranch)
{
    return run([&] {
        if (!path || !bmin3 || !bmax3 || !voxel3 || !shape3 || !phi || !n3 || !witness3)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL argument");
        if (stride > 0 && (!ccount || !cphi || !cn3 || !cbranch))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "candidate arrays are required when stride > 0");
        DenseHost d{};
        for (int a = 0; a < 3; ++a)
        {
            d.spec.bmin[a] = bmin3[a];
            d.spec.bmax[a] = bmax3[a];
            d.spec.voxel[a] = voxel3[a];
            d.spec.n[a] = shape3[a];
        }
        d.phi = phi;
        d.n3 = n3;
        d.h6 = h6;
        d.wit3 = witness3;
        d.stride = stride > 0 ? stride : 0;
        d.cc = ccount;
        d.cphi = cphi;
        d.cn3 = cn3;
        d.cbr = cbranch;
        std::ofstream out(path, std::ios::binary);
        if (!out.is_open()) throw Error(SDFO_ERR_IO, std::string("Cannot open file for writing: ") + path);
        writeDense(d, out);
        out.flush();
        if (!out.good()) throw Error(SDFO_ERR_IO, std::string("Failed to write ") + path);
    });
}

int sdfo_save_sparse_host(const char* path, const double* bmin3, const double* bmax3,
                          const double* voxel3, const int32_t* shape3, int32_t block_dim,
                          int32_t halo, double band, int32_t has_hessian, int32_t stride,
                          int32_t n_blocks, const int32_t* bijk, const float* phi, const float* n3,
                          const float* h6, const uint8_t* ccount, const float* cphi,
                          const float* cn3, const int32_t* cbranch)
{
    return run([&] {
        if (!path || !bmin3 || !bmax3 || !voxel3 || !shape3)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL argument");
        if (n_blocks < 0 || block_dim <= 0)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "Sparse SDF blockDim must be positive");
        if (n_blocks > 0 && (!bijk || !phi || !n3))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "block arrays are required");
        if (stride > 0 && n_blocks > 0 && (!ccount || !cphi || !cn3 || !cbranch))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "candidate arrays are required when stride > 0");
        SparseHost d{};
        for (int a = 0; a < 3; ++a)
        {
            d.spec.bmin[a] = bmin3[a];
            d.spec.bmax[a] = bmax3[a];
            d.spec.voxel[a] = voxel3[a];
            d.spec.n[a] = shape3[a];
        }
        d.bd = block_dim;
        d.halo = halo;
        d.nb = n_blocks;
        d.band = band;
        d.hasHessian = has_hessian != 0;
        d.stride = stride > 0 ? stride : 0;
        d.bijk = bijk;
        d.phi = phi;
        d.n3 = n3;
        d.h6 = h6;
        d.cc = ccount;
        d.cphi = cphi;
        d.cn3 = cn3;
        d.cbr = cbranch;
        std::ofstream out(path, std::ios::binary);
        if (!out.is_open()) throw Error(SDFO_ERR_IO, std::string("Cannot open file for writing: ") + path);
        writeSparse(d, out);
        out.flush();
        if (!out.good()) throw Error(SDFO_ERR_IO, std::string("Failed to write ") + path);
    });
}

int sdfo_grid_from_sparse(const double* bmin3, const double* bmax3, const double* voxel3,
                          const int32_t* shape3, int32_t block_dim, int32_t halo, double band,
                          int32_t has_hessian, int32_t stride, int32_t n_blocks,
                          const int32_t* bijk, const float* phi, const float* n3, const float* h6,
                          const uint8_t* ccount, const float* cphi, const float* cn3,
                          const int32_t* cbranch, sdfo_grid** out)
{
    return run([&] {
        if (!out || !bmin3 || !bmax3 || !voxel3 || !shape3)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "NULL argument");
        *out = nullptr;
        if (block_dim <= 0 || block_dim > 64)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "Sparse SDF blockDim must be in 1..64");
        if (n_blocks < 0 || stride < 0 || stride > 8)
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "invalid block count or candidate stride");
        const size_t bvc = size_t(block_dim) * block_dim * block_dim, cnt = size_t(n_blocks) * bvc;
        if (n_blocks > 0 && (!bijk || !phi || !n3 || (has_hessian && !h6)))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "block arrays are required");
        if (stride > 0 && n_blocks > 0 && (!ccount || !cphi || !cn3 || !cbranch))
            throw Error(SDFO_ERR_INVALID_ARGUMENT, "candidate arrays are required when stride > 0");
        auto g = std::make_unique<sdfo_grid>();
        g->sparse = true;
        for (int a = 0; a < 3; ++a)
        {
            g->spec.bmin[a] = bmin3[a];
            g->spec.bmax[a] = bmax3[a];
            g->spec.voxel[a] = voxel3[a];
            g->spec.n[a] = shape3[a];
            if (shape3[a] <= 0) throw Error(SDFO_ERR_INVALID_ARGUMENT, "empty grid");
        }
        g->blockDim = block_dim;
        g->halo = halo;
        g->band = band;
        g->hasHessian = has_hessian != 0;
        g->stride = stride;
        g->nb = n_blocks;
        const size_t S = size_t(stride);
        auto vec = [](auto* p, size_t n) {
            using T = std::remove_const_t<std::remove_pointer_t<decltype(p)>>;
            return p ? std::vector<T>(p, p + n) : std::vector<T>();
        };
        uploadSparse(*g, vec(bijk, size_t(n_blocks) * 3), vec(phi, cnt), vec(n3, 3 * cnt),
                     has_hessian ? vec(h6, 6 * cnt) : std::vector<float>(),
                     stride > 0 ? vec(ccount, cnt) : std::vector<uint8_t>(),
                     stride > 0 ? vec(cphi, cnt * S) : std::vector<float>(),
                     stride > 0 ? vec(cn3, 3 * cnt * S) : std::vector<float>(),
                     stride > 0 ? vec(cbranch, cnt * S) : std::vector<int32_t>());
        *out = g.release();
    });
}

} // extern "C"
