// "SDFO" 0.1.0 binary files (SdfIO.cpp:60-552): byte-identical writer and a
// reader with the reference's validation rules and messages.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
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

// saveSdf (SdfIO.cpp:93-183).
void saveDense(const sdfo_grid& g, std::ofstream& out)
{
    const size_t N = size_t(g.spec.total());
    const bool contact = g.stride > 0;
    Header h = makeHeader(g.spec, kFlagHessian | (contact ? kFlagContact : 0u));
    h.put<uint32_t>(Header::kReserved + 0, contact ? kLayoutDenseContact : kLayoutDense);
    if (contact) h.put<uint32_t>(Header::kReserved + 4, uint32_t(g.stride));
    out.write(reinterpret_cast<const char*>(h.b), 128);
    auto wr = [&](const std::vector<double>& v) {
        out.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(double)));
    };
    wr(download(g.phi, N));
    wr(download(g.n3, 3 * N));
    wr(download(g.h6, 6 * N));
    wr(download(g.witness3, 3 * N));
    if (contact)
    {
        const size_t NS = N * size_t(g.stride);
        const std::vector<uint8_t> c8 = download(g.ccount, N);
        std::vector<uint32_t> c32(c8.begin(), c8.end());
        out.write(reinterpret_cast<const char*>(c32.data()), std::streamsize(N * 4));
        wr(download(g.cphi, NS));
        wr(download(g.cn3, 3 * NS));
        const std::vector<int32_t> br = download(g.cbranch, NS);
        out.write(reinterpret_cast<const char*>(br.data()), std::streamsize(NS * 4));
    }
}

// saveSparseSdf (SdfIO.cpp:185-275).
void saveSparse(const sdfo_grid& g, std::ofstream& out)
{
    const int bd = g.blockDim;
    const size_t bvc = size_t(bd) * bd * bd;
    const bool contact = g.stride > 0 && g.nb > 0;
    Header h = makeHeader(g.spec, kFlagSparse | (g.hasHessian ? kFlagHessian : 0u) |
                                      (contact ? kFlagContact : 0u));
    h.put<uint32_t>(Header::kReserved + 0, contact ? kLayoutSparseContact : kLayoutSparse);
    h.put<uint32_t>(Header::kReserved + 4, uint32_t(bd));
    h.put<uint32_t>(Header::kReserved + 8, uint32_t(g.halo));
    h.put<uint32_t>(Header::kReserved + 12, uint32_t(g.nb));
    h.put<double>(Header::kReserved + 16, g.band);
    if (contact) h.put<uint32_t>(Header::kReserved + 24, uint32_t(g.stride));
    out.write(reinterpret_cast<const char*>(h.b), 128);
    const size_t cnt = size_t(g.nb) * bvc;
    const std::vector<int32_t> bijk = download(g.bijk, size_t(g.nb) * 3);
    const std::vector<float> phi = download(g.sphi, cnt);
    const std::vector<float> n3 = download(g.sn3, 3 * cnt);
    std::vector<float> h6;
    if (g.hasHessian) h6 = g.sh6 ? download(g.sh6, 6 * cnt) : std::vector<float>(6 * cnt, 0.0f);
    const size_t S = size_t(g.stride);
    std::vector<uint8_t> cc;
    std::vector<float> cphi, cn3;
    std::vector<int32_t> cbr;
    if (contact)
    {
        cc = download(g.sccount, cnt);
        cphi = download(g.scphi, cnt * S);
        cn3 = download(g.scn3, 3 * cnt * S);
        cbr = download(g.scbranch, cnt * S);
    }
    for (int b = 0; b < g.nb; ++b)
    {
        int32_t bh[4] = {bijk[3 * b], bijk[3 * b + 1], bijk[3 * b + 2], 0};
        out.write(reinterpret_cast<const char*>(bh), 16);
        out.write(reinterpret_cast<const char*>(&phi[b * bvc]), std::streamsize(bvc * 4));
        out.write(reinterpret_cast<const char*>(&n3[3 * b * bvc]), std::streamsize(3 * bvc * 4));
        if (g.hasHessian)
            out.write(reinterpret_cast<const char*>(&h6[6 * b * bvc]), std::streamsize(6 * bvc * 4));
        if (contact)
        {
            out.write(reinterpret_cast<const char*>(&cc[b * bvc]), std::streamsize(bvc));
            out.write(reinterpret_cast<const char*>(&cphi[b * bvc * S]), std::streamsize(bvc * S * 4));
            out.write(reinterpret_cast<const char*>(&cn3[3 * b * bvc * S]),
                      std::streamsize(3 * bvc * S * 4));
            out.write(reinterpret_cast<const char*>(&cbr[b * bvc * S]), std::streamsize(bvc * S * 4));
        }
    }
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
    SDFO_HIP(hipGetDevice(&g->device));
    g->nb = int(nblocks);
    g->nbx = (g->spec.n[0] + bd - 1) / bd;
    g->nby = (g->spec.n[1] + bd - 1) / bd;
    g->nbz = (g->spec.n[2] + bd - 1) / bd;
    g->bijk = upload(*g, bijk);
    g->sphi = upload(*g, phi);
    g->sn3 = upload(*g, n3);
    if (g->hasHessian) g->sh6 = upload(*g, h6);
    if (contact)
    {
        g->sccount = upload(*g, cc);
        g->scphi = upload(*g, cphi);
        g->scn3 = upload(*g, cn3);
        g->scbranch = upload(*g, cbr);
    }
    finishSparse(*g, bijk, nullptr);
    SDFO_HIP(hipDeviceSynchronize());
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

} // extern "C"
