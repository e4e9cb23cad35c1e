// Exercises the source-compatible C++ API (include/SdfOracle) on the GPU.
// Usage: sdf_api_check <dir>.  Reads <dir>/points.bin (int64 n, n x 3 doubles),
// <dir>/mesh.obj and <dir>/surface.txt written by tests/test_gpu_cpp_api.py;
// writes SDFO files and query results there for that test to compare with
// the Python/C-ABI path.  Exit code 0 when its own checks pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "SdfOracle/KdTree.h"
#include "SdfOracle/NexDynAdapter.h"
#include "SdfOracle/SdfBuilder.h"
#include "SdfOracle/SdfIO.h"
#include "SdfOracle/SdfOracle.h"
#include "SdfOracle/SdfSurfaceMesh.h"

namespace {

int failures = 0;

void expect(bool ok, const std::string& what)
{
    if (!ok)
    {
        ++failures;
        std::cerr << "CHECK FAILED: " << what << "\n";
    }
}

std::vector<sdf::Vector3d> readPoints(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    int64_t n = 0;
    in.read(reinterpret_cast<char*>(&n), 8);
    std::vector<double> v(size_t(3 * n));
    in.read(reinterpret_cast<char*>(v.data()), std::streamsize(v.size() * 8));
    std::vector<sdf::Vector3d> p(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) p[size_t(i)] = sdf::Vector3d(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    return p;
}

// Per point: gap, normal(3), witness(3), valid, candidate count, then up to 8
// candidates of (gap, normal(3), witness(3), branch) -> 9 + 8 * 8 doubles.
void writeResults(const std::string& path, const std::vector<sdf::SdfQueryResult>& r)
{
    std::ofstream out(path, std::ios::binary);
    for (const auto& q : r)
    {
        std::vector<double> rec(9 + 8 * 8, 0.0);
        rec[0] = q.gap;
        for (int a = 0; a < 3; ++a)
        {
            rec[size_t(1 + a)] = q.gapNormal[a];
            rec[size_t(4 + a)] = q.witnessPoint[a];
        }
        rec[7] = q.valid ? 1.0 : 0.0;
        rec[8] = double(q.candidates.size());
        for (size_t c = 0; c < q.candidates.size() && c < 8; ++c)
        {
            double* o = &rec[9 + 8 * c];
            o[0] = q.candidates[c].gap;
            for (int a = 0; a < 3; ++a)
            {
                o[1 + a] = q.candidates[c].gapNormal[a];
                o[4 + a] = q.candidates[c].witnessPoint[a];
            }
            o[7] = q.candidates[c].branchId;
        }
        out.write(reinterpret_cast<const char*>(rec.data()), std::streamsize(rec.size() * 8));
    }
}

const char* kModeNames[5] = {"FirstOrder", "SecondOrder", "Trilinear", "Tricubic", "ContactAware"};

} // namespace

// "loaders <file> <out>": host-only, no device use.  Parses <file> (OBJ when
// it ends in .obj, else the contact-aware surface format) and writes int64
// counts (vertices, triangles, normals) then the doubles / ints, or the
// exception text prefixed by "ERROR: ".
int dumpMesh(const std::string& file, const std::string& outPath)
{
    std::ofstream out(outPath, std::ios::binary);
    try
    {
        std::vector<sdf::Vector3d> v;
        std::vector<std::array<int, 3>> t;
        std::vector<std::array<sdf::Vector3d, 3>> n;
        if (file.size() > 4 && file.substr(file.size() - 4) == ".obj")
        {
            const sdf::TriangleMesh m = sdf::loadObjFile(file);
            v = m.vertices;
            t = m.triangles;
        }
        else
        {
            const sdf::CornerNormalTriangleMesh m = sdf::loadContactAwareSurfaceFile(file);
            v = m.vertices;
            t = m.triangles;
            n = m.cornerNormals;
        }
        const int64_t c[3] = {int64_t(v.size()), int64_t(t.size()), int64_t(n.size())};
        out.write(reinterpret_cast<const char*>(c), 24);
        auto put = [&](double x) { out.write(reinterpret_cast<const char*>(&x), 8); };
        for (const auto& p : v)
            for (int a = 0; a < 3; ++a) put(p[a]);
        for (const auto& q : t) out.write(reinterpret_cast<const char*>(q.data()), 12);
        for (const auto& q : n)
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) put(q[size_t(k)][a]);
    }
    catch (const std::exception& e)
    {
        out << "ERROR: " << e.what();
    }
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 4 && std::string(argv[1]) == "loaders") return dumpMesh(argv[2], argv[3]);
    if (argc < 2)
    {
        std::cerr << "usage: sdf_api_check <dir> | sdf_api_check loaders <file> <out>\n";
        return 2;
    }
    const std::string dir = argv[1];
    try
    {
        // ---- dense build from an OBJ, save, reload, query -----------------------
        const sdf::TriangleMesh mesh = sdf::loadObjFile(dir + "/mesh.obj");
        sdf::BuilderParams bp;
        bp.targetResolution = 32;
        bp.maxResolution = 32;
        std::vector<std::string> log;
        bp.logCallback = [&](const std::string& m) { log.push_back(m); };
        const sdf::SdfData data = sdf::SdfBuilder::build(mesh, bp);
        expect(data.spec.totalVoxels() == int(data.phi0.size()), "dense sizes");
        expect(!log.empty() && log.back() == "[SdfBuilder] Done.", "log callback");
        expect(sdf::saveSdf(data, dir + "/dense.sdf"), "saveSdf");
        expect(sdf::isValidSdf(dir + "/dense.sdf") && !sdf::isSparseSdf(dir + "/dense.sdf"), "file predicates");
        const sdf::SdfData back = sdf::loadSdf(dir + "/dense.sdf");
        expect(back.phi0 == data.phi0 && back.n0 == data.n0 && back.witness0 == data.witness0, "dense round trip");
        const std::vector<sdf::Vector3d> pts = readPoints(dir + "/points.bin");
        const sdf::StandaloneSdfOracle oracle(data);
        expect(oracle.isValid() && !oracle.isSparse(), "oracle state");
        for (int m = 0; m < 4; ++m)
        {
            const auto r = oracle.queryBatch(pts, sdf::QueryMode(m));
            writeResults(dir + "/dense_" + kModeNames[m] + ".bin", r);
            // single-point forms agree with the batch
            const sdf::SdfQueryResult one = oracle.query(pts[0], sdf::QueryMode(m));
            const sdf::SdfFastQueryResult f = oracle.queryFast(pts[0], sdf::QueryMode(m));
            expect(one.gap == r[0].gap && one.valid == r[0].valid, "query == queryBatch");
            expect(f.gap == r[0].gap && f.valid == r[0].valid &&
                       f.candidateCount == int(r[0].candidates.size()),
                   "queryFast == queryBatch");
        }
        const sdf::SdfOracle loaded = sdf::SdfOracle::loadFromFile(dir + "/dense.sdf");
        expect(loaded.queryBatch(pts, sdf::QueryMode::Tricubic)[3].gap ==
                   oracle.queryBatch(pts, sdf::QueryMode::Tricubic)[3].gap,
               "loadFromFile queries");

        // ---- exact mesh queries through sdf::KdTree ---------------------------------
        sdf::KdTree tree;
        expect(tree.triangleCount() == 0 && tree.findNearest(pts[0]).triangleIndex == -1, "empty KdTree");
        tree.build(mesh.vertices, mesh.triangles);
        expect(tree.triangleCount() == int(mesh.triangles.size()), "KdTree triangleCount");
        const std::vector<sdf::KdTree::SignedResult> sr = tree.findSignedNearestBatch(pts);
        const std::vector<sdf::ClosestPointResult> nr = tree.findNearestBatch(pts);
        {
            std::ofstream out(dir + "/kdtree.bin", std::ios::binary);
            for (const auto& r : sr)
            {
                const double rec[8] = {r.signedDistance, r.closestPoint[0], r.closestPoint[1],
                                       r.closestPoint[2], r.normal[0], r.normal[1], r.normal[2],
                                       double(r.triangleIndex)};
                out.write(reinterpret_cast<const char*>(rec), sizeof(rec));
            }
        }
        bool same = true;
        for (size_t i = 0; i < pts.size(); ++i)
            same = same && nr[i].distance == std::abs(sr[i].signedDistance) &&
                   nr[i].triangleIndex == sr[i].triangleIndex && nr[i].point[0] == sr[i].closestPoint[0] &&
                   nr[i].faceNormal[2] == sr[i].normal[2];
        expect(same, "findNearest == |findSignedNearest|");
        const sdf::KdTree::SignedResult one = tree.findSignedNearest(pts[1]);
        expect(one.signedDistance == sr[1].signedDistance && one.triangleIndex == sr[1].triangleIndex,
               "findSignedNearest == batch");

        // ---- sparse build, save, reload, query -----------------------------------
        sdf::SparseBuilderParams sp;
        sp.targetResolution = 32;
        sp.maxResolution = 32;
        sp.bandVoxels = 3;
        const sdf::SparseSdfData sd = sdf::SdfBuilder::buildSparse(mesh, sp);
        expect(!sd.empty() && sd.blockLookup.size() == sd.blocks.size(), "sparse blocks");
        expect(sdf::saveSparseSdf(sd, dir + "/sparse.sdf"), "saveSparseSdf");
        const sdf::SparseSdfData sback = sdf::loadSparseSdf(dir + "/sparse.sdf");
        expect(sback.blocks.size() == sd.blocks.size() && sback.blocks[0].phi0 == sd.blocks[0].phi0, "sparse round trip");
        const sdf::SdfOracle soracle(sd);
        for (int m = 0; m < 4; ++m)
            writeResults(dir + "/sparse_" + kModeNames[m] + ".bin", soracle.queryBatch(pts, sdf::QueryMode(m)));

        // ---- contact-aware builds from the surface file ---------------------------
        const sdf::CornerNormalTriangleMesh cm = sdf::loadContactAwareSurfaceFile(dir + "/surface.txt");
        sdf::ContactAwareBuilderParams cp;
        cp.targetResolution = 24;
        cp.maxResolution = 24;
        const sdf::SdfData ca = sdf::SdfBuilder::buildContactAware(cm, cp);
        expect(ca.hasContactCandidates(), "dense candidates");
        expect(sdf::saveSdf(ca, dir + "/ca.sdf"), "save contact-aware");
        const sdf::SdfOracle caOracle(ca);
        writeResults(dir + "/ca_ContactAware.bin", caOracle.queryBatch(pts, sdf::QueryMode::ContactAware));
        cp.bandVoxels = 2;
        const sdf::SparseSdfData sca = sdf::SdfBuilder::buildSparseContactAware(cm, cp);
        expect(sca.hasContactCandidates(), "sparse candidates");
        expect(sdf::saveSparseSdf(sca, dir + "/sca.sdf"), "save sparse contact-aware");
        writeResults(dir + "/sca_ContactAware.bin",
                     sdf::SdfOracle(sca).queryBatch(pts, sdf::QueryMode::ContactAware));

        // ---- errors keep the reference's exceptions ---------------------------------
        bool threw = false;
        try
        {
            sdf::SdfBuilder::build(std::vector<sdf::Vector3d>{}, std::vector<std::array<int, 3>>{});
        }
        catch (const std::runtime_error& e)
        {
            threw = std::string(e.what()) == "Empty mesh provided to SdfBuilder";
        }
        expect(threw, "empty mesh exception");
        threw = false;
        try
        {
            sdf::loadSdf(dir + "/missing.sdf");
        }
        catch (const std::runtime_error& e)
        {
            threw = std::string(e.what()).find("Cannot open SdfOracle data file") == 0;
        }
        expect(threw, "missing file exception");
        expect(!sdf::SdfOracle().query(pts[0]).valid, "empty oracle");
    }
    catch (const std::exception& e)
    {
        std::cerr << "exception: " << e.what() << "\n";
        return 1;
    }
    std::cout << (failures ? "FAILED" : "OK") << " (" << failures << " failures)\n";
    return failures ? 1 : 0;
}
