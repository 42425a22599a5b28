// sanitize_host.cpp -- the library's host-side code under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by tests/sanitize/Makefile, run by tests/test_sanitize.py).
//
// Drives, on seeded random inputs and on edge cases (one voxel, one plane, no pores, every
// voxel a pore, boxes of one voxel along an axis), the code that turns the reference's files
// into device tables: the sparse numbering and pull neighbour lists (build_sparse_geometry),
// their packed device form (pack_neighbours, which decodes every entry back and compares), the
// wrap-round connectivity test, the reference's rank split, the face-rule resolution
// (build_boundaries), the colour-gradient kernels' LDS staging tables (build_colour_windows:
// every slot a kernel can read holds the node it expects), and the INPUT / geometry / phase /
// checkpoint / VTK readers and writers.
// No device is touched.  Exits non-zero on a wrong result; the sanitizers abort on a memory or
// undefined-behaviour error.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "colour_win.h"
#include "handle.h"
#include "host_io.h"
#include "lattice.h"

using namespace lbm;

namespace {

int failures = 0;

void expect(bool ok, const char *what, int nx, int ny, int nz)
{
    if (!ok) {
        std::fprintf(stderr, "FAIL %s on %dx%dx%d\n", what, nx, ny, nz);
        ++failures;
    }
}

std::vector<int32_t> random_solid(int nx, int ny, int nz, double porosity, std::mt19937 &rng)
{
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<int32_t> s((size_t)nx * ny * nz);
    for (auto &v : s) v = u(rng) < porosity ? 0 : 1;
    return s;
}

// The staged colour-gradient kernels' tables: runs packed from slot 0 within the slots and the
// node range; every slot maps to a node; in a block that reads without a search, each node's own
// slot and, for every column whose upstream site is not solid, the slot id + d[j] hold exactly
// that node.
void check_colour_windows(const SparseGeometry &g, int nx, int ny, int nz)
{
    const int64_t grid = (g.N + CW_BLOCK - 1) / CW_BLOCK;
    for (int slots : {1024, 1536, 2560}) {
        std::vector<ColourWin> wins;
        const double fast = build_colour_windows(g, grid, slots, wins);
        bool ok = wins.size() == (size_t)grid;
        int64_t nfast = 0;
        for (int64_t b = 0; b < grid && ok; ++b) {
            const ColourWin &w = wins[(size_t)b];
            int32_t used = 0;
            for (int k = 0; k < CW_RUNS; ++k) {
                ok = ok && w.off[k] == used && w.len[k] >= 0 &&
                     (w.len[k] == 0 || (w.lo[k] >= 0 && (int64_t)w.lo[k] + w.len[k] <= g.N));
                used += w.len[k];
            }
            ok = ok && used <= slots;
            for (int32_t s = 0; s < slots && ok; ++s) {
                const int32_t id = cw_slot_node(w, s);
                ok = id >= 0 && id < g.N;
            }
            if (!ok || w.fast != CW_ALL) continue;
            ++nfast;
            const int64_t n1 = std::min<int64_t>(g.N, (b + 1) * CW_BLOCK);
            for (int64_t n = b * CW_BLOCK; n < n1 && ok; ++n) {
                const int32_t own = (int32_t)n + w.own;
                ok = own >= 0 && own < used && cw_slot_node(w, own) == (int32_t)n;
                for (int j = 0; j < Q - 1 && ok; ++j) {
                    const uint32_t v = g.nbr[(size_t)j * g.P + n];
                    if (v & BB_FLAG) continue;
                    const int32_t s = (int32_t)v + w.d[j];
                    ok = s >= 0 && s < used && cw_slot_node(w, s) == (int32_t)v;
                }
            }
        }
        expect(ok, "colour staging tables", nx, ny, nz);
        expect(ok && std::fabs((double)nfast / (double)grid - fast) < 1e-12, "colour staging fast fraction", nx, ny, nz);
    }
}

void check_geometry(int nx, int ny, int nz, const std::vector<int32_t> &solid)
{
    SparseGeometry g;
    std::string err;
    int64_t pores = 0;
    for (int32_t v : solid) pores += v == 0;
    const int rc = build_sparse_geometry(nx, ny, nz, solid.data(), g, err);
    if (pores == 0) {
        expect(rc == LBM_ERR_ARG, "no-pore geometry refused", nx, ny, nz);
        return;
    }
    expect(rc == LBM_OK && g.N == pores, "node count", nx, ny, nz);
    if (rc != LBM_OK) return;
    // every link names a node, or the node itself flagged for bounce back
    bool links_ok = true;
    for (int q = 1; q < Q; ++q)
        for (int64_t n = 0; n < g.N; ++n) {
            const uint32_t v = g.nbr[(size_t)((q - 1) * g.P + n)];
            links_ok &= (v & BB_FLAG) ? (int64_t)(v & ~BB_FLAG) == n : (int64_t)v < g.N;
        }
    expect(links_ok, "neighbour entries", nx, ny, nz);
    PackedNbr packed;
    expect(pack_neighbours(g, packed, err) == LBM_OK, "packed neighbour table decodes back", nx, ny, nz);
    for (int a = 0; a < 3; ++a) (void)pores_wrap_along(g, a);
    check_colour_windows(g, nx, ny, nz);
    for (int k = 1; k <= std::min(nx, 5); ++k) {
        std::vector<int64_t> ends;
        reference_rank_ends(g, k, ends);
        bool mono = (int)ends.size() == k && ends.back() == g.N;
        for (int r = 1; r < k; ++r) mono &= ends[(size_t)r - 1] <= ends[(size_t)r];
        expect(mono, "rank split", nx, ny, nz);
    }
    // face rules for every Sub_BC and a spread of face sets
    for (int sub = 0; sub < 4; ++sub)
        for (int mask = 1; mask < 64; mask += 7) {
            lbm_sp_bc bc{};
            bc.sub_bc = sub;
            for (int f = 0; f < 6; ++f) {
                bc.pre[f] = (mask >> f) & 1;
                bc.p[f] = 1.0 + 0.001 * f;
                bc.vel[f] = (sub < 2 && f == 5 - (mask % 6)) ? 1 : 0;
                bc.v[f] = 0.01;
                if (bc.vel[f]) bc.pre[f] = 0;
            }
            BoundaryTables t;
            const int r = build_boundaries(g, bc, t, err);
            expect(r == LBM_OK || r == LBM_ERR_ARG, "face rules", nx, ny, nz);
            if (r == LBM_OK) {
                bool sorted = true;
                for (size_t e = 1; e < t.saved.size(); ++e) sorted &= t.saved[e - 1] < t.saved[e];
                expect(sorted && t.wave.size() == (size_t)(g.P / 64), "face tables", nx, ny, nz);
            }
        }
}

std::string tmp_path(const char *name)
{
    const char *d = std::getenv("SANITIZE_TMP");
    return std::string(d ? d : "/tmp") + "/" + name;
}

void check_files(std::mt19937 &rng)
{
    const int nx = 7, ny = 5, nz = 4;
    std::string err;
    const auto solid = random_solid(nx, ny, nz, 0.6, rng);
    // binary geometry round trip
    {
        const std::string p = tmp_path("geom.bin");
        std::ofstream(p, std::ios::binary).write((const char *)solid.data(), (std::streamsize)(solid.size() * 4));
        std::vector<int32_t> back;
        expect(lbmhost::read_geometry(p, nx, ny, nz, 1, back, err) && back == solid, "binary geometry", nx, ny, nz);
        std::vector<int32_t> bad;
        expect(!lbmhost::read_geometry(p, nx + 1, ny, nz, 1, bad, err), "short binary geometry refused", nx, ny, nz);
    }
    // ASCII geometry, x fastest, with stray tokens
    {
        const std::string p = tmp_path("geom.dat");
        std::ofstream f(p);
        f << "header 7\n";
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) f << solid[(size_t)((i * ny + j) * nz + k)] << (i % 3 ? " " : "\n");
        f.close();
        std::vector<int32_t> back;
        expect(lbmhost::read_geometry(p, nx, ny, nz, 0, back, err) && back == solid, "ASCII geometry", nx, ny, nz);
    }
    // phase field clamp
    {
        const std::string p = tmp_path("phase.dat");
        std::ofstream f(p);
        for (int v = 0; v < nx * ny * nz; ++v) f << (v % 3 == 0 ? 2.5 : -0.25) << " ";
        f.close();
        std::vector<double> psi;
        expect(lbmhost::read_phase(p, nx, ny, nz, psi, err) && psi.size() == (size_t)(nx * ny * nz) && psi[0] == 1.0,
               "phase field", nx, ny, nz);
    }
    // checkpoint round trip with the padding record
    {
        const int64_t count = 33;
        std::vector<double> a((size_t)count * 19), b(a.size());
        for (size_t i = 0; i < a.size(); ++i) a[i] = 0.001 * (double)i;
        const std::string p = lbmhost::checkpoint_name(tmp_path(""), "f", 12);
        expect(lbmhost::write_checkpoint(p, a.data(), count, 19, err) && lbmhost::read_checkpoint(p, b.data(), count, 19, err) &&
                   a == b,
               "checkpoint", nx, ny, nz);
        expect(!lbmhost::read_checkpoint(p, b.data(), count + 1, 19, err), "short checkpoint refused", nx, ny, nz);
    }
    // VTK writers and the velocity profile
    {
        std::vector<double> u((size_t)nx * ny * nz * 3), rho((size_t)nx * ny * nz, 1.0);
        for (size_t i = 0; i < u.size(); ++i) u[i] = 1e-4 * (double)(i % 17);
        expect(lbmhost::write_vtk_geometry(tmp_path("g.vtk"), nx, ny, nz, solid.data(), err) &&
                   lbmhost::write_vtk_velocity(tmp_path("v.vtk"), nx, ny, nz, u.data(), err) &&
                   lbmhost::write_vtk_density(tmp_path("d.vtk"), nx, ny, nz, rho.data(), err) &&
                   lbmhost::write_velocity_profile(tmp_path("p.out"), nx, ny, nz, solid.data(), u.data(), err),
               "VTK and profile writers", nx, ny, nz);
    }
    // INPUT files: a truncated file is refused, a missing one too
    {
        lbmhost::SpInput in;
        const std::string p = tmp_path("INPUT_short.dat");
        std::ofstream(p) << "title\ngeometry.bin :file\n7 5 4 :n\n";
        expect(!lbmhost::read_sp_input(p, in, err), "truncated INPUT refused", nx, ny, nz);
        expect(!lbmhost::read_sp_input(tmp_path("no_such_INPUT.dat"), in, err), "missing INPUT refused", nx, ny, nz);
        lbmhost::CgInput cin;
        expect(!lbmhost::read_cg_input(p, cin, err), "truncated CG INPUT refused", nx, ny, nz);
    }
    int box[6];
    bool partial = false;
    const int par[3] = {1, 1, 1}, px[2] = {5, 1}, py[2] = {3, 0}, pz[2] = {3, 1};
    lbmhost::perm_box(par, px, py, pz, nx, ny, nz, box, partial);
    const double por = lbmhost::box_porosity(solid, nx, ny, nz, box);
    expect(partial && por >= 0.0 && por <= 1.0, "permeability box", nx, ny, nz);
}

}  // namespace

int main()
{
    std::mt19937 rng(20261203u);
    const int sizes[][3] = {{1, 1, 1}, {1, 1, 5}, {1, 4, 1}, {3, 1, 1}, {2, 2, 2}, {5, 4, 3},
                            {9, 7, 6}, {17, 13, 11}, {3, 40, 2}, {70, 3, 3}};
    for (const auto &s : sizes) {
        for (double por : {0.0, 0.15, 0.5, 0.85, 1.0})
            check_geometry(s[0], s[1], s[2], random_solid(s[0], s[1], s[2], por, rng));
    }
    // a pack wide enough for waves whose neighbour spread needs the wide table
    check_geometry(64, 120, 100, random_solid(64, 120, 100, 0.7, rng));
    // malformed arguments
    {
        SparseGeometry g;
        std::string err;
        std::vector<int32_t> one(1, 0);
        expect(build_sparse_geometry(0, 1, 1, one.data(), g, err) == LBM_ERR_ARG, "zero size refused", 0, 1, 1);
        expect(build_sparse_geometry(1, 1, 1, nullptr, g, err) == LBM_ERR_ARG, "null geometry refused", 1, 1, 1);
    }
    check_files(rng);
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("sanitized host checks passed\n");
    return 0;
}
