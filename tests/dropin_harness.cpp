// Harness for HandlerPO members of the HIP drop-in that the reference CLI
// cannot reach on a one-GPU host: HandleOrientationsToLocalGpuOnDevice with
// a pack-cache token (the reference passes one only from its multi-GPU
// theta-group scheduler, src/scattering/TracerPOTotal.cpp:6567-6693) and the
// theta probe DiffractThetasGpu (no caller in the reference).
//
// The prepared beams come from a flat batch file written by the test
// (tests/test_dropin_members_gpu.py) from a golden fixture: the C ABI's
// mbs_beam_batch arrays, turned back into PreparedOrientation records here.
// The handler is set up as src/main.cpp sets it up for a hexagonal column
// (the particle only supplies the refractive index these members read).
//
// usage: dropin_harness probe <batch> <out> nphi theta_rad...
//          DiffractThetasGpu on the batch; writes the M11 values (doubles)
//        dropin_harness token <batch> <out>
//          OnDevice on grid B without a token; with token 9 on grid A, then
//          on grid B (the cached batch); with token 9 and another start on
//          grid B and its tokenless twin; writes the four grid-B results
//          (doubles, [nAz][nZen+1][16] each) and prints how each call ran
#include "HandlerPO.h"
#include "Hexagonal.h"
#include "TracerPO.h"
#include "Tracks.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace
{
struct Batch
{
    int32_t nOrient = 0, nBeams = 0;
    int64_t nVert = 0;
    double wave = 0, nRe = 0, nIm = 0;
    std::vector<int64_t> orientOffsets, vertOffsets;
    std::vector<double> weights, scalars, vx, vy, syx, sxy;
    std::vector<int32_t> nv, ext;
    std::vector<uint8_t> flags;
};

template <class T>
void Read(FILE *f, std::vector<T> &v, size_t n)
{
    v.resize(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n)
    {
        std::fprintf(stderr, "short batch file\n");
        std::exit(2);
    }
}

Batch Load(const char *path)
{
    FILE *f = std::fopen(path, "rb");
    if (!f)
    {
        std::fprintf(stderr, "cannot open %s\n", path);
        std::exit(2);
    }
    char magic[8];
    Batch b;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "MBSBAT01", 8) != 0
        || std::fread(&b.nOrient, 4, 1, f) != 1 || std::fread(&b.nBeams, 4, 1, f) != 1
        || std::fread(&b.nVert, 8, 1, f) != 1 || std::fread(&b.wave, 8, 1, f) != 1
        || std::fread(&b.nRe, 8, 1, f) != 1 || std::fread(&b.nIm, 8, 1, f) != 1)
    {
        std::fprintf(stderr, "bad batch header\n");
        std::exit(2);
    }
    Read(f, b.orientOffsets, b.nOrient + 1);
    Read(f, b.weights, b.nOrient);
    Read(f, b.scalars, (size_t)b.nBeams * 36);
    Read(f, b.nv, b.nBeams);
    Read(f, b.ext, b.nBeams);
    Read(f, b.vertOffsets, (size_t)b.nBeams + 1);
    Read(f, b.vx, b.nVert);
    Read(f, b.vy, b.nVert);
    Read(f, b.syx, b.nVert);
    Read(f, b.sxy, b.nVert);
    Read(f, b.flags, b.nVert);
    std::fclose(f);
    return b;
}

// The PreparedOrientation records the drop-in packs, from the flat arrays
// (the inverse of its Pack; the raw-Jones fields of the multi-size path are
// left empty).
std::vector<PreparedOrientation> Prepared(const Batch &b)
{
    std::vector<PreparedOrientation> out(b.nOrient);
    for (int o = 0; o < b.nOrient; ++o)
    {
        out[o].sinZenith = b.weights[o];
        for (int64_t i = b.orientOffsets[o]; i < b.orientOffsets[o + 1]; ++i)
        {
            const double *s = &b.scalars[(size_t)i * 36];
            PreparedBeam pb;
            pb.horAx = s[0]; pb.horAy = s[1]; pb.horAz = s[2];
            pb.verAx = s[3]; pb.verAy = s[4]; pb.verAz = s[5];
            pb.bdx = s[6]; pb.bdy = s[7]; pb.bdz = s[8];
            pb.cenx = s[9]; pb.ceny = s[10]; pb.cenz = s[11];
            pb.beam_area = s[12];
            pb.pNTx = s[13]; pb.pNTy = s[14]; pb.pNTz = s[15];
            pb.pNPx = s[16]; pb.pNPy = s[17]; pb.pNPz = s[18];
            pb.pnxDTx = s[19]; pb.pnxDTy = s[20]; pb.pnxDTz = s[21];
            pb.pnxDPx = s[22]; pb.pnxDPy = s[23]; pb.pnxDPz = s[24];
            pb.jp00r = s[25]; pb.jp00i = s[26]; pb.jp01r = s[27]; pb.jp01i = s[28];
            pb.jp10r = s[29]; pb.jp10i = s[30]; pb.jp11r = s[31]; pb.jp11i = s[32];
            pb.normx = s[33]; pb.normy = s[34]; pb.normz = s[35];
            pb.isExternal = b.ext[i] != 0;
            BeamEdgeData &ed = pb.edgeData;
            ed.nVertices = b.nv[i];
            ed.valid = true;
            for (int v = 0; v < ed.nVertices; ++v)
            {
                const int64_t k = b.vertOffsets[i] + v;
                ed.x[v] = b.vx[k];
                ed.y[v] = b.vy[k];
                ed.slope_yx[v] = b.syx[k];
                ed.slope_xy[v] = b.sxy[k];
                ed.edge_valid_x[v] = (b.flags[k] & 1) != 0;
                ed.edge_valid_y[v] = (b.flags[k] & 2) != 0;
            }
            out[o].beams.push_back(pb);
        }
    }
    return out;
}

std::vector<double> Cells(const Arr2D &m, int nAz, int nZen)
{
    std::vector<double> g((size_t)nAz * (nZen + 1) * 16);
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
            std::memcpy(&g[((size_t)p * (nZen + 1) + t) * 16], m.RawCell(p, t),
                        16 * sizeof(double));
    return g;
}
}  // namespace

int main(int argc, char **argv)
{
    if (argc < 4)
    {
        std::fprintf(stderr, "usage: %s probe|token <batch> <out> [...]\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const Batch b = Load(argv[2]);
    const std::vector<PreparedOrientation> prepared = Prepared(b);
    const bool isAbs = b.nIm != 0.0;
    Hexagonal particle(complex(b.nRe, isAbs ? b.nIm : 0.0), 10.0, 20.0);
    TracerPO tracer(&particle, 1, "harness");
    tracer.m_scattering->m_wave = b.wave;
    Tracks tracks;
    const int nTheta = 45;
    HandlerPO handler(&particle, &tracer.m_incidentLight, nTheta, b.wave);
    handler.isCoh = true;
    handler.SetFullOnly(true);
    handler.SetTracks(&tracks);
    handler.SetScattering(tracer.m_scattering);
    handler.SetAbsorptionAccounting(isAbs);
    FILE *out = std::fopen(argv[3], "wb");
    if (!out)
        return 2;
    if (mode == "probe")
    {
        if (argc < 6)
            return 2;
        const int nPhi = std::atoi(argv[4]);
        handler.SetScatteringSphere(ScatteringRange(0.0, M_PI, nPhi, nTheta));
        std::vector<double> theta;
        for (int i = 5; i < argc; ++i)
            theta.push_back(std::atof(argv[i]));
        std::vector<double> m11;
        if (!handler.DiffractThetasGpu(prepared, theta.data(), (int)theta.size(), m11))
            return 1;
        std::fwrite(m11.data(), sizeof(double), m11.size(), out);
        std::fclose(out);
        std::printf("probe: %zu thetas\n", m11.size());
        return 0;
    }
    if (mode != "token")
        return 2;
    const ScatteringRange gridA(0.0, M_PI, 24, 90), gridB(0.0, M_PI / 2, 16, nTheta);
    const int n = b.nOrient;
    auto run = [&](const ScatteringRange &g, int start, int count, unsigned long long token,
                   const char *label, bool keep) {
        handler.SetScatteringSphere(g);
        const int nAz = g.nAzimuth, nZen = g.nZenith;
        Arr2D m(nAz + 1, nZen + 1, 4, 4), mns(nAz + 1, nZen + 1, 4, 4);
        m.ClearArr();
        mns.ClearArr();
        std::fprintf(stderr, "call %s\n", label);
        if (!handler.HandleOrientationsToLocalGpuOnDevice(prepared, start, count, 0, m, mns, 1.0,
                                                          0.0, token))
            std::exit(1);
        if (keep)
        {
            const std::vector<double> cells = Cells(m, nAz, nZen);
            std::fwrite(cells.data(), sizeof(double), cells.size(), out);
        }
    };
    run(gridB, 0, n, 0, "fresh-B", true);
    run(gridA, 0, n, 9, "token-A", false);
    run(gridB, 0, n, 9, "token-B", true);
    run(gridB, 1, n - 1, 9, "token-B-start1", true);
    run(gridB, 1, n - 1, 0, "fresh-B-start1", true);
    std::fclose(out);
    std::printf("token: 4 grids of %d x %d cells\n", gridB.nAzimuth, gridB.nZenith + 1);
    return 0;
}
