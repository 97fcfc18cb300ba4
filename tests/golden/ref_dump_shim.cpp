// Fixture writer for the PO diffraction hot path.
//
// This translation unit is linked into a build of the reference MBS-fast CLI
// in which src/handler/HandlerPO.cpp and src/main.cpp are compiled with
// -DUSE_CUDA, so that the GPU member functions of HandlerPO
// (src/handler/HandlerPO.h:98-146) and the GpuSupport queries
// (src/cuda/GpuSupport.h:30-34) are left for another unit to define.  Here
// they are defined to (1) serialise the PreparedOrientation batches the
// reference's own orientation-average driver hands to its GPU backend
// (src/scattering/TracerPOTotal.cpp:5915-5941) and (2) compute the expected
// result with the reference's own CPU routines HandleBeamsToLocal +
// AddToMuellerLocal (src/handler/HandlerPO.cpp:2561-2796, :2888-2907), which
// is exactly what the CPU backend does at TracerPOTotal.cpp:5954-5995.
//
// The binary stream it writes (path in $MBS_FIXTURE_DUMP) is converted to
// .npz fixtures by tests/golden/make_fixtures.py.  This file is only built in
// the container that holds the reference sources; no test or product code
// links it.
#include "HandlerPO.h"
#include "GpuSupport.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace
{
FILE *g_dump = nullptr;
bool g_headerDone = false;

FILE *DumpFile()
{
    if (g_dump)
        return g_dump;
    const char *path = std::getenv("MBS_FIXTURE_DUMP");
    if (!path || !*path)
    {
        std::fprintf(stderr, "fixture dump: MBS_FIXTURE_DUMP is not set\n");
        std::exit(3);
    }
    g_dump = std::fopen(path, "wb");
    if (!g_dump)
    {
        std::fprintf(stderr, "fixture dump: cannot open %s\n", path);
        std::exit(3);
    }
    return g_dump;
}

void PutI32(int v) { std::fwrite(&v, sizeof v, 1, DumpFile()); }
void PutF64(double v) { std::fwrite(&v, sizeof v, 1, DumpFile()); }
void PutTag(const char *tag) { std::fwrite(tag, 1, 4, DumpFile()); }
}  // namespace

// ---- GpuSupport (src/cuda/GpuSupport.h:30-34) ------------------------------
bool CheckGpuRuntime(GpuDeviceInfo &info, std::string &error)
{
    info.deviceId = 0;
    info.name = "fixture-dump";
    info.totalGlobalMem = 1LL << 40;
    info.freeGlobalMem = 1LL << 40;
    info.visibleDeviceCount = 1;
    error.clear();
    return true;
}

bool QueryActiveGpuMemory(long long &freeBytes, long long &totalBytes,
                          std::string &error)
{
    freeBytes = 1LL << 40;
    totalBytes = 1LL << 40;
    error.clear();
    return true;
}

int VisibleGpuDeviceCount() { return 1; }

std::string FormatGpuInfo(const GpuDeviceInfo &) { return "fixture-dump"; }

// ---- serialisation helpers --------------------------------------------------
static void WriteHeader(const ScatteringRange &sphere, double waveIndex,
                        double eps1, const complex &complWave,
                        const complex &invComplWave, bool legacySign,
                        bool noShadow)
{
    if (g_headerDone)
        return;
    g_headerDone = true;
    std::fwrite("MBSFIX01", 1, 8, DumpFile());
    const int nAz = sphere.nAzimuth;
    const int nZen = sphere.nZenith;
    PutI32(nAz);
    PutI32(nZen);
    for (int t = 0; t <= nZen; ++t)
        PutF64(sphere.GetZenith(t));
    PutF64(sphere.azinuthStep);
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
        {
            PutF64(sphere.vf[p][t].x);
            PutF64(sphere.vf[p][t].y);
            PutF64(sphere.vf[p][t].z);
        }
    PutF64(waveIndex);
    PutF64(eps1);
    PutF64(real(complWave));
    PutF64(imag(complWave));
    PutF64(real(invComplWave));
    PutF64(imag(invComplWave));
    PutI32(legacySign ? 1 : 0);
    PutI32(noShadow ? 1 : 0);
}

static void WriteOrientation(const PreparedOrientation &po)
{
    PutTag("ORNT");
    PutF64(po.sinZenith);
    PutI32((int)po.beams.size());
    for (const PreparedBeam &pb : po.beams)
    {
        const BeamEdgeData &ed = pb.edgeData;
        const int nv = ed.nVertices;
        PutI32(nv);
        PutI32(pb.isExternal ? 1 : 0);
        PutI32(ed.valid ? 1 : 0);
        const double s[36] = {
            pb.horAx, pb.horAy, pb.horAz, pb.verAx, pb.verAy, pb.verAz,
            pb.bdx, pb.bdy, pb.bdz, pb.cenx, pb.ceny, pb.cenz,
            pb.beam_area,
            pb.pNTx, pb.pNTy, pb.pNTz, pb.pNPx, pb.pNPy, pb.pNPz,
            pb.pnxDTx, pb.pnxDTy, pb.pnxDTz, pb.pnxDPx, pb.pnxDPy, pb.pnxDPz,
            pb.jp00r, pb.jp00i, pb.jp01r, pb.jp01i,
            pb.jp10r, pb.jp10i, pb.jp11r, pb.jp11i,
            pb.normx, pb.normy, pb.normz};
        std::fwrite(s, sizeof(double), 36, DumpFile());
        for (int v = 0; v < nv; ++v)
        {
            PutF64(ed.x[v]);
            PutF64(ed.y[v]);
            // Slopes of invalid edges are never read by the reference;
            // they are written as zero so the fixture is deterministic.
            PutF64(ed.edge_valid_x[v] ? ed.slope_yx[v] : 0.0);
            PutF64(ed.edge_valid_y[v] ? ed.slope_xy[v] : 0.0);
            unsigned char f = (ed.edge_valid_x[v] ? 1 : 0)
                            | (ed.edge_valid_y[v] ? 2 : 0);
            std::fwrite(&f, 1, 1, DumpFile());
        }
    }
}

// MBS_FIXTURE_SCALING=1: after a batch's orientations, the per-beam data
// from which a size-rescaled beam set is re-derived (the multi-size pack
// rule, src/cuda/GpuDiffraction.cu:273-358): the absorption coefficient,
// then per beam the raw Jones matrix (8 doubles), the path (optical path of
// an external beam, projected length of an internal one), the sign (-1 for
// external beams and internal ones with an odd number of interactions) and
// the absorption path lengths (count, then values).
static void WriteScaling(const std::vector<PreparedOrientation> &prepared, int start, int end,
                         double cAbs)
{
    PutTag("SCAL");
    PutF64(cAbs);
    for (int i = start; i < end; ++i)
        for (const PreparedBeam &pb : prepared[i].beams)
        {
            const complex raw[4] = {pb.origBeam.J.m11, pb.origBeam.J.m12, pb.origBeam.J.m21,
                                    pb.origBeam.J.m22};
            for (int k = 0; k < 4; ++k)
            {
                PutF64(real(raw[k]));
                PutF64(imag(raw[k]));
            }
            PutF64(pb.isExternal ? pb.origBeam.opticalPath : pb.info.projLenght);
            double sign = pb.isExternal ? -1.0 : 1.0;
            if (!pb.isExternal && (pb.origBeam.nActs & 1))
                sign = -sign;
            PutF64(sign);
            PutI32((int)pb.absorptionPaths.size());
            for (double p : pb.absorptionPaths)
                PutF64(p);
        }
}

static void WriteMueller(const char *tag, const Arr2D &m, int nAz, int nZen)
{
    PutTag(tag);
    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
        {
            const double *c = m.RawCell(p, t);
            std::fwrite(c, sizeof(double), 16, DumpFile());
        }
}

// ---- HandlerPO GPU members (src/handler/HandlerPO.h:98-146) -----------------
bool HandlerPO::HandleOrientationsToLocalGpu(
    const std::vector<PreparedOrientation> &prepared, int start, int count,
    Arr2D &localM, Arr2D &localM_noshadow, double scale, double waveIndex)
{
    if (scale != 1.0 || waveIndex != 0.0)
    {
        std::fprintf(stderr, "fixture dump: scaled batches are not dumped\n");
        return false;
    }
    const int nAz = m_sphere.nAzimuth;
    const int nZen = m_sphere.nZenith;
    WriteHeader(m_sphere, m_waveIndex, m_eps1, m_complWave, m_invComplWave,
                m_legacySign, ComputeNoShadow());

    Arr2D batchM(nAz + 1, nZen + 1, 4, 4);
    Arr2D batchNs(nAz + 1, nZen + 1, 4, 4);
    batchM.ClearArr();
    batchNs.ClearArr();
    std::vector<Arr2DC> localJ, localJns;
    Arr2DC tmp(nAz + 1, nZen + 1, 2, 2);
    tmp.ClearArr();
    localJ.push_back(tmp);
    localJns.push_back(tmp);

    // MBS_FIXTURE_NO_EXPECTED=1: beams only (the expected grids are left
    // zero), for callers that need the reference's beam sets but not the
    // reference's CPU result for them.
    const char *noExpected = std::getenv("MBS_FIXTURE_NO_EXPECTED");
    const bool expected = !(noExpected && noExpected[0] == '1');
    const int end = std::min((int)prepared.size(), start + count);
    PutTag("BTCH");
    PutI32(end - start);
    for (int i = start; i < end; ++i)
    {
        const PreparedOrientation &po = prepared[i];
        WriteOrientation(po);
        if (!expected)
            continue;
        if (!po.beams.empty())
            HandleBeamsToLocal(po, batchM, localJ, &localJns);
        AddToMuellerLocal(localJ, po.sinZenith, batchM, nAz, nZen);
        AddToMuellerLocal(localJns, po.sinZenith, batchNs, nAz, nZen);
        localJ[0].ClearArr();
        localJns[0].ClearArr();
    }
    const char *scaling = std::getenv("MBS_FIXTURE_SCALING");
    if (scaling && scaling[0] == '1')
        WriteScaling(prepared, start, end, AbsorptionCoefficient());
    WriteMueller("MFUL", batchM, nAz, nZen);
    WriteMueller("MNSH", batchNs, nAz, nZen);
    std::fflush(DumpFile());

    for (int p = 0; p < nAz; ++p)
        for (int t = 0; t <= nZen; ++t)
        {
            localM.insert(p, t, batchM(p, t));
            if (ComputeNoShadow())
                localM_noshadow.insert(p, t, batchNs(p, t));
        }
    return true;
}

bool HandlerPO::HandleOrientationsToLocalGpu(
    const std::vector<PreparedOrientation> &prepared, Arr2D &localM,
    Arr2D &localM_noshadow)
{
    return HandleOrientationsToLocalGpu(prepared, 0, (int)prepared.size(),
                                        localM, localM_noshadow, 1.0, 0.0);
}

bool HandlerPO::HandleOrientationsToLocalGpuOnDevice(
    const std::vector<PreparedOrientation> &prepared, int start, int count,
    int, Arr2D &localM, Arr2D &localM_noshadow, double scale,
    double waveIndex, unsigned long long)
{
    return HandleOrientationsToLocalGpu(prepared, start, count, localM,
                                        localM_noshadow, scale, waveIndex);
}

bool HandlerPO::HandleBeamsToLocalGpu(const PreparedOrientation &prepared,
                                      Arr2D &localM, Arr2D &localM_noshadow)
{
    std::vector<PreparedOrientation> one(1, prepared);
    return HandleOrientationsToLocalGpu(one, 0, 1, localM, localM_noshadow,
                                        1.0, 0.0);
}

bool HandlerPO::HandleOrientationsToLocalGpuFftPhi(
    const std::vector<PreparedOrientation> &, int, int, Arr2D &, Arr2D &,
    double, double)
{
    return false;
}

bool HandlerPO::HandleOrientationsToLocalGpuMultiK(
    const std::vector<PreparedOrientation> &, int, int,
    const std::vector<double> &, double, std::vector<Arr2D> &)
{
    return false;
}

bool HandlerPO::HandleOrientationsToLocalGpuFftPhiMultiK(
    const std::vector<PreparedOrientation> &, int, int,
    const std::vector<double> &, double, std::vector<Arr2D> &)
{
    return false;
}

bool HandlerPO::DiffractThetasGpu(const std::vector<PreparedOrientation> &,
                                  const double *, int, std::vector<double> &)
{
    return false;
}

int HandlerPO::SelectGpuOrientationBatchSize(
    const std::vector<PreparedOrientation> &, int, int maxCount) const
{
    return maxCount > 0 ? maxCount : 1;
}
