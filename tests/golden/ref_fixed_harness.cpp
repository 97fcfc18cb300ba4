// Fixed-orientation fixture harness.
//
// The reference never sends a fixed orientation to its GPU backend
// (src/tracer/TracerPO.cpp:47-68 calls the CPU HandlerPO::HandleBeams).  This
// harness sets the reference objects up the way the fixed-orientation branch of
// src/main.cpp:2000-2043 does with --cutoff-profile off (main.cpp:117-137,
// :178-196), traces one orientation, runs HandlerPO::PrepareBeams
// (src/handler/HandlerPO.cpp:1625-1810) and hands the prepared orientation to
// HandlerPO::HandleBeamsToLocalGpu, which in this build is the fixture writer
// of ref_dump_shim.cpp.  make_fixtures.py checks the expected Mueller grid
// written here against the .dat of the unmodified reference binary run on the
// same command line, which pins this harness to the real fixed path.
//
// usage: ref_fixed_harness L D n_re n_im wavelength reflections beta gamma
//                          theta0 theta1 nphi ntheta
#include "HandlerPO.h"
#include "Hexagonal.h"
#include "TracerPO.h"
#include "Tracks.h"

#include <cstdio>
#include <cstdlib>
#include <limits>
#include <vector>

int main(int argc, char **argv)
{
    if (argc != 13)
    {
        std::fprintf(stderr, "usage: %s L D n_re n_im wavelength reflections "
                             "beta gamma theta0 theta1 nphi ntheta\n", argv[0]);
        return 2;
    }
    const double height = std::atof(argv[1]);
    const double diameter = std::atof(argv[2]);
    const double nRe = std::atof(argv[3]);
    const double nIm = std::atof(argv[4]);
    const double wave = std::atof(argv[5]);
    const int reflections = std::atoi(argv[6]);
    const double betaDeg = std::atof(argv[7]);
    const double gammaDeg = std::atof(argv[8]);
    const double theta0 = std::atof(argv[9]);
    const double theta1 = std::atof(argv[10]);
    const int nPhi = std::atoi(argv[11]);
    const int nTheta = std::atoi(argv[12]);

    const bool isAbs = nIm != 0.0;
    const complex refr(nRe, isAbs ? nIm : 0.0);
    Hexagonal particle(refr, diameter, height);

    TracerPO tracer(&particle, reflections, "fixture");
    tracer.m_scattering->m_wave = wave;
    tracer.m_scattering->restriction = std::numeric_limits<double>::infinity();
    tracer.m_scattering->m_traceCutoffJRel = 0.0;
    tracer.m_scattering->m_traceCutoffAreaRel = 0.0;
    tracer.m_scattering->m_traceCutoffImportanceRel = 0.0;

    Tracks tracks;
    HandlerPO handler(&particle, &tracer.m_incidentLight, nTheta, wave);
    handler.isCoh = true;
    handler.SetFullOnly(false);
    handler.m_targetEps = 0.0;
    handler.m_beamCutoffJRel = 0.0;
    handler.m_beamCutoffAreaRel = 0.0;
    handler.m_beamCutoffImportanceRel = 0.0;
    ScatteringRange grid(DegToRad(theta0), DegToRad(theta1), nPhi, nTheta);
    handler.SetScatteringSphere(grid);
    handler.SetTracks(&tracks);
    handler.SetScattering(tracer.m_scattering);
    handler.SetAbsorptionAccounting(isAbs);

    const double b = DegToRad(betaDeg);
    const double g = DegToRad(gammaDeg);
    particle.Rotate(b, g, 0);
    std::vector<Beam> beams;
    tracer.m_scattering->ScatterLight(b, g, beams);

    PreparedOrientation prepared;
    handler.PrepareBeams(beams, 1.0, prepared);
    std::fprintf(stderr, "traced %zu beams, prepared %zu\n", beams.size(),
                 prepared.beams.size());

    Arr2D m(nPhi + 1, nTheta + 1, 4, 4), mns(nPhi + 1, nTheta + 1, 4, 4);
    m.ClearArr();
    mns.ClearArr();
    return handler.HandleBeamsToLocalGpu(prepared, m, mns) ? 0 : 1;
}
