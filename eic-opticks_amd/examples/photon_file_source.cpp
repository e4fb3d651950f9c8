// photon_file_source.cpp -- the GPUPhotonFileSource application (src/GPUPhotonFileSource.cpp,
// src/GPUPhotonFileSource.h) without Geant4: photons read from a text file are simulated on
// the GPU through the oxr facade (include/oxr_g4cx.hpp). It prints the lines the
// reference's test script reads ("Loaded N photons from FILE", "Opticks: NumHits:  N") and
// writes opticks_hits_output.txt in the reference's format.
//
//   photon_file_source -g <CSGFoundry folder> -p <photons.txt> [-s seed] [-o hits.txt]
//
// -g names a CSGFoundry folder (the reference takes GDML and converts it through Geant4;
// oxr.gdml + oxr.cfio write such a folder from GDML). -s is accepted for command-line
// compatibility: in the reference it seeds the host-side Geant4 run, not the GPU streams.
#include "oxr_g4cx.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

static int usage(const char* msg) {
    std::cerr << msg << "\nusage: photon_file_source -g <CSGFoundry folder> -p <photons.txt> [-s seed] [-o hits.txt]\n";
    return 1;
}

int main(int argc, char** argv) {
    std::string cfbase, photon_file, out_file = "opticks_hits_output.txt";
    long seed = -1;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&](const char* what) -> const char* {
            if (i + 1 >= argc) { std::cerr << what << " needs a value\n"; std::exit(1); }
            return argv[++i];
        };
        if (a == "-g" || a == "--geometry") cfbase = next("-g");
        else if (a == "-p" || a == "--photons") photon_file = next("-p");
        else if (a == "-s" || a == "--seed") seed = std::atol(next("-s"));
        else if (a == "-o" || a == "--output") out_file = next("-o");
        else if (a == "-h" || a == "--help") return usage("");
        else return usage(("unknown argument " + a).c_str());
    }
    if (photon_file.empty()) return usage("Required argument not used: -p");
    if (cfbase.empty()) return usage("Required argument not used: -g");
    if (seed >= 0) std::cout << "Random seed set to: " << seed << std::endl;
    try {
        oxr::G4CXOpticks::SetGeometry(cfbase.c_str());
        std::vector<oxr::sphoton> photons = oxr::load_photons_txt(photon_file);
        if (photons.empty()) {
            std::cerr << "ERROR: no photons loaded from " << photon_file << std::endl;
            return 1;
        }
        std::cout << "Loaded " << photons.size() << " photons from " << photon_file << std::endl;
        oxr::SEvt* sev = oxr::SEvt::Get_EGPU();
        sev->SetInputPhoton(photons);
        oxr::G4CXOpticks* gx = oxr::G4CXOpticks::Get();
        const int eventID = 0;
        gx->simulate(eventID, false);
        unsigned num_hits = sev->GetNumHit(0);
        std::cout << "Opticks: NumHits:  " << num_hits << std::endl;
        std::ofstream os(out_file);
        if (!os.is_open()) {
            std::cerr << "Error opening output file!" << std::endl;
        } else {
            for (unsigned idx = 0; idx < num_hits; idx++) {
                oxr::sphoton hit;
                sev->getHit(hit, idx);
                os << double(hit.time) << " " << double(hit.wavelength) << "  "
                   << "(" << double(hit.pos[0]) << ", " << double(hit.pos[1]) << ", " << double(hit.pos[2]) << ")  "
                   << "(" << double(hit.mom[0]) << ", " << double(hit.mom[1]) << ", " << double(hit.mom[2]) << ")  "
                   << "(" << double(hit.pol[0]) << ", " << double(hit.pol[1]) << ", " << double(hit.pol[2]) << ")"
                   << std::endl;
            }
        }
        gx->reset(eventID);
        std::cout << "Opticks: TotalHits:  " << num_hits << std::endl;
        oxr::G4CXOpticks::Finalize();
    } catch (const std::exception& e) {
        std::cerr << "ERROR: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
