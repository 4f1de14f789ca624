// The canonical point-cloud call sequence (the reference's
// Example/examPointCloud/examPointCloud.cpp:54-132), unchanged at the call
// sites: readConfig -> loadPointCloud -> SetMode -> generateHologram ->
// saveAsOhc -> encoding(PHASE) -> normalize -> save -> release.
//
//   examPointCloud <config.xml> <points.ply> <out_prefix> [rs|fresnel] [mode]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "ophPointCloud.h"

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s config.xml points.ply out_prefix [rs|fresnel] [mode]\n", argv[0]);
        return 2;
    }
    const std::string diff = argc > 4 ? argv[4] : "rs";
    const unsigned mode = argc > 5 ? (unsigned)std::strtoul(argv[5], nullptr, 0) : MODE_GPU;

    ophPointCloud* Hologram = new ophPointCloud();
    if (!Hologram->readConfig(argv[1])) {
        std::fprintf(stderr, "Fail to load Config file : %s\n", argv[1]);
        return 1;
    }
    if (Hologram->loadPointCloud(argv[2]) < 0) {
        std::fprintf(stderr, "Fail to load Point Cloud file : %s\n", argv[2]);
        return 1;
    }
    Hologram->SetMode(mode);
    const Real t = Hologram->generateHologram(diff == "fresnel" ? ophPointCloud::PC_DIFF_FLAG::PC_DIFF_FRESNEL
                                                                : ophPointCloud::PC_DIFF_FLAG::PC_DIFF_RS);
    if (t <= 0.0) {
        std::fprintf(stderr, "generateHologram failed: %s\n", Hologram->lastError().c_str());
        return 1;
    }
    std::printf("generateHologram: %.6f s\n", t);
    const std::string out = argv[3];
    if (!Hologram->saveAsOhc((out + ".ohc").c_str())) return 1;
    Hologram->encoding(Hologram->ENCODE_PHASE);
    Hologram->normalize();
    if (!Hologram->save((out + ".bmp").c_str())) return 1;
    Hologram->release();
    return 0;
}
