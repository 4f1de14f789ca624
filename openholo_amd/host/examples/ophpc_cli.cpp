// Command-line generator with the reference CLI's flags
// (Colab/OpenholoGeneration.cpp:119-195, point-cloud and WRP algorithms):
//   -a 0|4  -c config.xml  -i points.ply  -e encode  -m 0|1  -p 0|1  -f 0|1
// -a 0: point cloud, -a 4: WRP; -m 0: MODE_FLOAT|MODE_FASTMATH, -m 1:
// MODE_DOUBLE; -p 1: MODE_GPU (both -p values run on the GPU here); -f 0: RS,
// 1: Fresnel.  Writes <out>/PointCloud_<R-S|Fresnel>_<Phase|...>.bmp like
// ex_PointCloud, or <out>/WRP_<Phase|...>.bmp like ex_WRP (:100-117).
// Extensions used by the test suite:
//   -o DIR            output directory (default Result)
//   --vtx FILE        raw 56-byte Vertex records instead of -i
//   --mfma bf16x1|bf16x3|fp16x3   matrix-core variant (Fresnel: all three; RS: the BF16
//                                 ones, for the points whose tile expansion holds)
//   --dump-field F    raw complex128 fields of all channels (row-major)
//   --dump-enc F      raw float64 encoded buffers
//   --dump-u8 F       raw uint8 normalised buffers
//   --ohc F           also save the complex field as OHC
//   --no-generate     stop after loading (with --dump-points / --dump-config)
//   --dump-points F   raw 56-byte Vertex records as loaded
//   --dump-config     print the parsed configuration as one JSON line
//   --dump-wrp F      (WRP) raw complex128 WRP plane of the last channel
//   --passband P      (point cloud) call encoding(encode, P), the two-argument
//                     overload (-e 7: single-side-band encodeHologram)
//   --load-ohc F      only load the OHC file F (Openholo::loadAsOhc), print its
//                     configuration as one JSON line, write --dump-field
//   --resize W,H      (point cloud) after the first hologram, setResolution(W,H)
//                     and generate again; encoding and the outputs then use it
//   --dump-field2 F   raw complex128 fields of the second hologram
//   --conj-field      conjugate complex_H through getComplexField() after
//                     generation (and after --dump-field), before encoding
//   --set-pixel-number W,H  call setPixelNumber(W,H) alone (no setResolution)
//                     after loading, then generateHologram, encoding and
//                     normalize; print each call's lastError() and exit 3
//                     when all three refused the stale host buffers
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "ophPointCloud.h"
#include "ophWRP.h"

static const char* encode_name(unsigned e)
{
    switch (e) {
    case 0: return "Phase";
    case 1: return "Amplitude";
    case 2: return "Real";
    case 3: return "Imaginary";
    case 7: return "SSB";
    default: return "Unknown";
    }
}

static bool write_raw(const std::string& path, const void* p, size_t bytes)
{
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(p, 1, bytes, fp) == bytes;
    std::fclose(fp);
    return ok;
}

// ex_WRP (Colab/OpenholoGeneration.cpp:100-117): readConfig, loadPointCloud,
// generateHologram, encoding, normalize, save (waveNum x 8 bits per pixel)
static int run_wrp(const std::string& config, const std::string& input, unsigned mode, unsigned encode,
                   const std::string& outdir, const std::string& dump_points, const std::string& dump_field,
                   const std::string& dump_enc, const std::string& dump_u8, const std::string& dump_wrp)
{
    ophWRP* h = new ophWRP();
    auto fail = [&]() {
        h->release();
        return 1;
    };
    h->SetMode(mode);
    if (!h->readConfig(config.c_str())) return fail();
    if (h->loadPointCloud(input.c_str()) < 0) return fail();
    if (!dump_points.empty() &&
        !write_raw(dump_points, h->getPointCloudModel(), sizeof(Vertex) * h->getNumberOfPoints()))
        return fail();
    h->generateHologram();
    if (!h->lastError().empty()) {
        std::fprintf(stderr, "generateHologram failed: %s\n", h->lastError().c_str());
        return fail();
    }
    const uint nch = h->getContext().waveNum;
    const long long npx = (long long)h->getContext().pixel_number[_X] * h->getContext().pixel_number[_Y];
    if (!dump_field.empty()) {
        FILE* fp = std::fopen(dump_field.c_str(), "wb");
        for (uint c = 0; c < nch; ++c) std::fwrite(h->getComplexField()[c], 16, npx, fp);
        std::fclose(fp);
    }
    if (!dump_wrp.empty() && !write_raw(dump_wrp, h->getWRPBuff(), 16 * (size_t)npx)) return fail();
    h->encoding(encode);
    h->normalize();
    if (!dump_enc.empty()) {
        FILE* fp = std::fopen(dump_enc.c_str(), "wb");
        for (uint c = 0; c < nch; ++c) std::fwrite(h->getEncodedBuffer()[c], 8, npx, fp);
        std::fclose(fp);
    }
    if (!dump_u8.empty()) {
        FILE* fp = std::fopen(dump_u8.c_str(), "wb");
        for (uint c = 0; c < nch; ++c) std::fwrite(h->getNormalizedBuffer()[c], 1, npx, fp);
        std::fclose(fp);
    }
    ::mkdir(outdir.c_str(), 0755);
    char path[1024];
    std::snprintf(path, sizeof path, "%s/WRP_%s.bmp", outdir.c_str(), encode_name(encode));
    const bool ok = h->save(path, (uint8_t)(nch * 8));
    h->release();
    return ok ? 0 : 1;
}

int main(int argc, char** argv)
{
    std::string config, input, vtx, outdir = "Result", dump_field, dump_enc, dump_u8, ohc, dump_points, dump_wrp,
        load_ohc, dump_field2;
    int resize_w = 0, resize_h = 0, spn_w = 0, spn_h = 0;
    bool no_generate = false, dump_config = false, two_arg_encoding = false, conj_field = false;
    unsigned passband = 0;
    int alg = 0, flag = 0;
    unsigned encode = 0, mode = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const char* v = i + 1 < argc ? argv[i + 1] : "";
        if (a == "-a") alg = std::atoi(v), ++i;
        else if (a == "-c") config = v, ++i;
        else if (a == "-i") input = v, ++i;
        else if (a == "-e") encode = (unsigned)std::atoi(v), ++i;
        else if (a == "-m") {
            if (std::atoi(v) == 0) mode |= MODE_FLOAT | MODE_FASTMATH;
            else mode |= MODE_DOUBLE;
            ++i;
        } else if (a == "-p") {
            if (std::atoi(v) == 1) mode |= MODE_GPU;
            ++i;
        } else if (a == "-f") flag = std::atoi(v), ++i;
        else if (a == "-o") outdir = v, ++i;
        else if (a == "--vtx") vtx = v, ++i;
        else if (a == "--mfma") {
            const std::string m = v;
            mode |= m == "bf16x1" ? MODE_MFMA_BF16X1 : m == "bf16x3" ? MODE_MFMA_BF16X3 : MODE_MFMA_FP16X3;
            ++i;
        } else if (a == "--dump-field") dump_field = v, ++i;
        else if (a == "--dump-enc") dump_enc = v, ++i;
        else if (a == "--dump-u8") dump_u8 = v, ++i;
        else if (a == "--ohc") ohc = v, ++i;
        else if (a == "--dump-points") dump_points = v, ++i;
        else if (a == "--no-generate") no_generate = true;
        else if (a == "--dump-config") dump_config = true;
        else if (a == "--dump-wrp") dump_wrp = v, ++i;
        else if (a == "--load-ohc") load_ohc = v, ++i;
        else if (a == "--resize") {
            if (std::sscanf(v, "%d,%d", &resize_w, &resize_h) != 2 || resize_w < 1 || resize_h < 1) {
                std::fprintf(stderr, "--resize takes W,H\n");
                return 1;
            }
            ++i;
        } else if (a == "--set-pixel-number") {
            if (std::sscanf(v, "%d,%d", &spn_w, &spn_h) != 2 || spn_w < 1 || spn_h < 1) {
                std::fprintf(stderr, "--set-pixel-number takes W,H\n");
                return 1;
            }
            ++i;
        } else if (a == "--dump-field2") dump_field2 = v, ++i;
        else if (a == "--conj-field") conj_field = true;
        else if (a == "--passband") passband = (unsigned)std::atoi(v), two_arg_encoding = true, ++i;
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    if (!load_ohc.empty()) {
        ophPointCloud* h = new ophPointCloud();
        const bool ok = h->loadAsOhc(load_ohc.c_str());
        if (ok) {
            OphConfig& c = h->getContext();
            std::printf("{\"pn\": [%d, %d], \"pp\": [%.17g, %.17g], \"lambdas\": [", c.pixel_number[_X],
                        c.pixel_number[_Y], c.pixel_pitch[_X], c.pixel_pitch[_Y]);
            for (uint w = 0; w < c.waveNum; ++w) std::printf("%s%.17g", w ? ", " : "", c.wave_length[w]);
            std::printf("]}\n");
            if (!dump_field.empty()) {
                const long long npx = (long long)c.pixel_number[_X] * c.pixel_number[_Y];
                FILE* fp = std::fopen(dump_field.c_str(), "wb");
                for (uint w = 0; fp && w < c.waveNum; ++w) std::fwrite(h->getComplexField()[w], 16, npx, fp);
                if (fp) std::fclose(fp);
            }
        }
        h->release();
        return ok ? 0 : 1;
    }
    if (alg == 4)
        return run_wrp(config, input, mode, encode, outdir, dump_points, dump_field, dump_enc, dump_u8, dump_wrp);
    if (alg != 0) {
        std::fprintf(stderr, "only the point-cloud (-a 0) and WRP (-a 4) algorithms are provided\n");
        return 2;
    }
    ophPointCloud* h = new ophPointCloud();
    auto fail = [&]() {
        h->release();
        return 1;
    };
    h->SetMode(mode);
    if (!h->readConfig(config.c_str())) return fail();
    if (!vtx.empty()) {
        FILE* fp = std::fopen(vtx.c_str(), "rb");
        if (!fp) return fail();
        std::fseek(fp, 0, SEEK_END);
        const long bytes = std::ftell(fp);
        std::fseek(fp, 0, SEEK_SET);
        std::vector<Vertex> v(bytes / sizeof(Vertex));
        if (std::fread(v.data(), sizeof(Vertex), v.size(), fp) != v.size()) {
            std::fclose(fp);
            return fail();
        }
        std::fclose(fp);
        h->setNumberOfPoints(v.size());
        h->setPointCloudModel(v.data());
    } else if (h->loadPointCloud(input.c_str()) < 0) {
        return fail();
    }
    if (dump_config) {
        OphConfig& c = h->getContext();
        vec3 sc;
        h->getScale(sc);
        std::printf("{\"pn\": [%d, %d], \"pp\": [%.17g, %.17g], \"lambdas\": [", c.pixel_number[_X],
                    c.pixel_number[_Y], c.pixel_pitch[_X], c.pixel_pitch[_Y]);
        for (uint w = 0; w < c.waveNum; ++w) std::printf("%s%.17g", w ? ", " : "", c.wave_length[w]);
        std::printf("], \"scale\": [%.17g, %.17g, %.17g], \"distance\": %.17g, \"rotate\": %d, \"merge\": %d, "
                    "\"flip\": %d, \"field_length\": %.17g, \"points\": %llu}\n",
                    sc[0], sc[1], sc[2], h->getDistance(), (int)h->getImageConfig().rotate,
                    (int)h->getImageConfig().merge, h->getImageConfig().flip, h->getFieldLength(),
                    (unsigned long long)h->getNumberOfPoints());
    }
    if (!dump_points.empty()) {
        FILE* fp = std::fopen(dump_points.c_str(), "wb");
        if (!fp) return fail();
        std::fwrite(h->getPointCloudModel(), sizeof(Vertex), h->getNumberOfPoints(), fp);
        std::fclose(fp);
    }
    if (no_generate) {
        h->release();
        return 0;
    }
    if (spn_w > 0) {
        h->setPixelNumber(spn_w, spn_h);
        int refused = 0;
        const Real t = h->generateHologram((uint)flag);
        std::printf("generateHologram: %g s, lastError: %s\n", t, h->lastError().c_str());
        refused += t == 0.0 && h->lastError().find("host buffers do not match") != std::string::npos;
        h->encoding(encode);
        std::printf("encoding lastError: %s\n", h->lastError().c_str());
        refused += h->lastError().rfind("encoding: host buffers do not match", 0) == 0;
        h->normalize();
        std::printf("normalize lastError: %s\n", h->lastError().c_str());
        refused += h->lastError().rfind("normalize: host buffers do not match", 0) == 0;
        h->release();
        return refused == 3 ? 3 : 1;
    }
    const Real t = h->generateHologram((uint)flag);
    if (t <= 0.0) {
        std::fprintf(stderr, "generateHologram failed: %s\n", h->lastError().c_str());
        return fail();
    }
    std::printf("generateHologram: %.6f s\n", t);
    const uint nch = h->getContext().waveNum;
    long long npx = (long long)h->getContext().pixel_number[_X] * h->getContext().pixel_number[_Y];
    if (!ohc.empty() && !h->saveAsOhc(ohc.c_str())) return fail();
    if (!dump_field.empty()) {
        FILE* fp = std::fopen(dump_field.c_str(), "wb");
        for (uint c = 0; c < nch; ++c) std::fwrite(h->getComplexField()[c], 16, npx, fp);
        std::fclose(fp);
    }
    if (resize_w > 0) {
        // a second hologram at another resolution in the same process (the
        // host planes and device buffers are re-allocated); encoding and the
        // outputs below then use it
        h->setResolution(ivec2(resize_w, resize_h));
        const Real t2 = h->generateHologram((uint)flag);
        if (t2 <= 0.0) {
            std::fprintf(stderr, "generateHologram after --resize failed: %s\n", h->lastError().c_str());
            return fail();
        }
        std::printf("generateHologram (%dx%d): %.6f s\n", resize_w, resize_h, t2);
        npx = (long long)resize_w * resize_h;
        if (!dump_field2.empty()) {
            FILE* fp = std::fopen(dump_field2.c_str(), "wb");
            for (uint c = 0; c < nch; ++c) std::fwrite(h->getComplexField()[c], 16, npx, fp);
            std::fclose(fp);
        }
    }
    if (conj_field) {
        Complex<Real>** H = h->getComplexField();
        for (uint c = 0; c < nch; ++c)
            for (long long i = 0; i < npx; ++i) H[c][i][_IM] = -H[c][i][_IM];
    }
    if (two_arg_encoding) h->encoding(encode, passband);
    else h->encoding(encode);
    if (!h->lastError().empty()) {
        std::fprintf(stderr, "encoding failed: %s\n", h->lastError().c_str());
        return fail();
    }
    h->normalize();
    if (!dump_enc.empty()) {
        FILE* fp = std::fopen(dump_enc.c_str(), "wb");
        for (uint c = 0; c < nch; ++c) std::fwrite(h->getEncodedBuffer()[c], 8, npx, fp);
        std::fclose(fp);
    }
    if (!dump_u8.empty()) {
        FILE* fp = std::fopen(dump_u8.c_str(), "wb");
        for (uint c = 0; c < nch; ++c) std::fwrite(h->getNormalizedBuffer()[c], 1, npx, fp);
        std::fclose(fp);
    }
    ::mkdir(outdir.c_str(), 0755);
    char path[1024];
    std::snprintf(path, sizeof path, "%s/PointCloud_%s_%s.bmp", outdir.c_str(), flag == 0 ? "R-S" : "Fresnel",
                  encode_name(encode));
    const bool ok = h->save(path, (uint8_t)(nch * 8));
    h->release();
    return ok ? 0 : 1;
}
