// ophPointCloud: point-cloud CGH on the MI355X through libopenholo_amd.
#include "ophPointCloud.h"

#include <chrono>
#include <cstring>
#include <map>
#include <vector>

#include "../../../include/openholo_amd.h"

ophPointCloud::ophPointCloud(void) : ophGen() {}

ophPointCloud::ophPointCloud(const char* pc_file, const char* cfg_file) : ophGen()
{
    if (loadPointCloud(pc_file) == -1) OPH_LOG("<FAILED> Load point cloud data file('%s')", pc_file);
    if (!readConfig(cfg_file)) OPH_LOG("<FAILED> Load config specification data file('%s')", cfg_file);
}

ophPointCloud::~ophPointCloud(void) {}

void ophPointCloud::setPointCloudModel(Vertex* vertex)
{
    // ophgen/src/ophPointCloud.h:577-584: no-op below one point, deep copy
    if (pc_data_.n_points < 1) return;
    delete[] pc_data_.vertices;
    pc_data_.vertices = new Vertex[pc_data_.n_points];
    std::memcpy(pc_data_.vertices, vertex, sizeof(Vertex) * pc_data_.n_points);
    ++model_version;
}

void ophPointCloud::getFocalLength(Real* lens_in, Real* lens_out, Real* lens_eye_piece)
{
    if (lens_in) *lens_in = pc_config_.focal_length_lens_in;
    if (lens_out) *lens_out = pc_config_.focal_length_lens_out;
    if (lens_eye_piece) *lens_eye_piece = pc_config_.focal_length_lens_eye_piece;
}

int ophPointCloud::loadPointCloud(const char* pc_file)
{
    const int n = ophGen::loadPointCloud(pc_file, &pc_data_);
    ++model_version;
    return n;
}

bool ophPointCloud::readConfig(const char* fname)
{
    // ophgen/src/ophPointCloud.cpp:83-149: common tags, then ScaleX/Y/Z, Distance
    if (!ophGen::readConfig(fname)) return false;
    bool ok = true;
    std::FILE* fp = std::fopen(fname, "rb");
    if (!fp) return false;
    std::string text;
    char buf[4096];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, r);
    std::fclose(fp);
    auto tag = [&](const char* name, Real* v) {
        const std::string open = std::string("<") + name + ">", close = std::string("</") + name + ">";
        const size_t a = text.find(open);
        const size_t b = a == std::string::npos ? a : text.find(close, a);
        if (a == std::string::npos || b == std::string::npos ||
            std::sscanf(text.substr(a + open.size(), b - a - open.size()).c_str(), "%lf", v) != 1) {
            OPH_LOG("<FAILED> Not found node : '%s' (Double)\n", name);
            ok = false;
        }
    };
    tag("ScaleX", &pc_config_.scale[_X]);
    tag("ScaleY", &pc_config_.scale[_Y]);
    tag("ScaleZ", &pc_config_.scale[_Z]);
    tag("Distance", &pc_config_.distance);
    initialize();
    return ok;
}

Real ophPointCloud::generateHologram(uint diff_flag)
{
    // ophgen/src/ophPointCloud.cpp:151-191: validate, reset, generate, return seconds
    const auto begin = std::chrono::system_clock::now();
    if (diff_flag != PC_DIFF_RS && diff_flag != PC_DIFF_FRESNEL) {
        OPH_LOG("<FAILED> Wrong parameters.");
        return 0.0;
    }
    m_lastError.clear();
    resetBuffer();
    if (!genCghPointCloudGPU(diff_flag)) return 0.0;
    m_nProgress = 0;
    return std::chrono::duration<Real>(std::chrono::system_clock::now() - begin).count();
}

bool ophPointCloud::genCghPointCloudGPU(uint diff_flag)
{
    // replaces ophgen/src/ophPointCloud_GPU.cpp:179-412: vertices stay resident
    // on the device between calls; one diffraction launch per channel
    if (!hostPlanesReady("generateHologram")) return false;
    ohx_ctx* c = device();
    if (!c) return false;
    const uint nch = context_.waveNum;
    if (!ensureDeviceBuffers()) return false;
    const ulonglong n = pc_data_.n_points;
    std::vector<Vertex> vw;
    const Vertex* src = pc_data_.vertices;
    if (is_ViewingWindow && n) {
        vw.assign(src, src + n);
        transVW((int)n, vw.data(), vw.data());
        src = vw.data();
    }
    if (n && (is_ViewingWindow || device_version != model_version || d_vertices_n != n)) {
        if (d_vertices_n < n) {
            ohx_free(c, d_vertices);
            d_vertices = nullptr;
            d_vertices_n = 0;
            const int st = ohx_malloc(c, sizeof(Vertex) * n, &d_vertices);
            if (st) {
                fail("vertex allocation", st);
                return false;
            }
        }
        const int st = ohx_memcpy_h2d(c, d_vertices, src, sizeof(Vertex) * n);
        if (st) {
            fail("vertex upload", st);
            return false;
        }
        d_vertices_n = n;
        device_version = is_ViewingWindow ? 0 : model_version;
    }
    // channel ch's field is read back while channel ch + 1 computes
    auto launch = [&](uint ch) {
        const Real lambda = context_.wave_length[ch];
        context_.k = (2 * M_PI) / lambda;
        ohx_pc_params p{};
        p.pn_x = context_.pixel_number[_X];
        p.pn_y = context_.pixel_number[_Y];
        p.offset_x = context_.offset[_X];
        p.offset_y = context_.offset[_Y];
        p.pp_x = context_.pixel_pitch[_X];
        p.pp_y = context_.pixel_pitch[_Y];
        for (int a = 0; a < 3; ++a) p.scale[a] = pc_config_.scale[a];
        p.distance = pc_config_.distance;
        p.lambda = lambda;
        p.channel = (int32_t)ch;
        p.use_phase = 0;
        int st = ohx_field_zero(c, d_field[ch], d_pixels);
        if (!st) st = ohx_pc_diffract(c, d_vertices, (int64_t)n, d_field[ch], &p, diff_flag, m_mode);
        return st;
    };
    // each channel's copy is ordered after its own launch only, so it runs
    // while the next channel's kernels do
    int st = OHX_OK;
    for (uint ch = 0; !st && ch < nch; ++ch) {
        st = launch(ch);
        if (!st) st = ohx_memcpy_d2h_async_after(c, complex_H[ch], d_field[ch], 16 * (size_t)d_pixels);
    }
    const int st_copy = ohx_copy_synchronize(c);
    if (!st) st = st_copy;
    m_nProgress = 100;
    if (st) {
        (void)ohx_synchronize(c);
        fail("diffraction", st);
        return false;
    }
    return true;
}

void ophPointCloud::encoding(unsigned int ENCODE_FLAG) { ophGen::encoding(ENCODE_FLAG); }

void ophPointCloud::encoding(unsigned int ENCODE_FLAG, unsigned int SSB_PASSBAND)
{
    // ophgen/src/ophPointCloud.cpp:247-251
    (void)SSB_PASSBAND;
    if (ENCODE_FLAG == ENCODE_SSB) encodeHologram();
    else if (ENCODE_FLAG == ENCODE_OFFSSB) OPH_LOG("<FAILED> ENCODE_OFFSSB is not provided.\n");
}

void ophPointCloud::encodeHologram(vec2 band_limit, vec2 spectrum_shift)
{
    // ophgen/src/ophPointCloud.cpp:193-240 (the FFT round trip in closed form
    // on the GPU; complex_H is left unchanged)
    (void)band_limit;
    if (complex_H == nullptr) {
        OPH_LOG("<FAILED> Not found diffracted data.");
        return;
    }
    if (!hostPlanesReady("encodeHologram")) return;
    const int pnX = context_.pixel_number[_X], pnY = context_.pixel_number[_Y];
    m_vecEncodeSize = ivec2(pnX, pnY);
    context_.ss[_X] = pnX * context_.pixel_pitch[_X];
    context_.ss[_Y] = pnY * context_.pixel_pitch[_Y];
    if (!uploadFields()) return;
    ohx_ctx* c = device();
    for (uint ch = 0; ch < context_.waveNum; ++ch) {
        int st = ohx_encode_ssb(c, d_field[ch], pnX, pnY, context_.pixel_pitch[_X], context_.pixel_pitch[_Y],
                                spectrum_shift[_X], spectrum_shift[_Y], d_enc[ch]);
        if (!st) st = ohx_memcpy_d2h(c, m_lpEncoded[ch], d_enc[ch], 8 * (size_t)d_pixels);
        if (st) {
            fail("encodeHologram", st);
            return;
        }
    }
}

void ophPointCloud::ophFree(void)
{
    ophGen::ophFree();
    if (d_vertices) {
        ohx_ctx* c = device();
        if (c) ohx_free(c, d_vertices);
        d_vertices = nullptr;
        d_vertices_n = 0;
    }
    delete[] pc_data_.vertices;
    pc_data_.vertices = nullptr;
}
