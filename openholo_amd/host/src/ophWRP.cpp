// ophWRP: WRP hologram generation on the MI355X through libopenholo_amd.
#include "ophWRP.h"

#include <cstdio>
#include <cstring>
#include <string>

#include "../../../include/openholo_amd.h"

ophWRP::ophWRP(void) : ophGen() {}

ophWRP::~ophWRP(void) {}

void ophWRP::autoScaling()
{
    // ophgen/src/ophWRP.cpp:84-93: zmax over the scaled depths, from MIN_DOUBLE
    zmax_ = 2.2250738585072014e-308;
    for (ulonglong i = 0; i < obj_.n_points; ++i) {
        const Real z = obj_.vertices[i].point.pos[_Z] * wrp_config_.scale[_Z];
        if (zmax_ < z) zmax_ = z;
    }
}

int ophWRP::loadPointCloud(const char* pc_file)
{
    return ophGen::loadPointCloud(pc_file, &obj_);
}

bool ophWRP::readConfig(const char* fname)
{
    // ophgen/src/ophWRP.cpp:145-235: common tags, then the WRP tags
    if (!ophGen::readConfig(fname)) return false;
    bool ok = true;
    std::FILE* fp = std::fopen(fname, "rb");
    if (!fp) return false;
    std::string text;
    char buf[4096];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, r);
    std::fclose(fp);
    auto value = [&](const char* name, const char* fmt, void* out, const char* kind) {
        const std::string open = std::string("<") + name + ">", close = std::string("</") + name + ">";
        const size_t a = text.find(open);
        const size_t b = a == std::string::npos ? a : text.find(close, a);
        if (a == std::string::npos || b == std::string::npos ||
            std::sscanf(text.substr(a + open.size(), b - a - open.size()).c_str(), fmt, out) != 1) {
            OPH_LOG("<FAILED> Not found node : '%s' (%s) \n", name, kind);
            ok = false;
        }
    };
    value("FieldLength", "%lf", &wrp_config_.fieldLength, "Double");
    value("ScaleX", "%lf", &wrp_config_.scale[_X], "Double");
    value("ScaleY", "%lf", &wrp_config_.scale[_Y], "Double");
    value("ScaleZ", "%lf", &wrp_config_.scale[_Z], "Double");
    value("Distance", "%lf", &wrp_config_.propagation_distance, "Double");
    value("LocationOfWRP", "%lf", &wrp_config_.wrp_location, "Double");
    value("NumOfWRP", "%d", &wrp_config_.num_wrp, "Integer");
    initialize();
    return ok;
}

void ophWRP::generateHologram(void)
{
    // ophgen/src/ophWRP.cpp:437-463: reset, scale, generate every channel
    m_lastError.clear();
    resetBuffer();
    autoScaling();
    calculateWRP();
}

bool ophWRP::calculateWRP(void)
{
    ohx_ctx* c = device();
    if (!c) return false;
    const uint nch = context_.waveNum;
    if (!hostPlanesReady("WRP generateHologram")) return false;
    if (GetRandomPhase()) {
        m_lastError = "the WRP random phase is not supported";
        OPH_LOG("<FAILED> %s\n", m_lastError.c_str());
        return false;
    }
    if (!ensureDeviceBuffers()) return false;
    const ulonglong n = obj_.n_points;
    if (n && d_vertices_n != n) {
        ohx_free(c, d_vertices);
        d_vertices = nullptr;
        d_vertices_n = 0;
        const int st = ohx_malloc(c, sizeof(Vertex) * n, &d_vertices);
        if (st) {
            fail("vertex allocation", st);
            return false;
        }
        d_vertices_n = n;
    }
    if (n) {
        const int st = ohx_memcpy_h2d(c, d_vertices, obj_.vertices, sizeof(Vertex) * n);
        if (st) {
            fail("vertex upload", st);
            return false;
        }
    }
    const long long npx = (long long)context_.pixel_number[_X] * context_.pixel_number[_Y];
    if (d_wrp_px != npx) {
        ohx_free(c, d_wrp);
        d_wrp = nullptr;
        void* p = nullptr;
        const int st = ohx_malloc(c, 16 * (size_t)npx, &p);
        if (st) {
            fail("WRP allocation", st);
            return false;
        }
        d_wrp = (double*)p;
        d_wrp_px = npx;
        delete[] p_wrp_;
        p_wrp_ = new Complex<Real>[npx];
    }
    // channel ch's field is read back while channel ch + 1 is generated
    auto launch = [&](uint ch) {
        const Real lambda = context_.wave_length[ch];
        context_.k = (2 * M_PI) / lambda;
        ohx_wrp_params p{};
        p.pn_x = context_.pixel_number[_X];
        p.pn_y = context_.pixel_number[_Y];
        p.pp_x = context_.pixel_pitch[_X];
        p.pp_y = context_.pixel_pitch[_Y];
        for (int a = 0; a < 3; ++a) p.scale[a] = wrp_config_.scale[a];
        p.lambda = lambda;
        p.channel = (int32_t)ch;
        p.random_phase = 0;
        p.wrp_location = wrp_config_.wrp_location;
        p.propagation_distance = wrp_config_.propagation_distance;
        return ohx_wrp_generate(c, d_vertices, (int64_t)n, &p, d_field[ch], d_wrp);
    };
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
        fail("WRP generation", st);
        return false;
    }
    // the plane of the last channel, as the reference leaves in p_wrp_
    st = ohx_memcpy_d2h(c, p_wrp_, d_wrp, 16 * (size_t)npx);
    if (st) {
        fail("WRP download", st);
        return false;
    }
    return true;
}

Complex<Real>** ophWRP::calculateMWRP(void)
{
    OPH_LOG("<FAILED> multiple WRPs are not supported\n");
    return nullptr;
}

void ophWRP::ophFree(void)
{
    ophGen::ophFree();
    ohx_ctx* c = device();
    if (c) {
        if (d_vertices) ohx_free(c, d_vertices);
        if (d_wrp) ohx_free(c, d_wrp);
    }
    d_vertices = nullptr;
    d_vertices_n = 0;
    d_wrp = nullptr;
    d_wrp_px = 0;
    delete[] p_wrp_;
    p_wrp_ = nullptr;
    delete[] obj_.vertices;
    obj_.vertices = nullptr;
}
