// ophGen: configuration, buffers, point-cloud loading, encoding and 8-bit
// normalisation of the drop-in; all numerical work goes through
// libopenholo_amd's C ABI on the GPU.
#include "ophGen.h"

#include <algorithm>
#include <cmath>
#include <cctype>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

#include "../../../include/openholo_amd.h"

// ---------------------------------------------------------------------------
// device session: one context on GPU 0 for the process (the reference keeps a
// cudaWrapper singleton, ophgen/src/cudaWrapper.h:50-59)
ohx_ctx* ophGen::device()
{
    static std::mutex mu;
    static ohx_ctx* ctx = nullptr;
    static int status = 1;
    std::lock_guard<std::mutex> lk(mu);
    if (status == 1) {
        status = ohx_ctx_create(0, &ctx);
        if (status != OHX_OK) {
            ctx = nullptr;
            OPH_LOG("<FAILED> no usable MI355X (gfx950) device: %s\n", ohx_status_string(status));
        }
    }
    if (!ctx) m_lastError = "no usable gfx950 device";
    return ctx;
}

void ophGen::fail(const char* what, int status)
{
    ohx_ctx* c = device();
    m_lastError = std::string(what) + ": " + ohx_status_string(status) + (c ? std::string(" (") + ohx_last_error(c) + ")" : "");
    OPH_LOG("<FAILED> %s\n", m_lastError.c_str());
}

ophGen::ophGen(void) {}

ophGen::~ophGen(void) {}

// ---------------------------------------------------------------------------
void ophGen::initialize(void)
{
    // ophgen/src/ophGen.cpp:146-205: one complex field, one encoded and one
    // normalised buffer per channel, zero-filled
    const uint nch = context_.waveNum;
    const long long n = (long long)resCfg[_X] * resCfg[_Y];
    if (nch < 1 || nch > 3 || resCfg[_X] < 1 || resCfg[_Y] < 1 || n > (1LL << 28)) {
        // the buffers stay as they were (the upstream code would allocate garbage sizes)
        OPH_LOG("<FAILED> initialize: need 1..3 wavelengths and 1..2^28 pixels\n");
        m_lastError = "initialize: need 1..3 wavelengths and 1..2^28 pixels";
        return;
    }
    if (m_lpEncoded) {
        for (uint i = 0; i < m_nOldChannel; ++i) delete[] m_lpEncoded[i];
        delete[] m_lpEncoded;
    }
    if (m_lpNormalized) {
        for (uint i = 0; i < m_nOldChannel; ++i) delete[] m_lpNormalized[i];
        delete[] m_lpNormalized;
    }
    allocComplexField(nch, (uint64_t)n);
    m_lpEncoded = new Real*[nch];
    m_lpNormalized = new uchar*[nch];
    for (uint i = 0; i < nch; ++i) {
        m_lpEncoded[i] = new Real[n]();
        m_lpNormalized[i] = new uchar[n]();
    }
    m_nOldChannel = nch;
    m_vecEncodeSize = context_.pixel_number;
}

void ophGen::resetBuffer()
{
    // ophgen/src/ophGen.cpp:801-814 (the complex field is overwritten whole by
    // the next generation, so only the derived buffers are cleared here)
    const long long n2 = (long long)m_vecEncodeSize[_X] * m_vecEncodeSize[_Y];
    for (uint ch = 0; ch < context_.waveNum && ch < m_nOldChannel; ++ch) {
        if (m_lpEncoded && m_lpEncoded[ch]) std::memset(m_lpEncoded[ch], 0, sizeof(Real) * n2);
        if (m_lpNormalized && m_lpNormalized[ch]) std::memset(m_lpNormalized[ch], 0, n2);
    }
}

void ophGen::setResolution(ivec2 resolution)
{
    resCfg = resolution;
    if (context_.pixel_number[_X] != resolution[_X] || context_.pixel_number[_Y] != resolution[_Y]) {
        setPixelNumber(resolution);
        initialize();
    }
}

void ophGen::transVW(int nVertex, Vertex* dst, Vertex* src)
{
    const Real fl = m_dFieldLength;
    for (int i = 0; i < nVertex; ++i)
        for (int a = 0; a < 3; ++a)
            dst[i].point.pos[a] = -fl * src[i].point.pos[a] / (src[i].point.pos[a] - fl);
}

// ---------------------------------------------------------------------------
// XML configuration (tag names of ophgen/src/ophGen.cpp:244-327)
namespace {
// direct children <Tag>text</Tag> of the document's first element; comments,
// declarations and processing instructions skipped; first occurrence wins
bool xml_children(const char* fname, std::map<std::string, std::string>& out)
{
    std::ifstream f(fname, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    size_t i = 0;
    auto skip_misc = [&](size_t& p) {
        for (;;) {
            while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
            if (s.compare(p, 4, "<!--") == 0) {
                const size_t e = s.find("-->", p + 4);
                if (e == std::string::npos) return false;
                p = e + 3;
            } else if (s.compare(p, 2, "<?") == 0 || s.compare(p, 2, "<!") == 0) {
                const size_t e = s.find('>', p);
                if (e == std::string::npos) return false;
                p = e + 1;
            } else {
                return true;
            }
        }
    };
    if (!skip_misc(i) || i >= s.size() || s[i] != '<') return false;
    size_t e = s.find('>', i);
    if (e == std::string::npos) return false;
    std::string root = s.substr(i + 1, e - i - 1);
    root = root.substr(0, root.find_first_of(" \t\r\n/"));
    i = e + 1;
    const std::string close_root = "</" + root;
    for (;;) {
        // text between children is ignored
        while (i < s.size() && s[i] != '<') ++i;
        if (!skip_misc(i)) return false;
        if (i >= s.size()) return false;
        if (s.compare(i, close_root.size(), close_root) == 0) return true;
        e = s.find('>', i);
        if (e == std::string::npos) return false;
        std::string tag = s.substr(i + 1, e - i - 1);
        const bool selfclose = !tag.empty() && tag.back() == '/';
        tag = tag.substr(0, tag.find_first_of(" \t\r\n/"));
        i = e + 1;
        if (selfclose) {
            out.emplace(tag, "");
            continue;
        }
        const std::string close = "</" + tag + ">";
        const size_t c = s.find(close, i);
        if (c == std::string::npos) return false;
        out.emplace(tag, s.substr(i, c - i));
        i = c + close.size();
    }
}

std::string trim(const std::string& v)
{
    size_t a = 0, b = v.size();
    while (a < b && std::isspace((unsigned char)v[a])) ++a;
    while (b > a && std::isspace((unsigned char)v[b - 1])) --b;
    return v.substr(a, b - a);
}
bool q_int(const std::map<std::string, std::string>& m, const char* k, int* v)
{
    auto it = m.find(k);
    if (it == m.end()) return false;
    return std::sscanf(trim(it->second).c_str(), "%d", v) == 1;
}
bool q_double(const std::map<std::string, std::string>& m, const char* k, double* v)
{
    auto it = m.find(k);
    if (it == m.end()) return false;
    return std::sscanf(trim(it->second).c_str(), "%lf", v) == 1;
}
bool q_bool(const std::map<std::string, std::string>& m, const char* k, bool* v)
{
    auto it = m.find(k);
    if (it == m.end()) return false;
    std::string t = trim(it->second);
    for (auto& ch : t) ch = (char)std::tolower(ch);
    if (t == "true" || t == "1") { *v = true; return true; }
    if (t == "false" || t == "0") { *v = false; return true; }
    int iv;
    if (std::sscanf(t.c_str(), "%d", &iv) == 1) { *v = iv != 0; return true; }
    return false;
}
}  // namespace

bool ophGen::readConfig(const char* fname)
{
    if (!checkExtension(fname, ".xml")) {
        OPH_LOG("<FAILED> Wrong file ext.\n");
        return false;
    }
    std::map<std::string, std::string> m;
    if (!xml_children(fname, m)) {
        OPH_LOG("<FAILED> Loading file (%s)\n", fname);
        return false;
    }
    bool ok = true;
    int nWave = 1;
    if (!q_int(m, "SLM_WaveNum", &nWave)) {
        OPH_LOG("<FAILED> Not found node : 'SLM_WaveNum' (Integer)\n");
        ok = false;
    }
    // a channel indexes Vertex::color[3], so 1..3 wavelengths
    if (nWave < 1 || nWave > 3) {
        OPH_LOG("<FAILED> SLM_WaveNum must be 1..3 (got %d)\n", nWave);
        return false;
    }
    setWaveNum(nWave);
    for (int i = 1; i <= nWave; ++i) {
        char name[32];
        std::snprintf(name, sizeof name, "SLM_WaveLength_%d", i);
        if (!q_double(m, name, &context_.wave_length[i - 1])) {
            OPH_LOG("<FAILED> Not found node : '%s' (Double)\n", name);
            ok = false;
        }
    }
    if (!q_int(m, "SLM_PixelNumX", &context_.pixel_number[_X])) { OPH_LOG("<FAILED> Not found node : 'SLM_PixelNumX' (Integer)\n"); ok = false; }
    if (!q_int(m, "SLM_PixelNumY", &context_.pixel_number[_Y])) { OPH_LOG("<FAILED> Not found node : 'SLM_PixelNumY' (Integer)\n"); ok = false; }
    if (!q_double(m, "SLM_PixelPitchX", &context_.pixel_pitch[_X])) { OPH_LOG("<FAILED> Not found node : 'SLM_PixelPitchX' (Double)\n"); ok = false; }
    if (!q_double(m, "SLM_PixelPitchY", &context_.pixel_pitch[_Y])) { OPH_LOG("<FAILED> Not found node : 'SLM_PixelPitchY' (Double)\n"); ok = false; }
    // the buffers are allocated from these: 1..2^20 pixels per axis, at most 2^28 in all
    const int pnx = context_.pixel_number[_X], pny = context_.pixel_number[_Y];
    if (pnx < 1 || pny < 1 || pnx > (1 << 20) || pny > (1 << 20) || (long long)pnx * pny > (1LL << 28)) {
        OPH_LOG("<FAILED> SLM_PixelNumX/Y out of range (%d x %d)\n", pnx, pny);
        context_.pixel_number = ivec2(0, 0);
        return false;
    }
    if (!(context_.pixel_pitch[_X] > 0) || !(context_.pixel_pitch[_Y] > 0) ||
        !std::isfinite(context_.pixel_pitch[_X]) || !std::isfinite(context_.pixel_pitch[_Y])) {
        OPH_LOG("<FAILED> SLM_PixelPitchX/Y must be positive and finite\n");
        return false;
    }
    if (!q_bool(m, "IMG_Rotation", &imgCfg.rotate)) imgCfg.rotate = false;
    if (!q_bool(m, "IMG_Merge", &imgCfg.merge)) imgCfg.merge = false;
    if (!q_int(m, "IMG_Flip", &imgCfg.flip)) imgCfg.flip = 0;
    if (!q_bool(m, "DoublePrecision", &context_.bUseDP)) context_.bUseDP = true;
    if (!q_double(m, "ShiftX", &context_.shift[_X])) context_.shift[_X] = 0.0;
    if (!q_double(m, "ShiftY", &context_.shift[_Y])) context_.shift[_Y] = 0.0;
    if (!q_double(m, "ShiftZ", &context_.shift[_Z])) context_.shift[_Z] = 0.0;
    if (!q_double(m, "FieldLength", &m_dFieldLength)) m_dFieldLength = 0.0;
    if (!q_bool(m, "RandomPhase", &m_bRandomPhase)) m_bRandomPhase = false;
    if (!q_int(m, "NumOfStream", &m_nStream)) m_nStream = 1;
    context_.ss[_X] = context_.pixel_number[_X] * context_.pixel_pitch[_X];
    context_.ss[_Y] = context_.pixel_number[_Y] * context_.pixel_pitch[_Y];
    resCfg = context_.pixel_number;
    return ok;
}

// ---------------------------------------------------------------------------
// PLY point clouds (semantics of Openholo/src/PLYparser.cpp:142-449):
// x, y, z as stored (binary float or ASCII decimal); colour = red/255 computed
// in single precision; 'phase' kept if present; every colour becomes 0.5 when
// the file has no 'color' element.  Properties this loader does not use are
// skipped with their true size (the upstream binary reader does not advance
// past them).
namespace {
struct PlyProp {
    std::string name, type, list_count_type;
    bool list = false;
};
struct PlyElem {
    std::string name;
    long long count = 0;
    std::vector<PlyProp> props;
};
int type_size(const std::string& t)
{
    if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
    if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
    if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32") return 4;
    if (t == "double" || t == "float64" || t == "Real") return 8;
    return 0;
}
double read_bin(std::istream& f, const std::string& t, bool big)
{
    unsigned char b[8] = {0};
    const int n = type_size(t);
    f.read((char*)b, n);
    if (big) std::reverse(b, b + n);
    if (t == "char" || t == "int8") return (double)(int8_t)b[0];
    if (t == "uchar" || t == "uint8") return (double)b[0];
    if (t == "short" || t == "int16") { int16_t v; std::memcpy(&v, b, 2); return v; }
    if (t == "ushort" || t == "uint16") { uint16_t v; std::memcpy(&v, b, 2); return v; }
    if (t == "int" || t == "int32") { int32_t v; std::memcpy(&v, b, 4); return v; }
    if (t == "uint" || t == "uint32") { uint32_t v; std::memcpy(&v, b, 4); return v; }
    if (t == "float" || t == "float32") { float v; std::memcpy(&v, b, 4); return v; }
    double v;
    std::memcpy(&v, b, 8);
    return v;
}
}  // namespace

int ophGen::loadPointCloud(const char* pc_file, OphPointCloudData* pc)
{
    if (!pc_file || !pc) return -1;
    std::string path(pc_file);
    if (path.find(".ply") == std::string::npos && path.find(".PLY") == std::string::npos) path += ".ply";
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        OPH_LOG("<FAILED> Loading ply file.\n");
        return -1;
    }
    std::string line, tok;
    std::getline(f, line);
    {
        std::istringstream ls(line);
        ls >> tok;
        if (tok != "ply" && tok != "PLY") return -1;
    }
    bool binary = false, big = false;
    std::vector<PlyElem> elems;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        if (!(ls >> tok)) continue;
        if (tok == "format") {
            std::string fmt;
            ls >> fmt;
            binary = fmt != "ascii";
            big = fmt == "binary_big_endian";
        } else if (tok == "element") {
            PlyElem e;
            ls >> e.name >> e.count;
            elems.push_back(e);
        } else if (tok == "property") {
            if (elems.empty()) return -1;
            PlyProp p;
            std::string t;
            ls >> t;
            if (t == "list") {
                p.list = true;
                ls >> p.list_count_type >> p.type >> p.name;
            } else {
                p.type = t;
                ls >> p.name;
            }
            elems.back().props.push_back(p);
        } else if (tok == "end_header") {
            break;
        }
    }
    // counts the rest of the file cannot hold are rejected before allocating
    // (a record takes at least its fixed-size fields in binary, or one
    // character and a separator per property in ASCII)
    {
        const std::streampos body = f.tellg();
        f.seekg(0, std::ios::end);
        const long long rest = (long long)(f.tellg() - body);
        f.seekg(body);
        long long need = 0;
        for (const PlyElem& e : elems) {   // the elements up to and including 'vertex'
            if (e.count < 0) {
                OPH_LOG("<FAILED> negative element count in ply header.\n");
                return -1;
            }
            long long rec = 0;
            for (const PlyProp& p : e.props) rec += binary ? type_size(p.list ? p.list_count_type : p.type) : 2;
            rec = std::max(rec, 1LL);
            if (e.count > (rest - need) / rec + 1) {
                OPH_LOG("<FAILED> ply element count exceeds the file.\n");
                return -1;
            }
            need += e.count * rec;
            if (e.name == "vertex") break;
        }
    }
    int ev = -1;
    bool has_color_elem = false;
    for (size_t i = 0; i < elems.size(); ++i) {
        if (elems[i].name == "vertex") ev = (int)i;
        if (elems[i].name == "color") has_color_elem = true;
    }
    if (ev < 0) return -1;
    auto pidx = [&](const char* n) {
        for (size_t j = 0; j < elems[ev].props.size(); ++j)
            if (elems[ev].props[j].name == n) return (int)j;
        return -1;
    };
    const int ix = pidx("x"), iy = pidx("y"), iz = pidx("z");
    if (ix < 0 || iy < 0 || iz < 0) return -1;
    for (int j : {ix, iy, iz})
        if (binary && type_size(elems[ev].props[j].type) == 0) {
            OPH_LOG("<FAILED> unknown ply coordinate type '%s'.\n", elems[ev].props[j].type.c_str());
            return -1;
        }
    int ir = pidx("red"), ig = pidx("green"), ib = pidx("blue");
    if (ir < 0 || ig < 0 || ib < 0) {
        ir = pidx("diffuse_red");
        ig = pidx("diffuse_green");
        ib = pidx("diffuse_blue");
    }
    const int iph = pidx("phase");
    const long long n = elems[ev].count;
    delete[] pc->vertices;
    pc->vertices = new Vertex[n];
    std::memset(pc->vertices, 0, sizeof(Vertex) * n);
    pc->n_points = (ulonglong)n;
    pc->n_colors = 3;
    std::vector<double> vals;
    for (size_t ei = 0; ei < elems.size(); ++ei) {
        const PlyElem& e = elems[ei];
        for (long long k = 0; k < e.count; ++k) {
            vals.assign(e.props.size(), 0.0);
            if (binary) {
                if (!f) break;   // truncated file: the remaining records stay zero
                for (size_t j = 0; j < e.props.size(); ++j) {
                    const PlyProp& p = e.props[j];
                    if (p.list) {
                        const double cnt = read_bin(f, p.list_count_type, big);
                        for (long long q = 0; f && q < cnt; ++q) read_bin(f, p.type, big);
                    } else {
                        vals[j] = read_bin(f, p.type, big);
                    }
                }
            } else {
                if (!std::getline(f, line)) break;
                std::istringstream ls(line);
                for (size_t j = 0; j < e.props.size(); ++j) {
                    std::string v;
                    if (!(ls >> v)) break;
                    if (e.props[j].list) break;   // lists only occur outside 'vertex' here
                    vals[j] = std::strtod(v.c_str(), nullptr);
                }
            }
            if ((int)ei != ev) continue;
            Vertex& vx = pc->vertices[k];
            // binary coordinates arrive as stored (float); ASCII ones as decimals
            vx.point.pos[_X] = vals[ix];
            vx.point.pos[_Y] = vals[iy];
            vx.point.pos[_Z] = vals[iz];
            if (ir >= 0) {
                // integer colour channels (values outside int's range, possible only
                // in a corrupt file, read as 0)
                auto chan = [](double c) { return (c > -2147483649.0 && c < 2147483648.0) ? (int)c : 0; };
                vx.color.color[0] = (Real)((float)chan(vals[ir]) / 255.f);
                vx.color.color[1] = (Real)((float)chan(vals[ig]) / 255.f);
                vx.color.color[2] = (Real)((float)chan(vals[ib]) / 255.f);
            }
            if (iph >= 0) vx.phase = (Real)(float)vals[iph];
        }
        if ((int)ei == ev) break;   // nothing after the vertex element is used
    }
    if (!has_color_elem)
        for (long long k = 0; k < n; ++k)
            pc->vertices[k].color.color[0] = pc->vertices[k].color.color[1] = pc->vertices[k].color.color[2] = 0.5;
    return (int)n;
}

// ---------------------------------------------------------------------------
bool ophGen::ensureDeviceBuffers()
{
    ohx_ctx* c = device();
    if (!c) return false;
    const uint nch = context_.waveNum;
    const long long npx = (long long)context_.pixel_number[_X] * context_.pixel_number[_Y];
    if (d_field && d_channels == nch && d_pixels == npx) return true;
    auto drop = [&]() {
        for (uint i = 0; i < d_channels; ++i) {
            if (d_field) ohx_free(c, d_field[i]);
            if (d_enc) ohx_free(c, d_enc[i]);
            if (d_u8) ohx_free(c, d_u8[i]);
        }
        delete[] d_field;
        delete[] d_enc;
        delete[] d_u8;
        d_field = nullptr;
        d_enc = nullptr;
        d_u8 = nullptr;
        d_channels = 0;
    };
    drop();
    d_field = new double*[nch]();
    d_enc = new double*[nch]();
    d_u8 = new uint8_t*[nch]();
    d_channels = nch;
    d_pixels = npx;
    for (uint i = 0; i < nch; ++i) {
        int st = ohx_malloc(c, 16 * (size_t)npx, (void**)&d_field[i]);
        if (!st) st = ohx_malloc(c, 8 * (size_t)npx, (void**)&d_enc[i]);
        if (!st) st = ohx_malloc(c, (size_t)npx, (void**)&d_u8[i]);
        if (st) {
            fail("device buffer allocation", st);
            drop();
            return false;
        }
    }
    return true;
}

bool ophGen::hostPlanesReady(const char* what)
{
    const uint nch = context_.waveNum;
    const uint64_t npx = (uint64_t)context_.pixel_number[_X] * (uint64_t)context_.pixel_number[_Y];
    if (nch == 0 || nch > 3 || !complex_H || m_nOldChannel != nch || complex_H_planes < nch ||
        context_.pixel_number[_X] < 1 || context_.pixel_number[_Y] < 1 || complex_H_pixels != npx) {
        m_lastError = std::string(what) +
                      ": host buffers do not match the configured channels and pixel number "
                      "(call setResolution or readConfig/initialize after changing them)";
        OPH_LOG("<FAILED> %s\n", m_lastError.c_str());
        return false;
    }
    return true;
}

bool ophGen::uploadFields()
{
    // complex_H is always taken from the host: callers may have edited it
    // through getComplexField() since it was generated (the reference encodes
    // from the host field, ophgen/src/ophGen.cpp:816-849)
    if (!complex_H || !ensureDeviceBuffers()) return false;
    ohx_ctx* c = device();
    for (uint ch = 0; ch < context_.waveNum; ++ch) {
        const int st = ohx_memcpy_h2d(c, d_field[ch], complex_H[ch], 16 * (size_t)d_pixels);
        if (st) {
            fail("complex field upload", st);
            return false;
        }
    }
    return true;
}

void ophGen::encoding(unsigned int ENCODE_FLAG)
{
    // ophgen/src/ophGen.cpp:816-849, on the GPU
    if (ENCODE_FLAG > ENCODE_IMAGINARY) {
        OPH_LOG("<FAILED> WRONG PARAMETERS.\n");
        return;
    }
    if (!hostPlanesReady("encoding")) return;
    m_vecEncodeSize = context_.pixel_number;
    if (!uploadFields()) return;
    ohx_ctx* c = device();
    for (uint ch = 0; ch < context_.waveNum; ++ch) {
        int st = ohx_encode(c, d_field[ch], d_pixels, d_enc[ch], ENCODE_FLAG);
        if (!st) st = ohx_memcpy_d2h(c, m_lpEncoded[ch], d_enc[ch], 8 * (size_t)d_pixels);
        if (st) {
            fail("encoding", st);
            return;
        }
    }
}

void ophGen::normalize(void)
{
    // ophgen/src/ophGen.cpp:654-687: global min/max over all channels, on the GPU.
    // The encoded buffers are taken from the host (callers may edit them
    // through getEncodedBuffer()).
    if (!hostPlanesReady("normalize") || !ensureDeviceBuffers()) return;
    ohx_ctx* c = device();
    const uint nch = context_.waveNum;
    for (uint ch = 0; ch < nch; ++ch) {
        const int st = ohx_memcpy_h2d(c, d_enc[ch], m_lpEncoded[ch], 8 * (size_t)d_pixels);
        if (st) {
            fail("encoded buffer upload", st);
            return;
        }
    }
    std::vector<const double*> e(d_enc, d_enc + nch);
    std::vector<uint8_t*> o(d_u8, d_u8 + nch);
    int st = ohx_normalize_u8(c, e.data(), (int)nch, d_pixels, o.data());
    for (uint ch = 0; !st && ch < nch; ++ch) st = ohx_memcpy_d2h(c, m_lpNormalized[ch], d_u8[ch], (size_t)d_pixels);
    if (st) fail("normalize", st);
}

bool ophGen::save(const char* fname, uint8_t bitsperpixel, uchar* src, uint px, uint py)
{
    // ophgen/src/ophGen.cpp:689-753
    if (!fname) return false;
    const uint nch = context_.waveNum;
    ivec2 p((int)px, (int)py);
    if (px == 0 && py == 0) p = context_.pixel_number;
    std::string file(fname);
    std::replace(file.begin(), file.end(), '\\', '/');
    const size_t slash = file.rfind('/');
    const std::string dir = slash == std::string::npos ? "" : file.substr(0, slash + 1);
    std::string name = slash == std::string::npos ? file : file.substr(slash + 1);
    if (file.rfind('.') == std::string::npos) name += ".bmp";
    const std::string full = dir + name;
    if (src) return saveAsImg(full.c_str(), bitsperpixel, src, p[_X], p[_Y]);
    if (nch == 1) return saveAsImg(full.c_str(), bitsperpixel, m_lpNormalized[0], p[_X], p[_Y]);
    if (nch != 3) return false;
    if (imgCfg.merge) {
        std::vector<uchar> merged((size_t)p[_X] * p[_Y] * 3, 0);
        for (uint i = 0; i < nch; ++i) mergeColor((int)i, p[_X], p[_Y], m_lpNormalized[i], merged.data());
        return saveAsImg(full.c_str(), bitsperpixel, merged.data(), p[_X], p[_Y]);
    }
    bool ok = true;
    for (uint i = 0; i < nch; ++i) {
        const std::string path = dir + std::to_string(i) + "_" + name;
        ok = saveAsImg(path.c_str(), (uint8_t)(bitsperpixel / nch), m_lpNormalized[i], p[_X], p[_Y]) && ok;
    }
    return ok;
}

bool ophGen::loadAsOhc(const char* fname)
{
    // ophgen/src/ophGen.cpp:769-785: the field plus empty encode buffers
    releaseBuffers();
    if (!Openholo::loadAsOhc(fname)) return false;
    const uint nch = context_.waveNum;
    const long long n = (long long)context_.pixel_number[_X] * context_.pixel_number[_Y];
    resCfg = context_.pixel_number;
    m_lpEncoded = new Real*[nch];
    m_lpNormalized = new uchar*[nch];
    for (uint i = 0; i < nch; ++i) {
        m_lpEncoded[i] = new Real[n]();
        m_lpNormalized[i] = new uchar[n]();
    }
    m_nOldChannel = nch;
    m_vecEncodeSize = context_.pixel_number;
    return true;
}

void ophGen::ophFree(void) { releaseBuffers(); }

void ophGen::releaseBuffers(void)
{
    ohx_ctx* c = d_field ? device() : nullptr;
    if (c) {
        for (uint i = 0; i < d_channels; ++i) {
            ohx_free(c, d_field[i]);
            ohx_free(c, d_enc[i]);
            ohx_free(c, d_u8[i]);
        }
    }
    delete[] d_field;
    delete[] d_enc;
    delete[] d_u8;
    d_field = nullptr;
    d_enc = nullptr;
    d_u8 = nullptr;
    d_channels = 0;
    if (m_lpEncoded) {
        for (uint i = 0; i < m_nOldChannel; ++i) delete[] m_lpEncoded[i];
        delete[] m_lpEncoded;
        m_lpEncoded = nullptr;
    }
    if (m_lpNormalized) {
        for (uint i = 0; i < m_nOldChannel; ++i) delete[] m_lpNormalized[i];
        delete[] m_lpNormalized;
        m_lpNormalized = nullptr;
    }
    freeComplexField();
    m_nOldChannel = 0;
}
