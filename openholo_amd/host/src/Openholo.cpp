// Openholo base: configuration, host complex field, BMP and OHC files.
#include "Openholo.h"

#include <cstring>
#include <fstream>
#include <memory>
#include <vector>

#include "openholo_amd.h"

Openholo::Openholo(void) {}

Openholo::~Openholo(void)
{
    freeComplexField();
    delete[] context_.wave_length;
    context_.wave_length = nullptr;
}

void Openholo::allocComplexField(uint planes, uint64_t npx)
{
    freeComplexField();
    complex_H = new Complex<Real>*[planes]();
    complex_H_planes = planes;
    complex_H_pixels = npx;
    complex_H_pinned = true;
    for (uint i = 0; i < planes; ++i) {
        void* p = nullptr;
        if (ohx_host_alloc(npx * sizeof(Complex<Real>), &p) != OHX_OK) {
            // no device (CPU-only use of the host classes): heap planes
            for (uint j = 0; j < i; ++j) {
                ohx_host_free(complex_H[j]);
                complex_H[j] = nullptr;
            }
            complex_H_pinned = false;
            break;
        }
        complex_H[i] = static_cast<Complex<Real>*>(p);
        std::uninitialized_value_construct_n(complex_H[i], npx);
    }
    if (!complex_H_pinned)
        for (uint i = 0; i < planes; ++i) complex_H[i] = new Complex<Real>[npx];
}

void Openholo::freeComplexField(void)
{
    if (complex_H) {
        for (uint i = 0; i < complex_H_planes; ++i) {
            if (complex_H_pinned) ohx_host_free(complex_H[i]);
            else delete[] complex_H[i];
        }
        delete[] complex_H;
    }
    complex_H = nullptr;
    complex_H_planes = 0;
    complex_H_pixels = 0;
    complex_H_pinned = false;
}

unsigned long Openholo::release(void)
{
    // Openholo/src/Base.h:88-95: frees on the call that finds no extra reference
    if (!refCnt) {
        ophFree();
        delete this;
        return 0;
    }
    return refCnt--;
}

void Openholo::setPixelNumber(ivec2 n)
{
    context_.pixel_number = n;
    context_.ss[_X] = n[_X] * context_.pixel_pitch[_X];
    context_.ss[_Y] = n[_Y] * context_.pixel_pitch[_Y];
}

void Openholo::setPixelPitch(vec2 p)
{
    context_.pixel_pitch = p;
    context_.ss[_X] = context_.pixel_number[_X] * p[_X];
    context_.ss[_Y] = context_.pixel_number[_Y] * p[_Y];
}

void Openholo::setWaveLength(Real w, const uint idx)
{
    if (context_.wave_length && idx < context_.waveNum) context_.wave_length[idx] = w;
}

void Openholo::setWaveNum(int num)
{
    context_.waveNum = num;
    delete[] context_.wave_length;
    context_.wave_length = new Real[num > 0 ? num : 1]();
}

bool Openholo::checkExtension(const char* fname, const char* ext)
{
    const std::string f(fname ? fname : ""), e(ext);
    const size_t p = f.rfind('.');
    if (p == std::string::npos) return false;
    std::string a = f.substr(p), b = e;
    for (auto& c : a) c = (char)std::tolower(c);
    for (auto& c : b) c = (char)std::tolower(c);
    return a == b;
}

bool Openholo::mergeColor(int idx, int width, int height, uchar* src, uchar* dst)
{
    // channel 0 (red wavelength) goes to byte 2 of each BGR pixel
    if (idx < 0 || idx > 2) return false;
    const long long n = (long long)width * height;
    const int slot = 2 - idx;
    for (long long i = 0; i < n; ++i) dst[i * 3 + slot] = src[i];
    return true;
}

namespace {
#pragma pack(push, 1)
struct BmpFileHeader {
    uint8_t sig[2];
    uint32_t file_size;
    uint32_t reserved;
    uint32_t pixel_offset;
};
struct BmpInfoHeader {
    uint32_t header_size;
    uint32_t width, height;
    uint16_t planes, bpp;
    uint32_t compression, image_size, ypm, xpm, colors, important;
};
#pragma pack(pop)

// flip / rotate an 8- or 24-bit image (rows padded to 4 bytes), out of place
void transform_rows(std::vector<uchar>& img, int w, int h, int ch, int flip, bool rotate)
{
    const size_t stride = ((size_t)w * ch + 3) & ~size_t(3);
    int mode = flip;
    if (rotate) mode ^= 3;   // 180 degree rotation == vertical + horizontal flip
    if (mode == 0) return;
    std::vector<uchar> out(img.size(), 0);
    for (int y = 0; y < h; ++y) {
        const int sy = (mode & 1) ? h - 1 - y : y;
        for (int x = 0; x < w; ++x) {
            const int sx = (mode & 2) ? w - 1 - x : x;
            std::memcpy(&out[y * stride + (size_t)x * ch], &img[sy * stride + (size_t)sx * ch], ch);
        }
    }
    img.swap(out);
}
}  // namespace

bool Openholo::saveAsImg(const char* fname, uint8_t bitsperpixel, uchar* src, int width, int height)
{
    if (!fname || !src || width <= 0 || height <= 0) return false;
    if (bitsperpixel != 8 && bitsperpixel != 24) {
        OPH_LOG("<FAILED> saveAsImg: %d bits per pixel is not supported\n", bitsperpixel);
        return false;
    }
    const int ch = bitsperpixel / 8;
    const size_t stride = ((size_t)width * ch + 3) & ~size_t(3);
    const size_t pixel_bytes = stride * height;
    const bool palette = bitsperpixel <= 8;
    const size_t header = sizeof(BmpFileHeader) + sizeof(BmpInfoHeader) + (palette ? 256 * 4 : 0);

    // source rows are unpadded (width * ch bytes); pad each to the BMP stride
    std::vector<uchar> img(pixel_bytes, 0);
    for (int y = 0; y < height; ++y)
        std::memcpy(&img[y * stride], src + (size_t)y * width * ch, (size_t)width * ch);
    transform_rows(img, width, height, ch, imgCfg.flip, imgCfg.rotate);

    BmpFileHeader fh{{'B', 'M'}, (uint32_t)(header + pixel_bytes), 0, (uint32_t)header};
    BmpInfoHeader ih{};
    ih.header_size = sizeof(BmpInfoHeader);
    ih.width = (uint32_t)width;
    ih.height = (uint32_t)height;
    ih.planes = 1;
    ih.bpp = bitsperpixel;
    ih.compression = 0;
    ih.image_size = (uint32_t)pixel_bytes;
    ih.colors = palette ? 256 : 0;
    FILE* fp = std::fopen(fname, "wb");
    if (!fp) return false;
    bool ok = std::fwrite(&fh, sizeof fh, 1, fp) == 1 && std::fwrite(&ih, sizeof ih, 1, fp) == 1;
    if (palette) {
        uint8_t pal[256 * 4];
        for (int i = 0; i < 256; ++i) {
            pal[4 * i] = pal[4 * i + 1] = pal[4 * i + 2] = (uint8_t)i;
            pal[4 * i + 3] = 0;
        }
        ok = ok && std::fwrite(pal, sizeof pal, 1, fp) == 1;
    }
    ok = ok && std::fwrite(img.data(), 1, img.size(), fp) == img.size();
    std::fclose(fp);
    return ok;
}

// ---- OHC ------------------------------------------------------------------
// Layout (Openholo/src/ImgCodecDefine.h:137-206, ImgCodecOhc.cpp:842-876,1057-1200):
// packed file header (24 B), field-info header (65 B), wavelength table
// (8 B each), then every channel's real plane followed by every channel's
// imaginary plane, each in the field buffer's row-major order, FP64.
namespace {
#pragma pack(push, 1)
struct OhcFileHeader {
    int8_t sig[2];
    uint64_t file_size;
    uint8_t ver_major, ver_minor;
    uint32_t reserved1, reserved2;
    uint32_t data_offset;
};
struct OhcFieldInfo {
    uint32_t header_size;
    uint32_t px_num_x, px_num_y;
    double px_pitch_x, px_pitch_y;
    uint8_t pitch_unit;
    uint32_t wavlen_num;
    uint8_t color_type, color_arrange, wavlen_unit, data_type, store, code_type, phase_coded;
    double phase_min, phase_max;
    uint64_t field_size;
    uint8_t compression;
};
#pragma pack(pop)
static_assert(sizeof(OhcFileHeader) == 24, "OHC file header");
static_assert(sizeof(OhcFieldInfo) == 65, "OHC field info header");
enum : uint8_t { UNIT_M = 1, CLR_MLT = 2, ARR_EACH = 2, ARR_SEQT = 1, DT_F32 = 9, DT_F64 = 10, STORE_DIRECT = 1,
                 CODE_RI = 2 };
}  // namespace

bool Openholo::saveAsOhc(const char* fname)
{
    if (!fname || !complex_H || context_.waveNum == 0) return false;
    std::string path(fname);
    if (!checkExtension(fname, ".ohc")) path += ".ohc";
    const uint32_t nw = context_.waveNum;
    const uint64_t npx = (uint64_t)context_.pixel_number[_X] * context_.pixel_number[_Y];
    if (nw > complex_H_planes || npx != complex_H_pixels) {
        OPH_LOG("<FAILED> saveAsOhc: the complex field does not match the context's size.\n");
        return false;
    }
    OhcFieldInfo fi{};
    fi.header_size = (uint32_t)(sizeof(OhcFieldInfo) + 8ull * nw);
    fi.px_num_x = (uint32_t)context_.pixel_number[_X];
    fi.px_num_y = (uint32_t)context_.pixel_number[_Y];
    fi.px_pitch_x = context_.pixel_pitch[_X];
    fi.px_pitch_y = context_.pixel_pitch[_Y];
    fi.pitch_unit = UNIT_M;
    fi.wavlen_num = nw;
    fi.color_type = CLR_MLT;
    fi.color_arrange = ARR_EACH;
    fi.wavlen_unit = UNIT_M;
    fi.data_type = DT_F64;
    fi.store = STORE_DIRECT;
    fi.code_type = CODE_RI;
    fi.phase_coded = 0;
    fi.phase_min = -1.0;
    fi.phase_max = 1.0;
    fi.field_size = npx * nw * 2 * sizeof(double);
    fi.compression = 0;
    OhcFileHeader fh{{'O', 'H'}, sizeof(OhcFileHeader) + fi.header_size + fi.field_size, 3, 0, 0, 0,
                     (uint32_t)(sizeof(OhcFileHeader) + fi.header_size)};
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) return false;
    f.write((const char*)&fh, sizeof fh);
    f.write((const char*)&fi, sizeof fi);
    for (uint32_t w = 0; w < nw; ++w) f.write((const char*)&context_.wave_length[w], sizeof(double));
    std::vector<double> plane(npx);
    for (int part = 0; part < 2; ++part)
        for (uint32_t w = 0; w < nw; ++w) {
            for (uint64_t i = 0; i < npx; ++i) plane[i] = complex_H[w][i][part];
            f.write((const char*)plane.data(), npx * sizeof(double));
        }
    return (bool)f;
}

bool Openholo::loadAsOhc(const char* fname)
{
    if (!fname) return false;
    std::string path(fname);
    if (!checkExtension(fname, ".ohc")) path += ".ohc";
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    OhcFileHeader fh{};
    OhcFieldInfo fi{};
    f.read((char*)&fh, sizeof fh);
    f.read((char*)&fi, sizeof fi);
    if (!f || fh.sig[0] != 'O' || fh.sig[1] != 'H' || fi.wavlen_num == 0 || fi.store != STORE_DIRECT ||
        fi.code_type != CODE_RI || (fi.data_type != DT_F64 && fi.data_type != DT_F32))
        return false;
    const uint32_t nw = fi.wavlen_num;
    const size_t es = fi.data_type == DT_F64 ? 8 : 4;
    // header sizes are checked against the file before anything is sized from them
    const std::streampos here = f.tellg();
    f.seekg(0, std::ios::end);
    const uint64_t file_bytes = (uint64_t)f.tellg();
    f.seekg(here);
    const uint64_t npx = (uint64_t)fi.px_num_x * fi.px_num_y;
    if (nw > 64 || fi.px_num_x < 1 || fi.px_num_y < 1 || fi.px_num_x > (1u << 20) || fi.px_num_y > (1u << 20) ||
        npx > (1ull << 28) || (uint64_t)fh.data_offset > file_bytes ||
        2 * npx * nw * es > file_bytes - (uint64_t)fh.data_offset) {
        OPH_LOG("<FAILED> OHC header sizes do not match the file.\n");
        return false;
    }
    std::vector<double> wl(nw);
    f.read((char*)wl.data(), 8 * nw);
    if (!f) return false;
    f.seekg(fh.data_offset);
    freeComplexField();
    setWaveNum((int)nw);
    for (uint32_t w = 0; w < nw; ++w) context_.wave_length[w] = wl[w];
    context_.pixel_pitch = vec2(fi.px_pitch_x, fi.px_pitch_y);
    setPixelNumber(ivec2((int)fi.px_num_x, (int)fi.px_num_y));
    context_.k = (2 * M_PI) / context_.wave_length[0];
    allocComplexField(nw, npx);
    std::vector<char> buf(npx * es);
    auto value = [es](const char* p) {
        if (es == 8) {
            double d;
            std::memcpy(&d, p, 8);
            return d;
        }
        float x;
        std::memcpy(&x, p, 4);
        return (double)x;
    };
    for (int part = 0; part < 2; ++part)
        for (uint32_t w = 0; w < nw; ++w) {
            if (fi.color_arrange == ARR_SEQT) {
                // values of all wavelengths interleaved per pixel
                for (uint64_t i = 0; i < npx; ++i) {
                    f.seekg(fh.data_offset + (std::streamoff)(part * npx * nw + i * nw + w) * es);
                    char v[8];
                    f.read(v, es);
                    complex_H[w][i][part] = value(v);
                }
            } else {
                f.read(buf.data(), buf.size());
                for (uint64_t i = 0; i < npx; ++i) complex_H[w][i][part] = value(buf.data() + i * es);
            }
        }
    return (bool)f;
}

void Openholo::ophFree(void)
{
    freeComplexField();
}
