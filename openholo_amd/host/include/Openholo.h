// Openholo: base of the drop-in class hierarchy (Openholo -> ophGen ->
// ophPointCloud).  Same public surface as the reference's
// Openholo/src/Openholo.h:145-490 for the members the point-cloud path uses:
// the OphConfig context, the host complex field complex_H, the setters, BMP
// and OHC output and reference-counted release().
#pragma once

#include <string>
#include <vector>

#include "oph_types.h"

struct OphConfig {
    bool bUseDP = true;        // use double precision (parsed, informational)
    ivec2 pixel_number;        // SLM_PixelNumX / SLM_PixelNumY
    ivec2 offset;              // start offset (default 0)
    vec2 pixel_pitch;          // SLM_PixelPitchX / SLM_PixelPitchY
    vec3 shift;                // ShiftX/Y/Z
    Real k = 0;                // 2 pi / lambda of the channel being generated
    vec2 ss;                   // pn * pp
    uint waveNum = 0;          // number of wavelengths
    Real* wave_length = nullptr;
};

struct ImageConfig {
    bool rotate = false;
    bool merge = false;
    int flip = 0;              // FLIP: NONE, VERTICAL, HORIZONTAL, BOTH
};

class Openholo {
public:
    explicit Openholo(void);

protected:
    virtual ~Openholo(void);

public:
    // reference counting (Openholo/src/Base.h:88-95)
    unsigned long addRef(void) { return ++refCnt; }
    unsigned long release(void);

    Complex<Real>** getComplexField(void) { return complex_H; }
    OphConfig& getContext(void) { return context_; }
    ImageConfig& getImageConfig(void) { return imgCfg; }

    void setPixelNumber(ivec2 n);
    void setPixelNumber(int width, int height) { setPixelNumber(ivec2(width, height)); }
    void setPixelPitch(vec2 p);
    void setPixelPitch(Real px, Real py) { setPixelPitch(vec2(px, py)); }
    void setWaveLength(Real w, const uint idx = 0);
    void setWaveNum(int num);
    void setOffset(ivec2 offset) { context_.offset = offset; }
    void setImageMerge(bool merge) { imgCfg.merge = merge; }
    void setImageRotate(bool rotate) { imgCfg.rotate = rotate; }
    void setImageFlip(int flip) { imgCfg.flip = flip; }

    // 8-bit grayscale (palette) or 24-bit BMP, rows in buffer order, 4-byte
    // row padding (Openholo/src/Openholo.cpp:135-249)
    virtual bool saveAsImg(const char* fname, uint8_t bitsperpixel, uchar* src, int width, int height);
    // "OH" complex-field file, FP64 real/imaginary, one plane per wavelength
    virtual bool saveAsOhc(const char* fname);
    virtual bool loadAsOhc(const char* fname);

protected:
    bool checkExtension(const char* fname, const char* ext);
    bool mergeColor(int idx, int width, int height, uchar* src, uchar* dst);
    virtual void ophFree(void);
    // complex_H as `planes` zeroed planes of npx pixels, page-locked when a
    // device is present (the per-channel read-back then overlaps the next
    // channel's kernels), else plain heap memory; frees the previous planes.
    void allocComplexField(uint planes, uint64_t npx);
    void freeComplexField(void);

    OphConfig context_;
    ImageConfig imgCfg;
    Complex<Real>** complex_H = nullptr;
    uint complex_H_planes = 0;
    uint64_t complex_H_pixels = 0;   // per plane
    bool complex_H_pinned = false;
    unsigned long refCnt = 0;
};
