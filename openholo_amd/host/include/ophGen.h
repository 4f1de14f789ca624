// ophGen: generator base of the drop-in (reference ophgen/src/ophGen.h:76-560).
// Owns the per-channel host buffers (complex field, encoded, 8-bit normalised),
// the XML configuration reader, the point-cloud loader and the encoders.
// Every computation runs on the MI355X through libopenholo_amd's C ABI; the
// host buffers keep the reference's layouts so callers read them unchanged.
#pragma once

#include <string>

#include "Openholo.h"

struct ohx_ctx;

// OphPointCloudConfig / OphPointCloudData (ophgen/src/ophGen.h:546-582)
struct OphPointCloudConfig {
    vec3 scale;
    Real distance = 0;
    int8_t* filter_shape_flag = nullptr;
    vec2 filter_width;
    Real focal_length_lens_in = 0, focal_length_lens_out = 0, focal_length_lens_eye_piece = 0;
    vec2 tilt_angle;
};

struct OphPointCloudData {
    ulonglong n_points = 0;
    int n_colors = 0;
    Vertex* vertices = nullptr;
};

class ophGen : public Openholo {
public:
    enum PRECISION { SINGLE, DOUBLE };
    enum ENCODE_FLAG {
        ENCODE_PHASE,
        ENCODE_AMPLITUDE,
        ENCODE_REAL,
        ENCODE_IMAGINARY,
        ENCODE_SIMPLENI,
        ENCODE_BURCKHARDT,
        ENCODE_TWOPHASE,
        ENCODE_SSB,
        ENCODE_OFFSSB,
        ENCODE_SIMPLEBINARY,
        ENCODE_EDBINARY
    };

    explicit ophGen(void);

protected:
    virtual ~ophGen(void);

public:
    Real** getEncodedBuffer(void) { return m_lpEncoded; }
    uchar** getNormalizedBuffer(void) { return m_lpNormalized; }

    void initialize(void);
    int loadPointCloud(const char* pc_file, OphPointCloudData* pc_data_);
    bool readConfig(const char* fname);

    // ENCODE_PHASE / AMPLITUDE / REAL / IMAGINARY of every channel, on the GPU
    void encoding(unsigned int ENCODE_FLAG);
    void normalize(void);
    bool save(const char* fname, uint8_t bitsperpixel = 8, uchar* src = nullptr, uint px = 0, uint py = 0);
    virtual bool loadAsOhc(const char* fname) override;

    void SetMode(unsigned int mode) { m_mode = mode; }
    unsigned int GetMode() { return m_mode; }
    void setPrecision(uint precision) { m_precision = precision; }
    uint getPrecision() { return m_precision; }
    void setResolution(ivec2 resolution);
    ivec2& getEncodeSize(void) { return m_vecEncodeSize; }
    Real getFieldLength() { return m_dFieldLength; }
    int getStream() { return m_nStream; }
    void SetRandomPhase(bool b) { m_bRandomPhase = b; }
    bool GetRandomPhase() { return m_bRandomPhase; }
    // viewing-window transform (ophgen/src/ophGen.cpp:1976-1985)
    void transVW(int nVertex, Vertex* dst, Vertex* src);

    // last device/library error (empty when none); the reference exits instead
    const std::string& lastError() const { return m_lastError; }

protected:
    void resetBuffer();
    void releaseBuffers(void);
    virtual void ophFree(void) override;

    // device session shared by all generators of this process (device 0)
    ohx_ctx* device();
    bool ensureDeviceBuffers();
    bool uploadFields();
    // complex_H and the encode buffers exist for every configured channel and
    // hold pixel_number[_X] * pixel_number[_Y] pixels each (setPixelNumber
    // without setResolution/initialize leaves them at the old size); sets
    // lastError() and returns false otherwise
    bool hostPlanesReady(const char* what);
    void fail(const char* what, int status);

    ivec2 m_vecEncodeSize;
    Real** m_lpEncoded = nullptr;
    uchar** m_lpNormalized = nullptr;
    uint m_nOldChannel = 0;
    uint m_precision = 0;
    Real m_dFieldLength = 0;
    int m_nStream = 1;
    unsigned int m_mode = 0;
    bool m_bRandomPhase = false;
    ivec2 resCfg;

    // device-resident copies of the per-channel buffers
    double** d_field = nullptr;
    double** d_enc = nullptr;
    uint8_t** d_u8 = nullptr;
    uint d_channels = 0;
    long long d_pixels = 0;
    std::string m_lastError;
};
