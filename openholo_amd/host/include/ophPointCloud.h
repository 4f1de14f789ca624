// ophPointCloud: the drop-in point-cloud hologram generator.
// Public API as in the reference (ophgen/src/ophPointCloud.h:547-697):
// PC_DIFF_RS / PC_DIFF_FRESNEL, setters, readConfig, loadPointCloud,
// generateHologram (returns elapsed seconds; 0.0 on a bad flag or failure),
// encoding, viewing window.  genCghPointCloudGPU drives libopenholo_amd
// (FP32 or FP64 direct sums, or the matrix-core Fresnel GEMM when a
// MODE_MFMA_* bit is set); there is no CPU path in the product.
#pragma once

#include "ophGen.h"

class ophPointCloud : public ophGen {
public:
    enum PC_DIFF_FLAG : int8_t { PC_DIFF_RS = 0x00, PC_DIFF_FRESNEL = 0x01 };

    explicit ophPointCloud(void);
    explicit ophPointCloud(const char* pc_file, const char* cfg_file);

protected:
    virtual ~ophPointCloud(void);

public:
    void setScale(Real sx, Real sy, Real sz) { pc_config_.scale = vec3(sx, sy, sz); }
    void setDistance(Real distance) { pc_config_.distance = distance; }
    void setFilterShapeFlag(int8_t* fsf) { pc_config_.filter_shape_flag = fsf; }
    void setFilterWidth(Real wx, Real wy) { pc_config_.filter_width = vec2(wx, wy); }
    void setFocalLength(Real lens_in, Real lens_out, Real lens_eye_piece)
    {
        pc_config_.focal_length_lens_in = lens_in;
        pc_config_.focal_length_lens_out = lens_out;
        pc_config_.focal_length_lens_eye_piece = lens_eye_piece;
    }
    void setTiltAngle(Real ax, Real ay) { pc_config_.tilt_angle = vec2(ax, ay); }
    // deep copy of n_points vertices (set the count first, as upstream)
    void setPointCloudModel(Vertex* vertex);
    void setNumberOfPoints(ulonglong n_points) { pc_data_.n_points = n_points; }

    void getScale(vec3& scale) { scale = pc_config_.scale; }
    Real getDistance(void) { return pc_config_.distance; }
    int8_t* getFilterShapeFlag(void) { return pc_config_.filter_shape_flag; }
    void getFilterWidth(vec2& w) { w = pc_config_.filter_width; }
    void getFocalLength(Real* lens_in, Real* lens_out, Real* lens_eye_piece);
    void getTiltAngle(vec2& a) { a = pc_config_.tilt_angle; }
    Vertex* getPointCloudModel() { return pc_data_.vertices; }
    ulonglong getNumberOfPoints() { return pc_data_.n_points; }

    // Whether generation runs on the host CPU (ophgen/src/ophPointCloud.h:615-623
    // reports the MODE_GPU bit).  This library has no CPU path: generateHologram
    // runs on the GPU whatever the mode, so this is always false.
    bool isCPU() { return false; }
    int loadPointCloud(const char* pc_file);
    bool readConfig(const char* cfg_file);
    Real generateHologram(uint diff_flag = PC_DIFF_RS);
    virtual void encoding(unsigned int ENCODE_FLAG);
    // ENCODE_SSB -> encodeHologram(); ENCODE_OFFSSB (ophGen's side-band
    // encoder) is not provided and fails (ophPointCloud.cpp:247-251)
    virtual void encoding(unsigned int ENCODE_FLAG, unsigned int SSB_PASSBAND);
    // single-side-band encoding of every channel on the GPU
    // (ophgen/src/ophPointCloud.cpp:193-240); band_limit is unused, as there
    void encodeHologram(vec2 band_limit = vec2(0.8, 0.5), vec2 spectrum_shift = vec2(0.0, 0.5));
    void setViewingWindow(bool is_ViewingWindow) { this->is_ViewingWindow = is_ViewingWindow; }
    uint* getProgress() { return &m_nProgress; }

private:
    bool genCghPointCloudGPU(uint diff_flag);
    virtual void ophFree(void) override;

    bool is_ViewingWindow = false;
    uint m_nProgress = 0;
    OphPointCloudConfig pc_config_;
    OphPointCloudData pc_data_;
    // device copy of the vertices, re-uploaded when the model changes
    void* d_vertices = nullptr;
    ulonglong d_vertices_n = 0;
    unsigned long long model_version = 1, device_version = 0;
};
