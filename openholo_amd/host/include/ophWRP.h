// ophWRP: the drop-in WRP (wavefront recording plane) generator.
// Public API as in the reference (ophgen/src/ophWRP.h:96-234): scale,
// WRP location, propagation distance, number of WRPs and field length
// (OphWRPConfig, ophgen/src/ophGen.h:646-660); loadPointCloud; readConfig
// (common tags, then FieldLength, ScaleX/Y/Z, Distance, LocationOfWRP,
// NumOfWRP); autoScaling; generateHologram; getWRPBuff.  generateHologram
// drives ohx_wrp_generate once per channel; there is no CPU path (both
// calculateWRPCPU and calculateWRPGPU run it).
#pragma once

#include "ophGen.h"

struct OphWRPConfig {
    Real fieldLength = 0;
    vec3 scale;
    int num_wrp = 0;
    Real wrp_location = 0;
    Real propagation_distance = 0;
};

class ophWRP : public ophGen {
public:
    explicit ophWRP(void);

protected:
    virtual ~ophWRP(void);

public:
    const vec3& getScale() { return wrp_config_.scale; }
    const Real& getLocation() { return wrp_config_.wrp_location; }
    const Real& getDistance() { return wrp_config_.propagation_distance; }
    const int& getNumOfWRP() { return wrp_config_.num_wrp; }
    const Real& getFieldLens() { return wrp_config_.fieldLength; }
    void setScale(vec3 scale) { wrp_config_.scale = scale; }
    void setLocation(Real location) { wrp_config_.wrp_location = location; }
    void setDistance(Real distance) { wrp_config_.propagation_distance = distance; }
    // scaled z maximum (zmax), as the reference's autoScaling computes it; the
    // scaled positions themselves are formed on the device
    void autoScaling();
    int getNumOfPoints() { return (int)obj_.n_points; }

    virtual int loadPointCloud(const char* pc_file);
    virtual bool readConfig(const char* fname);
    void calculateWRPCPU(void) { calculateWRP(); }
    void calculateWRPGPU(void) { calculateWRP(); }
    void generateHologram(void);
    // the reference's multiple-WRP path is unfinished (it stores into a null
    // list, ophgen/src/ophWRP.cpp:466-489): not supported, returns nullptr
    Complex<Real>** calculateMWRP(void);
    Complex<Real>* getWRPBuff(void) { return p_wrp_; }

    void setViewingWindow(bool is_ViewingWindow) { this->is_ViewingWindow = is_ViewingWindow; }
    uint* getProgress() { return &m_nProgress; }
    Vertex* getPointCloudModel() { return obj_.vertices; }
    ulonglong getNumberOfPoints() { return obj_.n_points; }

private:
    bool calculateWRP(void);
    virtual void ophFree(void) override;

    OphPointCloudData obj_;
    OphWRPConfig wrp_config_;
    Complex<Real>* p_wrp_ = nullptr;   // WRP of the last channel (pn_x * pn_y)
    bool is_ViewingWindow = false;
    Real zmax_ = 0;
    uint m_nProgress = 0;
    void* d_vertices = nullptr;
    ulonglong d_vertices_n = 0;
    double* d_wrp = nullptr;
    long long d_wrp_px = 0;
};
