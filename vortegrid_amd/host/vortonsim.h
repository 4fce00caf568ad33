// vortonsim.h -- C++ facade with VorteGrid's VortonSim / particle-operation API
// (Samples/MjgIntelFluidDemo_Part20: VortonFluid/vortonSim.h:137-623,
// VortonFluid/pclOpVortonSim.h:46-71, Particles/Operation/particleOperation.h:
// 27-44, pclOpAdvect.h, pclOpEvolve.h, pclOpFindBoundingBox.h) over the C-ABI
// in include/vortegrid.h.  Header-only; link with libvortegrid.so.
//
// Host-side particle arrays keep the reference's 72-byte Particle record, so
// an application that owns VECTOR<Particle> groups can swap its VortonSim /
// operators for these classes.  Each Operate/Update call stages the records on
// the device (one copy in, one copy out), runs the HIP path, and writes
// velocities back into the records.  Grids are materialised on the host on
// demand (GetVelocityGrid and friends).
#pragma once

#include <sys/stat.h>

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vortegrid.h"

namespace vgfacade {

inline void check(vg_status s, const char *what) {
    if (s != VG_OK) throw std::runtime_error(std::string(what) + ": " + vg_last_error());
}

// ------------------------------------------------------------------ math / particles
struct Vec3 {
    float x, y, z;
    Vec3() : x(0), y(0), z(0) {}
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    Vec3 operator+(const Vec3 &o) const { return Vec3(x + o.x, y + o.y, z + o.z); }
    Vec3 operator-(const Vec3 &o) const { return Vec3(x - o.x, y - o.y, z - o.z); }
    Vec3 operator*(float f) const { return Vec3(x * f, y * f, z * f); }
    Vec3 operator-() const { return Vec3(-x, -y, -z); }
};

// particle.h:425-446 (72 bytes, ENABLE_FIRE layout).
struct Particle {
    Vec3 mPosition;
    Vec3 mVelocity;
    Vec3 mOrientation;
    Vec3 mAngularVelocity;  // = vorticity / 2 for vortons
    float mDensity = 1.0f;
    float mSize = 0.0f;     // diameter
    int mBirthTime = 0;
    float mFuelFraction = 0.0f, mFlameFraction = 0.0f, mSmokeFraction = 0.0f;
    float GetRadius() const { return mSize * 0.5f; }
    void SetRadius(float r) { mSize = 2.0f * r; }
};
static_assert(sizeof(Particle) == 72, "Particle must keep the reference's 72-byte record");

// vorton.h:198-387: a Particle whose angular velocity is half its vorticity.
struct Vorton : Particle {
    Vorton() = default;
    Vorton(const Vec3 &pos, const Vec3 &vort, float radius) {
        mPosition = pos;
        mAngularVelocity = vort * 0.5f;
        SetRadius(radius);
    }
    Vec3 GetVorticity() const { return mAngularVelocity * 2.0f; }
    void SetVorticity(const Vec3 &w) { mAngularVelocity = w * 0.5f; }
};
static_assert(sizeof(Vorton) == 72, "Vorton adds no data to Particle");

template <typename T>
using VECTOR = std::vector<T>;

constexpr size_t kStride = sizeof(Particle) / sizeof(float);  // 18 floats
constexpr size_t kOffVel = offsetof(Particle, mVelocity) / sizeof(float);
constexpr size_t kOffAngVel = offsetof(Particle, mAngularVelocity) / sizeof(float);
constexpr size_t kOffSize = offsetof(Particle, mSize) / sizeof(float);
constexpr size_t kOffBirth = offsetof(Particle, mBirthTime) / sizeof(float);

// ------------------------------------------------------------------ grids
// UniformGridGeometry (uniformGrid.h:144-860).
class UniformGridGeometry {
  public:
    UniformGridGeometry() { std::memset(&g_, 0, sizeof g_); }
    explicit UniformGridGeometry(const vg_grid_desc &g) : g_(g) {}
    void DefineShape(size_t n, const Vec3 &mn, const Vec3 &mx, bool pow2) {
        const float a[3] = {mn.x, mn.y, mn.z}, b[3] = {mx.x, mx.y, mx.z};
        check(vg_define_shape(&g_, n, a, b, pow2 ? 1 : 0), "DefineShape");
    }
    unsigned GetNumPoints(unsigned i) const { return g_.np[i]; }
    unsigned GetNumCells(unsigned i) const { return g_.np[i] - 1; }
    unsigned GetGridCapacity() const { return g_.np[0] * g_.np[1] * g_.np[2]; }
    Vec3 GetMinCorner() const { return Vec3(g_.min[0], g_.min[1], g_.min[2]); }
    Vec3 GetExtent() const { return Vec3(g_.extent[0], g_.extent[1], g_.extent[2]); }
    Vec3 GetCellSpacing() const { return Vec3(g_.cell[0], g_.cell[1], g_.cell[2]); }
    Vec3 GetCellsPerExtent() const { return Vec3(g_.cpe[0], g_.cpe[1], g_.cpe[2]); }
    bool HasZeroExtent() const {
        return g_.extent[0] == 0.f && g_.extent[1] == 0.f && g_.extent[2] == 0.f;
    }
    size_t OffsetFromIndices(size_t ix, size_t iy, size_t iz) const {
        return ix + g_.np[0] * (iy + g_.np[1] * iz);
    }
    const vg_grid_desc &desc() const { return g_; }

  protected:
    vg_grid_desc g_;
};

// uniformGrid.h:87-95
struct Stats_Float {
    float mMean = 0.0f;
    float mStdDev = 0.0f;
    float mMin = FLT_MAX;
    float mMax = -FLT_MAX;
};

// UniformGrid<T> (uniformGrid.h:862-1785).  The values live either on the
// host (contents_) or, for grids a VortonSim fills, on the device: such a
// grid is "attached" to the device field and copies it to the host only on
// the first host access.  A non-const host access detaches it (the host copy
// may then be changed), so operators that read the grid see the host values.
template <class ItemT>
class UniformGrid : public UniformGridGeometry {
  public:
    UniformGrid() = default;
    explicit UniformGrid(const UniformGridGeometry &geom) : UniformGridGeometry(geom) {}
    // uniformGrid.h:1044-1060
    void Init(const ItemT &initialValue = ItemT()) {
        device_ = nullptr;
        contents_.assign(GetGridCapacity(), initialValue);
        host_valid_ = true;
    }
    // uniformGrid.h:1721-1725: drop the values and the shape
    void Clear() {
        device_ = nullptr;
        contents_.clear();
        host_valid_ = true;
        std::memset(&g_, 0, sizeof g_);
    }
    void CopyShape(const UniformGridGeometry &geom) { g_ = geom.desc(); }
    size_t Size() const { return device_ ? (size_t)GetGridCapacity() : contents_.size(); }
    bool Empty() const { return Size() == 0; }
    const ItemT &operator[](size_t o) const {
        materialise();
        return contents_[o];
    }
    ItemT &operator[](size_t o) {
        materialise();
        device_ = nullptr;
        return contents_[o];
    }
    const ItemT *data() const {
        materialise();
        return contents_.data();
    }
    ItemT *data() {
        materialise();
        device_ = nullptr;
        return contents_.data();
    }
    // Copy a packed device field of this geometry into the host grid now.
    void Download(const vg_grid_desc &g, const void *device) {
        AttachDevice(g, device);
        materialise();
    }
    // Take the geometry and refer to a device field (packed, x fastest);
    // the host copy is made on the first host access.
    void AttachDevice(const vg_grid_desc &g, const void *device) {
        g_ = g;
        device_ = device;
        host_valid_ = device == nullptr;
        if (!device) contents_.assign(GetGridCapacity(), ItemT());
    }
    // The device field this grid refers to, or null when its values are the
    // host copy's.
    const void *DeviceData() const { return device_; }
    // Wrap host values of the given geometry.
    void Assign(const vg_grid_desc &g, const ItemT *values) {
        g_ = g;
        device_ = nullptr;
        contents_.assign(values, values + GetGridCapacity());
        host_valid_ = true;
    }
    // uniformGridDiagnostics.cpp:209-314 (Vec3 grids only): OGLE brick of bytes.
    void GenerateBrickOfBytes(const char *strFilenameBase, unsigned uFrame) const;

  private:
    void materialise() const {
        if (host_valid_) return;
        contents_.resize(GetGridCapacity());
        if (!contents_.empty())
            check(vg_memcpy_dtoh(contents_.data(), device_, contents_.size() * sizeof(ItemT)),
                  "UniformGrid download");
        host_valid_ = true;
    }

    mutable std::vector<ItemT> contents_;
    mutable bool host_valid_ = true;
    const void *device_ = nullptr;
};

// NestedGrid<T> (nestedGrid.h:31-393): a pyramid of grids, layer 0 finest.
template <class ItemT>
class NestedGrid {
  public:
    size_t GetDepth() const { return layers_.size(); }
    UniformGrid<ItemT> &operator[](size_t l) { return layers_[l]; }
    const UniformGrid<ItemT> &operator[](size_t l) const { return layers_[l]; }
    // nestedGrid.h:142-146: decimations of layer l relative to layer l-1
    const unsigned *GetDecimations(size_t l) const { return dec_[l].d; }
    void Clear() {
        layers_.clear();
        dec_.clear();
    }
    void Resize(size_t depth) {
        layers_.resize(depth);
        dec_.resize(depth);
    }
    void SetDecimations(size_t l, const unsigned d[3]) {
        for (int a = 0; a < 3; ++a) dec_[l].d[a] = d[a];
    }

  private:
    struct Dec { unsigned d[3] = {0, 0, 0}; };
    std::vector<UniformGrid<ItemT>> layers_;
    std::vector<Dec> dec_;
};

// uniformGridDiagnostics.cpp:209-314 with ENFORCE_SYMMETRIC_RANGE off, as the
// reference builds it.  Writes Ogle/Vols/<base>{X,Y,Z,M}<frame>-<nx>x<ny>x<nz>.dat
// (one byte per point, x fastest, then "MIN %g MAX %g\n") and appends two
// lines to Ogle/<base>.ogle.  The range comes from ComputeStatistics
// (uniformGrid.h:1095-1107): Min2/Max2 (macros.h:61-64) on Vec3 keep the
// running extreme only when it is strictly below (above) the next value in
// all three components (vec3.h:389-392), a sequential, order-dependent scan
// that the bytes depend on.
template <>
inline void UniformGrid<Vec3>::GenerateBrickOfBytes(const char *strFilenameBase,
                                                    unsigned uFrame) const {
    materialise();
    if (contents_.empty()) return;
    Vec3 vMin = contents_[0], vMax = vMin;
    for (const Vec3 &r : contents_) {
        const bool lt = vMin.x < r.x && vMin.y < r.y && vMin.z < r.z;
        const bool gt = vMax.x > r.x && vMax.y > r.y && vMax.z > r.z;
        if (!lt) vMin = r;
        if (!gt) vMax = r;
    }
    auto max2 = [](float a, float b) { return a > b ? a : b; };
    const float ex = max2(-vMin.x, vMax.x), ey = max2(-vMin.y, vMax.y), ez = max2(-vMin.z, vMax.z);
    const float fMagMax = std::sqrt(ex * ex + ey * ey + ez * ez);
    const float range[3] = {max2(FLT_MIN, vMax.x - vMin.x), max2(FLT_MIN, vMax.y - vMin.y),
                            max2(FLT_MIN, vMax.z - vMin.z)};
    const unsigned nx = GetNumPoints(0), ny = GetNumPoints(1), nz = GetNumPoints(2);
    char names[4][512];
    const char axis[4] = {'X', 'Y', 'Z', 'M'};
    for (int k = 0; k < 4; ++k)
        std::snprintf(names[k], sizeof names[k], "Ogle/Vols/%s%c%05u-%ux%ux%u.dat",
                      strFilenameBase, axis[k], uFrame, nx, ny, nz);
    ::mkdir("Ogle", 0777);
    ::mkdir("Ogle/Vols", 0777);
    {
        char script[512];
        std::snprintf(script, sizeof script, "Ogle/%s.ogle", strFilenameBase);
        FILE *f = std::fopen(script, "a");
        if (!f) throw std::runtime_error("GenerateBrickOfBytes: cannot open the script file");
        std::fprintf(f, "# %s ranges: {%9.7g,%9.7g,%9.7g} to {%9.7g,%9.7g,%9.7g}\n",
                     strFilenameBase, vMin.x, vMin.y, vMin.z, vMax.x, vMax.y, vMax.z);
        std::fprintf(f, "%ux%ux%u %s %s %s\n", nx, ny, nz, names[0], names[1], names[2]);
        std::fclose(f);
    }
    const float fAlmost256 = 256.0f * (1.0f - FLT_EPSILON);
    std::vector<unsigned char> planes[4];
    for (auto &p : planes) p.resize(contents_.size());
    for (size_t o = 0; o < contents_.size(); ++o) {
        const Vec3 &v = contents_[o];
        const float fMag = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        const float t[3] = {(v.x - vMin.x) / range[0], (v.y - vMin.y) / range[1],
                            (v.z - vMin.z) / range[2]};
        for (int k = 0; k < 3; ++k) planes[k][o] = (unsigned char)(int)(t[k] * fAlmost256);
        planes[3][o] = (unsigned char)(int)(fAlmost256 * fMag / (fMagMax + FLT_MIN));
    }
    const float lo[4] = {vMin.x, vMin.y, vMin.z, 0.0f}, hi[4] = {vMax.x, vMax.y, vMax.z, fMagMax};
    for (int k = 0; k < 4; ++k) {
        FILE *f = std::fopen(names[k], "wb");
        if (!f) throw std::runtime_error("GenerateBrickOfBytes: cannot create a data file");
        std::fwrite(planes[k].data(), 1, planes[k].size(), f);
        std::fprintf(f, "MIN %g MAX %g\n", lo[k], hi[k]);
        std::fclose(f);
    }
}

// ------------------------------------------------------------------ device staging
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { vg_free(p_); }
    void *reserve(size_t bytes) {
        if (bytes > cap_) {
            vg_free(p_);
            p_ = nullptr;
            cap_ = 0;
            check(vg_malloc(&p_, bytes), "device allocation");
            cap_ = bytes;
        }
        return p_;
    }
    void *get() const { return p_; }

  private:
    void *p_ = nullptr;
    size_t cap_ = 0;
};

// ------------------------------------------------------------------ operators
// particleOperation.h:27-44
class IParticleOperation {
  public:
    virtual ~IParticleOperation() {}
    virtual IParticleOperation *Create() = 0;
    virtual IParticleOperation *Clone() = 0;
    virtual void Operate(VECTOR<Particle> &particles, float timeStep, unsigned uFrame) = 0;
};

#define VGFACADE_VIRTUAL_CONSTRUCTORS(ClassT)                         \
    IParticleOperation *Create() override { return new ClassT(); }    \
    IParticleOperation *Clone() override { return new ClassT(*this); }

// ------------------------------------------------------------------ VortonSim
// vortonSim.h:137-623, velocity-from-vorticity subset.
class VortonSim {
  public:
    // vortonSim.h:143-149, 256-264
    enum FluidSimulationTechniqueE {
        FLUID_SIM_VORTEX_PARTICLE_METHOD,
        FLUID_SIM_SMOOTHED_PARTICLE_HYDRODYNAMICS,
        FLUID_SIM_VPM_SPH_HYBRID,
        FLUID_SIM_NUM
    };
    enum InvestigationTermE {
        INVESTIGATE_ALL,
        INVESTIGATE_STRETCHING_TILTING,
        INVESTIGATE_BAROCLINIC,
        INVESTIGATE_VISCOUS_DIFFUSION,
        INVESTIGATE_THERMAL_DIFFUSION,
        INVESTIGATE_NONE
    };

    explicit VortonSim(float viscosity = 0.0f, float ambientFluidDensity = 1.0f,
                       int technique = VG_TECH_POISSON)
        : viscosity_(viscosity), density_(ambientFluidDensity) {
        vg_sim_config_default(&cfg_);
        cfg_.technique = technique;
    }
    // Copies keep the settings and drop the vortons and the device state
    // (vortonSim.h:281-295).
    VortonSim(const VortonSim &that) { copy_settings(that); }
    VortonSim &operator=(const VortonSim &that) {
        if (this != &that) {
            vg_sim_destroy(sim_);
            sim_ = nullptr;
            copy_settings(that);
        }
        return *this;
    }
    ~VortonSim() { vg_sim_destroy(sim_); }

    vg_sim_config &Config() { return cfg_; }

    // vortonSim.h:325-326, 422; defaults of vortonSim.cpp:392, 398.  Of the
    // vorticity terms, stretching/tilting and viscous diffusion are on this
    // path (baroclinic generation and heat diffusion are not).
    void SetFluidSimulationTechnique(FluidSimulationTechniqueE t) { fluidSimTechnique_ = t; }
    const FluidSimulationTechniqueE &GetFluidSimulationTechnique() const {
        return fluidSimTechnique_;
    }
    InvestigationTermE &GetInvestigationTerm() { return investigationTerm_; }
    // No reference counterpart: VG_PSE_REFERENCE_ORDER (default, the
    // reference's result bit for bit) or VG_PSE_SIMULTANEOUS (parallel).
    void SetPseOrder(int order) { pseOrder_ = order; }

    // vortonSim.h:346-375.  Viscosity feeds PSE diffusion; the ambient
    // density, gravity and heat capacity belong to the baroclinic and thermal
    // terms, which are not on this path, and are only kept.
    void SetViscosity(float viscosity) { viscosity_ = viscosity; }
    const float &GetViscosity() const { return viscosity_; }
    void SetAmbientDensity(float d) { density_ = d; }
    const float &GetAmbientDensity() const { return density_; }
    void SetSpecificHeatCapacity(float c) { heatCapacity_ = c; }
    const float &GetSpecificHeatCapacity() const { return heatCapacity_; }
    void SetGravitationalAcceleration(const Vec3 &g) { gravity_ = g; }
    const Vec3 &GetGravitationalAcceleration() { return gravity_; }

    void SetVortons(VECTOR<Vorton> *vortons) { mVortons = vortons; }
    VECTOR<Vorton> *GetVortons() { return mVortons; }
    const VECTOR<Vorton> *GetVortons() const { return mVortons; }

    // Initialize (vortonSim.cpp:933-992), velocity-path part: bounding box.
    void Initialize(bool /*systemTracksFluidSurfaceUsingSdf*/ = false) {
        FindBoundingBox();
        if (mVortons && !mVortons->empty()) {
            const Vec3 p = (*mVortons)[0].mPosition;
            UpdateBoundingBox(p, p, true);
        }
    }

    // vortonSim.cpp:1001-1008
    void FindBoundingBox() {
        ensure();
        stage_vortons();
        check(vg_sim_find_bounding_box(sim_), "FindBoundingBox");
        sync_bounds();
    }

    // vortonSim.cpp:1027-1096
    void UpdateBoundingBox(const Vec3 &minCorner, const Vec3 &maxCorner, bool bFinal) {
        ensure();
        const float a[3] = {minCorner.x, minCorner.y, minCorner.z};
        const float b[3] = {maxCorner.x, maxCorner.y, maxCorner.z};
        check(vg_sim_update_bounding_box(sim_, a, b, bFinal ? 1 : 0), "UpdateBoundingBox");
        sync_bounds();
        if (bFinal) {
            vg_grid_desc g;
            check(vg_sim_grid_template(sim_, &g), "GetGrid");
            grid_ = UniformGridGeometry(g);
        }
    }

    // vortonSim.cpp:4279-4344 (velocity-from-vorticity part of
    // UpdateVortexParticleMethod, :4117-4256): vorton velocities are written
    // into the records (DIRECT: direct velocity; otherwise grid velocity);
    // for the vortex particle method with stretching investigated
    // (:4208-4215), StretchAndTiltVortons then updates their angular
    // velocities from the same velocity grid, and DiffuseAndDissipateVorticityPSE
    // (:4238-4241) exchanges vorticity between neighbours with the viscosity.
    // The velocity grid object (GetVelocityGrid) keeps its address and refers
    // to the new device field afterwards.
    void Update(float timeStep, unsigned uFrame) {
        ensure();
        if (!mVortons || mVortons->empty()) return;
        stage_vortons();
        float *base = static_cast<float *>(dev_.get());
        check(vg_sim_update(sim_, timeStep, uFrame, base + kOffVel, kStride), "Update");
        if (fluidSimTechnique_ == FLUID_SIM_VORTEX_PARTICLE_METHOD &&
            (investigationTerm_ == INVESTIGATE_ALL ||
             investigationTerm_ == INVESTIGATE_STRETCHING_TILTING))
            check(vg_sim_stretch_tilt(sim_, timeStep, base + kOffAngVel, kStride),
                  "StretchAndTiltVortons");
        // :4238-4241 (every technique); with zero viscosity every exchange
        // and the dissipation are zero, so the pass is skipped.
        if (viscosity_ != 0.0f && (investigationTerm_ == INVESTIGATE_ALL ||
                                   investigationTerm_ == INVESTIGATE_VISCOUS_DIFFUSION))
            check(vg_sim_pse_diffuse(sim_, viscosity_, timeStep, base + kOffAngVel, kStride,
                                     reinterpret_cast<const int32_t *>(base + kOffBirth),
                                     kStride, pseOrder_),
                  "DiffuseAndDissipateVorticityPSE");
        check(vg_memcpy_dtoh(mVortons->data(), dev_.get(), mVortons->size() * sizeof(Vorton)),
              "download vortons");
        vg_grid_desc g;
        const float *f = nullptr;
        check(vg_sim_velocity_grid(sim_, &g, &f), "GetVelocityGrid");
        mVelGrid.AttachDevice(g, f);
        multigrids_stale_ = true;
    }

    void Clear() {
        vg_sim_destroy(sim_);
        sim_ = nullptr;
        mVortons = nullptr;
        mVelGrid.Clear();
        mVectorPotentialMultiGrid.Clear();
        mNegativeVorticityMultiGrid.Clear();
    }

    // vortonSim.h:333-343.  The grids are members whose addresses stay valid
    // across frames (operators may keep pointers to them); their values are
    // the device fields of the last Update, copied to the host on first host
    // access.
    const UniformGrid<Vec3> &GetVelocityGrid() const { return mVelGrid; }
    UniformGrid<Vec3> &GetVelocityGrid() { return mVelGrid; }
    const NestedGrid<Vec3> &GetNegativeVorticityMultiGrid() const {
        sync_multigrids();
        return mNegativeVorticityMultiGrid;
    }
    NestedGrid<Vec3> &GetNegativeVorticityMultiGrid() {
        sync_multigrids();
        return mNegativeVorticityMultiGrid;
    }
    const NestedGrid<Vec3> &GetVectorPotentialMultiGrid() const {
        sync_multigrids();
        return mVectorPotentialMultiGrid;
    }
    NestedGrid<Vec3> &GetVectorPotentialMultiGrid() {
        sync_multigrids();
        return mVectorPotentialMultiGrid;
    }
    // vortonSim.h:318: statistics of the Poisson residual.  The reference
    // tallies the last SOR pass's update magnitudes, and only in a serial
    // build (uniformGridMath.cpp:1173-1184; its threaded build leaves the
    // reset values).  Here they are the statistics over the interior points
    // of |f - L A| for the last POISSON update's vector potential A and
    // right-hand side f = -w, the 7-point Laplacian L in FP64, reported in
    // FP32.  Computed on the host from the two grids when called.
    const Stats_Float &GetPoissonResidualStats() {
        sync_multigrids();
        residualStats_ = Stats_Float();
        if (mVectorPotentialMultiGrid.GetDepth() == 0) return residualStats_;
        const UniformGrid<Vec3> &A = mVectorPotentialMultiGrid[0];
        const UniformGrid<Vec3> &F = mNegativeVorticityMultiGrid[0];
        const unsigned nx = A.GetNumPoints(0), ny = A.GetNumPoints(1), nz = A.GetNumPoints(2);
        const Vec3 h = A.GetCellSpacing();
        const double r2[3] = {1.0 / ((double)h.x * h.x), 1.0 / ((double)h.y * h.y),
                              1.0 / ((double)h.z * h.z)};
        const Vec3 *a = A.data();
        const Vec3 *f = F.data();
        double sum = 0.0, sum2 = 0.0, mn = DBL_MAX, mx = -DBL_MAX;
        size_t count = 0;
        const size_t sx = 1, sy = nx, sz = (size_t)nx * ny;
        for (unsigned iz = 1; iz + 1 < nz; ++iz)
            for (unsigned iy = 1; iy + 1 < ny; ++iy)
                for (unsigned ix = 1; ix + 1 < nx; ++ix) {
                    const size_t o = ix + sy * iy + sz * iz;
                    double r[3];
                    const float *c = &a[o].x;
                    for (int k = 0; k < 3; ++k) {
                        auto at = [&](size_t off) { return (double)(&a[off].x)[k]; };
                        const double lap = (at(o + sx) + at(o - sx) - 2.0 * c[k]) * r2[0] +
                                           (at(o + sy) + at(o - sy) - 2.0 * c[k]) * r2[1] +
                                           (at(o + sz) + at(o - sz) - 2.0 * c[k]) * r2[2];
                        r[k] = (double)(&f[o].x)[k] - lap;
                    }
                    const double m = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                    sum += m;
                    sum2 += m * m;
                    mn = m < mn ? m : mn;
                    mx = m > mx ? m : mx;
                    ++count;
                }
        if (count) {
            const double mean = sum / count;
            const double var = sum2 / count - mean * mean;
            residualStats_.mMean = (float)mean;
            residualStats_.mStdDev = (float)std::sqrt(var > 0.0 ? var : 0.0);
            residualStats_.mMin = (float)mn;
            residualStats_.mMax = (float)mx;
        }
        return residualStats_;
    }
    // No reference counterpart: the FP64 true relative residual
    // ||f - L A|| / ||f|| of the last POISSON update (device reduction).
    double GetPoissonResidual() const {
        double r = -1.0;
        if (sim_) check(vg_sim_poisson_residual(sim_, &r), "GetPoissonResidual");
        return r;
    }
    // Device view of the velocity grid (no copy), for device-side operators.
    const float *GetVelocityGridDevice(vg_grid_desc *g) const {
        const float *f = nullptr;
        check(vg_sim_velocity_grid(sim_, g, &f), "GetVelocityGrid");
        return f;
    }
    // vortonSim.h:405-411.  The two corners are adjacent in memory, as the
    // reference's are: PclOpPopulateVelocityGrid reads mBoundingBox[0..1]
    // through the address of GetMinCorner() (particleSystemConfiguration.cpp:140).
    const UniformGridGeometry &GetGrid() const { return grid_; }
    const Vec3 &GetMinCorner() const { return bounds_[0]; }
    const Vec3 &GetMaxCorner() const { return bounds_[1]; }

  private:
    void copy_settings(const VortonSim &that) {
        cfg_ = that.cfg_;
        viscosity_ = that.viscosity_;
        density_ = that.density_;
        heatCapacity_ = that.heatCapacity_;
        gravity_ = that.gravity_;
        fluidSimTechnique_ = that.fluidSimTechnique_;
        investigationTerm_ = that.investigationTerm_;
        pseOrder_ = that.pseOrder_;
        mVortons = nullptr;
    }
    void ensure() {
        if (!sim_) check(vg_sim_create(&sim_, &cfg_, nullptr), "VortonSim");
    }
    void stage_vortons() {
        const size_t n = mVortons ? mVortons->size() : 0;
        vg_vortons v;
        std::memset(&v, 0, sizeof v);
        v.stride = kStride;
        v.n = n;
        if (n) {
            dev_.reserve(n * sizeof(Vorton));
            check(vg_memcpy_htod(dev_.get(), mVortons->data(), n * sizeof(Vorton)),
                  "upload vortons");
            float *base = static_cast<float *>(dev_.get());
            v.pos = base;
            v.angvel = base + kOffAngVel;
            v.size = base + kOffSize;
        }
        check(vg_sim_set_vortons(sim_, &v), "SetVortons");
    }
    void sync_bounds() {
        float a[3], b[3];
        check(vg_sim_bounds(sim_, a, b), "bounds");
        bounds_[0] = Vec3(a[0], a[1], a[2]);
        bounds_[1] = Vec3(b[0], b[1], b[2]);
    }
    void sync_multigrids() const {
        if (!multigrids_stale_ || !sim_) return;
        NestedGrid<Vec3> *grids[2] = {&mVectorPotentialMultiGrid, &mNegativeVorticityMultiGrid};
        const int fields[2] = {VG_MG_VECTOR_POTENTIAL, VG_MG_NEGATIVE_VORTICITY};
        for (int k = 0; k < 2; ++k) {
            vg_grid_desc g;
            const float *f = nullptr;
            uint32_t depth = 0;
            grids[k]->Clear();
            if (vg_sim_multigrid(sim_, fields[k], 0, &g, &f, &depth) != VG_OK || depth == 0)
                continue;
            grids[k]->Resize(depth);
            for (uint32_t l = 0; l < depth; ++l) {
                check(vg_sim_multigrid(sim_, fields[k], l, &g, &f, nullptr),
                      "GetVectorPotentialMultiGrid");
                (*grids[k])[l].AttachDevice(g, f);
            }
            vg_grid_desc layers[16], g0;
            uint32_t dec[48] = {0}, nl = 0;
            check(vg_sim_multigrid(sim_, fields[k], 0, &g0, &f, nullptr), "multigrid");
            check(vg_nested_layers(&g0, layers, dec, 16, &nl), "NestedGrid");
            for (uint32_t l = 0; l < depth && l < nl; ++l) {
                const unsigned d[3] = {dec[3 * l], dec[3 * l + 1], dec[3 * l + 2]};
                grids[k]->SetDecimations(l, d);
            }
        }
        multigrids_stale_ = false;
    }

    vg_sim_config cfg_;
    vg_sim *sim_ = nullptr;
    VECTOR<Vorton> *mVortons = nullptr;
    DeviceBuffer dev_;
    UniformGridGeometry grid_;
    UniformGrid<Vec3> mVelGrid;
    mutable NestedGrid<Vec3> mVectorPotentialMultiGrid, mNegativeVorticityMultiGrid;
    mutable bool multigrids_stale_ = true;
    Stats_Float residualStats_;
    Vec3 bounds_[2];
    float viscosity_ = 0.0f, density_ = 1.0f, heatCapacity_ = 1.0f;
    Vec3 gravity_;
    FluidSimulationTechniqueE fluidSimTechnique_ = FLUID_SIM_VPM_SPH_HYBRID;
    InvestigationTermE investigationTerm_ = INVESTIGATE_ALL;
    int pseOrder_ = VG_PSE_REFERENCE_ORDER;
};

// pclOpVortonSim.h:46-71 / pclOpVortonSim.cpp:51-71.  mMinCorner and
// mMaxCorner point at the tracer group's bounding box (patched in by the
// particle system, particleSystemConfiguration.cpp:286-296).
class PclOpVortonSim : public IParticleOperation {
  public:
    PclOpVortonSim() = default;
    PclOpVortonSim(const PclOpVortonSim &that)
        : mVortonSim(that.mVortonSim), mMinCorner(that.mMinCorner), mMaxCorner(that.mMaxCorner) {}
    VGFACADE_VIRTUAL_CONSTRUCTORS(PclOpVortonSim)
    void Operate(VECTOR<Particle> &particles, float timeStep, unsigned uFrame) override {
        mVortonSim.SetVortons(reinterpret_cast<VECTOR<Vorton> *>(&particles));
        mVortonSim.FindBoundingBox();
        if (mMinCorner && mMaxCorner) {
            mVortonSim.UpdateBoundingBox(*mMinCorner, *mMaxCorner, true);
        } else if (!particles.empty()) {
            const Vec3 p = particles[0].mPosition;
            mVortonSim.UpdateBoundingBox(p, p, true);
        }
        mVortonSim.Update(timeStep, uFrame);
    }
    VortonSim mVortonSim;
    const Vec3 *mMinCorner = nullptr;
    const Vec3 *mMaxCorner = nullptr;
};

// Particle group staged on the device for the particle operators below.
class DeviceParticles {
  public:
    float *upload(const VECTOR<Particle> &p) {
        buf_.reserve(p.size() * sizeof(Particle));
        if (!p.empty())
            check(vg_memcpy_htod(buf_.get(), p.data(), p.size() * sizeof(Particle)), "upload");
        return static_cast<float *>(buf_.get());
    }
    void download(VECTOR<Particle> &p) {
        if (!p.empty())
            check(vg_memcpy_dtoh(p.data(), buf_.get(), p.size() * sizeof(Particle)), "download");
    }

  private:
    DeviceBuffer buf_;
};

// A grid's values as a device field: the field a VortonSim filled when the
// grid is still attached to it, else the host values uploaded to `scratch`.
inline const float *device_field(const UniformGrid<Vec3> &grid, DeviceBuffer &scratch) {
    if (grid.DeviceData()) return static_cast<const float *>(grid.DeviceData());
    const size_t bytes = grid.Size() * sizeof(Vec3);
    if (!bytes) return nullptr;
    scratch.reserve(bytes);
    check(vg_memcpy_htod(scratch.get(), grid.data(), bytes), "upload grid");
    return static_cast<const float *>(scratch.get());
}

// pclOpAdvect.cpp:132-152: velocity = interpolated grid velocity (gain 1)
// or velocity += gain (interpolated - velocity).  A grid with zero extent
// leaves the particles as they are; so does a null grid (the reference
// would dereference it).
inline void AssignVelocityFromField(VECTOR<Particle> &particles,
                                    const UniformGrid<Vec3> *velocityGrid, const float gain) {
    if (!velocityGrid || particles.empty() || velocityGrid->HasZeroExtent() ||
        velocityGrid->Empty())
        return;
    static thread_local DeviceParticles dev;
    static thread_local DeviceBuffer grid_scratch;
    const float *f = device_field(*velocityGrid, grid_scratch);
    float *base = dev.upload(particles);
    check(vg_assign_velocity_from_field(&velocityGrid->desc(), f, base, base + kOffVel, kStride,
                                        particles.size(), gain, nullptr),
          "AssignVelocityFromField");
    dev.download(particles);
}

// pclOpAdvect.h:19-33, pclOpAdvect.cpp:157-163
class PclOpAssignVelocityFromField : public IParticleOperation {
  public:
    PclOpAssignVelocityFromField() = default;
    PclOpAssignVelocityFromField(const PclOpAssignVelocityFromField &o)
        : mVelocityGrid(o.mVelocityGrid), mGridWeight(o.mGridWeight) {}
    VGFACADE_VIRTUAL_CONSTRUCTORS(PclOpAssignVelocityFromField)
    void Operate(VECTOR<Particle> &particles, float, unsigned) override {
        if (mVelocityGrid) AssignVelocityFromField(particles, mVelocityGrid, mGridWeight);
    }
    const UniformGrid<Vec3> *mVelocityGrid = nullptr;  // grid of velocity values
    float mGridWeight = 1.0f;  // share of the grid velocity assigned each frame
};

// pclOpEvolve.cpp:118-123
class PclOpEvolve : public IParticleOperation {
  public:
    VGFACADE_VIRTUAL_CONSTRUCTORS(PclOpEvolve)
    PclOpEvolve() = default;
    PclOpEvolve(const PclOpEvolve &) {}
    void Operate(VECTOR<Particle> &particles, float timeStep, unsigned) override {
        if (particles.empty()) return;
        float *base = dev_.upload(particles);
        check(vg_evolve(base, base + kOffVel, kStride, particles.size(), timeStep, nullptr),
              "Evolve");
        dev_.download(particles);
    }

  private:
    DeviceParticles dev_;
};

// pclOpFindBoundingBox.h:16-40, pclOpFindBoundingBox.cpp:20-148
class PclOpFindBoundingBox : public IParticleOperation {
  public:
    VGFACADE_VIRTUAL_CONSTRUCTORS(PclOpFindBoundingBox)
    PclOpFindBoundingBox() : mMinCorner(FLT_MAX, FLT_MAX, FLT_MAX), mMaxCorner(-mMinCorner) {}
    PclOpFindBoundingBox(const PclOpFindBoundingBox &o)
        : mMinCorner(o.mMinCorner), mMaxCorner(o.mMaxCorner) {}
    void Operate(VECTOR<Particle> &particles, float, unsigned) override {
        mMinCorner = Vec3(FLT_MAX, FLT_MAX, FLT_MAX);
        mMaxCorner = -mMinCorner;
        FindBoundingBox(particles, mMinCorner, mMaxCorner);
    }
    const Vec3 &GetMinCorner() const { return mMinCorner; }
    const Vec3 &GetMaxCorner() const { return mMaxCorner; }
    // Grows minCorner / maxCorner to contain every particle.
    static void FindBoundingBox(const VECTOR<Particle> &particles, Vec3 &minCorner,
                                Vec3 &maxCorner) {
        if (particles.empty()) return;
        static thread_local DeviceParticles dev;
        float mn[3] = {minCorner.x, minCorner.y, minCorner.z};
        float mx[3] = {maxCorner.x, maxCorner.y, maxCorner.z};
        float *base = dev.upload(particles);
        check(vg_find_bounding_box(base, kStride, particles.size(), mn, mx, nullptr),
              "FindBoundingBox");
        minCorner = Vec3(mn[0], mn[1], mn[2]);
        maxCorner = Vec3(mx[0], mx[1], mx[2]);
    }

  private:
    Vec3 mMinCorner, mMaxCorner;
};

namespace Particles {
// pclOpPopulateVelocityGrid.cpp:33-87: velocityGrid (its shape already
// defined) becomes the trilinear-weighted average of the particles'
// velocities.  On the device: the particles are binned into the grid's
// cells (an influence tree over the grid, the cell lists it records) and
// every grid point gathers from the 8 cells around it.
inline void PopulateVelocityGrid(UniformGrid<Vec3> &velocityGrid,
                                 const VECTOR<Particle> &particles) {
    const size_t cap = velocityGrid.GetGridCapacity();
    std::vector<Vec3> out(cap);
    if (cap && !particles.empty()) {
        static thread_local DeviceParticles dev;
        static thread_local DeviceBuffer field;
        float *base = dev.upload(particles);
        vg_vortons v;
        std::memset(&v, 0, sizeof v);
        v.pos = base;
        v.angvel = base + kOffVel;  // the tree's cluster sums are not used
        v.size = base + kOffSize;
        v.stride = kStride;
        v.n = particles.size();
        vg_tree *t = nullptr;
        check(vg_tree_create(&t), "PopulateVelocityGrid");
        field.reserve(cap * sizeof(Vec3));
        vg_status st = vg_tree_build(t, &velocityGrid.desc(), &v, nullptr);
        if (st == VG_OK)
            st = vg_tree_populate_velocity_grid(t, base + kOffVel, kStride,
                                                static_cast<float *>(field.get()), nullptr);
        if (st == VG_OK)
            st = vg_memcpy_dtoh(out.data(), field.get(), cap * sizeof(Vec3));
        vg_tree_destroy(t);
        check(st, "PopulateVelocityGrid");
    }
    velocityGrid.Assign(velocityGrid.desc(), out.data());
}
}  // namespace Particles

// pclOpPopulateVelocityGrid.h:16-33, pclOpPopulateVelocityGrid.cpp:90-136
class PclOpPopulateVelocityGrid : public IParticleOperation {
  public:
    VGFACADE_VIRTUAL_CONSTRUCTORS(PclOpPopulateVelocityGrid)
    PclOpPopulateVelocityGrid() = default;
    PclOpPopulateVelocityGrid(const PclOpPopulateVelocityGrid &o)
        : mVelocityGrid(o.mVelocityGrid), mBoundingBox(o.mBoundingBox),
          mUpdateBoundingBox(o.mUpdateBoundingBox) {}
    void Operate(VECTOR<Particle> &particles, float, unsigned) override {
        if (particles.empty() || !mVelocityGrid) return;
        if (!mBoundingBox || mUpdateBoundingBox) {
            // a new box holding the given one (if any) and the particles,
            // padded by the first particle's radius
            Vec3 minCorner = mBoundingBox ? mBoundingBox[0] : Vec3(FLT_MAX, FLT_MAX, FLT_MAX);
            Vec3 maxCorner = mBoundingBox ? mBoundingBox[1] : -Vec3(FLT_MAX, FLT_MAX, FLT_MAX);
            PclOpFindBoundingBox::FindBoundingBox(particles, minCorner, maxCorner);
            const float margin = particles[0].GetRadius();
            const Vec3 nudge(margin, margin, margin);
            minCorner = minCorner - nudge;
            maxCorner = maxCorner + nudge;
            mVelocityGrid->Clear();
            mVelocityGrid->DefineShape(particles.size(), minCorner, maxCorner, true);
        }
        Particles::PopulateVelocityGrid(*mVelocityGrid, particles);
    }
    UniformGrid<Vec3> *mVelocityGrid = nullptr;  // grid of velocity values
    Vec3 *mBoundingBox = nullptr;                // [min, max] corners, or null
    bool mUpdateBoundingBox = true;              // grow the box by the particles
};

}  // namespace vgfacade
