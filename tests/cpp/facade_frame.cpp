// Drives reference-ordered frames through the C++ facade
// (vortegrid_amd/host/vortonsim.h) on 72-byte Particle records, wired with
// the reference's own statements (particleSystemConfiguration.cpp):
//   :139-140  PopulateVelocityGrid writes the VortonSim's velocity grid,
//             bounded by the address of GetMinCorner()
//   :196      the tracers' AssignVelocityFromField reads that grid
//   :282-296  tracers run before vortons, and PclOpVortonSim reads the
//             tracer group's bounding box through GetMinCorner/GetMaxCorner
// Each frame: tracer group [AssignVelocityFromField, Evolve, FindBoundingBox],
// then vorton group [PclOpVortonSim, Evolve, PopulateVelocityGrid], every
// operator called through IParticleOperation::Operate.
//
// usage: facade_frame <technique> <in.bin> <out.bin> [frames [vpm|hybrid [viscosity]]]
//   technique: VG_TECH_* (0 direct, 1 tree, 2 Poisson)
//   frames: number of frames (default 1)
//   vpm: the vortex particle method (stretching and tilting included)
//        instead of the reference's default VPM-SPH hybrid
//   viscosity: VortonSim viscosity (default 0: no PSE diffusion)
//   in.bin : u64 nv, u64 nt, nv Particle records, nt Particle records
//   out.bin: per frame: vg_grid_desc + 3*cap floats of the VortonSim's
//            velocity grid as Update left it, vg_grid_desc + 3*cap floats
//            of the grid after PopulateVelocityGrid, then nv and nt records
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "vortegrid_amd/host/vortonsim.h"

using namespace vgfacade;

static void write_grid(FILE *o, const UniformGrid<Vec3> &grid) {
    const vg_grid_desc &g = grid.desc();
    std::fwrite(&g, sizeof g, 1, o);
    if (grid.Size()) std::fwrite(grid.data(), sizeof(Vec3), grid.Size(), o);
}

struct FluidVortonPclGrpInfo {
    std::unique_ptr<PclOpVortonSim> mPclOpVortonSim;
    std::unique_ptr<PclOpEvolve> mPclOpEvolve;
    std::unique_ptr<PclOpPopulateVelocityGrid> mPclOpPopulateVelocityGrid;
};

struct FluidTracerPclGrpInfo {
    std::unique_ptr<PclOpAssignVelocityFromField> mPclOpAssignVelocityFromField;
    std::unique_ptr<PclOpEvolve> mPclOpEvolve;
    std::unique_ptr<PclOpFindBoundingBox> mPclOpFindBoundingBox;
};

int main(int argc, char **argv) {
    if (argc < 4 || argc > 7) {
        std::fprintf(stderr, "usage: %s technique in.bin out.bin [frames [vpm|hybrid [viscosity]]]\n",
                     argv[0]);
        return 2;
    }
    const int technique = std::atoi(argv[1]);
    const int frames = argc >= 5 ? std::atoi(argv[4]) : 1;
    const bool vpm = argc >= 6 && std::string(argv[5]) == "vpm";
    const float viscosity = argc == 7 ? (float)std::atof(argv[6]) : 0.0f;
    FILE *f = std::fopen(argv[2], "rb");
    if (!f) return 3;
    uint64_t nv = 0, nt = 0;
    if (std::fread(&nv, 8, 1, f) != 1 || std::fread(&nt, 8, 1, f) != 1) return 3;
    VECTOR<Particle> vortons(nv), tracers(nt);
    if (std::fread(vortons.data(), sizeof(Particle), nv, f) != nv) return 3;
    if (std::fread(tracers.data(), sizeof(Particle), nt, f) != nt) return 3;
    std::fclose(f);
    const float dt = 1.0f / 30.0f;
    FILE *o = std::fopen(argv[3], "wb");
    if (!o) return 4;
    try {
        FluidVortonPclGrpInfo vortonPclGrpInfo;
        FluidTracerPclGrpInfo tracerPclGrpInfo;
        // CreateVortonParticleGroup (particleSystemConfiguration.cpp:43-143)
        vortonPclGrpInfo.mPclOpVortonSim.reset(new PclOpVortonSim());
        vortonPclGrpInfo.mPclOpVortonSim->mVortonSim = VortonSim(0.0f, 1.0f, technique);
        vortonPclGrpInfo.mPclOpVortonSim->mVortonSim.SetVortons(
            reinterpret_cast<VECTOR<Vorton> *>(&vortons));
        vortonPclGrpInfo.mPclOpVortonSim->mVortonSim.SetViscosity(viscosity);
        vortonPclGrpInfo.mPclOpVortonSim->mVortonSim.SetAmbientDensity(1.0f);
        if (vpm)
            vortonPclGrpInfo.mPclOpVortonSim->mVortonSim.SetFluidSimulationTechnique(
                VortonSim::FLUID_SIM_VORTEX_PARTICLE_METHOD);
        vortonPclGrpInfo.mPclOpEvolve.reset(new PclOpEvolve());
        vortonPclGrpInfo.mPclOpPopulateVelocityGrid.reset(new PclOpPopulateVelocityGrid());
        vortonPclGrpInfo.mPclOpPopulateVelocityGrid->mVelocityGrid = & vortonPclGrpInfo.mPclOpVortonSim->mVortonSim.GetVelocityGrid() ;
        vortonPclGrpInfo.mPclOpPopulateVelocityGrid->mBoundingBox  = const_cast< Vec3 * >( & vortonPclGrpInfo.mPclOpVortonSim->mVortonSim.GetMinCorner() ) ;
        // CreateTracerParticleGroup (:148-268)
        tracerPclGrpInfo.mPclOpAssignVelocityFromField.reset(new PclOpAssignVelocityFromField());
        tracerPclGrpInfo.mPclOpAssignVelocityFromField->mVelocityGrid    =  & vortonPclGrpInfo.mPclOpVortonSim->mVortonSim.GetVelocityGrid() ;
        tracerPclGrpInfo.mPclOpEvolve.reset(new PclOpEvolve());
        tracerPclGrpInfo.mPclOpFindBoundingBox.reset(new PclOpFindBoundingBox());
        // CreateFluidParticleSystem (:274-299): the vorton group reads the
        // tracer box (the reference patches these addresses in through
        // IndirectAddress; here they are assigned directly)
        vortonPclGrpInfo.mPclOpVortonSim->mMinCorner = & tracerPclGrpInfo.mPclOpFindBoundingBox->GetMinCorner() ;
        vortonPclGrpInfo.mPclOpVortonSim->mMaxCorner = & tracerPclGrpInfo.mPclOpFindBoundingBox->GetMaxCorner() ;
        std::vector<IParticleOperation *> tracerGroup = {
            tracerPclGrpInfo.mPclOpAssignVelocityFromField.get(), tracerPclGrpInfo.mPclOpEvolve.get(),
            tracerPclGrpInfo.mPclOpFindBoundingBox.get()};
        std::vector<IParticleOperation *> vortonGroup = {
            vortonPclGrpInfo.mPclOpVortonSim.get(), vortonPclGrpInfo.mPclOpEvolve.get(),
            vortonPclGrpInfo.mPclOpPopulateVelocityGrid.get()};
        UniformGrid<Vec3> &velGrid = vortonPclGrpInfo.mPclOpVortonSim->mVortonSim.GetVelocityGrid();
        for (int frame = 0; frame < frames; ++frame) {
            // ParticleSystem::Update: tracers first, then vortons
            // (particleGroup.cpp:71-80 runs each group's operators in order)
            for (IParticleOperation *op : tracerGroup) op->Operate(tracers, dt, (unsigned)frame);
            // the VortonSim's own velocity grid, captured between its Update
            // and PopulateVelocityGrid, which redefines the same grid object
            vortonGroup[0]->Operate(vortons, dt, (unsigned)frame);
            write_grid(o, velGrid);
            for (size_t k = 1; k < vortonGroup.size(); ++k)
                vortonGroup[k]->Operate(vortons, dt, (unsigned)frame);
            write_grid(o, velGrid);
            std::fwrite(vortons.data(), sizeof(Particle), nv, o);
            std::fwrite(tracers.data(), sizeof(Particle), nt, o);
        }
        const VortonSim &sim = vortonPclGrpInfo.mPclOpVortonSim->mVortonSim;
        std::printf("facade frames ok: %d frames, grid %u x %u x %u, %llu vortons, %llu tracers, "
                    "min corner at %p, max corner at %p\n",
                    frames, sim.GetGrid().GetNumPoints(0), sim.GetGrid().GetNumPoints(1),
                    sim.GetGrid().GetNumPoints(2), (unsigned long long)nv,
                    (unsigned long long)nt, (const void *)&sim.GetMinCorner(),
                    (const void *)&sim.GetMaxCorner());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "facade error: %s\n", e.what());
        std::fclose(o);
        return 1;
    }
    std::fclose(o);
    return 0;
}
