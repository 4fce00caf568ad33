// Drives one reference-ordered frame through the C++ facade
// (vortegrid_amd/host/vortonsim.h) on 72-byte Particle records:
//   tracers: FindBoundingBox
//   vortons: PclOpVortonSim (bounding box includes the tracer box) -> Evolve
//   tracers: AssignVelocityFromField(grid) -> Evolve
// usage: facade_frame <technique> <in.bin> <out.bin> [vpm [viscosity]]
//   vpm: run the vortex particle method (stretching and tilting included)
//        instead of the reference's default VPM-SPH hybrid
//   viscosity: VortonSim viscosity (default 0: no PSE diffusion)
//   in.bin : u64 nv, u64 nt, nv Particle records, nt Particle records
//   out.bin: vg_grid_desc, grid (3*cap floats), nv records, nt records
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "vortegrid_amd/host/vortonsim.h"

using namespace vgfacade;

int main(int argc, char **argv) {
    if (argc < 4 || argc > 6) {
        std::fprintf(stderr, "usage: %s technique in.bin out.bin [vpm [viscosity]]\n", argv[0]);
        return 2;
    }
    const bool vpm = argc >= 5 && std::string(argv[4]) == "vpm";
    const float viscosity = argc == 6 ? (float)std::atof(argv[5]) : 0.0f;
    const int technique = std::atoi(argv[1]);
    FILE *f = std::fopen(argv[2], "rb");
    if (!f) return 3;
    uint64_t nv = 0, nt = 0;
    if (std::fread(&nv, 8, 1, f) != 1 || std::fread(&nt, 8, 1, f) != 1) return 3;
    VECTOR<Particle> vortons(nv), tracers(nt);
    if (std::fread(vortons.data(), sizeof(Particle), nv, f) != nv) return 3;
    if (std::fread(tracers.data(), sizeof(Particle), nt, f) != nt) return 3;
    std::fclose(f);
    const float dt = 1.0f / 30.0f;
    try {
        PclOpFindBoundingBox tracerBox;
        tracerBox.Operate(tracers, dt, 0);
        PclOpVortonSim sim;
        sim.mVortonSim = VortonSim(viscosity, 1.0f, technique);
        if (vpm) sim.mVortonSim.SetFluidSimulationTechnique(VortonSim::FLUID_SIM_VORTEX_PARTICLE_METHOD);
        sim.mMinCorner = &tracerBox.mMinCorner;
        sim.mMaxCorner = &tracerBox.mMaxCorner;
        sim.Operate(vortons, dt, 0);
        const UniformGrid<Vec3> &grid = sim.mVortonSim.GetVelocityGrid();
        PclOpEvolve evolve;
        evolve.Operate(vortons, dt, 0);
        PclOpAssignVelocityFromField assign;
        assign.mVortonSim = &sim.mVortonSim;
        assign.Operate(tracers, dt, 0);
        evolve.Operate(tracers, dt, 0);
        FILE *o = std::fopen(argv[3], "wb");
        if (!o) return 4;
        const vg_grid_desc &g = grid.desc();
        std::fwrite(&g, sizeof g, 1, o);
        std::fwrite(&grid[0], sizeof(Vec3), grid.Size(), o);
        std::fwrite(vortons.data(), sizeof(Particle), nv, o);
        std::fwrite(tracers.data(), sizeof(Particle), nt, o);
        std::fclose(o);
        std::printf("facade frame ok: grid %u x %u x %u, %llu vortons, %llu tracers\n",
                    g.np[0], g.np[1], g.np[2], (unsigned long long)nv, (unsigned long long)nt);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "facade error: %s\n", e.what());
        return 1;
    }
    return 0;
}
