// ref_e_driver.cpp -- test harness (built only into oracle/_ref, never shipped):
// drives the reference's own _E physics (assim/swSpE_hyp.cpp,
// swSpE_friction.cpp, swSpE_core.cpp, compiled from /root/reference) through
// the dry-run loop of build_checkpoint_records (assim/swSpE_checkpoint.cpp:
// 210-401, restated here because that file needs the Enzyme runtime), and
// prints the recorded schedule and the final state at full precision.
//   usage: ref_e_driver <case folder> <n_obs> <t_obs_1> ... <t_obs_n>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "swSpE.h"
#include "../src/common/swFile.h"
#include "../src/common/swMessage.h"
#include "../src/common/swIO.h"
#include "../src/common/geometry/swMesh.h"
#include "../src/sw2dSolver/common/swSolverParam.h"

struct Driver : public swSpE {
    void run(const std::vector<double> &obs) {
        z_cell.assign(cell_nb, 0.0);
        dz_dx.assign(cell_nb, 0.0);
        dz_dy.assign(cell_nb, 0.0);
        swept_data.assign(cell_nb, 0.0);
        if (friction_coeff_direction_1.size() != cell_nb) friction_coeff_direction_1.assign(cell_nb, 0.0);
        if (friction_coeff_direction_2.size() != cell_nb) friction_coeff_direction_2.assign(cell_nb, 0.0);
        if (friction_cos_alpha.size() != cell_nb) friction_cos_alpha.assign(cell_nb, 1.0);
        if (friction_sin_alpha.size() != cell_nb) friction_sin_alpha.assign(cell_nb, 0.0);
        bc_val_interpolated.assign(edge_bnd_nb, 0.0);
        phi_ptr = phi.data();
        computeEdgeFactors();
        t = t_0;
        dt = 0.0;  // the solver's dt member before any step (swSp.h:155); the MUSCL-EVR
                   // reconstruction of the first step reads it
        size_t oi = 0;
        size_t n = 0;
        while (t < t_max) {
            computeBC_E(t);
            double target = dt_max;
            if (oi < obs.size()) {
                double to = obs[oi] - t;
                if (to - 10e-6 > 0.0 && to < target) target = to;
            }
            hypComputeLRStates_E();
            double dtc = hypComputeTimeStep_E();
            double cur = (dtc < target) ? dtc : target;
            if ((t + cur) > t_max) cur = t_max - t;
            if (cur < 1e-6) cur = 1e-6;
            dt = cur;
            hypComputeFluxes_E();
            hypComputeBalance_E(cur);
            for (size_t i = 0; i < cell_nb; ++i)
                if (water_depth[i] < 0.) {
                    water_depth[i] = 0.;
                    unit_discharge_x[i] = 0.;
                    unit_discharge_y[i] = 0.;
                }
            if (model_friction != "none") sourceFrictionStrickler_E(cur);
            if (model_singular_head_loss != "none") sourceSingularHeadLoss_E(cur);
            t += cur;
            while (oi < obs.size() && t > obs[oi] + 1e-6) oi++;
            int trig = -1;
            if (oi < obs.size() && std::abs(t - obs[oi]) < 1e-6) trig = (int)oi++;
            std::printf("step %zu %.17e %d\n", n, cur, trig);
            ++n;
        }
        for (size_t i = 0; i < cell_nb; ++i)
            std::printf("cell %zu %.17e %.17e %.17e\n", i, water_depth[i], unit_discharge_x[i],
                        unit_discharge_y[i]);
    }
};

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    swIO *io = new swIO(argv[1]);
    io->createEmptyFolders(swFile::log_solver_file);
    swMessage::outfile = new std::ofstream(io->absolutePath(swFile::log_solver_file));
    swSolverParam *sp = new swSolverParam(io);
    if (sp->readParameters() != EXIT_SUCCESS) return 3;
    swMesh *mesh = new swMesh(io);
    if (mesh->build() != EXIT_SUCCESS) return 4;
    std::vector<double> obs;
    const int n = std::atoi(argv[2]);
    for (int k = 0; k < n && 3 + k < argc; ++k) obs.push_back(std::atof(argv[3 + k]));
    Driver *d = new Driver();
    d->setIOPointer(io);
    io->createEmptyOutputFolder(true);
    d->importAttributesFromSolverParam(sp);
    if (d->init(mesh) != EXIT_SUCCESS) return 5;
    d->run(obs);
    return 0;
}
