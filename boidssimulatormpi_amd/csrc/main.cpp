// main.cpp — the boidsim / boidsim_grid driver over the MI355X engine.
// Same flow, stdout and output files as the reference driver's non-MPI path
// (src/main.cpp:77-303): parse settings, setup + init, print the
// parameters, write the 50-byte .ptcl v0.2 header, then for t = 0..T-1
// write state t (positions; optional velocity file; optional 9 force floats
// after each position) and call update().  All output is little-endian.
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "args.h"
#include "boid_simulation.h"
#include "ptcl.h"

static const unsigned int FPS = 30;

static void write_header(std::ofstream& f, const Args& args, const BoidSimulation& sim)
{
    data_file_header_v02 header;
    header.header_length = sizeof(header);
    header.N = args.population;
    header.step = args.time_step;
    header.t_0 = sim.time_step;
    header.fps = FPS;
    header.x_min = 0.0f;
    header.x_max = (float)sim.field_size_X;
    header.y_min = 0.0f;
    header.y_max = (float)sim.field_size_Y;
    header.z_min = 0.0f;
    header.z_max = (float)sim.field_size_Z;
    f.write((const char*)&header, sizeof(header));
}

int main(int argc, char** argv)
{
    {
        int x = 1;
        if (*(char*)&x == 0) {
            std::cerr << "big-endian hosts are not supported" << std::endl;
            return -1;
        }
    }
    BoidSimulation* boid_sim = new BoidSimulation();
    Args args(argc, (const char**)argv);
    boid_sim->enable_force_output(args.is_force_data_output);
    boid_sim->setup(args.population, args.field_size,
                    args.separation_sight_distance, args.separation_sight_angle, args.separation_force_coefficient,
                    args.alignment_sight_distance, args.alignment_sight_angle, args.alignment_force_coefficient,
                    args.cohesion_sight_distance, args.cohesion_sight_angle, args.cohesion_force_coefficient,
                    args.velocity_max, args.velocity_min, args.initialization, args.random_seed);
    boid_sim->init();

    std::cout << "N: " << boid_sim->N << std::endl
              << "field size: " << boid_sim->field_size_X << "," << boid_sim->field_size_Y << ","
              << boid_sim->field_size_Z << std::endl
              << "#Separation" << std::endl
              << "force: " << boid_sim->separation.force_coefficient << std::endl
              << "area distance: " << boid_sim->separation.sight_distance << std::endl
              << "area angle: " << boid_sim->separation.sight_agnle << std::endl
              << "#Alignment" << std::endl
              << "force: " << boid_sim->alignment.force_coefficient << std::endl
              << "area distance: " << boid_sim->alignment.sight_distance << std::endl
              << "area angle: " << boid_sim->alignment.sight_agnle << std::endl
              << "#Cohesion" << std::endl
              << "force: " << boid_sim->cohesion.force_coefficient << std::endl
              << "area distance: " << boid_sim->cohesion.sight_distance << std::endl
              << "area angle: " << boid_sim->cohesion.sight_agnle << std::endl
              << "#Velocity" << std::endl
              << "min: " << boid_sim->velocity.min << std::endl
              << "max: " << boid_sim->velocity.max << std::endl;
    if (boid_sim->is_openmp_enabled())
        std::cout << "#OpenMP: enabled (max threads:" << boid_sim->get_max_threads() << ")" << std::endl;
    else
        std::cout << "#OpenMP: disabled" << std::endl;

    const bool is_output_v = !args.velocity_output_filename.empty();
    std::ofstream fout, v_fout;
    fout.open(args.output_filename.c_str(), std::ios::out | std::ios::binary | std::ios::trunc);
    if (!fout) {
        std::cerr << "Couldn't open: " << args.output_filename << std::endl;
        std::exit(-1);
    }
    write_header(fout, args, *boid_sim);
    if (is_output_v) {
        v_fout.open(args.velocity_output_filename.c_str(), std::ios::out | std::ios::binary | std::ios::trunc);
        if (!v_fout) {
            std::cerr << "Couldn't open: " << args.velocity_output_filename << std::endl;
            std::exit(-1);
        }
        write_header(v_fout, args, *boid_sim);
    }

    std::cout << "simulation start." << std::endl;
    float x, y, z, c[9];
    for (unsigned int t = 0; t < args.time_step; t++) {
        for (unsigned int i = 0; i < boid_sim->N; i++) {
            boid_sim->get(i, &x, &y, &z);
            fout.write((const char*)&x, sizeof(x));
            fout.write((const char*)&y, sizeof(y));
            fout.write((const char*)&z, sizeof(z));
            if (is_output_v) {
                boid_sim->get_v(i, &x, &y, &z);
                v_fout.write((const char*)&x, sizeof(x));
                v_fout.write((const char*)&y, sizeof(y));
                v_fout.write((const char*)&z, sizeof(z));
            }
            if (args.is_force_data_output) {
                boid_sim->get_force(i, &c[0], &c[1], &c[2], &c[3], &c[4], &c[5], &c[6], &c[7], &c[8]);
                fout.write((const char*)c, sizeof(c));
            }
        }
        fout.flush();
        if (is_output_v) v_fout.flush();
        std::cout << "  " << t << "/" << args.time_step << std::endl << std::flush;
        boid_sim->update();
    }
    std::cout << "simulation finished." << std::endl;
    fout.close();
    if (is_output_v) v_fout.close();
    delete boid_sim;
    return 0;
}
