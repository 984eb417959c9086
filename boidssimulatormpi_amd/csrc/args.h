// args.h — command line and .ini settings of the reference driver
// (src/args.h:11-38), parsed without Boost.  Same fields, same option
// names, same precedence: the .ini file first, then the command-line
// simulation options override it (src/args.cpp:78-153).
#ifndef BOIDS_AMD_ARGS_H
#define BOIDS_AMD_ARGS_H

#include <string>

class Args {
public:
    Args(int argc, const char** argv);
    std::string setting_filename;
    std::string output_filename;
    std::string velocity_output_filename;
    bool is_parallel_output = false;
    bool is_force_data_output = false;
    bool is_output_init_state = false;

    unsigned int time_step = 0;

    unsigned int population = 0;
    double field_size = 0;
    double separation_sight_distance = 0;
    double separation_sight_angle = 0;
    double separation_force_coefficient = 0;
    double alignment_sight_distance = 0;
    double alignment_sight_angle = 0;
    double alignment_force_coefficient = 0;
    double cohesion_sight_distance = 0;
    double cohesion_sight_angle = 0;
    double cohesion_force_coefficient = 0;
    double velocity_max = 0;
    double velocity_min = 0;
    std::string initialization;
    int random_seed = 0;
};

#endif
