// ptcl.h — the reference's file header and parameter records.
// Byte layouts must match src/dtype.h:8-49 exactly: the .ptcl readers of
// the reference's tools (tools/swarm_util.py:6, "<4sBBIIIIIffffff") parse
// these 50 bytes.  Field names are kept (including `sight_agnle`) so code
// written against the reference compiles unchanged.
#ifndef BOIDS_AMD_PTCL_H
#define BOIDS_AMD_PTCL_H

#pragma pack(push, 1)
typedef struct {
    char format_name[4] = {'P', 'T', 'C', 'L'};
    unsigned char major_version = 0;
    unsigned char minor_version = 2;
    unsigned int header_length;
    unsigned int N;
    unsigned int step;
    unsigned int t_0;
    unsigned int fps;
    float x_min;
    float x_max;
    float y_min;
    float y_max;
    float z_min;
    float z_max;
} data_file_header_v02;
#pragma pack(pop)

static_assert(sizeof(data_file_header_v02) == 50, ".ptcl v0.2 header is 50 bytes");

typedef struct {
    double sight_distance;
    double sight_agnle;
    double force_coefficient;
} interaction_parameters;

typedef struct {
    double max;
    double min;
} velocity_parameters;

#endif
