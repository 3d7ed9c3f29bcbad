// The reference's ci-test hBN run (ci-test/inputs/hBN.json.in), key by key,
// as edus_sim_params for the C++ callers of the C ABI.
#pragma once
#include <cstdio>
#include <cstring>

#include "edus_gpu.h"

inline void hbn_set_str(char* dst, size_t cap, const char* src) { std::snprintf(dst, cap, "%s", src); }

inline edus_sim_params hbn_params(const char* tb_seedname)
{
    edus_sim_params p;
    std::memset(&p, 0, sizeof(p));
    hbn_set_str(p.tb_file, sizeof(p.tb_file), tb_seedname);
    p.grid[0] = 50;
    p.grid[1] = 50;
    p.grid[2] = 1;
    p.filledbands = 1;
    p.dt = 0.01;
    hbn_set_str(p.dt_units, sizeof(p.dt_units), "femtoseconds");
    p.initialtime = 0.0;
    hbn_set_str(p.initialtime_units, sizeof(p.initialtime_units), "femtoseconds");
    p.finaltime = 1.0;
    hbn_set_str(p.finaltime_units, sizeof(p.finaltime_units), "femtoseconds");
    p.printresolution = 1;
    p.printresolution_pulse = 1;
    hbn_set_str(p.solver, sizeof(p.solver), "RK");
    p.order = 4;
    p.epsilon = 2.0;
    p.nr0 = 1;
    p.r0[0] = 10.0;
    hbn_set_str(p.r0_units, sizeof(p.r0_units), "angstrom");
    p.coulomb_center_quirk = 1;
    p.nlasers = 1;
    edus_laser_params& L = p.lasers[0];
    L.intensity = 1e10;
    hbn_set_str(L.intensity_units, sizeof(L.intensity_units), "wcm2");
    L.frequency = 10.0;
    hbn_set_str(L.frequency_units, sizeof(L.frequency_units), "electronvolt");
    hbn_set_str(L.wavelength_units, sizeof(L.wavelength_units), "nm");
    L.polarization[0] = 1.0;
    L.polarization[1] = 1.0;
    L.cycles = 1.0;
    L.t0 = 0.0;
    hbn_set_str(L.t0_units, sizeof(L.t0_units), "fs");
    p.precision = EDUS_PREC_FP64;
    p.layout = EDUS_LAYOUT_FULL;
    return p;
}
