// tdmd.hpp -- the drop-in C++ interface of the single-GPU step path, over
// the C-ABI of tdmd_hip.h (link libtdmd_hip.so). Define exactly one of
// TDMD_PRECISION_MIXED / TDMD_PRECISION_FP64 before including it.
#pragma once

#include "core/box.hpp"
#include "core/system_state.hpp"
#include "core/device_buffer.hpp"
#include "core/types.hpp"
#include "domain/device_cell_list.hpp"
#include "integrator/device_nose_hoover.hpp"
#include "integrator/device_velocity_verlet.hpp"
#include "integrator/device_velocity_verlet_zone.hpp"
#include "io/dump.hpp"
#include "io/lammps_data.hpp"
#include "io/setfl.hpp"
#include "io/thermo.hpp"
#include "neighbors/device_neighbor_list.hpp"
#include "potentials/device_eam.hpp"
#include "potentials/device_morse.hpp"
#include "potentials/device_morse_zone.hpp"
#include "potentials/setfl_potential.hpp"
#include "domain/zone_partition.hpp"
#include "scheduler/fast_pipeline_scheduler.hpp"
#include "scheduler/pipeline_scheduler.hpp"
#include "scheduler/zone.hpp"
