// rsg_big.h -- host interface of the kernel-per-phase path for single large
// members (rsg_big.hip), used by the engine in rsg_capi.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "rsg_layout.h"
#include "rsg_prep.h"

struct RsgBig;

// State of one large member whose workspace (hM, device pointers) is already
// carved and staged on stream s; dM: device memory for the member struct.
// sc0: its initial scalars.  The member runs on a stream of its own, so the
// engine may run several members concurrently from separate host threads
// (each thread must have the device current).  maxLevels: 0 coarsens like the reference (at most 20
// levels, down to 50 rows); k > 0 stops at k levels once the coarsest level
// can be inverted on the matrix cores (at most 16,384 rows).  graphs: replay
// the p' V-cycles from hipGraphs.  comp: null for the incompressible solve;
// else the compressible problem's data (temperature BCs, gas constants,
// alphaT, initial T / rho / mu; its field pointers are filled here) and the
// member runs the compressible SIMPLE iteration.  Returns null (and err) on a
// device error.
RsgBig* rsg_big_create(const RsgMemberDev& hM, RsgMemberDev* dM, const RsgSettingsDev& set,
                       const rsg::PrepMesh& pm, const RsgScalars& sc0, hipStream_t s, int maxLevels, bool graphs,
                       const RsgCompDev* comp, std::string& err);
// Runs up to `iters` outer iterations (fewer if the member finishes); out gets
// the member's scalars.  Returns with the member's stream drained: a finished
// member's fields, history, scalars and observables are on the device for
// collection.
int rsg_big_run(RsgBig* B, int iters, RsgScalars& out, std::string& err);
void rsg_big_destroy(RsgBig* B);
// A compressible member's temperature, density and mechanical pressure
// (cells) and temperature boundary values (nB), and its energy-equation inner
// iterations; any pointer may be null.  Synchronous.
int rsg_big_comp_fields(RsgBig* B, double* T, double* rho, double* pMech, double* Tb, double* energyIters,
                        std::string& err);
